// C++ class API of the meshless fractional-step path: the names, members and
// signatures of MeMPhyS's header_files/class.hpp (PARAMETERS 58-94, POINTS
// 96-140, CLOUD 142-169, SOLVER 171-217, FRACTIONAL_STEP_1 246-270), backed by
// the gfx950 HIP library through the C ABI of include/memphys_hip.h.
//
// Differences a caller sees:
//  * Eigen is not used: memphys::VectorXd / memphys::SparseMatrixCSR stand in
//    for Eigen::VectorXd / Eigen::SparseMatrix<double, RowMajor> (same element
//    access, size(), Zero()).
//  * POINTS can be built from the seeded box generator (POINTS::box) as well
//    as from a Gmsh file.
//  * FRACTIONAL_STEP_1 gains single_timestep_3d (the reference's step is 2D).
//  * Errors keep the reference convention: printf("ERROR from ...") then
//    throw std::bad_exception().
#pragma once
#include <cstdint>
#include <cstdio>
#include <exception>
#include <string>
#include <vector>

struct mp_ctx;

namespace memphys {

class VectorXd {
  public:
    VectorXd() = default;
    explicit VectorXd(size_t n, double v = 0.0) : d_(n, v) {}
    static VectorXd Zero(size_t n) { return VectorXd(n, 0.0); }
    size_t size() const { return d_.size(); }
    long rows() const { return (long)d_.size(); }
    double& operator[](size_t i) { return d_[i]; }
    const double& operator[](size_t i) const { return d_[i]; }
    double& operator()(size_t i) { return d_[i]; }
    const double& operator()(size_t i) const { return d_[i]; }
    double* data() { return d_.data(); }
    const double* data() const { return d_.data(); }
    void resize(size_t n) { d_.resize(n); }

  private:
    std::vector<double> d_;
};

class MatrixXi {
  public:
    MatrixXi() = default;
    void resize(int r, int c) {
        r_ = r;
        c_ = c;
        d_.assign((size_t)r * c, 0);
    }
    int& operator()(int r, int c) { return d_[(size_t)r * c_ + c]; }
    int operator()(int r, int c) const { return d_[(size_t)r * c_ + c]; }
    int rows() const { return r_; }
    int cols() const { return c_; }
    const int* data() const { return d_.data(); }

  private:
    int r_ = 0, c_ = 0;
    std::vector<int> d_;
};

// Row-major CSR operator.  Products go to the GPU library when the operator
// is one of the cloud's stencil operators (op >= 0).
class SparseMatrixCSR {
  public:
    int n = 0;
    std::vector<int> row_ptr, col;
    std::vector<double> val;
    mp_ctx* ctx = nullptr;
    int op = -1;
    int rows_mask = 0;  // 0 all rows, 1 boundary rows only, 2 interior rows only
    const std::vector<bool>* boundary = nullptr;
    VectorXd operator*(const VectorXd& x) const;
    long nonZeros() const { return (long)val.size(); }
};

[[noreturn]] void fail(const char* where, const std::string& what);
void check(mp_ctx* ctx, int rc, const char* where);

}  // namespace memphys

using memphys::VectorXd;

class PARAMETERS {
  public:
    int cloud_size;
    double cloud_size_multiplier;
    int phs_deg;
    int poly_deg;
    int num_poly_terms;
    int dimension;
    memphys::MatrixXi polynomial_term_exponents;
    std::string meshfile;
    std::string output_file_prefix;
    std::string restart_tecplot_file;
    double max_dx, min_dx, avg_dx;
    std::vector<int> periodic_bc_index;
    double dt;
    double rho = -10.0, mu = rho;
    double steady_tolerance, solver_tolerance, Courant, precond_droptol;
    int nt, euclid_precond_level_hypre, gmres_kdim, n_iter;
    std::string solver_type;  // eigen_ilu_bicgstab / hypre_ilu_gmres / eigen_direct map to the GPU Krylov solve

    double cloud_id_timer = 0, rcm_timer = 0, cloud_misc_timer = 0, points_timer = 0,
           grad_laplace_coeff_timer = 0, factoring_timer = 0, solve_timer = 0, total_timer = 0;
    int nt_actual = 0;
    std::vector<double> rel_res_log, abs_res_log, regul_alpha_log, steady_error_log;
    std::vector<int> n_iter_actual;

    double grad_x_eigval_real = 0, grad_x_eigval_imag = 0, grad_y_eigval_real = 0,
           grad_y_eigval_imag = 0, grad_z_eigval_real = 0, grad_z_eigval_imag = 0,
           laplace_eigval_real = 0, laplace_eigval_imag = 0;

    // GPU-side choices (not in the reference's CSV).  They are read from an
    // optional settings file of "key,value" lines: the file named by the
    // environment variable MEMPHYS_GPU_SETTINGS, else gpu_settings.csv beside
    // the parameter file (read_gpu_settings lists the keys).
    int precision = 0;       // 0 FP64, 1 FP32-operator inner solve with FP64 defect correction
    int device = 0;
    int preconditioner = 0;  // 0 Jacobi, 1 multigrid V-cycle (enum mp_precond)
    int guess = 0;           // 0 previous pressure, 1 learned, 2 extrapolation, 3 projection
    int guess_order = 4;     // pressures the extrapolation goes through (2..4)
    int guess_window = 20;   // solutions the projection combines (1..24)
    int guess_base = 3;      // the classical guess the learned one corrects (2 or 3)
    std::string guess_params_file;  // packed float32 network parameters (guess = 1)
    int graphs = 1;          // replay the Krylov chunks from hipGraphs
    double inner_tolerance = 1e-3;  // precision 1: inner relative reduction per cycle
    std::string gpu_settings_file;  // the settings file read, if any

  public:
    PARAMETERS(std::string parameter_file, std::string gmsh_file);
    // Synthetic clouds: the dimension is given instead of read from a mesh.
    PARAMETERS(std::string parameter_file, int dimension);
    void read_calc_parameters(std::string parameter_file);
    void verify_parameters();
    // Keys: precision (fp64 | mixed), preconditioner (jacobi | multigrid),
    // guess (previous | extrapolate | project | learned), guess_params (file),
    // extrapolation_order (2 | 3 | 4), projection_window (1..24),
    // learned_base (extrapolate | project), graphs (0 | 1), device (int),
    // inner_tolerance (double).  Unknown keys and values are errors.
    void read_gpu_settings(std::string settings_file);
    void calc_cloud_num_points();
    void get_problem_dimension_msh();
    void calc_polynomial_term_exponents();
    // Explicit time step from the largest-magnitude eigenvalues of the operators
    // (parameters.cpp:269-296), computed on the GPU.
    void calc_dt(memphys::SparseMatrixCSR& grad_x, memphys::SparseMatrixCSR& grad_y,
                 memphys::SparseMatrixCSR& grad_z, memphys::SparseMatrixCSR& laplacian, double u0,
                 double v0, double w0, double alpha);
};

class POINTS {
  public:
    std::vector<bool> corner_edge_vertices;
    int nv = 0;
    int nv_original = 0;
    std::vector<double> xyz_min, xyz_max, xyz_length;
    std::vector<double> xyz;
    std::vector<double> xyz_original;
    std::vector<std::vector<int>> elem_vert_original;  // connectivity of the original mesh
    std::vector<int> iv_original_nearest_vert;         // [nv_original]: node holding each original vertex
    std::vector<double> normal;
    std::vector<bool> boundary_flag;
    std::vector<int> bc_tag;
    std::vector<int> stencil_size;  // optional per-node stencil sizes (empty: cloud_size)
    memphys::SparseMatrixCSR grad_x_matrix_EIGEN, grad_y_matrix_EIGEN, grad_z_matrix_EIGEN,
        laplacian_matrix_EIGEN;
    memphys::SparseMatrixCSR grad_x_matrix_EIGEN_boundary, grad_x_matrix_EIGEN_internal;
    memphys::SparseMatrixCSR grad_y_matrix_EIGEN_boundary, grad_y_matrix_EIGEN_internal;
    memphys::SparseMatrixCSR grad_z_matrix_EIGEN_boundary, grad_z_matrix_EIGEN_internal;
    memphys::SparseMatrixCSR laplacian_matrix_EIGEN_boundary, laplacian_matrix_EIGEN_internal;

  public:
    POINTS(PARAMETERS& parameters);  // reads parameters.meshfile (Gmsh 2.2)
    POINTS() = default;
    // Seeded jittered box cloud on [0,1]^dim (n lattice points per axis);
    // k_layer: stencil sizes of the face layer and the two layers under it.
    static POINTS box(PARAMETERS& parameters, int n, double jitter, uint64_t seed,
                      const int k_layer[3]);
};

class CLOUD {
  public:
    std::vector<int> nb_points_row;
    std::vector<int> nb_points_col;
    std::vector<double> grad_x_coeff;
    std::vector<double> grad_y_coeff;
    std::vector<double> grad_z_coeff;
    std::vector<double> laplacian_coeff;
    std::vector<double> cond_num_RBF;
    double cond_num_RBF_max = 0, cond_num_RBF_min = 0, cond_num_RBF_avg = 0;
    std::vector<int> rcm_points_order;
    mp_ctx* ctx = nullptr;  // GPU context holding this cloud

  public:
    CLOUD(POINTS& points, PARAMETERS& parameters);
    ~CLOUD();
    CLOUD(const CLOUD&) = delete;
    CLOUD& operator=(const CLOUD&) = delete;
    void calc_cloud_points_fast(POINTS& points, PARAMETERS& parameters);
    void re_order_points_reverse_cuthill_mckee(POINTS& points, PARAMETERS& parameters);
    void calc_iv_original_nearest_vert(POINTS& points, PARAMETERS& parameters);
    void calc_charac_dx(POINTS& points, PARAMETERS& parameters);
    void calc_grad_laplace_coeffs(POINTS& points, PARAMETERS& parameters);
    void EIGEN_set_grad_laplace_matrix(POINTS& points, PARAMETERS& parameters);
    void EIGEN_set_grad_laplace_matrix_separate(POINTS& points, PARAMETERS& parameters);
};

class SOLVER {
  public:
    void init(POINTS& points, CLOUD& cloud, PARAMETERS& parameters,
              std::vector<bool>& dirichlet_flag, double unsteady_term_coeff, double conv_term_coeff,
              double diff_term_coeff, bool log_flag);
    void general_solve(POINTS& points, PARAMETERS& parameters, VectorXd& field_new,
                       VectorXd& field_old, VectorXd& rhs);
    ~SOLVER();
    std::string solver_type;
    int n_iter = 0;
    int gmres_kdim = 0;             // GMRES restart length (hypre_ilu_gmres)
    double solver_tolerance = 0;
    std::vector<bool> dirichlet_flag_1;
    int system_size = 0;
    bool log_flag_1 = true;
    mp_ctx* ctx = nullptr;  // own context: the cloud is uploaded again with its weights
};

class FRACTIONAL_STEP_1 {
  public:
    FRACTIONAL_STEP_1(POINTS& points, CLOUD& cloud, PARAMETERS& parameters,
                      std::vector<bool>& u_dirichlet_flag1, std::vector<bool>& v_dirichlet_flag1,
                      std::vector<bool>& p_dirichlet_flag1, int temporal_order1);
    // 3D form (no reference counterpart: the reference's step is 2D only).
    FRACTIONAL_STEP_1(POINTS& points, CLOUD& cloud, PARAMETERS& parameters,
                      std::vector<bool>& u_dirichlet_flag1, std::vector<bool>& v_dirichlet_flag1,
                      std::vector<bool>& w_dirichlet_flag1, std::vector<bool>& p_dirichlet_flag1,
                      int temporal_order1);
    void check_bc(POINTS& points, PARAMETERS& parameters);
    void single_timestep_2d(POINTS& points, CLOUD& cloud, PARAMETERS& parameters, VectorXd& u_new,
                            VectorXd& v_new, VectorXd& p_new, VectorXd& u_old, VectorXd& v_old,
                            VectorXd& p_old, int it1);
    // with a body force per unit volume (class.hpp:253, fractional_step_1.cpp:65-76)
    void single_timestep_2d(POINTS& points, CLOUD& cloud, PARAMETERS& parameters, VectorXd& u_new,
                            VectorXd& v_new, VectorXd& p_new, VectorXd& u_old, VectorXd& v_old,
                            VectorXd& p_old, VectorXd& body_force_x, VectorXd& body_force_y,
                            int it1);
    void single_timestep_3d(POINTS& points, CLOUD& cloud, PARAMETERS& parameters, VectorXd& u_new,
                            VectorXd& v_new, VectorXd& w_new, VectorXd& p_new, VectorXd& u_old,
                            VectorXd& v_old, VectorXd& w_old, VectorXd& p_old, int it1);
    void single_timestep_3d(POINTS& points, CLOUD& cloud, PARAMETERS& parameters, VectorXd& u_new,
                            VectorXd& v_new, VectorXd& w_new, VectorXd& p_new, VectorXd& u_old,
                            VectorXd& v_old, VectorXd& w_old, VectorXd& p_old,
                            VectorXd& body_force_x, VectorXd& body_force_y,
                            VectorXd& body_force_z, int it1);

    std::vector<bool> u_dirichlet_flag, v_dirichlet_flag, w_dirichlet_flag, p_dirichlet_flag;
    bool p_bc_full_neumann = true;
    int temporal_order = -1, it = 0;
    int system_size = 0;

  private:
    void create(POINTS& points, CLOUD& cloud, PARAMETERS& parameters);
    void step(POINTS& points, CLOUD& cloud, PARAMETERS& parameters, VectorXd* U_new[3],
              VectorXd& p_new, VectorXd* U_old[3], VectorXd& p_old, VectorXd* F[3], int it1);
    mp_ctx* ctx = nullptr;
};

// Output and restart files in the reference's formats
// (postprocessing_functions.cpp:52-152, 597-682, 765-805).
void write_iteration_details(PARAMETERS& parameters);
void write_simulation_details(POINTS& points, CLOUD& cloud, PARAMETERS& parameters);
void write_tecplot_steady_variables(POINTS& points, PARAMETERS& parameters,
                                    std::vector<std::string>& variable_names,
                                    std::vector<VectorXd*>& variable_pointers);
void read_restart_tecplot_file(POINTS& points, PARAMETERS& parameters, std::string filename,
                               std::vector<std::string>& tecplot_variable_names,
                               std::vector<VectorXd*>& variable_pointers);

// Continuity residual of postprocessing_functions.cpp:192-195 / 248-251:
// out = {max_int, mean_int, max_bnd, mean_bnd, max_all, mean_all}.
void calc_continuity_residual(POINTS& points, CLOUD& cloud, PARAMETERS& parameters,
                              VectorXd& u, VectorXd& v, VectorXd* w, double out[6]);
