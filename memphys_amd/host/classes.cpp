// C++ class layer (class.hpp names and signatures) over the HIP C ABI.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>

#include "../../include/memphys_hip.h"
#include "cloud_setup.hpp"
#include "memphys.hpp"

namespace memphys {

void fail(const char* where, const std::string& what) {
    printf("\n\nERROR from %s %s\n\n", where, what.c_str());
    fflush(stdout);
    throw std::bad_exception();
}

void check(mp_ctx* ctx, int rc, const char* where) {
    if (rc != MP_OK) fail(where, std::string("GPU library: ") + mp_last_error(ctx));
}

VectorXd SparseMatrixCSR::operator*(const VectorXd& x) const {
    if (!ctx || op < 0) fail("SparseMatrixCSR::operator*", "matrix is not bound to a GPU cloud");
    if ((int)x.size() < n) fail("SparseMatrixCSR::operator*", "vector shorter than the matrix");
    VectorXd y(n);
    check(ctx, mp_operator_apply(ctx, op, x.data(), y.data()), "SparseMatrixCSR::operator*");
    if (rows_mask && boundary)
        for (int i = 0; i < n; ++i)
            if ((*boundary)[i] != (rows_mask == 1)) y[i] = 0.0;
    return y;
}

}  // namespace memphys

using memphys::check;
using memphys::fail;

// ---------------------------------------------------------------- PARAMETERS
namespace {
std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}
}  // namespace

namespace {
// The GPU settings file of a parameter file: MEMPHYS_GPU_SETTINGS if set,
// else gpu_settings.csv in the parameter file's directory when it exists.
std::string gpu_settings_path(const std::string& parameter_file) {
    if (const char* e = std::getenv("MEMPHYS_GPU_SETTINGS")) return e;
    size_t sl = parameter_file.find_last_of('/');
    std::string dir = sl == std::string::npos ? "" : parameter_file.substr(0, sl + 1);
    std::string p = dir + "gpu_settings.csv";
    std::ifstream f(p);
    return f ? p : "";
}
}  // namespace

void PARAMETERS::read_gpu_settings(std::string settings_file) {
    std::ifstream f(settings_file);
    if (!f) fail("PARAMETERS::read_gpu_settings", "cannot open " + settings_file);
    gpu_settings_file = settings_file;
    std::string line;
    while (std::getline(f, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        size_t c = line.find(',');
        if (c == std::string::npos)
            fail("PARAMETERS::read_gpu_settings", "expected key,value: " + line);
        std::string key = trim(line.substr(0, c)), val = trim(line.substr(c + 1));
        auto bad = [&]() {
            fail("PARAMETERS::read_gpu_settings", "bad value for " + key + ": " + val);
        };
        try {
            if (key == "precision") {
                if (val == "fp64") precision = 0;
                else if (val == "mixed") precision = 1;
                else bad();
            } else if (key == "preconditioner") {
                if (val == "jacobi") preconditioner = 0;
                else if (val == "multigrid") preconditioner = 1;
                else bad();
            } else if (key == "guess") {
                if (val == "previous") guess = 0;
                else if (val == "learned") guess = 1;
                else if (val == "extrapolate") guess = 2;
                else if (val == "project") guess = 3;
                else bad();
            } else if (key == "guess_params") {
                guess_params_file = val;
            } else if (key == "extrapolation_order") {
                if (val == "2" || val == "3" || val == "4") guess_order = std::stoi(val);
                else bad();
            } else if (key == "projection_window") {
                guess_window = std::stoi(val);
                if (guess_window < 1 || guess_window > 24) bad();
            } else if (key == "learned_base") {
                if (val == "extrapolate") guess_base = 2;
                else if (val == "project") guess_base = 3;
                else bad();
            } else if (key == "graphs") {
                if (val == "0" || val == "1") graphs = std::stoi(val);
                else bad();
            } else if (key == "device") {
                device = std::stoi(val);
            } else if (key == "inner_tolerance") {
                inner_tolerance = std::stod(val);
                if (!(inner_tolerance > 0.0 && inner_tolerance < 1.0)) bad();
            } else {
                fail("PARAMETERS::read_gpu_settings", "unknown key " + key);
            }
        } catch (const std::invalid_argument&) {
            bad();
        } catch (const std::out_of_range&) {
            bad();
        }
    }
    if (guess == 1 && guess_params_file.empty())
        fail("PARAMETERS::read_gpu_settings", "guess learned needs guess_params");
}

PARAMETERS::PARAMETERS(std::string parameter_file, std::string gmsh_file) {
    meshfile = gmsh_file;
    std::ifstream f(meshfile);
    if (!f) fail("PARAMETERS::PARAMETERS", "cannot open " + meshfile);
    read_calc_parameters(parameter_file);
    verify_parameters();
    if (std::string g = gpu_settings_path(parameter_file); !g.empty()) read_gpu_settings(g);
    get_problem_dimension_msh();
    calc_cloud_num_points();
    calc_polynomial_term_exponents();
}

PARAMETERS::PARAMETERS(std::string parameter_file, int dim) {
    meshfile = "box.msh";
    read_calc_parameters(parameter_file);
    verify_parameters();
    if (std::string g = gpu_settings_path(parameter_file); !g.empty()) read_gpu_settings(g);
    if (dim != 2 && dim != 3) fail("PARAMETERS::PARAMETERS", "dimension must be 2 or 3");
    dimension = dim;
    calc_cloud_num_points();
    calc_polynomial_term_exponents();
}

// parameters.cpp:42-80: twelve positional "name,value" lines.
void PARAMETERS::read_calc_parameters(std::string parameter_file) {
    std::ifstream f(parameter_file);
    if (!f) fail("PARAMETERS::read_calc_parameters", "cannot open " + parameter_file);
    std::vector<std::string> vals;
    std::string line;
    while (vals.size() < 12 && std::getline(f, line)) {
        size_t c = line.find(',');
        if (c == std::string::npos) continue;
        vals.push_back(trim(line.substr(c + 1)));
    }
    if (vals.size() < 12) fail("PARAMETERS::read_calc_parameters", "expected 12 name,value lines");
    try {
        poly_deg = std::stoi(vals[0]);
        phs_deg = std::stoi(vals[1]);
        cloud_size_multiplier = std::stod(vals[2]);
        nt = std::stoi(vals[3]);
        Courant = std::stod(vals[4]);
        solver_type = vals[5];
        steady_tolerance = std::stod(vals[6]);
        solver_tolerance = std::stod(vals[7]);
        euclid_precond_level_hypre = std::stoi(vals[8]);
        gmres_kdim = std::stoi(vals[9]);
        precond_droptol = std::stod(vals[10]);
        n_iter = std::stoi(vals[11]);
    } catch (const std::exception&) {
        fail("PARAMETERS::read_calc_parameters", "malformed value in " + parameter_file);
    }
}

// parameters.cpp:82-120
void PARAMETERS::verify_parameters() {
    if (poly_deg < 2 || poly_deg > 15)
        fail("PARAMETERS::verify_parameters",
             "poly_deg should be in range [2, 15]; current value: " + std::to_string(poly_deg));
    if (phs_deg != 3 && phs_deg != 5 && phs_deg != 7 && phs_deg != 9 && phs_deg != 11)
        fail("PARAMETERS::verify_parameters",
             "phs_deg should be 3, 5, 7, 9, or 11; current value: " + std::to_string(phs_deg));
    std::transform(solver_type.begin(), solver_type.end(), solver_type.begin(), ::tolower);
    if (solver_type != "hypre_ilu_gmres" && solver_type != "eigen_direct" &&
        solver_type != "eigen_ilu_bicgstab" && solver_type != "sor")
        fail("PARAMETERS::verify_parameters",
             "solver_type should be either hypre_ilu_gmres, eigen_ilu_bicgstab or eigen_direct or "
             "sor; current value: " + solver_type);
    size_t dot = meshfile.find_last_of('.');
    if (dot == std::string::npos || dot + 1 == meshfile.size() || meshfile.substr(dot + 1) != "msh")
        fail("PARAMETERS::verify_parameters", "extension of meshfile should be msh: " + meshfile);
    output_file_prefix = meshfile;
    size_t sl = output_file_prefix.find_last_of('/');
    if (sl != std::string::npos) output_file_prefix = output_file_prefix.substr(sl + 1);
    dot = output_file_prefix.find_last_of('.');
    if (dot != std::string::npos) output_file_prefix = output_file_prefix.substr(0, dot);
    output_file_prefix += "_polydeg_" + std::to_string(poly_deg);
}

// parameters.cpp:122-130
void PARAMETERS::calc_cloud_num_points() {
    num_poly_terms = memphys::num_poly_terms(dimension, poly_deg);
    cloud_size = (int)std::ceil(cloud_size_multiplier * num_poly_terms);
}

// parameters.cpp:132-215: 3D when any volume element type appears.
void PARAMETERS::get_problem_dimension_msh() {
    std::ifstream f(meshfile);
    if (!f) fail("PARAMETERS::get_problem_dimension_msh", "cannot open " + meshfile);
    std::string tok;
    bool fmt = false;
    while (f >> tok)
        if (tok == "$MeshFormat") {
            fmt = true;
            break;
        }
    if (!fmt) fail("PARAMETERS::get_problem_dimension_msh", "$MeshFormat not found in " + meshfile);
    std::string v1, v2, v3;
    f >> v1 >> v2 >> v3;
    if (v1 != "2.2" || v2 != "0" || v3 != "8")
        fail("PARAMETERS::get_problem_dimension_msh", "MeshFormat should be 2.2 0 8 in " + meshfile);
    bool el = false;
    while (f >> tok)
        if (tok == "$Elements") {
            el = true;
            break;
        }
    if (!el) fail("PARAMETERS::get_problem_dimension_msh", "$Elements not found in " + meshfile);
    int ncv = 0;
    f >> ncv;
    dimension = 2;
    std::string line;
    std::getline(f, line);
    for (int i = 0; i < ncv && std::getline(f, line); ++i) {
        std::istringstream ls(line);
        int id = 0, type = 0;
        ls >> id >> type;
        if ((type >= 4 && type <= 7) || (type >= 11 && type <= 14) || (type >= 17 && type <= 19) ||
            (type >= 29 && type <= 31)) {
            dimension = 3;
            break;
        }
    }
}

void PARAMETERS::calc_polynomial_term_exponents() {
    std::vector<int> e = memphys::polynomial_exponents(dimension, poly_deg);
    polynomial_term_exponents.resize(num_poly_terms, dimension);
    for (int r = 0; r < num_poly_terms; ++r)
        for (int c = 0; c < dimension; ++c) polynomial_term_exponents(r, c) = e[r * dimension + c];
}

// parameters.cpp:269-296: dt = Courant * 2 / (u0 |l_x| + v0 |l_y| (+ w0 |l_z|) + alpha |l_L|)
// with l_* the largest-magnitude eigenvalue of each operator (forward Euler).
void PARAMETERS::calc_dt(memphys::SparseMatrixCSR& grad_x, memphys::SparseMatrixCSR& grad_y,
                         memphys::SparseMatrixCSR& grad_z, memphys::SparseMatrixCSR& laplacian,
                         double u0, double v0, double w0, double alpha) {
    auto eig = [&](memphys::SparseMatrixCSR& m, double& re, double& im) {
        if (!m.ctx || m.op < 0 || m.rows_mask != 0)
            fail("PARAMETERS::calc_dt", "operator matrix is not bound to the GPU cloud");
        memphys::check(m.ctx, mp_operator_eigenvalue(m.ctx, m.op, 1e-12, 500, &re, &im),
                       "PARAMETERS::calc_dt");
    };
    eig(grad_x, grad_x_eigval_real, grad_x_eigval_imag);
    eig(grad_y, grad_y_eigval_real, grad_y_eigval_imag);
    if (dimension == 3) eig(grad_z, grad_z_eigval_real, grad_z_eigval_imag);
    eig(laplacian, laplace_eigval_real, laplace_eigval_imag);
    printf("PARAMETERS::calc_dt Eigenvalue with largest magnitude: grad_x: (%g, %g)\n",
           grad_x_eigval_real, grad_x_eigval_imag);
    printf("PARAMETERS::calc_dt Eigenvalue with largest magnitude: grad_y: (%g, %g)\n",
           grad_y_eigval_real, grad_y_eigval_imag);
    if (dimension == 3)
        printf("PARAMETERS::calc_dt Eigenvalue with largest magnitude: grad_z: (%g, %g)\n",
               grad_z_eigval_real, grad_z_eigval_imag);
    printf("PARAMETERS::calc_dt Eigenvalue with largest magnitude: laplacian: (%g, %g)\n",
           laplace_eigval_real, laplace_eigval_imag);
    auto mag = [](double re, double im) { return std::sqrt(re * re + im * im); };
    double ev = u0 * mag(grad_x_eigval_real, grad_x_eigval_imag);
    ev += v0 * mag(grad_y_eigval_real, grad_y_eigval_imag);
    if (dimension == 3) ev += w0 * mag(grad_z_eigval_real, grad_z_eigval_imag);
    ev += alpha * mag(laplace_eigval_real, laplace_eigval_imag);
    dt = 2.0 / ev;
    dt = dt * Courant;
    printf("PARAMETERS::calc_dt Courant: %g, dt: %g seconds\n\n", Courant, dt);
}

// ---------------------------------------------------------------- POINTS
POINTS POINTS::box(PARAMETERS& parameters, int n, double jitter, uint64_t seed,
                   const int k_layer[3]) {
    const int kl[3] = {k_layer ? k_layer[0] : parameters.cloud_size,
                       k_layer ? k_layer[1] : parameters.cloud_size,
                       k_layer ? k_layer[2] : parameters.cloud_size};
    memphys::BoxCloud c = memphys::generate_box_cloud(parameters.dimension, n, jitter, seed,
                                                      parameters.cloud_size, kl);
    POINTS p;
    p.nv = p.nv_original = c.nv;
    p.xyz = c.xyz;
    p.xyz_original = c.xyz;
    p.normal = c.normal;
    p.boundary_flag.assign(c.boundary.begin(), c.boundary.end());
    p.corner_edge_vertices.assign(c.nv, false);
    p.bc_tag.assign(c.nv, -1);
    p.stencil_size = c.stencil_size;
    const int d = parameters.dimension;
    p.xyz_min.assign(d, 0.0);
    p.xyz_max.assign(d, 1.0);
    p.xyz_length.assign(d, 1.0);
    return p;
}

// points.cpp:27-72: read the Gmsh 2.2 mesh, boundary flags, vertex normals,
// delete corner/edge vertices, bounding box.
POINTS::POINTS(PARAMETERS& parameters) {
    clock_t t1 = clock();
    memphys::GmshCloud g;
    try {
        g = memphys::read_gmsh22(parameters.meshfile);
    } catch (const std::exception& e) {
        fail("POINTS::POINTS", e.what());
    }
    if (g.dim != parameters.dimension)
        fail("POINTS::POINTS", "mesh dimension differs from PARAMETERS::dimension");
    nv = g.nv;
    nv_original = g.nv_original;
    xyz = g.xyz;
    xyz_original = g.xyz_original;
    elem_vert_original = g.elem_vert_original;
    normal = g.normal;
    boundary_flag.assign(g.boundary.begin(), g.boundary.end());
    corner_edge_vertices.assign(g.nv_original, true);
    for (int o : g.original) corner_edge_vertices[o] = false;
    bc_tag.assign(nv, -1);
    const int d = g.dim;
    xyz_min.assign(d, 0.0);
    xyz_max.assign(d, 0.0);
    for (int k = 0; k < d; ++k) {
        xyz_min[k] = xyz_max[k] = xyz[k];
        for (int i = 0; i < nv; ++i) {
            xyz_min[k] = std::min(xyz_min[k], xyz[(size_t)i * d + k]);
            xyz_max[k] = std::max(xyz_max[k], xyz[(size_t)i * d + k]);
        }
    }
    xyz_length.resize(d);
    for (int k = 0; k < d; ++k) xyz_length[k] = xyz_max[k] - xyz_min[k];
    parameters.points_timer = (double)(clock() - t1) / CLOCKS_PER_SEC;
    printf("POINTS::POINTS original mesh nv: %d, after deleting corners nv: %d\n", nv_original, nv);
}

// ---------------------------------------------------------------- CLOUD
CLOUD::CLOUD(POINTS& points, PARAMETERS& parameters) {
    clock_t t1 = clock();
    calc_cloud_points_fast(points, parameters);
    parameters.cloud_id_timer = (double)(clock() - t1) / CLOCKS_PER_SEC;
    t1 = clock();
    re_order_points_reverse_cuthill_mckee(points, parameters);
    parameters.rcm_timer = (double)(clock() - t1) / CLOCKS_PER_SEC;
    t1 = clock();
    calc_iv_original_nearest_vert(points, parameters);
    calc_charac_dx(points, parameters);
    parameters.cloud_misc_timer = (double)(clock() - t1) / CLOCKS_PER_SEC;
    calc_grad_laplace_coeffs(points, parameters);
    EIGEN_set_grad_laplace_matrix(points, parameters);
    EIGEN_set_grad_laplace_matrix_separate(points, parameters);
}

CLOUD::~CLOUD() {
    if (ctx) mp_ctx_destroy(ctx);
}

// cloud.cpp:570-648 on the GPU (mp_knn_stencils)
void CLOUD::calc_cloud_points_fast(POINTS& points, PARAMETERS& parameters) {
    const int nv = points.nv, dim = parameters.dimension;
    std::vector<int> k(nv, parameters.cloud_size);
    if ((int)points.stencil_size.size() == nv) k = points.stencil_size;
    std::vector<uint8_t> b(nv);
    for (int i = 0; i < nv; ++i) b[i] = points.boundary_flag[i] ? 1 : 0;
    if (!ctx && mp_ctx_create(parameters.device, &ctx) != MP_OK)
        fail("CLOUD::calc_cloud_points_fast", "cannot create a GPU context");
    long total = 0;
    for (int i = 0; i < nv; ++i) total += std::max(k[i], 0);
    nb_points_row.assign(nv + 1, 0);
    nb_points_col.assign(total, 0);
    check(ctx, mp_knn_stencils(ctx, nv, dim, points.xyz.data(), b.data(), k.data(),
                               nb_points_row.data(), nb_points_col.data()),
          "CLOUD::calc_cloud_points_fast");
}

// cloud.cpp:315-331 with the order from the GPU (mp_rcm_order); the arrays are
// permuted on the host as re_order_points does (cloud.cpp:338-439)
void CLOUD::re_order_points_reverse_cuthill_mckee(POINTS& points, PARAMETERS& parameters) {
    const int nv = points.nv, dim = parameters.dimension;
    rcm_points_order.assign(nv, 0);
    check(ctx, mp_rcm_order(ctx, nv, nb_points_row.data(), nb_points_col.data(),
                            rcm_points_order.data()),
          "CLOUD::re_order_points_reverse_cuthill_mckee");
    std::vector<uint8_t> b(nv);
    for (int i = 0; i < nv; ++i) b[i] = points.boundary_flag[i] ? 1 : 0;
    std::vector<int> extra = points.bc_tag;
    if ((int)extra.size() != nv) extra.assign(nv, -1);
    memphys::apply_order(nv, dim, rcm_points_order, points.xyz, points.normal, b, extra,
                         nb_points_row, nb_points_col);
    for (int i = 0; i < nv; ++i) points.boundary_flag[i] = b[i] != 0;
    points.bc_tag = extra;
    if ((int)points.stencil_size.size() == nv)
        for (int i = 0; i < nv; ++i) points.stencil_size[i] = nb_points_row[i + 1] - nb_points_row[i];
}

// cloud.cpp:281-313
void CLOUD::calc_charac_dx(POINTS& points, PARAMETERS& parameters) {
    const int dim = parameters.dimension;
    parameters.avg_dx = 0.0;
    parameters.max_dx = 0.0;
    parameters.min_dx = 1e20;
    for (int iv = 0; iv < points.nv; ++iv) {
        double local = 1e20;
        for (int t = nb_points_row[iv]; t < nb_points_row[iv + 1]; ++t) {
            int j = nb_points_col[t];
            if (j == iv) continue;
            double s = 0.0;
            for (int a = 0; a < dim; ++a) {
                double d = points.xyz[dim * iv + a] - points.xyz[dim * j + a];
                s += d * d;
            }
            local = std::min(local, std::sqrt(s));
        }
        parameters.avg_dx += local;
        parameters.min_dx = std::min(parameters.min_dx, local);
        parameters.max_dx = std::max(parameters.max_dx, local);
    }
    parameters.avg_dx /= (double)points.nv;
}

// cloud.cpp:207-279 on the GPU (mp_weights_assemble)
void CLOUD::calc_grad_laplace_coeffs(POINTS& points, PARAMETERS& parameters) {
    clock_t t1 = clock();
    const int nv = points.nv, dim = parameters.dimension;
    if (!ctx && mp_ctx_create(parameters.device, &ctx) != MP_OK)
        fail("CLOUD::calc_grad_laplace_coeffs", "cannot create a GPU context");
    std::vector<uint8_t> b(nv);
    for (int i = 0; i < nv; ++i) b[i] = points.boundary_flag[i] ? 1 : 0;
    check(ctx, mp_cloud_upload(ctx, nv, dim, points.xyz.data(), b.data(), points.normal.data(),
                               nb_points_row.data(), nb_points_col.data()),
          "CLOUD::calc_grad_laplace_coeffs");
    const size_t W = nb_points_col.size();
    grad_x_coeff.assign(W, 0.0);
    grad_y_coeff.assign(W, 0.0);
    grad_z_coeff.assign(dim == 3 ? W : 0, 0.0);
    laplacian_coeff.assign(W, 0.0);
    cond_num_RBF.assign(nv, 0.0);
    check(ctx, mp_weights_assemble(ctx, parameters.phs_deg, parameters.poly_deg,
                                   parameters.polynomial_term_exponents.data(), MP_PREC_FP64,
                                   grad_x_coeff.data(), grad_y_coeff.data(),
                                   dim == 3 ? grad_z_coeff.data() : nullptr, laplacian_coeff.data(),
                                   cond_num_RBF.data()),
          "CLOUD::calc_grad_laplace_coeffs");
    cond_num_RBF_max = *std::max_element(cond_num_RBF.begin(), cond_num_RBF.end());
    cond_num_RBF_min = *std::min_element(cond_num_RBF.begin(), cond_num_RBF.end());
    cond_num_RBF_avg = std::accumulate(cond_num_RBF.begin(), cond_num_RBF.end(), 0.0) / nv;
    parameters.grad_laplace_coeff_timer = (double)(clock() - t1) / CLOCKS_PER_SEC;
}

namespace {
void bind_operator(memphys::SparseMatrixCSR& m, const CLOUD& c, const std::vector<double>& v, int op,
          int mask, const POINTS& p) {
    m.n = (int)c.nb_points_row.size() - 1;
    m.row_ptr.clear();
    m.col.clear();
    m.val.clear();
    m.row_ptr.push_back(0);
    for (int i = 0; i < m.n; ++i) {
        bool keep = mask == 0 || (mask == 1) == (bool)p.boundary_flag[i];
        if (keep)
            for (int t = c.nb_points_row[i]; t < c.nb_points_row[i + 1]; ++t) {
                m.col.push_back(c.nb_points_col[t]);
                m.val.push_back(v[t]);
            }
        m.row_ptr.push_back((int)m.col.size());
    }
    m.ctx = c.ctx;
    m.op = op;
    m.rows_mask = mask;
    m.boundary = &p.boundary_flag;
}
}  // namespace

// cloud.cpp:168-205
void CLOUD::EIGEN_set_grad_laplace_matrix(POINTS& points, PARAMETERS& parameters) {
    bind_operator(points.grad_x_matrix_EIGEN, *this, grad_x_coeff, MP_GX, 0, points);
    bind_operator(points.grad_y_matrix_EIGEN, *this, grad_y_coeff, MP_GY, 0, points);
    if (parameters.dimension == 3) bind_operator(points.grad_z_matrix_EIGEN, *this, grad_z_coeff, MP_GZ, 0, points);
    bind_operator(points.laplacian_matrix_EIGEN, *this, laplacian_coeff, MP_LAP, 0, points);
}

// cloud.cpp:89-166
void CLOUD::EIGEN_set_grad_laplace_matrix_separate(POINTS& points, PARAMETERS& parameters) {
    bind_operator(points.grad_x_matrix_EIGEN_boundary, *this, grad_x_coeff, MP_GX, 1, points);
    bind_operator(points.grad_x_matrix_EIGEN_internal, *this, grad_x_coeff, MP_GX, 2, points);
    bind_operator(points.grad_y_matrix_EIGEN_boundary, *this, grad_y_coeff, MP_GY, 1, points);
    bind_operator(points.grad_y_matrix_EIGEN_internal, *this, grad_y_coeff, MP_GY, 2, points);
    if (parameters.dimension == 3) {
        bind_operator(points.grad_z_matrix_EIGEN_boundary, *this, grad_z_coeff, MP_GZ, 1, points);
        bind_operator(points.grad_z_matrix_EIGEN_internal, *this, grad_z_coeff, MP_GZ, 2, points);
    }
    bind_operator(points.laplacian_matrix_EIGEN_boundary, *this, laplacian_coeff, MP_LAP, 1, points);
    bind_operator(points.laplacian_matrix_EIGEN_internal, *this, laplacian_coeff, MP_LAP, 2, points);
}

// ---------------------------------------------------------------- SOLVER
// solver.cpp:14-79
void SOLVER::init(POINTS& points, CLOUD& cloud, PARAMETERS& parameters,
                  std::vector<bool>& dirichlet_flag, double unsteady_term_coeff,
                  double conv_term_coeff, double diff_term_coeff, bool log_flag) {
    solver_tolerance = parameters.solver_tolerance;
    solver_type = parameters.solver_type;
    std::transform(solver_type.begin(), solver_type.end(), solver_type.begin(), ::tolower);
    // sor (solver.cpp:243-281) is a sequential Gauss-Seidel sweep with no GPU
    // form here: refused with the message the reference prints for a type it
    // cannot run (solver.cpp:73-77)
    if (solver_type != "hypre_ilu_gmres" && solver_type != "eigen_direct" &&
        solver_type != "eigen_ilu_bicgstab")
        fail("SOLVER::SOLVER", "solver_type should be either hypre_ilu_gmres, eigen_ilu_bicgstab or "
                               "eigen_direct; current value: " + solver_type);
    n_iter = parameters.n_iter;
    gmres_kdim = parameters.gmres_kdim;
    dirichlet_flag_1 = dirichlet_flag;
    log_flag_1 = log_flag;
    const int nv = points.nv, dim = parameters.dimension;
    if ((int)dirichlet_flag.size() != nv) fail("SOLVER::init", "dirichlet_flag must have nv entries");
    if (ctx) mp_ctx_destroy(ctx);
    ctx = nullptr;
    if (mp_ctx_create(parameters.device, &ctx) != MP_OK) fail("SOLVER::init", "no GPU context");
    std::vector<uint8_t> b(nv), d(nv);
    for (int i = 0; i < nv; ++i) {
        b[i] = points.boundary_flag[i] ? 1 : 0;
        d[i] = dirichlet_flag[i] ? 1 : 0;
    }
    check(ctx, mp_cloud_upload(ctx, nv, dim, points.xyz.data(), b.data(), points.normal.data(),
                               cloud.nb_points_row.data(), cloud.nb_points_col.data()),
          "SOLVER::init");
    check(ctx, mp_weights_set(ctx, cloud.grad_x_coeff.data(), cloud.grad_y_coeff.data(),
                              dim == 3 ? cloud.grad_z_coeff.data() : nullptr,
                              cloud.laplacian_coeff.data()),
          "SOLVER::init");
    clock_t t1 = clock();
    check(ctx, mp_solver_init(ctx, d.data(), unsteady_term_coeff, conv_term_coeff, diff_term_coeff,
                              solver_tolerance, n_iter, parameters.precision),
          "SOLVER::init");
    if (solver_type == "hypre_ilu_gmres")  // GMRES(gmres_kdim), solver.cpp:140-144
        check(ctx, mp_solver_krylov(ctx, MP_KRYLOV_GMRES, gmres_kdim), "SOLVER::init");
    parameters.factoring_timer = (double)(clock() - t1) / CLOCKS_PER_SEC;
    check(ctx, mp_solver_system_size(ctx, &system_size), "SOLVER::init");
}

SOLVER::~SOLVER() {
    if (ctx) mp_ctx_destroy(ctx);
}

// solver.cpp:283-363 (BiCGSTAB branch): field_new = solve(rhs) from field_old;
// field_old[nv] = field_new[nv] in the bordered case.
void SOLVER::general_solve(POINTS& points, PARAMETERS& parameters, VectorXd& field_new,
                           VectorXd& field_old, VectorXd& rhs) {
    (void)points;
    if (!ctx) fail("SOLVER::general_solve", "init was not called");
    if ((int)rhs.size() != system_size || (int)field_old.size() != system_size)
        fail("SOLVER::general_solve", "vectors must have system_size entries");
    if ((int)field_new.size() != system_size) field_new = VectorXd::Zero(system_size);
    check(ctx, mp_solver_solve(ctx, rhs.data(), field_new.data(), field_old.data()),
          "SOLVER::general_solve");
    mp_step_stats st;
    check(ctx, mp_stats_get(ctx, &st), "SOLVER::general_solve");
    if (log_flag_1) {
        parameters.rel_res_log.push_back(st.rel_res);
        parameters.abs_res_log.push_back(st.abs_res);
        parameters.n_iter_actual.push_back(st.iterations);
        if (system_size == (int)dirichlet_flag_1.size() + 1) parameters.regul_alpha_log.push_back(st.alpha);
    }
}

// ---------------------------------------------------------------- FRACTIONAL_STEP_1
FRACTIONAL_STEP_1::FRACTIONAL_STEP_1(POINTS& points, CLOUD& cloud, PARAMETERS& parameters,
                                     std::vector<bool>& u1, std::vector<bool>& v1,
                                     std::vector<bool>& p1, int temporal_order1) {
    if (parameters.dimension != 2)
        fail("FRACTIONAL_STEP_1::FRACTIONAL_STEP_1", "3D clouds need the w_dirichlet_flag form");
    temporal_order = temporal_order1;
    u_dirichlet_flag = u1;
    v_dirichlet_flag = v1;
    p_dirichlet_flag = p1;
    create(points, cloud, parameters);
}

FRACTIONAL_STEP_1::FRACTIONAL_STEP_1(POINTS& points, CLOUD& cloud, PARAMETERS& parameters,
                                     std::vector<bool>& u1, std::vector<bool>& v1,
                                     std::vector<bool>& w1, std::vector<bool>& p1,
                                     int temporal_order1) {
    temporal_order = temporal_order1;
    u_dirichlet_flag = u1;
    v_dirichlet_flag = v1;
    w_dirichlet_flag = w1;
    p_dirichlet_flag = p1;
    create(points, cloud, parameters);
}

// fractional_step_1.cpp:78-102
void FRACTIONAL_STEP_1::check_bc(POINTS& points, PARAMETERS& parameters) {
    (void)points;
    int su = std::accumulate(u_dirichlet_flag.begin(), u_dirichlet_flag.end(), 0);
    if (su == 0)
        fail("FRACTIONAL_STEP_1::check_bc", "Setting u_dirichlet_flag to full Neumann BC is not "
                                            "permitted; sum of u_dirichlet_flag: 0");
    int sv = std::accumulate(v_dirichlet_flag.begin(), v_dirichlet_flag.end(), 0);
    if (sv == 0)
        fail("FRACTIONAL_STEP_1::check_bc", "Setting v_dirichlet_flag to full Neumann BC is not "
                                            "permitted; sum of v_dirichlet_flag: 0");
    if (parameters.dimension == 3) {
        int sw = std::accumulate(w_dirichlet_flag.begin(), w_dirichlet_flag.end(), 0);
        if (sw == 0)
            fail("FRACTIONAL_STEP_1::check_bc", "Setting w_dirichlet_flag to full Neumann BC is "
                                                "not permitted; sum of w_dirichlet_flag: 0");
    }
    if (parameters.rho < 0 || parameters.mu < 0)
        fail("FRACTIONAL_STEP_1::check_bc", "Some parameters are not set; parameters.rho: " +
                                                std::to_string(parameters.rho) +
                                                ", parameters.mu: " + std::to_string(parameters.mu));
    if (temporal_order != 1 && temporal_order != 2)
        fail("FRACTIONAL_STEP_1::check_bc", "temporal_order should be either '1' or '2'; current "
                                            "value: " + std::to_string(temporal_order));
}

void FRACTIONAL_STEP_1::create(POINTS& points, CLOUD& cloud, PARAMETERS& parameters) {
    check_bc(points, parameters);
    ctx = cloud.ctx;
    const int nv = points.nv;
    std::vector<uint8_t> fl[4];
    const std::vector<bool>* src[4] = {&u_dirichlet_flag, &v_dirichlet_flag, &w_dirichlet_flag,
                                       &p_dirichlet_flag};
    for (int q = 0; q < 4; ++q) {
        fl[q].assign(nv, 0);
        if ((int)src[q]->size() == nv)
            for (int i = 0; i < nv; ++i) fl[q][i] = (*src[q])[i] ? 1 : 0;
        else if (q != 2 || parameters.dimension == 3)
            fail("FRACTIONAL_STEP_1::FRACTIONAL_STEP_1", "dirichlet flags must have nv entries");
    }
    p_bc_full_neumann = true;
    for (int i = 0; i < nv; ++i)
        if (points.boundary_flag[i] && p_dirichlet_flag[i]) p_bc_full_neumann = false;
    mp_fs_params prm{};
    prm.rho = parameters.rho;
    prm.mu = parameters.mu;
    prm.dt = parameters.dt;
    prm.temporal_order = temporal_order;
    prm.solver_tolerance = parameters.solver_tolerance;
    prm.n_iter = parameters.n_iter;
    prm.precision = parameters.precision;
    prm.inner_tolerance = parameters.inner_tolerance;
    prm.check_every = 8;
    prm.preconditioner = parameters.preconditioner;
    std::string st = parameters.solver_type;
    std::transform(st.begin(), st.end(), st.begin(), ::tolower);
    if (st == "sor")  // solver.cpp:243-281 has no GPU form (see SOLVER::init)
        fail("FRACTIONAL_STEP_1::FRACTIONAL_STEP_1",
             "solver_type should be either hypre_ilu_gmres, eigen_ilu_bicgstab or eigen_direct; "
             "current value: " + st);
    if (st == "hypre_ilu_gmres") {  // GMRES(gmres_kdim), solver.cpp:140-144
        prm.krylov = MP_KRYLOV_GMRES;
        prm.restart = parameters.gmres_kdim;
    }
    clock_t t1 = clock();
    check(ctx, mp_fs_create(ctx, &prm, fl[0].data(), fl[1].data(),
                            parameters.dimension == 3 ? fl[2].data() : nullptr, fl[3].data()),
          "FRACTIONAL_STEP_1::FRACTIONAL_STEP_1");
    parameters.factoring_timer = (double)(clock() - t1) / CLOCKS_PER_SEC;
    check(ctx, mp_solver_system_size(ctx, &system_size), "FRACTIONAL_STEP_1::FRACTIONAL_STEP_1");
    check(ctx, mp_graph_mode(ctx, parameters.graphs), "FRACTIONAL_STEP_1::FRACTIONAL_STEP_1");
    check(ctx, mp_guess_order(ctx, parameters.guess_order), "FRACTIONAL_STEP_1::FRACTIONAL_STEP_1");
    check(ctx, mp_guess_window(ctx, parameters.guess_window), "FRACTIONAL_STEP_1::FRACTIONAL_STEP_1");
    check(ctx, mp_guess_base(ctx, parameters.guess_base), "FRACTIONAL_STEP_1::FRACTIONAL_STEP_1");
    if (parameters.guess == MP_GUESS_LEARNED) {
        std::vector<float> w(mp_guess_param_count());
        std::ifstream f(parameters.guess_params_file, std::ios::binary);
        if (!f || !f.read(reinterpret_cast<char*>(w.data()), sizeof(float) * w.size()))
            fail("FRACTIONAL_STEP_1::FRACTIONAL_STEP_1",
                 "cannot read " + std::to_string(w.size()) + " float32 parameters from " +
                     parameters.guess_params_file);
        check(ctx, mp_guess_mode(ctx, MP_GUESS_LEARNED), "FRACTIONAL_STEP_1::FRACTIONAL_STEP_1");
        check(ctx, mp_guess_params(ctx, 1, w.data()), "FRACTIONAL_STEP_1::FRACTIONAL_STEP_1");
    } else {
        check(ctx, mp_guess_mode(ctx, parameters.guess), "FRACTIONAL_STEP_1::FRACTIONAL_STEP_1");
    }
}

void FRACTIONAL_STEP_1::step(POINTS& points, CLOUD& cloud, PARAMETERS& parameters,
                             VectorXd* U_new[3], VectorXd& p_new, VectorXd* U_old[3],
                             VectorXd& p_old, VectorXd* F[3], int it1) {
    (void)cloud;
    it = it1;
    const int nv = points.nv, dim = parameters.dimension;
    if (F) {
        for (int q = 0; q < dim; ++q)
            if ((int)F[q]->size() != nv)
                fail("FRACTIONAL_STEP_1::single_timestep", "body force size != nv");
        check(ctx, mp_fs_body_force(ctx, F[0]->data(), F[1]->data(), dim == 3 ? F[2]->data() : nullptr),
              "FRACTIONAL_STEP_1::single_timestep");
    } else {
        check(ctx, mp_fs_body_force(ctx, nullptr, nullptr, nullptr),
              "FRACTIONAL_STEP_1::single_timestep");
    }
    if ((int)p_new.size() != system_size) p_new = VectorXd::Zero(system_size);
    if ((int)p_old.size() != system_size) p_old = VectorXd::Zero(system_size);
    for (int q = 0; q < dim; ++q) {
        if ((int)U_old[q]->size() != nv) fail("FRACTIONAL_STEP_1::single_timestep", "field size != nv");
        check(ctx, mp_fields_set(ctx, q, U_old[q]->data()), "FRACTIONAL_STEP_1::single_timestep");
    }
    check(ctx, mp_fields_set(ctx, MP_P, p_old.data()), "FRACTIONAL_STEP_1::single_timestep");
    check(ctx, mp_fs_step(ctx, 1, it1), "FRACTIONAL_STEP_1::single_timestep");
    for (int q = 0; q < dim; ++q) {
        if ((int)U_new[q]->size() != nv) *U_new[q] = VectorXd::Zero(nv);
        check(ctx, mp_fields_get(ctx, q, U_new[q]->data()), "FRACTIONAL_STEP_1::single_timestep");
    }
    check(ctx, mp_fields_get(ctx, MP_P, p_new.data()), "FRACTIONAL_STEP_1::single_timestep");
    if (p_bc_full_neumann) p_old[nv] = p_new[nv];  // general_solve's side effect (solver.cpp:361-362)
    mp_step_stats st;
    check(ctx, mp_stats_get(ctx, &st), "FRACTIONAL_STEP_1::single_timestep");
    parameters.rel_res_log.push_back(st.rel_res);
    parameters.abs_res_log.push_back(st.abs_res);
    parameters.n_iter_actual.push_back(st.iterations);
    if (p_bc_full_neumann) parameters.regul_alpha_log.push_back(st.alpha);
}

// fractional_step_1.cpp:52-63
void FRACTIONAL_STEP_1::single_timestep_2d(POINTS& points, CLOUD& cloud, PARAMETERS& parameters,
                                           VectorXd& u_new, VectorXd& v_new, VectorXd& p_new,
                                           VectorXd& u_old, VectorXd& v_old, VectorXd& p_old,
                                           int it1) {
    VectorXd* Un[3] = {&u_new, &v_new, nullptr};
    VectorXd* Uo[3] = {&u_old, &v_old, nullptr};
    step(points, cloud, parameters, Un, p_new, Uo, p_old, nullptr, it1);
}

// fractional_step_1.cpp:65-76 (calc_vel_hat with the body force, 104-118)
void FRACTIONAL_STEP_1::single_timestep_2d(POINTS& points, CLOUD& cloud, PARAMETERS& parameters,
                                           VectorXd& u_new, VectorXd& v_new, VectorXd& p_new,
                                           VectorXd& u_old, VectorXd& v_old, VectorXd& p_old,
                                           VectorXd& body_force_x, VectorXd& body_force_y,
                                           int it1) {
    VectorXd* Un[3] = {&u_new, &v_new, nullptr};
    VectorXd* Uo[3] = {&u_old, &v_old, nullptr};
    VectorXd* F[3] = {&body_force_x, &body_force_y, nullptr};
    step(points, cloud, parameters, Un, p_new, Uo, p_old, F, it1);
}

void FRACTIONAL_STEP_1::single_timestep_3d(POINTS& points, CLOUD& cloud, PARAMETERS& parameters,
                                           VectorXd& u_new, VectorXd& v_new, VectorXd& w_new,
                                           VectorXd& p_new, VectorXd& u_old, VectorXd& v_old,
                                           VectorXd& w_old, VectorXd& p_old, int it1) {
    if (parameters.dimension != 3) fail("FRACTIONAL_STEP_1::single_timestep_3d", "2D cloud");
    VectorXd* Un[3] = {&u_new, &v_new, &w_new};
    VectorXd* Uo[3] = {&u_old, &v_old, &w_old};
    step(points, cloud, parameters, Un, p_new, Uo, p_old, nullptr, it1);
}

void FRACTIONAL_STEP_1::single_timestep_3d(POINTS& points, CLOUD& cloud, PARAMETERS& parameters,
                                           VectorXd& u_new, VectorXd& v_new, VectorXd& w_new,
                                           VectorXd& p_new, VectorXd& u_old, VectorXd& v_old,
                                           VectorXd& w_old, VectorXd& p_old,
                                           VectorXd& body_force_x, VectorXd& body_force_y,
                                           VectorXd& body_force_z, int it1) {
    if (parameters.dimension != 3) fail("FRACTIONAL_STEP_1::single_timestep_3d", "2D cloud");
    VectorXd* Un[3] = {&u_new, &v_new, &w_new};
    VectorXd* Uo[3] = {&u_old, &v_old, &w_old};
    VectorXd* F[3] = {&body_force_x, &body_force_y, &body_force_z};
    step(points, cloud, parameters, Un, p_new, Uo, p_old, F, it1);
}

// postprocessing_functions.cpp:192-195 / 248-251 with calc_max_l1_error
void calc_continuity_residual(POINTS& points, CLOUD& cloud, PARAMETERS& parameters, VectorXd& u,
                              VectorXd& v, VectorXd* w, double out[6]) {
    (void)cloud;
    VectorXd r = points.grad_x_matrix_EIGEN * u;
    VectorXd ry = points.grad_y_matrix_EIGEN * v;
    for (size_t i = 0; i < r.size(); ++i) r[i] = r[i] + ry[i];
    if (parameters.dimension == 3) {
        if (!w) fail("calc_continuity_residual", "3D needs w");
        VectorXd rz = points.grad_z_matrix_EIGEN * (*w);
        for (size_t i = 0; i < r.size(); ++i) r[i] = r[i] + rz[i];
    }
    double mi = 0, si = 0, mb = 0, sb = 0;
    int ni = 0, nb = 0;
    for (int i = 0; i < points.nv; ++i) {
        double e = std::fabs(r[i]);
        if (points.boundary_flag[i]) {
            nb++;
            sb += e;
            mb = std::max(mb, e);
        } else {
            ni++;
            si += e;
            mi = std::max(mi, e);
        }
    }
    out[0] = mi;
    out[1] = ni ? si / ni : 0.0;
    out[2] = mb;
    out[3] = nb ? sb / nb : 0.0;
    out[4] = std::max(mi, mb);
    out[5] = (si + sb) / points.nv;
}
