// Lid-driven cavity through the C++ class API, in the pattern of the
// reference's example drivers (examples/navier_stokes_2D/*.cpp): PARAMETERS ->
// POINTS -> CLOUD -> boundary conditions -> FRACTIONAL_STEP_1 -> time loop
// with old <- new.
//
// usage: lid_driven_cavity <parameters.csv> <dim> <n> <steps> <dt> <Re> <out.bin> [g]
// out.bin: int32 nv, dim, system_size; then xyz (nv*dim), u, v, (w), p (float64).
// With g, every step runs the body-force overload with f_x = 0.5 g sin(pi y),
// f_y = -g (and f_z = 0 in 3D).  GPU options (precision, preconditioner,
// guess) come from the settings file PARAMETERS finds (MEMPHYS_GPU_SETTINGS).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../memphys_amd/host/memphys.hpp"

int main(int argc, char** argv) {
    if (argc < 8) {
        fprintf(stderr, "usage: %s parameters.csv dim n steps dt Re out.bin [g]\n", argv[0]);
        return 2;
    }
    const bool gravity = argc > 8;
    const double g = gravity ? atof(argv[8]) : 0.0;
    const int dim = atoi(argv[2]), n = atoi(argv[3]), steps = atoi(argv[4]);
    const double dt = atof(argv[5]), Re = atof(argv[6]);
    try {
        PARAMETERS parameters(argv[1], dim);
        parameters.rho = 1.0;
        parameters.mu = 1.0 / Re;  // lid speed 1, cavity side 1
        parameters.dt = dt;
        const int k_layer[3] = {dim == 3 ? 60 : parameters.cloud_size,
                                dim == 3 ? 55 : parameters.cloud_size,
                                dim == 3 ? 50 : parameters.cloud_size};
        POINTS points = POINTS::box(parameters, n, dim == 3 ? 0.3 : 0.25, 20240601, k_layer);
        CLOUD cloud(points, parameters);
        const int nv = points.nv;
        VectorXd u_new = VectorXd::Zero(nv), v_new = u_new, w_new = u_new, p_new;
        std::vector<bool> u_dir(nv, true), v_dir(nv, true), w_dir(nv, true), p_dir(nv, false);
        for (int iv = 0; iv < nv; ++iv)
            if (points.boundary_flag[iv] && std::fabs(points.xyz[dim * iv + dim - 1] - 1.0) < 1e-12)
                u_new[iv] = 1.0;  // lid
        VectorXd u_old = u_new, v_old = v_new, w_old = w_new, p_old;
        FRACTIONAL_STEP_1* fs =
            (dim == 2) ? new FRACTIONAL_STEP_1(points, cloud, parameters, u_dir, v_dir, p_dir, 2)
                       : new FRACTIONAL_STEP_1(points, cloud, parameters, u_dir, v_dir, w_dir, p_dir, 2);
        p_new = VectorXd::Zero(fs->system_size);
        p_old = p_new;
        VectorXd fx = VectorXd::Zero(nv), fy = VectorXd::Zero(nv), fz = VectorXd::Zero(nv);
        for (int iv = 0; iv < nv; ++iv) {
            fx[iv] = 0.5 * g * std::sin(std::acos(-1.0) * points.xyz[dim * iv + 1]);
            fy[iv] = -g;
        }
        for (int it = 0; it < steps; ++it) {
            if (dim == 2 && gravity)
                fs->single_timestep_2d(points, cloud, parameters, u_new, v_new, p_new, u_old, v_old,
                                       p_old, fx, fy, it);
            else if (dim == 2)
                fs->single_timestep_2d(points, cloud, parameters, u_new, v_new, p_new, u_old, v_old,
                                       p_old, it);
            else if (gravity)
                fs->single_timestep_3d(points, cloud, parameters, u_new, v_new, w_new, p_new, u_old,
                                       v_old, w_old, p_old, fx, fy, fz, it);
            else
                fs->single_timestep_3d(points, cloud, parameters, u_new, v_new, w_new, p_new, u_old,
                                       v_old, w_old, p_old, it);
            parameters.nt_actual = it + 1;
            u_old = u_new;
            v_old = v_new;
            w_old = w_new;
            p_old = p_new;
        }
        double res[6];
        calc_continuity_residual(points, cloud, parameters, u_new, v_new,
                                 dim == 3 ? &w_new : nullptr, res);
        printf("nv %d steps %d solver %s precision %d preconditioner %d guess %d last PPE iterations %d rel_res %.3e continuity max %.3e "
               "mean %.3e cond max %.3e dx avg %.5f\n",
               nv, steps, parameters.solver_type.c_str(), parameters.precision,
               parameters.preconditioner, parameters.guess,
               parameters.n_iter_actual.empty() ? 0 : parameters.n_iter_actual.back(),
               parameters.rel_res_log.empty() ? 0.0 : parameters.rel_res_log.back(), res[4], res[5],
               cloud.cond_num_RBF_max, parameters.avg_dx);
        FILE* f = fopen(argv[7], "wb");
        if (!f) return 3;
        int hdr[3] = {nv, dim, fs->system_size};
        fwrite(hdr, sizeof(int), 3, f);
        fwrite(points.xyz.data(), sizeof(double), points.xyz.size(), f);
        fwrite(u_new.data(), sizeof(double), nv, f);
        fwrite(v_new.data(), sizeof(double), nv, f);
        if (dim == 3) fwrite(w_new.data(), sizeof(double), nv, f);
        fwrite(p_new.data(), sizeof(double), p_new.size(), f);
        fclose(f);
        // the solver log of every step (postprocessing_functions.cpp:52-72)
        parameters.output_file_prefix = argv[7];
        write_iteration_details(parameters);
        delete fs;
    } catch (const std::exception&) {
        return 1;
    }
    return 0;
}
