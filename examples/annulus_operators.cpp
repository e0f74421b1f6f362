// The reference's set-up stack on a Gmsh mesh through the C++ class API:
// PARAMETERS(csv, msh) -> POINTS (Gmsh reader) -> CLOUD (kNN, RCM, GPU weights)
// -> PARAMETERS::calc_dt (GPU eigenvalues).  Prints the quantities the
// reference writes to its simulation_details.csv
// (postprocessing_functions.cpp:52-152), one "key,value" per line.
//
// With an output prefix and a tecplot file it also reads the file's u_new,
// v_new, p_new and T_new onto the nodes (read_restart_tecplot_file), writes
// them back (write_tecplot_steady_variables) and writes the
// simulation_details.csv (write_simulation_details), with the fluid of the
// reference's sample run (rho = 10, mu = 0.03).
//
// usage: annulus_operators <parameters.csv> <mesh.msh> <u0> <v0> <alpha> [prefix tecplot.plt]
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../memphys_amd/host/memphys.hpp"

int main(int argc, char** argv) {
    if (argc < 6) {
        fprintf(stderr, "usage: %s parameters.csv mesh.msh u0 v0 alpha\n", argv[0]);
        return 2;
    }
    try {
        PARAMETERS parameters(argv[1], argv[2]);
        POINTS points(parameters);
        CLOUD cloud(points, parameters);
        parameters.calc_dt(points.grad_x_matrix_EIGEN, points.grad_y_matrix_EIGEN,
                           points.grad_z_matrix_EIGEN, points.laplacian_matrix_EIGEN,
                           atof(argv[3]), atof(argv[4]), 0.0, atof(argv[5]));
        int nb = 0;
        for (int iv = 0; iv < points.nv; ++iv) nb += points.boundary_flag[iv] ? 1 : 0;
        printf("nv,%d\nnv_boundary,%d\nnv_internal,%d\n", points.nv, nb, points.nv - nb);
        printf("dx_max,%.16g\ndx_min,%.16g\ndx_avg,%.16g\n", parameters.max_dx, parameters.min_dx,
               parameters.avg_dx);
        printf("cloud_size,%d\n", parameters.cloud_size);
        printf("RBF_cond_num_max,%.16g\nRBF_cond_num_min,%.16g\nRBF_cond_num_avg,%.16g\n",
               cloud.cond_num_RBF_max, cloud.cond_num_RBF_min, cloud.cond_num_RBF_avg);
        printf("grad_x_eigval,%.16g,%.16g\n", parameters.grad_x_eigval_real,
               parameters.grad_x_eigval_imag);
        printf("grad_y_eigval,%.16g,%.16g\n", parameters.grad_y_eigval_real,
               parameters.grad_y_eigval_imag);
        printf("laplace_eigval,%.16g,%.16g\n", parameters.laplace_eigval_real,
               parameters.laplace_eigval_imag);
        printf("timestep_dt,%.16g\n", parameters.dt);
        if (argc >= 8) {
            parameters.output_file_prefix = argv[6];
            parameters.rho = 10.0;
            parameters.mu = 0.03;
            VectorXd u(points.nv), v(points.nv), p(points.nv), T(points.nv);
            std::vector<std::string> names = {"u_new", "v_new", "p_new", "T_new"};
            std::vector<VectorXd*> fields = {&u, &v, &p, &T};
            read_restart_tecplot_file(points, parameters, argv[7], names, fields);
            write_tecplot_steady_variables(points, parameters, names, fields);
            write_simulation_details(points, cloud, parameters);
        }
    } catch (const std::exception&) {
        return 1;
    }
    return 0;
}
