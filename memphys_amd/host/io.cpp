// Output and restart files of the class layer, in the reference's formats:
// write_iteration_details / write_simulation_details
// (postprocessing_functions.cpp:52-152), write_tecplot_steady_variables
// (postprocessing_functions.cpp:765-805), read_restart_tecplot_file
// (postprocessing_functions.cpp:597-682), and the map from the mesh's original
// vertices to the kept nodes (CLOUD::calc_iv_original_nearest_vert,
// cloud.cpp:50-86).  Numbers are printed with "%.16g" as there, so a file
// written here can be compared with one the reference wrote.
#include <unistd.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "memphys.hpp"

using memphys::fail;

namespace {

// Index of the nearest of `pts` (n points, dim coordinates, AoS) to each query
// point, ties to the lower index, over a uniform grid of about two points per
// cell searched ring by ring (exact: a ring r+1 point is at least r cells away).
std::vector<int> nearest_points(int dim, const std::vector<double>& pts, int n,
                                const std::vector<double>& q, int nq) {
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) {
        lo[a] = hi[a] = pts[a];
        for (int i = 0; i < n; ++i) {
            lo[a] = std::min(lo[a], pts[(size_t)i * dim + a]);
            hi[a] = std::max(hi[a], pts[(size_t)i * dim + a]);
        }
    }
    double vol = 1.0;
    for (int a = 0; a < dim; ++a) vol *= std::max(hi[a] - lo[a], 1e-300);
    double cell = std::pow(2.0 * vol / (double)n, 1.0 / dim);
    if (!(cell > 0.0) || !std::isfinite(cell)) cell = 1.0;
    int nc[3] = {1, 1, 1};
    long ncell = 1;
    for (;;) {
        ncell = 1;
        for (int a = 0; a < dim; ++a) {
            nc[a] = std::max(1, (int)std::ceil((hi[a] - lo[a]) / cell));
            ncell *= nc[a];
        }
        if (ncell <= 4L * n + 64) break;
        cell *= 1.5;
    }
    auto axis = [&](double x, int a) {
        const int c = (int)std::floor((x - lo[a]) / cell);
        return c < 0 ? 0 : (c >= nc[a] ? nc[a] - 1 : c);
    };
    std::vector<int> start(ncell + 1, 0), cid(n), items(n);
    for (int i = 0; i < n; ++i) {
        int lin = 0, mul = 1;
        for (int a = 0; a < dim; ++a) {
            lin += axis(pts[(size_t)i * dim + a], a) * mul;
            mul *= nc[a];
        }
        cid[i] = lin;
        start[lin + 1]++;
    }
    for (long c = 0; c < ncell; ++c) start[c + 1] += start[c];
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int i = 0; i < n; ++i) items[fill[cid[i]]++] = i;
    const int max_ring = std::max(nc[0], std::max(nc[1], nc[2]));
    std::vector<int> out(nq, -1);
    for (int k = 0; k < nq; ++k) {
        const double* x = q.data() + (size_t)k * dim;
        int cq[3] = {0, 0, 0};
        for (int a = 0; a < dim; ++a) cq[a] = axis(x[a], a);
        double best = std::numeric_limits<double>::infinity();
        int bi = -1;
        for (int r = 0; r <= max_ring; ++r) {
            const int zr = dim == 3 ? r : 0;
            for (int dz = -zr; dz <= zr; ++dz)
                for (int dy = -r; dy <= r; ++dy)
                    for (int dx = -r; dx <= r; ++dx) {
                        const int d[3] = {dx, dy, dz};
                        int cheb = 0, lin = 0, mul = 1;
                        bool inside = true;
                        for (int a = 0; a < dim; ++a) {
                            const int c = cq[a] + d[a];
                            cheb = std::max(cheb, std::abs(d[a]));
                            inside = inside && c >= 0 && c < nc[a];
                            lin += c * mul;
                            mul *= nc[a];
                        }
                        if (cheb != r || !inside) continue;
                        for (int t = start[lin]; t < start[lin + 1]; ++t) {
                            const int i = items[t];
                            double d2 = 0.0;
                            for (int a = 0; a < dim; ++a) {
                                const double dd = pts[(size_t)i * dim + a] - x[a];
                                d2 = d2 + dd * dd;
                            }
                            if (d2 < best || (d2 == best && i < bi)) {
                                best = d2;
                                bi = i;
                            }
                        }
                    }
            const double bound = (double)r * cell * (1.0 - 1e-9);
            if (bi >= 0 && best < bound * bound) break;
        }
        out[k] = bi;
    }
    return out;
}

std::string user_name() {
    char name[LOGIN_NAME_MAX + 1] = {0};
    if (getlogin_r(name, sizeof(name)) == 0) return name;
    const char* env = getenv("USER");
    return env ? env : "";
}

}  // namespace

// cloud.cpp:50-86: a kept vertex maps to its node (through the RCM order); a
// deleted one (corner/edge or hanging vertex) to the nearest node, the last
// of equally near nodes winning as there.
void CLOUD::calc_iv_original_nearest_vert(POINTS& points, PARAMETERS& parameters) {
    const int dim = parameters.dimension;
    points.iv_original_nearest_vert.assign(points.nv_original, -1);
    int offset = 0;
    for (int iv0 = 0; iv0 < points.nv_original; ++iv0) {
        if (points.corner_edge_vertices[iv0]) {
            const double* x0 = points.xyz_original.data() + (size_t)dim * iv0;
            double dist_square = std::numeric_limits<double>::infinity();
            for (int iv1 = 0; iv1 < points.nv; ++iv1) {
                double temp = 0.0;
                for (int a = 0; a < dim; ++a) {
                    const double d = points.xyz[(size_t)dim * iv1 + a] - x0[a];
                    temp = temp + d * d;
                }
                if (dist_square >= temp) {
                    dist_square = temp;
                    points.iv_original_nearest_vert[iv0] = iv1;
                }
            }
            offset++;
        } else {
            const int pre = iv0 - offset;  // index before the RCM reorder
            points.iv_original_nearest_vert[iv0] =
                rcm_points_order.empty() ? pre : rcm_points_order[pre];
        }
    }
}

void write_iteration_details(PARAMETERS& parameters) {
    const std::string output_file = parameters.output_file_prefix + "_iteration_details.csv";
    FILE* file = fopen(output_file.c_str(), "w");
    if (!file) fail("write_iteration_details", "cannot open " + output_file);
    const size_t n = parameters.n_iter_actual.size();
    if (parameters.rel_res_log.size() < n || parameters.abs_res_log.size() < n)
        fail("write_iteration_details", "solver logs are shorter than n_iter_actual");
    if (parameters.regul_alpha_log.empty()) {
        fprintf(file, "it,n_iter_actual,rel_res_log,abs_res_log\n");
        for (size_t it = 0; it < n; ++it)
            fprintf(file, "%i,%i,%.16g,%.16g\n", (int)it, parameters.n_iter_actual[it],
                    parameters.rel_res_log[it], parameters.abs_res_log[it]);
    } else {
        if (parameters.regul_alpha_log.size() < n)
            fail("write_iteration_details", "regul_alpha_log is shorter than n_iter_actual");
        fprintf(file, "it,n_iter_actual,rel_res_log,abs_res_log,regul_alpha_log\n");
        for (size_t it = 0; it < n; ++it)
            fprintf(file, "%i,%i,%.16g,%.16g,%.16g\n", (int)it, parameters.n_iter_actual[it],
                    parameters.rel_res_log[it], parameters.abs_res_log[it],
                    parameters.regul_alpha_log[it]);
    }
    fclose(file);
    printf("\nwrite_iteration_details wrote %s\n\n", output_file.c_str());
}

void write_simulation_details(POINTS& points, CLOUD& cloud, PARAMETERS& parameters) {
    char hostname[HOST_NAME_MAX + 1] = {0};
    gethostname(hostname, HOST_NAME_MAX);
    const std::string output_file = parameters.output_file_prefix + "_simulation_details.csv";
    FILE* file = fopen(output_file.c_str(), "w");
    if (!file) fail("write_simulation_details", "cannot open " + output_file);
    fprintf(file, "hostname,%s\n", hostname);
    fprintf(file, "username,%s\n", user_name().c_str());
    fprintf(file, "\n");

    fprintf(file, "Grid,Details\n");
    fprintf(file, "meshfile,name,,,,%s\n", parameters.meshfile.c_str());
    if (!parameters.restart_tecplot_file.empty())
        fprintf(file, "restart,file,,,,%s\n", parameters.restart_tecplot_file.c_str());
    fprintf(file, "Problem,dimension,%i\n", parameters.dimension);
    fprintf(file, "No. of,nodes,nv,%i\n", points.nv);
    int nv_boundary = 0;
    for (int iv = 0; iv < points.nv; ++iv)
        if (points.boundary_flag[iv]) nv_boundary++;
    fprintf(file, "No. of,nodes,nv_boundary,%i\n", nv_boundary);
    fprintf(file, "No. of,nodes,nv_internal,%i\n", points.nv - nv_boundary);
    fprintf(file, "Characteristic,dx_max,%.16g\n", parameters.max_dx);
    fprintf(file, "Characteristic,dx_min,%.16g\n", parameters.min_dx);
    fprintf(file, "Characteristic,dx_avg,%.16g\n", parameters.avg_dx);
    fprintf(file, "\n");

    fprintf(file, "RBF,Details\n");
    fprintf(file, "phs_deg,%i\n", parameters.phs_deg);
    fprintf(file, "poly_deg,%i\n", parameters.poly_deg);
    fprintf(file, "cloud_size_multiplier,%.16g\n", parameters.cloud_size_multiplier);
    fprintf(file, "num_poly_terms,%i\n", parameters.num_poly_terms);
    fprintf(file, "cloud_size,%i\n", parameters.cloud_size);
    fprintf(file, "RBF_cond_num_max,%.16g\n", cloud.cond_num_RBF_max);
    fprintf(file, "RBF_cond_num_min,%.16g\n", cloud.cond_num_RBF_min);
    fprintf(file, "RBF_cond_num_avg,%.16g\n", cloud.cond_num_RBF_avg);
    fprintf(file, "grad_x_eigval,%.16g,%.16g\n", parameters.grad_x_eigval_real,
            parameters.grad_x_eigval_imag);
    fprintf(file, "grad_y_eigval,%.16g,%.16g\n", parameters.grad_y_eigval_real,
            parameters.grad_y_eigval_imag);
    if (parameters.dimension == 3)
        fprintf(file, "grad_z_eigval,%.16g,%.16g\n", parameters.grad_z_eigval_real,
                parameters.grad_z_eigval_imag);
    fprintf(file, "laplace_eigval,%.16g,%.16g\n", parameters.laplace_eigval_real,
            parameters.laplace_eigval_imag);
    fprintf(file, "\n");

    fprintf(file, "Solver,Settings\n");
    fprintf(file, "nt,%i\n", parameters.nt);
    fprintf(file, "nt_actual,%i\n", parameters.nt_actual);
    fprintf(file, "timestep_dt,%.16g,seconds\n", parameters.dt);
    fprintf(file, "steady_tolerance,%.16g\n", parameters.steady_tolerance);
    fprintf(file, "solver_tolerance,%.16g\n", parameters.solver_tolerance);
    fprintf(file, "Courant,%.16g\n", parameters.Courant);
    fprintf(file, "solver_type,%s\n", parameters.solver_type.c_str());
    fprintf(file, "gmres_kdim,%i\n", parameters.gmres_kdim);
    fprintf(file, "euclid_precond_level_hypre,%i\n", parameters.euclid_precond_level_hypre);
    fprintf(file, "precond_droptol,%.16g\n", parameters.precond_droptol);
    fprintf(file, "n_iter,%i\n", parameters.n_iter);
    fprintf(file, "\n");

    fprintf(file, "Fluid and Flow,Details\n");
    fprintf(file, "mu,%.16g\n", parameters.mu);
    fprintf(file, "rho,%.16g\n", parameters.rho);
    fprintf(file, "\n");

    fprintf(file, "Run Times,(seconds)\n");
    fprintf(file, "points_timer,%.16g\n", parameters.points_timer);
    fprintf(file, "cloud_id_timer,%.16g\n", parameters.cloud_id_timer);
    fprintf(file, "rcm_timer,%.16g\n", parameters.rcm_timer);
    fprintf(file, "cloud_misc_timer,%.16g\n", parameters.cloud_misc_timer);
    fprintf(file, "grad_laplace_coeff_timer,%.16g\n", parameters.grad_laplace_coeff_timer);
    fprintf(file, "factoring_timer,%.16g\n", parameters.factoring_timer);
    fprintf(file, "solve_timer,%.16g\n", parameters.solve_timer);
    fprintf(file, "total_timer,%.16g\n", parameters.total_timer);
    fclose(file);
    printf("\nwrite_simulation_details wrote %s\n\n", output_file.c_str());
}

// Values at the original mesh's vertices (deleted ones take their nearest
// node's value) and the mesh's triangles (2D) or tetrahedra (3D), 1-based.
void write_tecplot_steady_variables(POINTS& points, PARAMETERS& parameters,
                                    std::vector<std::string>& variable_names,
                                    std::vector<VectorXd*>& variable_pointers) {
    const int dim = parameters.dimension;
    if ((int)points.iv_original_nearest_vert.size() != points.nv_original)
        fail("write_tecplot_steady_variables", "no original-vertex map (build a CLOUD first)");
    if (variable_names.size() != variable_pointers.size())
        fail("write_tecplot_steady_variables", "names and variables differ in number");
    for (VectorXd* v : variable_pointers)
        if (!v || (int)v->size() < points.nv)
            fail("write_tecplot_steady_variables", "a variable has fewer than nv entries");
    int ncv = 0;
    for (const auto& e : points.elem_vert_original)
        if ((int)e.size() == dim + 1) ncv++;
    const std::string output_file = parameters.output_file_prefix + "_tecplot.plt";
    FILE* file = fopen(output_file.c_str(), "w");
    if (!file) fail("write_tecplot_steady_variables", "cannot open " + output_file);
    fprintf(file, dim == 2 ? "VARIABLES=\"x\", \"y\"" : "VARIABLES=\"x\", \"y\", \"z\"");
    for (const std::string& name : variable_names) fprintf(file, ", \"%s\"", name.c_str());
    fprintf(file, "\n");
    fprintf(file, "ZONE T=\"%.16g\"  DATAPACKING=POINT, NODES=%i, ELEMENTS=%i, ZONETYPE=%s\n", 1.0,
            points.nv_original, ncv, dim == 2 ? "FETRIANGLE" : "FETETRAHEDRON");
    for (int iv0 = 0; iv0 < points.nv_original; ++iv0) {
        const int iv1 = points.iv_original_nearest_vert[iv0];
        for (int a = 0; a < dim; ++a)
            fprintf(file, "%.16g ", points.xyz_original[(size_t)dim * iv0 + a]);
        for (VectorXd* v : variable_pointers) fprintf(file, "%.16g ", (*v)[iv1]);
        fprintf(file, "\n");
    }
    for (const auto& e : points.elem_vert_original)
        if ((int)e.size() == dim + 1) {
            for (int v : e) fprintf(file, "%i ", v + 1);
            fprintf(file, "\n");
        }
    fclose(file);
    printf("\nwrite_tecplot_steady_variables wrote %s\n\n", output_file.c_str());
}

// Reads the named variables of a tecplot point file (as written above) and
// gives every node the values of the file's nearest vertex.  The header names
// a variable as "name", (or "name" when it is the last one).
void read_restart_tecplot_file(POINTS& points, PARAMETERS& parameters, std::string filename,
                               std::vector<std::string>& tecplot_variable_names,
                               std::vector<VectorXd*>& variable_pointers) {
    const int dim = parameters.dimension;
    FILE* file = fopen(filename.c_str(), "r");
    if (!file) fail("read_restart_tecplot_file", "file does not exist: " + filename);
    parameters.restart_tecplot_file = filename;
    if (tecplot_variable_names.size() != variable_pointers.size())
        fail("read_restart_tecplot_file", "names and variables differ in number");
    std::vector<char> buf(1 << 16);
    if (!fgets(buf.data(), (int)buf.size(), file)) {
        fclose(file);
        fail("read_restart_tecplot_file", "empty file: " + filename);
    }
    const std::string first = buf.data();
    std::istringstream ss(first);
    std::vector<std::string> tokens;
    for (std::string t; ss >> t;) tokens.push_back(t);
    std::vector<int> column;
    for (const std::string& name : tecplot_variable_names) {
        int at = -1;
        for (size_t t = 0; t < tokens.size(); ++t)
            if (tokens[t] == "\"" + name + "\"," ||
                (t + 1 == tokens.size() && tokens[t] == "\"" + name + "\""))
                at = (int)t;
        if (at < 0) {
            fclose(file);
            fail("read_restart_tecplot_file",
                 "unable to find variable: " + name + " in first line of file: " + filename);
        }
        column.push_back(at - dim);  // the first dim tokens are the coordinates
    }
    if (!fgets(buf.data(), (int)buf.size(), file)) {
        fclose(file);
        fail("read_restart_tecplot_file", "no second line in " + filename);
    }
    std::string second = buf.data();
    const size_t at = second.find("NODES");
    if (at == std::string::npos) {
        fclose(file);
        fail("read_restart_tecplot_file", "unable to find NODES in second line of file: " + filename);
    }
    size_t p = second.find_first_of("0123456789", at);
    const int nv_tecplot = p == std::string::npos ? 0 : atoi(second.c_str() + p);
    const int nvar = (int)tokens.size() - dim;
    if (nv_tecplot <= 0 || nvar <= 0) {
        fclose(file);
        fail("read_restart_tecplot_file", "no nodes or no variables in " + filename);
    }
    std::vector<double> xyz_t((size_t)nv_tecplot * dim), vals((size_t)nv_tecplot * nvar);
    for (int iv = 0; iv < nv_tecplot; ++iv) {
        for (int a = 0; a < dim; ++a)
            if (fscanf(file, "%lf ", &xyz_t[(size_t)iv * dim + a]) != 1) {
                fclose(file);
                fail("read_restart_tecplot_file", "truncated node data in " + filename);
            }
        for (int k = 0; k < nvar; ++k)
            if (fscanf(file, "%lf ", &vals[(size_t)iv * nvar + k]) != 1) {
                fclose(file);
                fail("read_restart_tecplot_file", "truncated node data in " + filename);
            }
    }
    fclose(file);
    const std::vector<int> nearest = nearest_points(dim, xyz_t, nv_tecplot, points.xyz, points.nv);
    for (size_t k = 0; k < variable_pointers.size(); ++k) {
        VectorXd& v = *variable_pointers[k];
        if ((int)v.size() < points.nv) v.resize(points.nv);
        for (int iv = 0; iv < points.nv; ++iv)
            v[iv] = vals[(size_t)nearest[iv] * nvar + column[k]];
    }
}
