// Host-side set-up of a meshless point cloud: seeded jittered box clouds,
// k-nearest-neighbour stencil selection, reverse Cuthill-McKee ordering and
// the graded monomial exponent table.  Everything here is plain C++17 on the
// CPU; the per-step work lives in the HIP library behind include/memphys_hip.h.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace memphys {

// std::mt19937_64 with an explicit 53-bit mapping to [0, 1): u = (x >> 11) * 2^-53.
// (std::uniform_real_distribution differs across standard libraries.)
struct Rng53 {
    explicit Rng53(uint64_t seed);
    double uniform();  // [0, 1)
    uint64_t next();
  private:
    uint64_t mt[312];
    int idx;
};

// A box cloud on [0,1]^dim built from an n^dim lattice (h = 1/(n-1)).
//  * interior lattice nodes are jittered by U(-jitter*h, +jitter*h) per axis,
//    drawn in lattice order (axis 0 fastest), one draw per axis;
//  * face nodes stay on their faces, unjittered, with outward unit normals;
//  * corner (2D) and edge/corner (3D) nodes are removed, as the reference
//    removes them from Gmsh clouds (points.cpp:323-352);
//  * stencil size per node: k_layer[L] for lattice layer L = 0 (face), 1, 2
//    (L = min over axes of min(i, n-1-i)), k_interior otherwise.
struct BoxCloud {
    int dim = 0;
    int nv = 0;
    std::vector<double> xyz;       // nv*dim, AoS
    std::vector<uint8_t> boundary; // nv
    std::vector<double> normal;    // nv*dim (zero on interior nodes)
    std::vector<int> stencil_size; // nv
    std::vector<int> layer;        // nv (lattice layer index)
};

BoxCloud generate_box_cloud(int dim, int n, double jitter, uint64_t seed,
                            int k_interior, const int k_layer[3]);

// Stencil selection, following CLOUD::calc_cloud_points_fast (cloud.cpp:570-648):
//  interior node: its k nearest nodes among all nodes (itself first, distance 0);
//  boundary node: itself, then its k-1 nearest *interior* nodes.
// Neighbours are sorted by (squared distance, index); exact ties, which the
// reference leaves to kd-tree traversal order, are broken by the lower index.
// k may vary per node.  Returns CSR (row_ptr nv+1, col).
void knn_stencils(int nv, int dim, const double* xyz, const uint8_t* boundary,
                  const int* k_per_node, std::vector<int>& row_ptr, std::vector<int>& col,
                  int n_threads);

// Reverse Cuthill-McKee order over the directed stencil graph, following
// cuthill_mckee_ordering / reverse_cuthill_mckee_ordering
// (general_functions.cpp:807-879): start at the first node of minimum degree,
// visit neighbours in stencil order, append each newly visited set sorted by
// (degree, index), then reverse.  order[old] = new.  Nodes the start cannot
// reach are handled by restarting at the lowest-index unvisited node of minimum
// degree (the reference indexes past the end in that case).
void rcm_order(int nv, const int* row_ptr, const int* col, std::vector<int>& order);

// Applies order (old -> new) to point arrays and CSR as CLOUD::re_order_points
// does (cloud.cpp:338-439): rows move, column ids are renumbered, the
// within-row neighbour order is preserved.
void apply_order(int nv, int dim, const std::vector<int>& order,
                 std::vector<double>& xyz, std::vector<double>& normal,
                 std::vector<uint8_t>& boundary, std::vector<int>& extra_int,
                 std::vector<int>& row_ptr, std::vector<int>& col);

// A Gmsh 2.2 ASCII mesh read as POINTS does (points.cpp:27-72, 323-406,
// 498-794): node coordinates (z dropped in 2D); boundary nodes = nodes of the
// boundary elements (2D: 2-node lines; 3D: 3-node triangles, 4-node quads);
// outward unit normals averaged over a node's boundary elements, each element
// normal oriented away from the centroid of the first volume element at its
// first node; corner points (2D: 1-node elements), edge nodes (3D: nodes of
// 2-node lines) and nodes of no element are deleted.  The dimension is 3
// when any volume element type is present (parameters.cpp:132-215).
struct GmshCloud {
    int dim = 0;
    int nv = 0, nv_original = 0, nelem = 0;
    std::vector<double> xyz;        // nv*dim, AoS
    std::vector<uint8_t> boundary;  // nv
    std::vector<double> normal;     // nv*dim (zero on interior nodes)
    std::vector<int> original;      // nv: index of each kept node in the file's node list
    std::vector<double> xyz_original;  // nv_original*dim: every node of the file, file order
    // elements as POINTS::read_elem_vert_complete_msh keeps them (points.cpp:568-645):
    // 2D lines and triangles, 3D triangles and tetrahedra, node indices in file order
    std::vector<std::vector<int>> elem_vert_original;
};
GmshCloud read_gmsh22(const std::string& path);

// Graded monomial table of PARAMETERS::calc_polynomial_term_exponents
// (parameters.cpp:217-267), row-major [m x dim].
std::vector<int> polynomial_exponents(int dim, int poly_deg);
int num_poly_terms(int dim, int poly_deg);

}  // namespace memphys
