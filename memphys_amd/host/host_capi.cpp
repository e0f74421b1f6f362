// extern "C" entry points of the host set-up library (libmemphys_host.so), for
// callers that are not C++ (the Python tests and bench drive set-up through
// these; C++ callers use the class API in memphys.hpp).
#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

#include "../csrc/amg_build.hpp"
#include "../csrc/eig_host.hpp"
#include "cloud_setup.hpp"

extern "C" {

int mph_box_cloud_count(int dim, int n) {
    if ((dim != 2 && dim != 3) || n < 4) return -1;
    if (dim == 2) return n * n - 4;
    return n * n * n - 12 * (n - 2) - 8;
}

// Fills caller arrays sized by mph_box_cloud_count: xyz/normal nv*dim,
// boundary/ksize/layer nv.  k_layer: stencil sizes of lattice layers 0..2.
int mph_box_cloud_fill(int dim, int n, double jitter, uint64_t seed, int k_interior,
                       const int* k_layer, double* xyz, uint8_t* boundary, double* normal,
                       int* ksize, int* layer) {
    try {
        memphys::BoxCloud c = memphys::generate_box_cloud(dim, n, jitter, seed, k_interior, k_layer);
        std::memcpy(xyz, c.xyz.data(), sizeof(double) * c.xyz.size());
        std::memcpy(normal, c.normal.data(), sizeof(double) * c.normal.size());
        std::memcpy(boundary, c.boundary.data(), c.boundary.size());
        std::memcpy(ksize, c.stencil_size.data(), sizeof(int) * c.stencil_size.size());
        std::memcpy(layer, c.layer.data(), sizeof(int) * c.layer.size());
        return c.nv;
    } catch (const std::exception&) {
        return -1;
    }
}

// First `count` outputs of the 53-bit uniform stream (for generator checks).
int mph_rng_uniform(uint64_t seed, int count, double* out, uint64_t* raw) {
    memphys::Rng53 a(seed), b(seed);
    for (int i = 0; i < count; ++i) {
        if (out) out[i] = a.uniform();
        if (raw) raw[i] = b.next();
    }
    return 0;
}

// row_ptr: nv+1, col: sum(k) (caller allocated).
int mph_knn(int nv, int dim, const double* xyz, const uint8_t* boundary, const int* k_per_node,
            int* row_ptr, int* col, int n_threads) {
    try {
        std::vector<int> rp, c;
        memphys::knn_stencils(nv, dim, xyz, boundary, k_per_node, rp, c, n_threads);
        std::memcpy(row_ptr, rp.data(), sizeof(int) * rp.size());
        std::memcpy(col, c.data(), sizeof(int) * c.size());
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int mph_rcm(int nv, const int* row_ptr, const int* col, int* order) {
    try {
        std::vector<int> o;
        memphys::rcm_order(nv, row_ptr, col, o);
        std::memcpy(order, o.data(), sizeof(int) * o.size());
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// Applies order (old -> new) in place to xyz, normal (nv*dim), boundary (nv),
// extra (nv ints, may be NULL) and the CSR.
int mph_apply_order(int nv, int dim, const int* order, double* xyz, double* normal,
                    uint8_t* boundary, int* extra, int* row_ptr, int* col) {
    try {
        std::vector<int> o(order, order + nv);
        std::vector<double> x(xyz, xyz + (size_t)nv * dim), nr(normal, normal + (size_t)nv * dim);
        std::vector<uint8_t> b(boundary, boundary + nv);
        std::vector<int> e;
        if (extra) e.assign(extra, extra + nv);
        std::vector<int> rp(row_ptr, row_ptr + nv + 1), c(col, col + row_ptr[nv]);
        memphys::apply_order(nv, dim, o, x, nr, b, e, rp, c);
        std::memcpy(xyz, x.data(), sizeof(double) * x.size());
        std::memcpy(normal, nr.data(), sizeof(double) * nr.size());
        std::memcpy(boundary, b.data(), b.size());
        if (extra) std::memcpy(extra, e.data(), sizeof(int) * e.size());
        std::memcpy(row_ptr, rp.data(), sizeof(int) * rp.size());
        std::memcpy(col, c.data(), sizeof(int) * c.size());
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int mph_poly_exponents(int dim, int poly_deg, int* out) {
    try {
        std::vector<int> e = memphys::polynomial_exponents(dim, poly_deg);
        std::memcpy(out, e.data(), sizeof(int) * e.size());
        return (int)(e.size() / dim);
    } catch (const std::exception&) {
        return -1;
    }
}

// Gmsh 2.2 mesh read as POINTS(PARAMETERS&) does: first call with xyz == NULL
// returns the node count and dimension; the second fills xyz/normal (nv*dim),
// boundary (nv) and original (nv, index in the file's node list; may be NULL).
int mph_gmsh_read(const char* path, int* dim, int* nv, double* xyz, uint8_t* boundary,
                  double* normal, int* original) {
    try {
        memphys::GmshCloud g = memphys::read_gmsh22(path);
        if (dim) *dim = g.dim;
        if (nv) *nv = g.nv;
        if (xyz) {
            std::memcpy(xyz, g.xyz.data(), sizeof(double) * g.xyz.size());
            std::memcpy(boundary, g.boundary.data(), g.boundary.size());
            std::memcpy(normal, g.normal.data(), sizeof(double) * g.normal.size());
            if (original) std::memcpy(original, g.original.data(), sizeof(int) * g.original.size());
        }
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// First coarsening step of the multigrid set-up the HIP library runs
// (amg_build.hpp), for CPU checks: aggregates (n entries), the smoothed
// prolongator P (n x n_agg, truncated at `trunc`) and the Galerkin operator
// R A P (n_agg square, sparsified at `drop`; 0 keeps every entry) as
// CSR into caller buffers of the given capacities.  Returns 0, -1 on error, or
// 1 when a capacity is too small (p_ptr[n] / ac_ptr[n_agg] then hold the sizes
// needed when they fit).
int mph_amg_first_level(int n, const int* ptr, const int* col, const double* val, int k,
                        double omega_p, double trunc, double drop, int* agg, int* n_agg,
                        int* p_ptr, int* p_col,
                        double* p_val, long p_cap, int* ac_ptr, int* ac_col, double* ac_val,
                        long ac_cap) {
    try {
        mp_amg::Csr A;
        A.nrows = A.ncols = n;
        A.ptr.assign(ptr, ptr + n + 1);
        A.col.assign(col, col + ptr[n]);
        A.val.assign(val, val + ptr[n]);
        std::vector<int> a;
        int na = 0;
        mp_amg::aggregate(A, k, a, na);
        mp_amg::Csr P = mp_amg::smoothed_prolongator(A, a, na, omega_p, trunc);
        mp_amg::Csr R = mp_amg::transpose(P);
        mp_amg::Csr Ac = mp_amg::multiply(R, mp_amg::multiply(A, P, 0), 0);
        if (drop > 0.0) Ac = mp_amg::sparsify(Ac, drop);
        std::memcpy(agg, a.data(), sizeof(int) * n);
        *n_agg = na;
        std::memcpy(p_ptr, P.ptr.data(), sizeof(int) * (n + 1));
        std::memcpy(ac_ptr, Ac.ptr.data(), sizeof(int) * (na + 1));
        if (P.nnz() > p_cap || Ac.nnz() > ac_cap) return 1;
        std::memcpy(p_col, P.col.data(), sizeof(int) * P.nnz());
        std::memcpy(p_val, P.val.data(), sizeof(double) * P.nnz());
        std::memcpy(ac_col, Ac.col.data(), sizeof(int) * Ac.nnz());
        std::memcpy(ac_val, Ac.val.data(), sizeof(double) * Ac.nnz());
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// Eigenvalues (re, im: m each) of an upper Hessenberg matrix H (row-major
// m x m) and the unit eigenvector (vre, vim) of eigenvalue number `which`, as
// the GPU Arnoldi iteration computes them.
int mph_hessenberg_eig(int m, const double* H, double* re, double* im, int which, double* vre,
                       double* vim) {
    try {
        std::vector<double> h(H, H + (size_t)m * m);
        std::vector<mp_eig::cplx> ev = mp_eig::hessenberg_eigenvalues(h, m);
        for (int i = 0; i < m; ++i) {
            re[i] = ev[i].real();
            im[i] = ev[i].imag();
        }
        if (vre && which >= 0 && which < m) {
            std::vector<mp_eig::cplx> y = mp_eig::hessenberg_eigenvector(h, m, ev[which]);
            for (int i = 0; i < m; ++i) {
                vre[i] = y[i].real();
                vim[i] = y[i].imag();
            }
        }
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// Top-left n x n block of the inverse of [[A, 1], [1^T, 0]] (the coarsest
// multigrid solve), row-major into G.
int mph_amg_bordered_inverse(int n, const int* ptr, const int* col, const double* val,
                             double* G) {
    try {
        mp_amg::Csr A;
        A.nrows = A.ncols = n;
        A.ptr.assign(ptr, ptr + n + 1);
        A.col.assign(col, col + ptr[n]);
        A.val.assign(val, val + ptr[n]);
        std::vector<double> g = mp_amg::bordered_inverse(A);
        std::memcpy(G, g.data(), sizeof(double) * g.size());
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// Inverse of A (the coarsest multigrid solve of a system with Dirichlet
// rows), row-major into G.
int mph_amg_dense_inverse(int n, const int* ptr, const int* col, const double* val, double* G) {
    try {
        mp_amg::Csr A;
        A.nrows = A.ncols = n;
        A.ptr.assign(ptr, ptr + n + 1);
        A.col.assign(col, col + ptr[n]);
        A.val.assign(val, val + ptr[n]);
        std::vector<double> g = mp_amg::dense_inverse(A);
        std::memcpy(G, g.data(), sizeof(double) * g.size());
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

}  // extern "C"
