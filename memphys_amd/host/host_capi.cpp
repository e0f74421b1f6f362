// extern "C" entry points of the host set-up library (libmemphys_host.so), for
// callers that are not C++ (the Python tests and bench drive set-up through
// these; C++ callers use the class API in memphys.hpp).
#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

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

}  // extern "C"
