// Host-side cloud set-up: generator, kNN stencils, RCM, monomial exponents.
#include "cloud_setup.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>
#include <unordered_map>
#include <numeric>
#include <queue>
#include <random>
#include <stdexcept>
#include <utility>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace memphys {

Rng53::Rng53(uint64_t seed) {
    // Seeding and generation as std::mt19937_64 (Matsumoto-Nishimura MT19937-64),
    // written out so the stream is fixed by this file alone.
    mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
    idx = 312;
}

uint64_t Rng53::next() {
    const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
    if (idx >= 312) {
        for (int i = 0; i < 312; ++i) {
            uint64_t x = (mt[i] & UM) | (mt[(i + 1) % 312] & LM);
            uint64_t xa = x >> 1;
            if (x & 1ULL) xa ^= 0xB5026F5AA96619E9ULL;
            mt[i] = mt[(i + 156) % 312] ^ xa;
        }
        idx = 0;
    }
    uint64_t x = mt[idx++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

double Rng53::uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }

BoxCloud generate_box_cloud(int dim, int n, double jitter, uint64_t seed, int k_interior,
                            const int k_layer[3]) {
    if (dim != 2 && dim != 3) throw std::invalid_argument("generate_box_cloud: dim must be 2 or 3");
    if (n < 4) throw std::invalid_argument("generate_box_cloud: n must be >= 4");
    BoxCloud c;
    c.dim = dim;
    const double h = 1.0 / (double)(n - 1);
    Rng53 rng(seed);
    const int nz = (dim == 3) ? n : 1;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int id[3] = {i, j, k};
                int n_ext = 0, L = n;
                for (int a = 0; a < dim; ++a) {
                    if (id[a] == 0 || id[a] == n - 1) ++n_ext;
                    L = std::min(L, std::min(id[a], n - 1 - id[a]));
                }
                if (n_ext >= 2) continue;  // corner (2D) or edge/corner (3D): removed
                double p[3], nrm[3] = {0.0, 0.0, 0.0};
                for (int a = 0; a < dim; ++a) p[a] = (id[a] == n - 1) ? 1.0 : (double)id[a] * h;
                if (n_ext == 0) {
                    for (int a = 0; a < dim; ++a) {
                        double u = rng.uniform();
                        p[a] = p[a] + (2.0 * u - 1.0) * jitter * h;
                    }
                } else {
                    for (int a = 0; a < dim; ++a) {
                        if (id[a] == 0) nrm[a] = -1.0;
                        if (id[a] == n - 1) nrm[a] = 1.0;
                    }
                }
                for (int a = 0; a < dim; ++a) {
                    c.xyz.push_back(p[a]);
                    c.normal.push_back(nrm[a]);
                }
                c.boundary.push_back(n_ext == 1 ? 1 : 0);
                c.layer.push_back(L);
                c.stencil_size.push_back(L < 3 ? k_layer[L] : k_interior);
            }
    c.nv = (int)c.boundary.size();
    return c;
}

namespace {

struct Grid {
    int dim = 0;
    double lo[3] = {0, 0, 0};
    double cell = 1.0;
    int nc[3] = {1, 1, 1};
    std::vector<int> start;  // cell -> first item
    std::vector<int> items;  // node ids sorted by cell

    int cell_of(const double* p, int* ci) const {
        int lin = 0, mul = 1;
        for (int a = 0; a < dim; ++a) {
            int c = (int)std::floor((p[a] - lo[a]) / cell);
            c = std::max(0, std::min(nc[a] - 1, c));
            ci[a] = c;
            lin += c * mul;
            mul *= nc[a];
        }
        return lin;
    }
};

Grid build_grid(int dim, const double* xyz, const std::vector<int>& ids) {
    Grid g;
    g.dim = dim;
    double hi[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) {
        g.lo[a] = INFINITY;
        hi[a] = -INFINITY;
    }
    for (int id : ids)
        for (int a = 0; a < dim; ++a) {
            g.lo[a] = std::min(g.lo[a], xyz[dim * id + a]);
            hi[a] = std::max(hi[a], xyz[dim * id + a]);
        }
    double vol = 1.0;
    for (int a = 0; a < dim; ++a) vol *= std::max(hi[a] - g.lo[a], 1e-300);
    // about two points per cell
    g.cell = std::pow(2.0 * vol / (double)std::max<size_t>(ids.size(), 1), 1.0 / dim);
    if (!(g.cell > 0)) g.cell = 1.0;
    size_t ncell = 1;
    for (int a = 0; a < dim; ++a) {
        g.nc[a] = std::max(1, (int)std::ceil((hi[a] - g.lo[a]) / g.cell));
        ncell *= (size_t)g.nc[a];
    }
    std::vector<int> cnt(ncell + 1, 0), cid(ids.size());
    int ci[3];
    for (size_t t = 0; t < ids.size(); ++t) {
        cid[t] = g.cell_of(xyz + dim * ids[t], ci);
        cnt[cid[t] + 1]++;
    }
    for (size_t c = 0; c < ncell; ++c) cnt[c + 1] += cnt[c];
    g.start = cnt;
    g.items.resize(ids.size());
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (size_t t = 0; t < ids.size(); ++t) g.items[fill[cid[t]]++] = ids[t];
    return g;
}

using Cand = std::pair<double, int>;  // (squared distance, node id)

// k nearest items of grid g to point q, sorted by (d2, id).
void grid_knn(const Grid& g, int dim, const double* xyz, const double* q, int k,
              std::vector<Cand>& heap) {
    heap.clear();
    int cq[3] = {0, 0, 0};
    g.cell_of(q, cq);
    auto cmp = [](const Cand& a, const Cand& b) { return a < b; };  // max-heap on (d2, id)
    int max_ring = 0;
    for (int a = 0; a < dim; ++a) max_ring = std::max(max_ring, g.nc[a]);
    for (int r = 0; r <= max_ring; ++r) {
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a) {
            lo[a] = (a < dim) ? cq[a] - r : 0;
            hi[a] = (a < dim) ? cq[a] + r : 0;
        }
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    int c3[3] = {x, y, z};
                    int cheb = 0;
                    bool inside = true;
                    for (int a = 0; a < dim; ++a) {
                        cheb = std::max(cheb, std::abs(c3[a] - cq[a]));
                        if (c3[a] < 0 || c3[a] >= g.nc[a]) inside = false;
                    }
                    if (cheb != r || !inside) continue;
                    int lin = 0, mul = 1;
                    for (int a = 0; a < dim; ++a) {
                        lin += c3[a] * mul;
                        mul *= g.nc[a];
                    }
                    for (int t = g.start[lin]; t < g.start[lin + 1]; ++t) {
                        int id = g.items[t];
                        const double* p = xyz + dim * id;
                        double d2 = 0.0;
                        for (int a = 0; a < dim; ++a) {
                            double d = p[a] - q[a];
                            d2 = d2 + d * d;
                        }
                        Cand cand(d2, id);
                        if ((int)heap.size() < k) {
                            heap.push_back(cand);
                            std::push_heap(heap.begin(), heap.end(), cmp);
                        } else if (cand < heap.front()) {
                            std::pop_heap(heap.begin(), heap.end(), cmp);
                            heap.back() = cand;
                            std::push_heap(heap.begin(), heap.end(), cmp);
                        }
                    }
                }
        // every unvisited item lies in a ring >= r+1, at distance >= r*cell
        if ((int)heap.size() == k) {
            double bound = (double)r * g.cell;
            if (heap.front().first < bound * bound) break;
        }
    }
    std::sort(heap.begin(), heap.end());
}

}  // namespace

void knn_stencils(int nv, int dim, const double* xyz, const uint8_t* boundary,
                  const int* k_per_node, std::vector<int>& row_ptr, std::vector<int>& col,
                  int n_threads) {
    std::vector<int> all(nv), interior;
    std::iota(all.begin(), all.end(), 0);
    for (int i = 0; i < nv; ++i)
        if (!boundary[i]) interior.push_back(i);
    for (int i = 0; i < nv; ++i) {
        int need = boundary[i] ? k_per_node[i] - 1 : k_per_node[i];
        size_t avail = boundary[i] ? interior.size() : all.size();
        if (k_per_node[i] < 1 || (size_t)need > avail)
            throw std::invalid_argument("knn_stencils: stencil size exceeds available nodes");
    }
    Grid g_all = build_grid(dim, xyz, all);
    Grid g_int = build_grid(dim, xyz, interior);
    row_ptr.assign(nv + 1, 0);
    for (int i = 0; i < nv; ++i) row_ptr[i + 1] = row_ptr[i] + k_per_node[i];
    col.assign(row_ptr[nv], -1);
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
#endif
    {
        std::vector<Cand> heap;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1024)
#endif
        for (int i = 0; i < nv; ++i) {
            const double* q = xyz + dim * i;
            int* out = col.data() + row_ptr[i];
            if (boundary[i]) {
                out[0] = i;
                grid_knn(g_int, dim, xyz, q, k_per_node[i] - 1, heap);
                for (size_t t = 0; t < heap.size(); ++t) out[1 + t] = heap[t].second;
            } else {
                grid_knn(g_all, dim, xyz, q, k_per_node[i], heap);
                // the node itself leads its stencil (distance 0); an exact duplicate
                // point of lower index is moved behind it
                size_t self = heap.size();
                for (size_t t = 0; t < heap.size(); ++t)
                    if (heap[t].second == i) self = t;
                if (self == heap.size()) {  // not found: only if k duplicates coincide
                    heap.pop_back();
                    heap.insert(heap.begin(), Cand(0.0, i));
                } else if (self != 0) {
                    Cand s = heap[self];
                    heap.erase(heap.begin() + self);
                    heap.insert(heap.begin(), s);
                }
                for (size_t t = 0; t < heap.size(); ++t) out[t] = heap[t].second;
            }
        }
    }
}

void rcm_order(int nv, const int* row_ptr, const int* col, std::vector<int>& order) {
    std::vector<int> degree(nv);
    for (int i = 0; i < nv; ++i) degree[i] = row_ptr[i + 1] - row_ptr[i];
    std::vector<char> visited(nv, 0);
    std::vector<int> numbering;
    numbering.reserve(nv);
    std::vector<std::pair<int, int>> new_set;
    size_t head = 0;
    while ((int)numbering.size() < nv) {
        // first node (lowest index) of minimum degree among the unvisited
        int start = -1;
        for (int i = 0; i < nv; ++i)
            if (!visited[i] && (start < 0 || degree[i] < degree[start])) start = i;
        numbering.push_back(start);
        visited[start] = 1;
        for (; head < numbering.size() && (int)numbering.size() < nv; ++head) {
            int v = numbering[head];
            new_set.clear();
            for (int t = row_ptr[v]; t < row_ptr[v + 1]; ++t) {
                int c = col[t];
                if (!visited[c]) {
                    visited[c] = 1;
                    new_set.push_back(std::make_pair(degree[c], c));
                }
            }
            std::sort(new_set.begin(), new_set.end());
            for (auto& pr : new_set) numbering.push_back(pr.second);
        }
        head = numbering.size();
    }
    order.assign(nv, -1);
    for (int i = 0; i < nv; ++i) order[numbering[i]] = i;
    for (int i = 0; i < nv; ++i) order[i] = nv - order[i] - 1;
}

void apply_order(int nv, int dim, const std::vector<int>& order, std::vector<double>& xyz,
                 std::vector<double>& normal, std::vector<uint8_t>& boundary,
                 std::vector<int>& extra_int, std::vector<int>& row_ptr, std::vector<int>& col) {
    std::vector<double> x2(xyz.size()), n2(normal.size());
    std::vector<uint8_t> b2(boundary.size());
    std::vector<int> e2(extra_int.size());
    for (int iv = 0; iv < nv; ++iv) {
        int nw = order[iv];
        for (int a = 0; a < dim; ++a) {
            x2[dim * nw + a] = xyz[dim * iv + a];
            if (!normal.empty()) n2[dim * nw + a] = normal[dim * iv + a];
        }
        b2[nw] = boundary[iv];
        if (!extra_int.empty()) e2[nw] = extra_int[iv];
    }
    std::vector<int> len(nv);
    for (int iv = 0; iv < nv; ++iv) len[order[iv]] = row_ptr[iv + 1] - row_ptr[iv];
    std::vector<int> rp(nv + 1, 0);
    for (int i = 0; i < nv; ++i) rp[i + 1] = rp[i] + len[i];
    std::vector<int> c2(col.size());
    for (int iv = 0; iv < nv; ++iv) {
        int nw = order[iv];
        int dst = rp[nw];
        for (int t = row_ptr[iv]; t < row_ptr[iv + 1]; ++t) c2[dst++] = order[col[t]];
    }
    xyz.swap(x2);
    if (!normal.empty()) normal.swap(n2);
    boundary.swap(b2);
    if (!extra_int.empty()) extra_int.swap(e2);
    row_ptr.swap(rp);
    col.swap(c2);
}

int num_poly_terms(int dim, int poly_deg) {
    if (dim == 2) return (poly_deg + 1) * (poly_deg + 2) / 2;
    return (poly_deg + 1) * (poly_deg + 2) * (poly_deg + 3) / 6;
}

std::vector<int> polynomial_exponents(int dim, int poly_deg) {
    if (dim != 2 && dim != 3) throw std::invalid_argument("polynomial_exponents: dim must be 2 or 3");
    const int m = num_poly_terms(dim, poly_deg);
    std::vector<int> e(m * dim, 0);
    auto E = [&](int r, int c) -> int& { return e[r * dim + c]; };
    if (poly_deg >= 1)
        for (int a = 0; a < dim; ++a) E(1 + a, a) = 1;
    int prev = 1;  // first row of the previous degree's block
    for (int deg = 2; deg <= poly_deg; ++deg) {
        int row = prev + ((dim == 2) ? deg : deg * (deg + 1) / 2);  // first row of this block
        int out = row;
        if (dim == 2) {
            for (int i = prev; i < prev + deg; ++i) {  // multiply each degree-(deg-1) term by x
                E(out, 0) = E(i, 0) + 1;
                E(out, 1) = E(i, 1);
                ++out;
            }
            E(out, 0) = 0;
            E(out, 1) = deg;
            prev = prev + deg;
        } else {
            int cnt = deg * (deg + 1) / 2;  // number of degree-(deg-1) terms
            for (int i = prev; i < prev + cnt; ++i) {  // times x
                E(out, 0) = E(i, 0) + 1;
                E(out, 1) = E(i, 1);
                E(out, 2) = E(i, 2);
                ++out;
            }
            for (int i = prev + deg * (deg - 1) / 2; i < prev + cnt; ++i) {  // x-free terms times y
                E(out, 0) = E(i, 0);
                E(out, 1) = E(i, 1) + 1;
                E(out, 2) = E(i, 2);
                ++out;
            }
            E(out, 0) = 0;
            E(out, 1) = 0;
            E(out, 2) = deg;
            prev = prev + cnt;
        }
    }
    return e;
}

// ---------------------------------------------------------------- Gmsh 2.2

namespace {

bool is_volume_type(int t) {
    return (t >= 4 && t <= 7) || (t >= 11 && t <= 14) || (t >= 17 && t <= 19) ||
           (t >= 29 && t <= 31);
}

int nodes_of_type(int t) {
    switch (t) {
        case 15: return 1;  // point
        case 1: return 2;   // line
        case 2: return 3;   // triangle
        case 3: return 4;   // quadrangle
        case 4: return 4;   // tetrahedron
        default: return -1;
    }
}

}  // namespace

GmshCloud read_gmsh22(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("read_gmsh22: cannot open " + path);
    std::string tok;
    while (f >> tok && tok != "$MeshFormat") {}
    if (tok != "$MeshFormat") throw std::runtime_error("read_gmsh22: $MeshFormat not found");
    std::string v1, v2, v3;
    f >> v1 >> v2 >> v3;
    if (v1 != "2.2" || v2 != "0" || v3 != "8")
        throw std::runtime_error("read_gmsh22: MeshFormat should be 2.2 0 8");
    while (f >> tok && tok != "$Nodes") {}
    if (tok != "$Nodes") throw std::runtime_error("read_gmsh22: $Nodes not found");
    int nn = 0;
    f >> nn;
    if (nn <= 0) throw std::runtime_error("read_gmsh22: no nodes");
    std::vector<double> x3((size_t)nn * 3);
    std::vector<int> ids(nn);
    std::unordered_map<int, int> index_of;
    for (int i = 0; i < nn; ++i) {
        f >> ids[i] >> x3[3 * i] >> x3[3 * i + 1] >> x3[3 * i + 2];
        index_of[ids[i]] = i;
    }
    if (!f) throw std::runtime_error("read_gmsh22: truncated $Nodes");
    while (f >> tok && tok != "$Elements") {}
    if (tok != "$Elements") throw std::runtime_error("read_gmsh22: $Elements not found");
    int ne = 0;
    f >> ne;
    struct Elem {
        int type;
        std::vector<int> v;
    };
    std::vector<Elem> elems;
    elems.reserve(ne);
    int dim = 2;
    std::string line;
    std::getline(f, line);
    for (int e = 0; e < ne; ++e) {
        if (!std::getline(f, line)) throw std::runtime_error("read_gmsh22: truncated $Elements");
        std::istringstream ls(line);
        int id = 0, type = 0, ntags = 0;
        ls >> id >> type >> ntags;
        for (int t = 0; t < ntags; ++t) {
            int tag;
            ls >> tag;
        }
        Elem el{type, {}};
        const int k = nodes_of_type(type);
        for (int j = 0; j < std::max(0, k); ++j) {
            int nid;
            ls >> nid;
            auto it = index_of.find(nid);
            if (it == index_of.end()) throw std::runtime_error("read_gmsh22: unknown node id");
            el.v.push_back(it->second);
        }
        if (is_volume_type(type)) dim = 3;
        elems.push_back(std::move(el));
    }
    // boundary and corner/edge flags (read_points_flag_msh)
    std::vector<uint8_t> bflag(nn, 0), drop(nn, 0);
    for (const Elem& el : elems) {
        if (dim == 2) {
            if (el.type == 1)
                for (int v : el.v) bflag[v] = 1;
            else if (el.type == 15)
                drop[el.v[0]] = 1;
        } else {
            if (el.type == 2 || el.type == 3)
                for (int v : el.v) bflag[v] = 1;
            else if (el.type == 1)
                for (int v : el.v) drop[v] = 1;
        }
    }
    // elements of the normal computation (read_elem_vert_complete_msh):
    // 2D: lines (boundary) and triangles (volume); 3D: triangles (boundary) and
    // tetrahedra (volume); points and (in 3D) lines are skipped
    std::vector<int> el_idx;
    std::vector<uint8_t> el_bnd;
    for (int e = 0; e < ne; ++e) {
        const int t = elems[e].type;
        if (t == 15 || (t == 1 && dim == 3)) continue;
        if (t == 1 || t == 2 || t == 4) {
            el_idx.push_back(e);
            el_bnd.push_back((t == 1 && dim == 2) || (t == 2 && dim == 3) ? 1 : 0);
        } else {
            throw std::runtime_error("read_gmsh22: unsupported element type " + std::to_string(t));
        }
    }
    const int ncv = (int)el_idx.size();
    std::vector<std::vector<int>> vert_cv(nn);
    for (int c = 0; c < ncv; ++c)
        for (int v : elems[el_idx[c]].v) vert_cv[v].push_back(c);
    for (int v = 0; v < nn; ++v)
        if (vert_cv[v].empty()) drop[v] = 1;  // hanging node
    auto X = [&](int v, int a) { return x3[3 * v + a]; };
    std::vector<double> en((size_t)ncv * 3, 0.0);
    for (int c = 0; c < ncv; ++c) {
        if (!el_bnd[c]) continue;
        const std::vector<int>& v = elems[el_idx[c]].v;
        double nrm[3] = {0, 0, 0};
        if (dim == 2) {
            nrm[0] = X(v[1], 1) - X(v[0], 1);
            nrm[1] = -(X(v[1], 0) - X(v[0], 0));
        } else {
            double a[3], b[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = X(v[1], k) - X(v[0], k);
                b[k] = X(v[2], k) - X(v[1], k);
            }
            nrm[0] = a[1] * b[2] - a[2] * b[1];
            nrm[1] = a[2] * b[0] - a[0] * b[2];
            nrm[2] = a[0] * b[1] - a[1] * b[0];
        }
        int inner = -1;
        for (int c2 : vert_cv[v[0]])
            if (!el_bnd[c2]) {
                inner = c2;
                break;
            }
        if (inner < 0)
            throw std::runtime_error("read_gmsh22: boundary element without a volume neighbour");
        const std::vector<int>& iv = elems[el_idx[inner]].v;
        double cen[3] = {0, 0, 0}, dir[3], dot = 0.0, mag = 0.0;
        for (int k = 0; k < dim; ++k) {
            for (int u : iv) cen[k] += X(u, k);
            cen[k] = cen[k] / (double)iv.size();
            dir[k] = X(v[1], k) - cen[k];
            dot += dir[k] * nrm[k];
        }
        if (dot < 0)
            for (int k = 0; k < dim; ++k) nrm[k] = -nrm[k];
        for (int k = 0; k < dim; ++k) mag += nrm[k] * nrm[k];
        mag = std::sqrt(mag);
        for (int k = 0; k < dim; ++k) en[3 * c + k] = nrm[k] / mag;
    }
    // node normals: mean of the boundary elements' normals, normalised
    std::vector<double> vn((size_t)nn * 3, 0.0);
    for (int v = 0; v < nn; ++v) {
        if (!bflag[v]) continue;
        int count = 0;
        for (int c : vert_cv[v])
            if (el_bnd[c]) {
                for (int k = 0; k < dim; ++k) vn[3 * v + k] += en[3 * c + k];
                count++;
            }
        double mag = 0.0;
        for (int k = 0; k < dim; ++k) {
            vn[3 * v + k] /= (double)count;
            mag += vn[3 * v + k] * vn[3 * v + k];
        }
        mag = std::sqrt(mag);
        for (int k = 0; k < dim; ++k) vn[3 * v + k] /= mag;
    }
    GmshCloud out;
    out.dim = dim;
    out.nv_original = nn;
    out.nelem = ncv;
    out.xyz_original.resize((size_t)nn * dim);
    for (int v = 0; v < nn; ++v)
        for (int k = 0; k < dim; ++k) out.xyz_original[(size_t)v * dim + k] = X(v, k);
    out.elem_vert_original.reserve(ncv);
    for (int c = 0; c < ncv; ++c) out.elem_vert_original.push_back(elems[el_idx[c]].v);
    for (int v = 0; v < nn; ++v) {
        if (drop[v]) continue;
        for (int k = 0; k < dim; ++k) {
            out.xyz.push_back(X(v, k));
            out.normal.push_back(vn[3 * v + k]);
        }
        out.boundary.push_back(bflag[v]);
        out.original.push_back(v);
    }
    out.nv = (int)out.boundary.size();
    return out;
}

}  // namespace memphys
