// C ABI of the meshless fractional-step path (include/memphys_hip.h).
//
// Device order.  Each point cloud (one, or every problem of a batch) gets a
// run of device rows starting on a BLOCK boundary and padded to whole blocks
// with one row to spare (the border unknown of a full-Neumann pressure
// system); padding rows have no stencil.  Inside a cloud the caller's node
// order (RCM, as CLOUD produces) is re-sorted into Morton (Z-order) of the
// coordinates and then by stencil length inside windows of SIGMA rows, and
// cut into slices of 64 rows; each slice stores its neighbours column-major
// (SELL-64).  Every field, weight and operator lives on the device in that
// order; host arrays crossing the ABI are permuted on the way in and out.
// Slots follow the SELL-64x4 numbering of mp_internal.hpp.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <stdexcept>
#include <string>
#include <vector>

#include "gnn.hpp"
#include "krylov.hpp"
#include "mp_internal.hpp"

using namespace mp;

struct mp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // cloud: nv real nodes in nprob problems; npad device rows (multiple of
    // BLOCK); perm: device row -> caller node (-1 on padding rows); iperm:
    // caller node -> device row
    bool has_cloud = false;
    int nv = 0, dim = 0, nslices = 0, npad = 0, total_slots = 0, max_len = 0;
    long nnz = 0;
    std::vector<int> perm, iperm, csr_slot, h_rowptr;
    int nprob = 1;
    std::vector<int> prob_off, prob_nv, prob_row0, prob_rows;  // caller offset, nodes, rows
    DevBuf<int> blk_prob;     // problem of each row block
    DevBuf<double> prob_prm;  // rho, mu, dt per problem (batch parameters)
    bool per_problem_prm = false;
    std::vector<long> prob_total_iters;  // Krylov iterations per problem since mp_fs_create
    std::vector<uint8_t> h_bnd;
    DevBuf<int> sptr, col, len, brows;
    // the column-sorted slot order of the FP32/BF16 products: 32-bit and
    // 16-bit columns, window bases, and slot_perm[sorted slot] = stencil slot
    DevBuf<int> col_s, cbase_s, slot_perm;
    DevBuf<uint16_t> col16_s;
    int nb = 0;
    DevBuf<double> xyz, normal;
    DevBuf<uint8_t> bnd;
    SellView sv{};   // stencil order (weights, FP64 operator)
    SellView svs{};  // column-sorted order (FP32 operator, multigrid level 0)

    // stencil weights (SELL slots)
    bool has_weights = false;
    DevBuf<double> w[4];
    DevBuf<double> vec_a, vec_b;

    // standalone SOLVER
    std::unique_ptr<PpeSolver> solver;
    DevBuf<uint8_t> solver_dir;

    // fractional step
    bool has_fs = false;
    mp_fs_params prm{};
    DevBuf<uint8_t> udir[3], pdir;
    std::unique_ptr<PpeSolver> ppe;
    DevBuf<double> U[2][3];
    DevBuf<double> upk, uhpk;  // velocity before / after the predictor, interleaved per node
    int cur = 0;
    DevBuf<double> P, b, uh[3], S[2];
    int s_cur = 0;
    mp_step_stats stats{};
    int guess_mode = MP_GUESS_PREVIOUS;
    int guess_order = 4;  // pressures the extrapolated guess goes through (2..4)
    int guess_base = MP_GUESS_PROJECT;  // the guess the learned network corrects
    // back-off of the learned guess: after a rejected guess the network is
    // skipped for guess_backoff steps (1, 2, 4, ... 64), an accepted one resets it
    int guess_backoff = 0, guess_skip = 0;
    DevBuf<double> div_partial;
    // pressure history and the learned guess
    DevBuf<double> P1, P2, P3, P4, Ptmp;  // pressures of the last four steps (P1 the latest)
    DevBuf<double> force;   // body force [3][npad] (mp_fs_body_force)
    bool has_force = false;
    int hist = 0;
    // projection guess: rings of the last PROJ_MAX pressures and their
    // products with the operator, the device record, entries held, window
    // (20 by default: over the 495 timed steps of the 1M-node cavity, window
    // 16 took 6.6 GMRES iterations per step and windows 20 / 22 / 24 took
    // 6.4 / 6.2 / 6.3, at rates within 1 % of each other)
    DevBuf<double> Hp, Hw, proj_partial;
    DevBuf<ProjState> pstate;
    int pcount = 0, proj_window = 20;
    std::unique_ptr<GuessNet> net;
    bool train = false;
    bool use_graphs = true;
    double last_loss = 0.0;
    float lr = 1e-3f;

    KernelTimer timers[5];
    double bytes[5] = {0, 0, 0, 0, 0};
    int weights_fp64_rows = 0;  // stencils the last assembly solved in FP64

    // The fixed parts of a timestep (before the solve: packing, predictor,
    // right-hand side, extrapolated guess; after it: pressure history,
    // corrector, boundary rows) replayed from hipGraphs.  The key holds the
    // part, the ping-pong indices, the Euler/AB2 switch and the buffers that
    // can change between steps; every call that reallocates a buffer the
    // captured kernels read drops the graphs first (release_step_graphs).
    struct StepGraph {
        uintptr_t key[8];
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> step_graphs;
    void release_step_graphs() {
        if (step_graphs.empty()) return;
        (void)hipStreamSynchronize(stream);  // (a replay may still be running)
        for (StepGraph& g : step_graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        step_graphs.clear();
    }
    ~mp_ctx() { release_step_graphs(); }
};

namespace {

int fail(mp_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

template <typename F>
int guarded(mp_ctx* c, const char* fn, F&& body) {
    if (!c) return MP_ERR_ARG;
    try {
        MP_HIP(hipSetDevice(c->device));
        return body();
    } catch (const HipError& e) {
        return fail(c, MP_ERR_HIP,
                    std::string(fn) + ": " + e.what + " -> " + hipGetErrorString(e.code));
    } catch (const std::exception& e) {
        return fail(c, MP_ERR_ARG, std::string(fn) + ": " + e.what());
    }
}

// Caller order -> device rows.  Node fields have nv entries; a pressure
// system (S != null) has nv node entries followed by the border unknowns of
// its bordered problems, in problem order.
void to_device_order(const mp_ctx* c, const double* src, std::vector<double>& dst,
                     const PpeSolver* S = nullptr) {
    dst.assign((size_t)c->npad + 64, 0.0);
    for (int o = 0; o < c->nv; ++o) dst[c->iperm[o]] = src[o];
    if (S) {
        int k = c->nv;
        for (const Seg& g : S->segs)
            if (g.border) dst[g.row0 + g.nv] = src[k++];
    }
}

void from_device_order(const mp_ctx* c, const std::vector<double>& src, double* dst,
                       const PpeSolver* S = nullptr) {
    for (int o = 0; o < c->nv; ++o) dst[o] = src[c->iperm[o]];
    if (S) {
        int k = c->nv;
        for (const Seg& g : S->segs)
            if (g.border) dst[k++] = src[g.row0 + g.nv];
    }
}

// The systems of a pressure operator with Dirichlet flags dir (caller order):
// a problem without a Dirichlet boundary node is bordered (solver.cpp:205-213).
std::vector<Seg> make_systems(const mp_ctx* c, const uint8_t* dir) {
    std::vector<Seg> out(c->nprob);
    for (int q = 0; q < c->nprob; ++q) {
        Seg& g = out[q];
        g.row0 = c->prob_row0[q];
        g.nv = c->prob_nv[q];
        g.blk0 = g.row0 / BLOCK;
        g.nblk = c->prob_rows[q] / BLOCK;
        g.border = 1;
        for (int o = c->prob_off[q]; o < c->prob_off[q] + c->prob_nv[q]; ++o)
            if (c->h_bnd[c->iperm[o]] && dir[o]) {
                g.border = 0;
                break;
            }
    }
    return out;
}

// The solve's statistics into the ABI record (totals are the caller's).
void put_stats(mp_step_stats& o, const SolveStats& st) {
    o.iterations = st.iterations;
    o.outer_cycles = st.outer_cycles;
    o.restarts = st.restarts;
    o.converged = st.converged;
    o.rel_res = st.rel_res;
    o.abs_res = st.abs_res;
    o.alpha = st.alpha;
    o.guess_rel_res = st.guess_rel_res;
    o.gmres_switches = st.gmres_switches;
    o.guess_rejected = st.guess_rejected;
}

void upload(const mp_ctx* c, double* dev, const std::vector<double>& h) {
    MP_HIP(hipMemcpyAsync(dev, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                          c->stream));
    MP_HIP(hipStreamSynchronize(c->stream));
}

void download(const mp_ctx* c, const double* dev, std::vector<double>& h) {
    h.resize((size_t)c->npad + 64);
    MP_HIP(hipMemcpyAsync(h.data(), dev, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                          c->stream));
    MP_HIP(hipStreamSynchronize(c->stream));
}

void alloc_vec(const mp_ctx* c, DevBuf<double>& b) {
    b.alloc((size_t)c->npad + 64);
    MP_HIP(hipMemsetAsync(b.p, 0, sizeof(double) * b.n, c->stream));
}

void slots_to_csr(const mp_ctx* c, const DevBuf<double>& dev, double* out) {
    std::vector<double> h(c->total_slots);
    MP_HIP(hipMemcpyAsync(h.data(), dev.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                          c->stream));
    MP_HIP(hipStreamSynchronize(c->stream));
    for (size_t t = 0; t < c->csr_slot.size(); ++t) out[t] = h[c->csr_slot[t]];
}

void csr_to_slots(const mp_ctx* c, const double* in, DevBuf<double>& dev) {
    std::vector<double> h(c->total_slots, 0.0);
    for (size_t t = 0; t < c->csr_slot.size(); ++t) h[c->csr_slot[t]] = in[t];
    MP_HIP(hipMemcpyAsync(dev.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                          c->stream));
    MP_HIP(hipStreamSynchronize(c->stream));
}

void upload_flags(const mp_ctx* c, const uint8_t* src, DevBuf<uint8_t>& dev) {
    std::vector<uint8_t> h((size_t)c->npad, 0);
    if (src)
        for (int o = 0; o < c->nv; ++o) h[c->iperm[o]] = src[o] ? 1 : 0;
    dev.alloc(h.size());
    MP_HIP(hipMemcpy(dev.p, h.data(), h.size(), hipMemcpyHostToDevice));
}

// The step keeps the projection's rings (projected guesses, or learned ones
// correcting a projection).
bool uses_projection(const mp_ctx* c) {
    return c->guess_mode == MP_GUESS_PROJECT ||
           (c->guess_mode == MP_GUESS_LEARNED && c->guess_base == MP_GUESS_PROJECT);
}

// Rings of the projection guess for the current rows, emptied.
void ensure_projection(mp_ctx* c) {
    const size_t ring = (size_t)PROJ_MAX * c->npad;
    if (c->Hp.n < ring) {
        c->release_step_graphs();  // (captured steps hold the buffers replaced here)
        c->Hp.alloc(ring);
        c->Hw.alloc(ring);
    }
    if (!c->pstate.p) c->pstate.alloc(1);
    if (c->proj_partial.n < proj_partial_size()) c->proj_partial.alloc(proj_partial_size());
    launch_proj_reset(c->stream, c->pstate.p);
    c->pcount = 0;
}

// The polynomial through the pressures of the last `order` steps (equal time
// steps), at the next step: 2 p1 - p2, 3 (p1 - p2) + p3, or 4 (p1 + p3) -
// 6 p2 - p4 (p1 the latest), on rows below n_ext; rows n_ext .. n-1 keep p1.
// The time step extrapolates every device row (node rows and border unknowns
// of every problem; padding rows hold zeros throughout).
__global__ __launch_bounds__(256) void k_extrapolate(int n_ext, int n, int order,
                                                     const double* __restrict__ p1,
                                                     const double* __restrict__ p2,
                                                     const double* __restrict__ p3,
                                                     const double* __restrict__ p4,
                                                     double* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (i >= n_ext)
        out[i] = p1[i];
    else if (order == 2)
        out[i] = 2.0 * p1[i] - p2[i];
    else if (order == 3)
        out[i] = 3.0 * (p1[i] - p2[i]) + p3[i];
    else
        out[i] = 4.0 * (p1[i] + p3[i]) - 6.0 * p2[i] - p4[i];
}

__global__ __launch_bounds__(256) void k_copy(int n, const double* __restrict__ a,
                                              double* __restrict__ b) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// (Re)builds the network for the current cloud and pressure operator, keeping
// its parameters when it already exists.
void ensure_net(mp_ctx* c) {
    std::vector<float> keep;
    if (c->net) {
        keep.resize(GNN_NPARAM);
        c->net->get_params(c->stream, keep.data());
    }
    c->net.reset(new GuessNet());
    SellView gv = c->sv;  // the node rows of the (single) problem
    gv.nv = c->nv;
    gv.nslices = (c->nv + SLICE - 1) / SLICE;
    c->net->setup(c->stream, gv, c->ppe->scaled32, c->total_slots);
    c->net->lr = c->lr;
    if (keep.empty())
        c->net->init_params(c->stream, 20240611ull);
    else
        c->net->set_params(c->stream, keep.data());
}

}  // namespace

// Scratch of mp_divergence_noalloc: the per-row-block partial sums of the
// continuity residual (4 doubles per block of 256 rows), so that entry point
// allocates nothing.  One copy per GPU (each device holds its own instance of
// a __device__ array), shared by every context on that GPU; its address on a
// device is taken once, with the device current, when the first context on it
// is created.  In host code the array's name is not a device address: only
// the recorded pointer is handed to kernels.
constexpr int NOALLOC_MAX_BLOCKS = MP_NOALLOC_MAX_ROWS / BLOCK;
static __device__ __align__(256) double scratch[4 * NOALLOC_MAX_BLOCKS];
namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_ptr[MAX_DEVICES] = {};

// Address of `scratch` on `device` (which must be current), recorded on first use.
double* noalloc_scratch_address(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
        scratch_ptr[device] = static_cast<double*>(ptr);
    }
    return scratch_ptr[device];
}

// out[0..5] = {max_int, mean_int, max_bnd, mean_bnd, max_all, mean_all} from
// the block partials, summed in block order by one thread (mp_divergence's
// order, so both entry points return the same bits).
__global__ void k_div_finish(const double* __restrict__ part, int g, int ni, int nb,
                             double* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double mi = 0, si = 0, mb = 0, sb = 0;
    for (int k = 0; k < g; ++k) {
        mi = (mi < part[k * 4 + 0]) ? part[k * 4 + 0] : mi;
        si += part[k * 4 + 1];
        mb = (mb < part[k * 4 + 2]) ? part[k * 4 + 2] : mb;
        sb += part[k * 4 + 3];
    }
    out[0] = mi;
    out[1] = ni ? si / ni : 0.0;
    out[2] = mb;
    out[3] = nb ? sb / nb : 0.0;
    out[4] = (mi < mb) ? mb : mi;
    out[5] = (si + sb) / (ni + nb);
}
}  // namespace

extern "C" {

int mp_ctx_create(int device, mp_ctx** out) {
    if (!out) return MP_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MP_ERR_HIP;
    if (device < 0 || device >= n) return MP_ERR_ARG;
    mp_ctx* c = new mp_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return MP_ERR_HIP;
    }
    if (!noalloc_scratch_address(device)) {  // the device is current
        (void)hipStreamDestroy(c->stream);
        delete c;
        return MP_ERR_HIP;
    }
    *out = c;
    return MP_OK;
}

int mp_ctx_destroy(mp_ctx* c) {
    if (!c) return MP_ERR_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& t : c->timers) t.release();
    c->solver.reset();
    c->ppe.reset();
    hipStream_t s = c->stream;
    delete c;
    if (s) (void)hipStreamDestroy(s);
    return MP_OK;
}

const char* mp_last_error(const mp_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mp_sync(mp_ctx* c) {
    return guarded(c, "mp_sync", [&]() -> int {
        MP_HIP(hipStreamSynchronize(c->stream));
        return MP_OK;
    });
}

}  // extern "C"

namespace {

// Uploads nprob clouds given back to back in caller order (problem q owns
// caller nodes [off_q, off_q + nvp[q]); its stencils may only reference its
// own nodes).  One problem is the ordinary single-cloud case.
// 16-bit column form of the SELL slots: per slice, the sorted distinct
// columns are covered greedily by windows of 2^CW_BITS ids; a slice needing
// at most CW_COUNT windows stores (window << CW_BITS) | (column - base) per
// slot, any other keeps its 32-bit columns (first base -1).
void compress_columns(const std::vector<int>& sptr, const std::vector<int>& col, int nslices,
                      std::vector<uint16_t>& c16, std::vector<int>& cb) {
    c16.assign(col.size(), 0);
    cb.assign((size_t)nslices * CW_COUNT, -1);
    std::vector<int> u;
    for (int sl = 0; sl < nslices; ++sl) {
        u.assign(col.begin() + sptr[sl], col.begin() + sptr[sl + 1]);
        if (u.empty()) continue;
        std::sort(u.begin(), u.end());
        int base[CW_COUNT];
        int nw = 0;
        for (size_t i = 0; i < u.size() && nw <= CW_COUNT;) {
            if (nw < CW_COUNT) base[nw] = u[i];
            ++nw;
            const int end = u[i] + (1 << CW_BITS);
            while (i < u.size() && u[i] < end) ++i;
        }
        if (nw > CW_COUNT) continue;
        for (int w = 0; w < CW_COUNT; ++w) cb[(size_t)sl * CW_COUNT + w] = base[std::min(w, nw - 1)];
        for (int t = sptr[sl]; t < sptr[sl + 1]; ++t) {
            int w = nw - 1;
            while (base[w] > col[t]) --w;
            c16[t] = (uint16_t)((w << CW_BITS) | (col[t] - base[w]));
        }
    }
}

int upload_clouds(mp_ctx* c, const char* fn, int nprob, const int* nvp, int dim,
                  const double* xyz, const uint8_t* boundary, const double* normal,
                  const int* row_ptr, const int* col_idx) {
    const std::string f(fn);
    c->release_step_graphs();  // (every buffer of the cloud is replaced)
    if (nprob < 1 || !nvp || (dim != 2 && dim != 3) || !xyz || !boundary || !normal || !row_ptr ||
        !col_idx)
        return fail(c, MP_ERR_ARG, f + ": bad arguments");
    long nvl = 0;
    for (int q = 0; q < nprob; ++q) {
        if (nvp[q] <= 0) return fail(c, MP_ERR_ARG, f + ": a problem has no nodes");
        nvl += nvp[q];
    }
    if (nvl > (1L << 30)) return fail(c, MP_ERR_ARG, f + ": too many nodes");
    const int nv = (int)nvl;
    if (row_ptr[0] != 0) return fail(c, MP_ERR_ARG, f + ": row_ptr[0] != 0");
    int maxlen = 0;
    {
        int q = 0, lo = 0;
        for (int i = 0; i < nv; ++i) {
            while (i >= lo + nvp[q]) lo += nvp[q++];
            const int l = row_ptr[i + 1] - row_ptr[i];
            if (l < 1 || l > MAX_STENCIL)
                return fail(c, MP_ERR_ARG,
                            f + ": row " + std::to_string(i) + " has " + std::to_string(l) +
                                " entries (allowed 1.." + std::to_string(MAX_STENCIL) + ")");
            if (col_idx[row_ptr[i]] != i)
                return fail(c, MP_ERR_ARG,
                            f + ": entry 0 of row " + std::to_string(i) + " is not the node itself");
            for (int t = row_ptr[i]; t < row_ptr[i + 1]; ++t)
                if (col_idx[t] < lo || col_idx[t] >= lo + nvp[q])
                    return fail(c, MP_ERR_ARG,
                                f + ": column index out of range (row " + std::to_string(i) + ")");
            maxlen = std::max(maxlen, l);
        }
    }
    c->has_cloud = c->has_weights = c->has_fs = false;
    c->solver.reset();
    c->ppe.reset();
    c->net.reset();
    c->nv = nv;
    c->dim = dim;
    c->max_len = maxlen;
    c->nprob = nprob;
    c->prob_off.assign(nprob, 0);
    c->prob_nv.assign(nvp, nvp + nprob);
    c->prob_row0.assign(nprob, 0);
    c->prob_rows.assign(nprob, 0);
    int rows = 0;
    for (int q = 0; q < nprob; ++q) {
        c->prob_off[q] = q ? c->prob_off[q - 1] + nvp[q - 1] : 0;
        c->prob_row0[q] = rows;
        c->prob_rows[q] = ((nvp[q] + 1 + BLOCK - 1) / BLOCK) * BLOCK;
        rows += c->prob_rows[q];
    }
    c->npad = rows;
    c->nslices = rows / SLICE;
    c->h_rowptr.assign(row_ptr, row_ptr + nv + 1);
    c->nnz = row_ptr[nv];
    c->per_problem_prm = false;
    // device order per problem: Morton (Z-order) of the coordinates, so that 64
    // consecutive rows are a compact cluster whose stencils overlap, then a
    // stable sort by decreasing stencil length inside windows of SIGMA rows
    c->perm.assign(rows, -1);
    c->iperm.assign(nv, 0);
    for (int q = 0; q < nprob; ++q) {
        const int o0 = c->prob_off[q], n = nvp[q];
        std::vector<int> ord(n);
        std::iota(ord.begin(), ord.end(), o0);
        double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int a = 0; a < dim; ++a) {
            lo[a] = hi[a] = xyz[(size_t)o0 * dim + a];
            for (int i = o0; i < o0 + n; ++i) {
                lo[a] = std::min(lo[a], xyz[(size_t)i * dim + a]);
                hi[a] = std::max(hi[a], xyz[(size_t)i * dim + a]);
            }
        }
        const int bits = (dim == 3) ? 10 : 16;
        const double levels = (double)(1u << bits);
        std::vector<uint64_t> key(n, 0);
        for (int i = 0; i < n; ++i) {
            uint64_t k = 0;
            for (int a = 0; a < dim; ++a) {
                const double span = hi[a] - lo[a];
                const double t = span > 0 ? (xyz[(size_t)(o0 + i) * dim + a] - lo[a]) / span : 0.0;
                const uint64_t g = (uint64_t)std::min(levels - 1.0, std::floor(t * levels));
                for (int bt = 0; bt < bits; ++bt) k |= ((g >> bt) & 1ull) << (bt * dim + a);
            }
            key[i] = k;
        }
        std::stable_sort(ord.begin(), ord.end(),
                         [&](int a, int b) { return key[a - o0] < key[b - o0]; });
        for (int w0 = 0; w0 < n; w0 += SIGMA) {
            const int w1 = std::min(n, w0 + SIGMA);
            std::stable_sort(ord.begin() + w0, ord.begin() + w1, [&](int a, int b) {
                return (row_ptr[a + 1] - row_ptr[a]) > (row_ptr[b + 1] - row_ptr[b]);
            });
        }
        for (int i = 0; i < n; ++i) {
            c->perm[c->prob_row0[q] + i] = ord[i];
            c->iperm[ord[i]] = c->prob_row0[q] + i;
        }
    }
    std::vector<int> hsptr(c->nslices + 1, 0), hlen(c->npad, 0);
    for (int d = 0; d < rows; ++d)
        if (c->perm[d] >= 0) hlen[d] = row_ptr[c->perm[d] + 1] - row_ptr[c->perm[d]];
    for (int sl = 0; sl < c->nslices; ++sl) {
        int wmax = 0;
        for (int l = 0; l < SLICE; ++l) wmax = std::max(wmax, hlen[sl * SLICE + l]);
        hsptr[sl + 1] = hsptr[sl] + sell_width4(wmax) * SLICE;
    }
    c->total_slots = hsptr[c->nslices];
    std::vector<int> hcol(c->total_slots, 0);
    c->csr_slot.assign(row_ptr[nv], 0);
    for (int sl = 0; sl < c->nslices; ++sl) {
        const int wmax = (hsptr[sl + 1] - hsptr[sl]) / SLICE;
        for (int l = 0; l < SLICE; ++l) {
            const int d = sl * SLICE + l;
            for (int j = 0; j < wmax; ++j) {
                const int slot = hsptr[sl] + sell_off(j) + l * SGROUP;
                if (j < hlen[d]) {
                    const int t = row_ptr[c->perm[d]] + j;
                    hcol[slot] = c->iperm[col_idx[t]];
                    c->csr_slot[t] = slot;
                } else {
                    hcol[slot] = d;  // padding slot: own row, value 0
                }
            }
        }
    }
    std::vector<double> hx((size_t)dim * c->npad, 0.0), hn((size_t)dim * c->npad, 0.0);
    c->h_bnd.assign(c->npad, 0);
    std::vector<int> hb;
    for (int d = 0; d < rows; ++d) {
        const int o = c->perm[d];
        if (o < 0) continue;
        for (int a = 0; a < dim; ++a) {
            hx[(size_t)a * c->npad + d] = xyz[(size_t)o * dim + a];
            hn[(size_t)a * c->npad + d] = normal[(size_t)o * dim + a];
        }
        c->h_bnd[d] = boundary[o] ? 1 : 0;
        if (boundary[o]) hb.push_back(d);
    }
    for (int d : hb)
        for (int t = row_ptr[c->perm[d]] + 1; t < row_ptr[c->perm[d] + 1]; ++t)
            if (boundary[col_idx[t]])
                return fail(c, MP_ERR_ARG,
                            f + ": boundary node " + std::to_string(c->perm[d]) +
                                " couples to boundary node " + std::to_string(col_idx[t]));
    std::vector<int> hbp(rows / BLOCK, 0);
    for (int q = 0; q < nprob; ++q)
        for (int b = 0; b < c->prob_rows[q] / BLOCK; ++b) hbp[c->prob_row0[q] / BLOCK + b] = q;
    c->nb = (int)hb.size();
    c->sptr.alloc(hsptr.size());
    // column-sorted copy: each row's slots reordered by increasing column,
    // so that a wavefront's gathers of one neighbour step fall closer
    // together; the FP32 operator and the multigrid's finest level use it
    std::vector<int> hcol_s(hcol), hperm(c->total_slots);
    std::iota(hperm.begin(), hperm.end(), 0);
    {
        const int nthr = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (int tn = 0; tn < nthr; ++tn)
            pool.emplace_back([&, tn]() {
                std::vector<std::pair<int, int>> e;
                for (int d = tn; d < rows; d += nthr) {
                    const int L = hlen[d];
                    if (L == 0) continue;
                    const int base = sell_base(hsptr.data(), d);
                    e.clear();
                    for (int j = 0; j < L; ++j)
                        e.emplace_back(hcol[base + sell_off(j)], base + sell_off(j));
                    std::stable_sort(e.begin(), e.end(),
                                     [](const std::pair<int, int>& a,
                                        const std::pair<int, int>& b) { return a.first < b.first; });
                    for (int j = 0; j < L; ++j) {
                        hcol_s[base + sell_off(j)] = e[j].first;
                        hperm[base + sell_off(j)] = e[j].second;
                    }
                }
            });
        for (auto& t : pool) t.join();
    }
    std::vector<uint16_t> hcol16;
    std::vector<int> hcb;
    compress_columns(hsptr, hcol_s, c->nslices, hcol16, hcb);
    c->col.alloc(hcol.size());
    c->col_s.alloc(hcol_s.size());
    c->slot_perm.alloc(hperm.size());
    c->col16_s.alloc(hcol16.size());
    c->cbase_s.alloc(hcb.size());
    c->len.alloc(hlen.size());
    c->brows.alloc(std::max<size_t>(1, hb.size()));
    c->xyz.alloc(hx.size());
    c->normal.alloc(hn.size());
    c->bnd.alloc(c->h_bnd.size());
    c->blk_prob.alloc(hbp.size());
    MP_HIP(hipMemcpy(c->sptr.p, hsptr.data(), sizeof(int) * hsptr.size(), hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(c->col.p, hcol.data(), sizeof(int) * hcol.size(), hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(c->col_s.p, hcol_s.data(), sizeof(int) * hcol_s.size(),
                     hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(c->slot_perm.p, hperm.data(), sizeof(int) * hperm.size(),
                     hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(c->col16_s.p, hcol16.data(), sizeof(uint16_t) * hcol16.size(),
                     hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(c->cbase_s.p, hcb.data(), sizeof(int) * hcb.size(), hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(c->len.p, hlen.data(), sizeof(int) * hlen.size(), hipMemcpyHostToDevice));
    if (!hb.empty())
        MP_HIP(hipMemcpy(c->brows.p, hb.data(), sizeof(int) * hb.size(), hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(c->xyz.p, hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(c->normal.p, hn.data(), sizeof(double) * hn.size(), hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(c->bnd.p, c->h_bnd.data(), c->h_bnd.size(), hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(c->blk_prob.p, hbp.data(), sizeof(int) * hbp.size(), hipMemcpyHostToDevice));
    c->sv.nv = rows;
    c->sv.nslices = c->nslices;
    c->sv.sptr = c->sptr.p;
    c->sv.col = c->col.p;
    c->svs = c->sv;
    c->svs.col = c->col_s.p;
    c->svs.col16 = c->col16_s.p;
    c->svs.cbase = c->cbase_s.p;
    c->sv.len = c->len.p;
    for (auto& wb : c->w) {
        wb.alloc(c->total_slots);
        MP_HIP(hipMemsetAsync(wb.p, 0, sizeof(double) * wb.n, c->stream));
    }
    alloc_vec(c, c->vec_a);
    alloc_vec(c, c->vec_b);
    c->div_partial.alloc((size_t)(rows / BLOCK) * 4);
    MP_HIP(hipStreamSynchronize(c->stream));
    c->has_cloud = true;
    return MP_OK;
}

}  // namespace

extern "C" {

int mp_cloud_upload(mp_ctx* c, int nv, int dim, const double* xyz, const uint8_t* boundary,
                    const double* normal, const int* row_ptr, const int* col_idx) {
    return guarded(c, "mp_cloud_upload", [&]() -> int {
        return upload_clouds(c, "mp_cloud_upload", 1, &nv, dim, xyz, boundary, normal, row_ptr,
                             col_idx);
    });
}

int mp_batch_upload(mp_ctx* c, int nprob, const int* nv_per_problem, int dim, const double* xyz,
                    const uint8_t* boundary, const double* normal, const int* row_ptr,
                    const int* col_idx) {
    return guarded(c, "mp_batch_upload", [&]() -> int {
        return upload_clouds(c, "mp_batch_upload", nprob, nv_per_problem, dim, xyz, boundary,
                             normal, row_ptr, col_idx);
    });
}

int mp_weights_assemble(mp_ctx* c, int phs_deg, int poly_deg, const int* exps, int precision_mode,
                        double* gx, double* gy, double* gz, double* lap, double* cond_out) {
    return guarded(c, "mp_weights_assemble", [&]() -> int {
        if (!c->has_cloud) return fail(c, MP_ERR_STATE, "mp_weights_assemble: no cloud uploaded");
        c->release_step_graphs();
        if (phs_deg < 3 || phs_deg > 11 || (phs_deg % 2) == 0)
            return fail(c, MP_ERR_ARG, "mp_weights_assemble: phs_deg must be 3, 5, 7, 9 or 11");
        int m = (c->dim == 2) ? (poly_deg + 1) * (poly_deg + 2) / 2
                              : (poly_deg + 1) * (poly_deg + 2) * (poly_deg + 3) / 6;
        if (poly_deg < 0 || m > MAX_POLY)
            return fail(c, MP_ERR_ARG, "mp_weights_assemble: poly_deg out of range");
        if (!exps) return fail(c, MP_ERR_ARG, "mp_weights_assemble: exps is NULL");
        int minlen = MAX_STENCIL + 1;
        for (int i = 0; i < c->nv; ++i) minlen = std::min(minlen, c->h_rowptr[i + 1] - c->h_rowptr[i]);
        if (minlen < m)
            return fail(c, MP_ERR_ARG,
                        "mp_weights_assemble: a stencil has fewer points than monomials");
        if (precision_mode != MP_PREC_FP64 && precision_mode != MP_PREC_MIXED)
            return fail(c, MP_ERR_ARG, "mp_weights_assemble: unknown precision_mode");
        DevBuf<int> de;
        de.alloc((size_t)m * c->dim);
        MP_HIP(hipMemcpy(de.p, exps, sizeof(int) * m * c->dim, hipMemcpyHostToDevice));
        DevBuf<double> dcond;
        if (cond_out) dcond.alloc(c->npad);
        c->timers[4].begin(c->stream);
        c->weights_fp64_rows = c->nv;
        DevBuf<int> wrows;
        auto row_len = [&](int d) { return c->h_rowptr[c->perm[d] + 1] - c->h_rowptr[c->perm[d]]; };
        if (cond_out) {
            // the condition estimate needs the LU factors: the LDS kernel keeps them
            launch_weights_fp64(c->stream, c->dim, phs_deg, m, de.p, c->sv, c->xyz.p, c->npad,
                                c->w[0].p, c->w[1].p, c->w[2].p, c->w[3].p, dcond.p, c->max_len);
        } else if (precision_mode == MP_PREC_FP64) {
            std::vector<int> rows, lens;
            for (int d = 0; d < c->npad; ++d) {
                if (c->perm[d] < 0) continue;
                rows.push_back(d);
                lens.push_back(row_len(d));
            }
            launch_weights_fp64_rows(c->stream, c->dim, phs_deg, m, de.p, c->sv, c->xyz.p, c->npad,
                                     c->w[0].p, c->w[1].p, c->w[2].p, c->w[3].p, rows, lens,
                                     wrows);
        } else {
            // FP32 LU + FP64 refinement for the shortest (interior) stencils; the
            // longer boundary / near-boundary stencils and any stencil whose
            // refinement stalls are solved in FP64.
            int kmin = MAX_STENCIL + 1;
            for (int o = 0; o < c->nv; ++o)
                kmin = std::min(kmin, c->h_rowptr[o + 1] - c->h_rowptr[o]);
            std::vector<int> mixed_rows, fp64_rows;
            // (high polynomial degrees, or systems the mixed kernel cannot hold
            // in LDS: everything in FP64)
            const bool mixed_ok = m <= WAVE_MAX_POLY && weights_mixed_lds(kmin, m) <= WEIGHTS_LDS_MAX;
            for (int d = 0; d < c->npad; ++d) {
                if (c->perm[d] < 0) continue;
                const int len = c->h_rowptr[c->perm[d] + 1] - c->h_rowptr[c->perm[d]];
                (mixed_ok && len == kmin ? mixed_rows : fp64_rows).push_back(d);
            }
            DevBuf<int> dm, dflag;
            dm.alloc(mixed_rows.size());
            dflag.alloc(c->npad);
            MP_HIP(hipMemsetAsync(dflag.p, 0, sizeof(int) * c->npad, c->stream));
            if (!mixed_rows.empty())
                MP_HIP(hipMemcpy(dm.p, mixed_rows.data(), sizeof(int) * mixed_rows.size(),
                                 hipMemcpyHostToDevice));
            launch_weights_mixed(c->stream, c->dim, phs_deg, m, de.p, c->sv, c->xyz.p, c->npad,
                                 c->w[0].p, c->w[1].p, c->w[2].p, c->w[3].p, dm.p,
                                 (int)mixed_rows.size(), dflag.p, kmin, 8);
            std::vector<int> hflag(c->npad);
            MP_HIP(hipMemcpyAsync(hflag.data(), dflag.p, sizeof(int) * c->npad,
                                  hipMemcpyDeviceToHost, c->stream));
            MP_HIP(hipStreamSynchronize(c->stream));
            for (int d : mixed_rows)
                if (hflag[d]) fp64_rows.push_back(d);
            c->weights_fp64_rows = (int)fp64_rows.size();
            if (!fp64_rows.empty()) {
                std::vector<int> lens;
                for (int d : fp64_rows) lens.push_back(row_len(d));
                launch_weights_fp64_rows(c->stream, c->dim, phs_deg, m, de.p, c->sv, c->xyz.p,
                                         c->npad, c->w[0].p, c->w[1].p, c->w[2].p, c->w[3].p,
                                         fp64_rows, lens, wrows);
            }
        }
        c->timers[4].end(c->stream);
        MP_HIP(hipStreamSynchronize(c->stream));
        c->timers[4].harvest();
        {
            double M = (double)(c->max_len + m);
            c->bytes[4] = (2.0 / 3.0) * M * M * M;  // flops per stencil (largest M)
        }
        c->has_weights = true;
        if (gx) slots_to_csr(c, c->w[0], gx);
        if (gy) slots_to_csr(c, c->w[1], gy);
        if (gz && c->dim == 3) slots_to_csr(c, c->w[2], gz);
        if (lap) slots_to_csr(c, c->w[3], lap);
        if (cond_out) {
            std::vector<double> h(c->npad);
            MP_HIP(hipMemcpy(h.data(), dcond.p, sizeof(double) * c->npad, hipMemcpyDeviceToHost));
            for (int d = 0; d < c->npad; ++d)
                if (c->perm[d] >= 0) cond_out[c->perm[d]] = h[d];
        }
        return MP_OK;
    });
}

int mp_weights_set(mp_ctx* c, const double* gx, const double* gy, const double* gz,
                   const double* lap) {
    return guarded(c, "mp_weights_set", [&]() -> int {
        if (!c->has_cloud) return fail(c, MP_ERR_STATE, "mp_weights_set: no cloud uploaded");
        c->release_step_graphs();
        if (!gx || !gy || !lap || (c->dim == 3 && !gz))
            return fail(c, MP_ERR_ARG, "mp_weights_set: missing operator");
        csr_to_slots(c, gx, c->w[0]);
        csr_to_slots(c, gy, c->w[1]);
        if (c->dim == 3) csr_to_slots(c, gz, c->w[2]);
        csr_to_slots(c, lap, c->w[3]);
        c->has_weights = true;
        return MP_OK;
    });
}

int mp_operator_apply(mp_ctx* c, int op, const double* x, double* y) {
    return guarded(c, "mp_operator_apply", [&]() -> int {
        if (!c->has_weights) return fail(c, MP_ERR_STATE, "mp_operator_apply: no weights");
        if (op < 0 || op > 3 || (op == MP_GZ && c->dim != 3) || !x || !y)
            return fail(c, MP_ERR_ARG, "mp_operator_apply: bad operator or pointer");
        std::vector<double> h;
        to_device_order(c, x, h);
        upload(c, c->vec_a.p, h);
        launch_spmv(c->stream, c->sv, c->w[op].p, c->vec_a.p, c->vec_b.p);
        download(c, c->vec_b.p, h);
        from_device_order(c, h, y);
        return MP_OK;
    });
}

int mp_knn_stencils(mp_ctx* c, int nv, int dim, const double* xyz, const uint8_t* boundary,
                    const int* k_per_node, int* row_ptr, int* col) {
    return guarded(c, "mp_knn_stencils", [&]() -> int {
        if (!xyz || !boundary || !k_per_node || !row_ptr || !col)
            return fail(c, MP_ERR_ARG, "mp_knn_stencils: null pointer");
        knn_stencils(c->stream, nv, dim, xyz, boundary, k_per_node, row_ptr, col);
        return MP_OK;
    });
}

int mp_rcm_order(mp_ctx* c, int nv, const int* row_ptr, const int* col, int* order) {
    return guarded(c, "mp_rcm_order", [&]() -> int {
        if (!row_ptr || !col || !order) return fail(c, MP_ERR_ARG, "mp_rcm_order: null pointer");
        rcm_order(c->stream, nv, row_ptr, col, order, nullptr);
        return MP_OK;
    });
}

int mp_operator_eigenvalue(mp_ctx* c, int op, double tol, int max_restarts, double* re,
                           double* im) {
    return guarded(c, "mp_operator_eigenvalue", [&]() -> int {
        if (!c->has_weights) return fail(c, MP_ERR_STATE, "mp_operator_eigenvalue: no weights");
        if (op < 0 || op > 3 || (op == MP_GZ && c->dim != 3) || !re || !im || !(tol > 0) ||
            max_restarts < 0)
            return fail(c, MP_ERR_ARG, "mp_operator_eigenvalue: bad arguments");
        largest_eigenvalue(c->stream, c->sv, c->w[op].p, 40, tol, max_restarts, re, im);
        return MP_OK;
    });
}

int mp_solver_init(mp_ctx* c, const uint8_t* dirichlet, double unsteady, double conv, double diff,
                   double tolerance, int n_iter, int precision) {
    return guarded(c, "mp_solver_init", [&]() -> int {
        if (!c->has_weights) return fail(c, MP_ERR_STATE, "mp_solver_init: no weights");
        if (!dirichlet || n_iter < 1 || !(tolerance > 0) ||
            (precision != MP_PREC_FP64 && precision != MP_PREC_MIXED))
            return fail(c, MP_ERR_ARG, "mp_solver_init: bad arguments");
        upload_flags(c, dirichlet, c->solver_dir);
        c->solver.reset(new PpeSolver());
        c->solver->svs = c->svs;
        c->solver->slot_perm = c->slot_perm.p;
        c->solver->setup(c->stream, c->sv, make_systems(c, dirichlet), c->dim, c->w[0].p,
                         c->w[1].p, c->w[2].p, c->w[3].p, c->normal.p, c->bnd.p,
                         c->solver_dir.p, unsteady, conv, diff);
        c->solver->tol = tolerance;
        c->solver->maxit = n_iter;
        c->solver->precision = precision;
        c->solver->spmv_timer.enabled = c->timers[0].enabled;
        c->solver->use_graphs = c->use_graphs;
        return MP_OK;
    });
}

int mp_solver_krylov(mp_ctx* c, int krylov, int restart) {
    return guarded(c, "mp_solver_krylov", [&]() -> int {
        if (!c->solver) return fail(c, MP_ERR_STATE, "mp_solver_krylov: mp_solver_init not called");
        if (krylov != MP_KRYLOV_BICGSTAB && krylov != MP_KRYLOV_GMRES)
            return fail(c, MP_ERR_ARG, "mp_solver_krylov: unknown Krylov method");
        if (restart < 0 || restart > GM_MAX)
            return fail(c, MP_ERR_ARG,
                        "mp_solver_krylov: GMRES restart length must be 1.." +
                            std::to_string(GM_MAX) + " (0 = 30)");
        c->solver->krylov = krylov;
        c->solver->restart_m = restart > 0 ? restart : 30;
        return MP_OK;
    });
}

int mp_solver_system_size(mp_ctx* c, int* n) {
    return guarded(c, "mp_solver_system_size", [&]() -> int {
        PpeSolver* s = c->solver ? c->solver.get() : c->ppe.get();
        if (!s || !n) return fail(c, MP_ERR_STATE, "mp_solver_system_size: no solver");
        *n = s->n_sys;
        return MP_OK;
    });
}

int mp_solver_solve(mp_ctx* c, const double* rhs, double* field_new, double* field_old) {
    return guarded(c, "mp_solver_solve", [&]() -> int {
        if (!c->solver) return fail(c, MP_ERR_STATE, "mp_solver_solve: mp_solver_init not called");
        if (!rhs || !field_new || !field_old) return fail(c, MP_ERR_ARG, "mp_solver_solve: NULL");
        PpeSolver& S = *c->solver;
        std::vector<double> hb, hx;
        to_device_order(c, rhs, hb, &S);
        to_device_order(c, field_old, hx, &S);
        upload(c, c->vec_a.p, hb);
        upload(c, c->vec_b.p, hx);
        SolveStats st = S.solve(c->stream, c->vec_a.p, c->vec_b.p);
        download(c, c->vec_b.p, hx);
        from_device_order(c, hx, field_new, &S);
        for (int k = c->nv; k < S.n_sys; ++k) field_old[k] = field_new[k];  // solver.cpp:361-362
        put_stats(c->stats, st);
        c->stats.total_iterations += st.iterations;
        c->stats.total_solves += 1;
        return MP_OK;
    });
}

int mp_fs_create(mp_ctx* c, const mp_fs_params* prm, const uint8_t* u_dir, const uint8_t* v_dir,
                 const uint8_t* w_dir, const uint8_t* p_dir) {
    return guarded(c, "mp_fs_create", [&]() -> int {
        if (!c->has_weights) return fail(c, MP_ERR_STATE, "mp_fs_create: no weights");
        c->release_step_graphs();
        if (!prm || !u_dir || !v_dir || !p_dir || (c->dim == 3 && !w_dir))
            return fail(c, MP_ERR_ARG, "mp_fs_create: NULL argument");
        // FRACTIONAL_STEP_1::check_bc, fractional_step_1.cpp:78-102
        const uint8_t* dirs[3] = {u_dir, v_dir, w_dir};
        for (int q = 0; q < c->dim; ++q)
            for (int pr = 0; pr < c->nprob; ++pr) {
                long sum = 0;
                for (int i = c->prob_off[pr]; i < c->prob_off[pr] + c->prob_nv[pr]; ++i)
                    sum += dirs[q][i] ? 1 : 0;
                if (sum == 0)
                    return fail(c, MP_ERR_ARG,
                                "mp_fs_create: full Neumann velocity BC is not permitted");
            }
        if (prm->rho < 0 || prm->mu < 0)
            return fail(c, MP_ERR_ARG, "mp_fs_create: rho and mu must be set (>= 0)");
        if (prm->temporal_order != 1 && prm->temporal_order != 2)
            return fail(c, MP_ERR_ARG, "mp_fs_create: temporal_order must be 1 or 2");
        if (!(prm->dt > 0) || !(prm->rho > 0) || !(prm->solver_tolerance > 0) || prm->n_iter < 1)
            return fail(c, MP_ERR_ARG, "mp_fs_create: dt, rho, tolerance and n_iter must be > 0");
        if (prm->precision != MP_PREC_FP64 && prm->precision != MP_PREC_MIXED)
            return fail(c, MP_ERR_ARG, "mp_fs_create: unknown precision");
        if (prm->preconditioner != MP_PRECOND_JACOBI && prm->preconditioner != MP_PRECOND_AMG)
            return fail(c, MP_ERR_ARG, "mp_fs_create: unknown preconditioner");
        if (prm->krylov != MP_KRYLOV_BICGSTAB && prm->krylov != MP_KRYLOV_GMRES)
            return fail(c, MP_ERR_ARG, "mp_fs_create: unknown Krylov method");
        if (prm->restart < 0 || prm->restart > GM_MAX)
            return fail(c, MP_ERR_ARG,
                        "mp_fs_create: GMRES restart length must be 1.." + std::to_string(GM_MAX) +
                            " (0 = 30)");
        c->prm = *prm;
        for (int q = 0; q < c->dim; ++q) upload_flags(c, dirs[q], c->udir[q]);
        if (c->dim == 2) upload_flags(c, nullptr, c->udir[2]);
        upload_flags(c, p_dir, c->pdir);
        c->ppe.reset(new PpeSolver());
        c->ppe->svs = c->svs;
        c->ppe->slot_perm = c->slot_perm.p;
        c->ppe->setup(c->stream, c->sv, make_systems(c, p_dir), c->dim, c->w[0].p, c->w[1].p,
                      c->w[2].p, c->w[3].p, c->normal.p, c->bnd.p, c->pdir.p, 0.0, 0.0, 1.0,
                      prm->preconditioner);
        c->per_problem_prm = false;
        c->ppe->tol = prm->solver_tolerance;
        c->ppe->maxit = prm->n_iter;
        c->ppe->precision = prm->precision;
        c->ppe->inner_tol = (prm->inner_tolerance > 0) ? prm->inner_tolerance : 1e-3;
        c->ppe->check_every = (prm->check_every > 0) ? prm->check_every : 8;
        c->ppe->krylov = prm->krylov;
        c->ppe->restart_m = (prm->restart > 0) ? prm->restart : 30;
        c->ppe->spmv_timer.enabled = c->timers[0].enabled;
        c->ppe->use_graphs = c->use_graphs;
        for (int k = 0; k < 2; ++k)
            for (int q = 0; q < 3; ++q) alloc_vec(c, c->U[k][q]);
        alloc_vec(c, c->P);
        alloc_vec(c, c->b);
        alloc_vec(c, c->P1);
        alloc_vec(c, c->P2);
        alloc_vec(c, c->P3);
        alloc_vec(c, c->P4);
        alloc_vec(c, c->Ptmp);
        c->hist = 0;
        if (uses_projection(c)) ensure_projection(c);  // (rows may have changed)
        if (c->net && c->nprob == 1) ensure_net(c);
        for (int q = 0; q < 3; ++q) alloc_vec(c, c->uh[q]);
        c->upk.alloc((size_t)4 * c->npad);
        c->uhpk.alloc((size_t)4 * c->npad);
        c->has_force = false;
        for (int k = 0; k < 2; ++k) {
            c->S[k].alloc((size_t)3 * c->npad + 64);
            MP_HIP(hipMemsetAsync(c->S[k].p, 0, sizeof(double) * c->S[k].n, c->stream));
        }
        MP_HIP(hipStreamSynchronize(c->stream));
        c->cur = 0;
        c->s_cur = 0;
        c->stats = mp_step_stats{};
        c->prob_total_iters.assign(c->nprob, 0);
        // algorithmic bytes per launch (real entries only)
        const double W = (double)c->nnz, N = (double)c->nv, D = (double)c->dim;
        c->bytes[1] = W * (4.0 + 8.0 * (D + 1.0)) + N * (4.0 + 8.0 * (D /*own*/ + D /*gather*/ +
                                                                  D /*uh*/ + D /*S*/ + D /*S_oo*/));
        c->bytes[2] = W * (4.0 + 8.0 * D) + N * (4.0 + 8.0 * (D /*uh*/ + 1.0 /*b*/));
        c->bytes[3] = W * (4.0 + 8.0 * D) + N * (4.0 + 8.0 * (1.0 + D /*uh*/ + D /*un*/));
        c->has_fs = true;
        return MP_OK;
    });
}

int mp_fields_set(mp_ctx* c, int which, const double* values) {
    return guarded(c, "mp_fields_set", [&]() -> int {
        if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_fields_set: mp_fs_create not called");
        if (!values || which < 0 || which > MP_SW ||
            ((which == MP_W || which == MP_SW) && c->dim != 3))
            return fail(c, MP_ERR_ARG, "mp_fields_set: bad field");
        std::vector<double> h;
        if (which == MP_P) {
            to_device_order(c, values, h, c->ppe.get());
            upload(c, c->P.p, h);
            c->hist = 0;  // the pressure history no longer matches
            c->pcount = 0;
        } else if (which >= MP_SU) {  // AB2 source history, component which - MP_SU
            to_device_order(c, values, h);
            h.resize(c->npad);
            upload(c, c->S[c->s_cur].p + (size_t)(which - MP_SU) * c->npad, h);
        } else {
            to_device_order(c, values, h);
            upload(c, c->U[c->cur][which].p, h);
        }
        return MP_OK;
    });
}

int mp_fs_body_force(mp_ctx* c, const double* fx, const double* fy, const double* fz) {
    return guarded(c, "mp_fs_body_force", [&]() -> int {
        if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_fs_body_force: mp_fs_create not called");
        c->release_step_graphs();
        if (!fx && !fy && !fz) {
            c->has_force = false;
            return MP_OK;
        }
        if (!fx || !fy || (c->dim == 3) != (fz != nullptr))
            return fail(c, MP_ERR_ARG, "mp_fs_body_force: give fx, fy (and fz in 3D), or none");
        if (c->force.n < (size_t)3 * c->npad + 64) c->force.alloc((size_t)3 * c->npad + 64);
        const double* f[3] = {fx, fy, fz};
        for (int q = 0; q < c->dim; ++q) {
            std::vector<double> h;
            to_device_order(c, f[q], h);
            h.resize(c->npad);
            upload(c, c->force.p + (size_t)q * c->npad, h);
        }
        c->has_force = true;
        return MP_OK;
    });
}

int mp_fields_get(mp_ctx* c, int which, double* values) {
    return guarded(c, "mp_fields_get", [&]() -> int {
        if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_fields_get: mp_fs_create not called");
        if (!values || which < 0 || which > MP_SW ||
            ((which == MP_W || which == MP_SW) && c->dim != 3))
            return fail(c, MP_ERR_ARG, "mp_fields_get: bad field");
        std::vector<double> h;
        if (which == MP_P) {
            download(c, c->P.p, h);
            from_device_order(c, h, values, c->ppe.get());
        } else if (which >= MP_SU) {
            download(c, c->S[c->s_cur].p + (size_t)(which - MP_SU) * c->npad, h);
            from_device_order(c, h, values);
        } else {
            download(c, c->U[c->cur][which].p, h);
            from_device_order(c, h, values);
        }
        return MP_OK;
    });
}

}  // extern "C"

namespace {

// Runs launch() on the context's stream: replayed from the graph cached under
// key, captured into one on first use; launched directly when graphs are off
// or the step kernels are being timed (events stay out of graphs).
template <typename F>
void run_step_part(mp_ctx* c, const uintptr_t (&key)[8], F&& launch) {
    if (!c->use_graphs || c->timers[1].enabled || c->timers[2].enabled || c->timers[3].enabled) {
        launch();
        return;
    }
    for (mp_ctx::StepGraph& g : c->step_graphs)
        if (std::memcmp(g.key, key, sizeof(key)) == 0) {
            MP_HIP(hipGraphLaunch(g.exec, c->stream));
            return;
        }
    hipGraph_t graph = nullptr;
    MP_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
        launch();
    } catch (...) {
        (void)hipStreamEndCapture(c->stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    MP_HIP(hipStreamEndCapture(c->stream, &graph));
    mp_ctx::StepGraph g{};
    std::memcpy(g.key, key, sizeof(key));
    const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MP_HIP(e);
    c->step_graphs.push_back(g);
    MP_HIP(hipGraphLaunch(g.exec, c->stream));
}

template <int DIM>
void fs_step_one(mp_ctx* c, int it) {
    const mp_fs_params& q = c->prm;
    const int o = c->cur, nw = c->cur ^ 1;
    DevBuf<double>(&Uo)[3] = c->U[o];
    DevBuf<double>(&Un)[3] = c->U[nw];
    const int euler = (q.temporal_order == 1 || it == 0) ? 1 : 0;
    double* S_new = c->S[c->s_cur ^ 1].p;
    const double* S_oo = c->S[c->s_cur].p;
    const int* bp = c->per_problem_prm ? c->blk_prob.p : nullptr;
    const double* pp = c->per_problem_prm ? c->prob_prm.p : nullptr;
    const double* force = c->has_force ? c->force.p : nullptr;
    // the classical guess: the mode's own, or the one the learned network corrects
    const bool learned_mode = c->guess_mode == MP_GUESS_LEARNED;
    const int base = learned_mode ? c->guess_base : c->guess_mode;
    const bool extrapolate = base == MP_GUESS_EXTRAPOLATE && c->hist >= 2;
    const bool project = base == MP_GUESS_PROJECT && c->pcount >= 1;
    const int kp = std::min(c->pcount, c->proj_window);
    const int order = std::min(c->guess_order, c->hist);  // (the history held so far)
    const uintptr_t key_pre[8] = {1, (uintptr_t)o, (uintptr_t)c->s_cur, (uintptr_t)euler,
                                  (uintptr_t)force,
                                  (uintptr_t)(project ? 16 + kp : extrapolate ? order : 0),
                                  (uintptr_t)bp, (uintptr_t)pp};
    run_step_part(c, key_pre, [&]() {
        // calc_vel_hat
        c->timers[1].begin(c->stream);
        launch_predictor<DIM>(c->stream, c->sv, c->w[0].p, c->w[1].p, c->w[2].p, c->w[3].p,
                              Uo[0].p, Uo[1].p, Uo[2].p, c->uh[0].p, c->uh[1].p, c->uh[2].p, S_new,
                              S_oo, c->npad, q.rho, q.mu, q.dt, euler, c->upk.p, c->uhpk.p, bp, pp,
                              force);
        c->timers[1].end(c->stream);
        // calc_pressure
        c->timers[2].begin(c->stream);
        launch_ppe_rhs<DIM>(c->stream, c->sv, c->w[0].p, c->w[1].p, c->w[2].p, c->uh[0].p,
                            c->uh[1].p, c->uh[2].p, c->uhpk.p, Uo[0].p, Uo[1].p, Uo[2].p, c->P.p,
                            c->bnd.p, c->pdir.p, c->normal.p, c->npad, q.rho, q.dt, c->b.p, bp, pp);
        c->timers[2].end(c->stream);
        if (extrapolate) {
            hipLaunchKernelGGL(k_extrapolate, dim3(c->npad / 256), dim3(256), 0, c->stream,
                               c->npad, c->npad, order, c->P1.p, c->P2.p, c->P3.p, c->P4.p,
                               c->P.p);
            MP_HIP(hipGetLastError());
        } else if (project) {
            launch_proj_guess(c->stream, kp, c->npad, c->Hp.p, c->Hw.p, (size_t)c->npad,
                              c->pstate.p, c->b.p, c->proj_partial.p, c->P.p);
        }
    });
    bool learned = false, skipped = false;
    if (learned_mode && c->hist >= 2 && (extrapolate || project)) {
        if (!c->train && c->guess_skip > 0) {
            --c->guess_skip;  // backing off: the classical guess alone
            skipped = true;
        } else {
            const double* r = c->ppe->residual_of(c->stream, c->b.p, c->P.p);
            c->net->features(c->stream, c->P1.p, c->P2.p, r, c->ppe->diag_inv, c->xyz.p, c->npad,
                             c->dim);
            c->net->forward(c->stream, c->train != 0);
            // the classical guess stays the second candidate: the solve starts
            // from it when the network's leaves a non-finite or far larger
            // FP64 residual
            hipLaunchKernelGGL(k_copy, dim3(c->npad / 256), dim3(256), 0, c->stream, c->npad,
                               c->P.p, c->Ptmp.p);
            MP_HIP(hipGetLastError());
            c->net->guess(c->stream, c->Ptmp.p, c->P.p);
            if (!c->train) {
                c->ppe->fallback_rr = c->ppe->last_residual_norm2(c->stream);
                c->ppe->fallback_x = c->Ptmp.p;
            }
            learned = true;
        }
    }
    c->ppe->target_margin = project ? 0.25 : 0.5;
    SolveStats st = c->ppe->solve(c->stream, c->b.p, c->P.p);
    if (learned && !c->train) {
        if (st.guess_rejected) {
            c->guess_backoff = c->guess_backoff ? std::min(2 * c->guess_backoff, 64) : 1;
            c->guess_skip = c->guess_backoff;
        } else {
            c->guess_backoff = 0;
        }
    }
    if (learned && c->train) {
        c->net->backward(c->stream, c->Ptmp.p, c->P.p);
        c->net->adam_step(c->stream);
        c->last_loss = c->net->loss_host(c->stream);
    }
    const int depth = c->guess_order;  // (history kept: as many pressures as the order uses)
    const bool keep_proj = base == MP_GUESS_PROJECT;
    const uintptr_t key_post[8] = {2, (uintptr_t)o, (uintptr_t)bp, (uintptr_t)pp,
                                   (uintptr_t)depth, (uintptr_t)keep_proj, 0, 0};
    run_step_part(c, key_post, [&]() {
        // pressure history for the extrapolated / learned guesses
        if (depth >= 4)
            hipLaunchKernelGGL(k_copy, dim3(c->npad / 256), dim3(256), 0, c->stream, c->npad,
                               c->P3.p, c->P4.p);
        if (depth >= 3)
            hipLaunchKernelGGL(k_copy, dim3(c->npad / 256), dim3(256), 0, c->stream, c->npad,
                               c->P2.p, c->P3.p);
        hipLaunchKernelGGL(k_copy, dim3(c->npad / 256), dim3(256), 0, c->stream, c->npad, c->P1.p,
                           c->P2.p);
        hipLaunchKernelGGL(k_copy, dim3(c->npad / 256), dim3(256), 0, c->stream, c->npad, c->P.p,
                           c->P1.p);
        MP_HIP(hipGetLastError());
        if (keep_proj)  // the solution and A p = b - (b - A p) as the newest entry
            launch_proj_store(c->stream, c->npad, c->P.p, c->b.p, c->ppe->final_residual(),
                              c->Hp.p, c->Hw.p, (size_t)c->npad, c->pstate.p);
        // calc_vel_corr
        c->timers[3].begin(c->stream);
        launch_corrector<DIM>(c->stream, c->sv, c->w[0].p, c->w[1].p, c->w[2].p, c->P.p,
                              c->uh[0].p, c->uh[1].p, c->uh[2].p, Un[0].p, Un[1].p, Un[2].p,
                              c->npad, q.rho, q.dt, bp, pp);
        c->timers[3].end(c->stream);
        launch_boundary_fix<DIM>(c->stream, c->sv, c->brows.p, c->nb, c->w[0].p, c->w[1].p,
                                 c->w[2].p, c->normal.p, c->npad, c->pdir.p, c->udir[0].p,
                                 c->udir[1].p, c->udir[2].p, Uo[0].p, Uo[1].p, Uo[2].p, Un[0].p,
                                 Un[1].p, Un[2].p);
    });
    c->hist = std::min(c->hist + 1, depth);
    if (keep_proj) c->pcount = std::min(c->pcount + 1, PROJ_MAX);
    // extras (AB2 history) and the driver's old <- new
    if (q.temporal_order == 2) c->s_cur ^= 1;
    c->cur = nw;
    put_stats(c->stats, st);
    c->stats.guess_skipped = skipped ? 1 : 0;
    c->stats.total_iterations += st.iterations;
    c->stats.total_solves += 1;
    c->prob_total_iters.resize(c->nprob, 0);
    for (int pr = 0; pr < c->nprob; ++pr) c->prob_total_iters[pr] += c->ppe->seg_stats[pr].iterations;
}

}  // namespace

extern "C" {

int mp_fs_step(mp_ctx* c, int n_steps, int it0) {
    return guarded(c, "mp_fs_step", [&]() -> int {
        if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_fs_step: mp_fs_create not called");
        if (n_steps < 0 || it0 < 0) return fail(c, MP_ERR_ARG, "mp_fs_step: negative count");
        for (int k = 0; k < n_steps; ++k) {
            if (c->dim == 2)
                fs_step_one<2>(c, it0 + k);
            else
                fs_step_one<3>(c, it0 + k);
        }
        MP_HIP(hipStreamSynchronize(c->stream));
        for (int t = 1; t <= 3; ++t) c->timers[t].harvest();
        return MP_OK;
    });
}

int mp_stats_get(mp_ctx* c, mp_step_stats* stats) {
    return guarded(c, "mp_stats_get", [&]() -> int {
        if (!stats) return fail(c, MP_ERR_ARG, "mp_stats_get: NULL");
        *stats = c->stats;
        return MP_OK;
    });
}

int mp_guess_mode(mp_ctx* c, int mode) {
    return guarded(c, "mp_guess_mode", [&]() -> int {
        if (mode != MP_GUESS_PREVIOUS && mode != MP_GUESS_LEARNED &&
            mode != MP_GUESS_EXTRAPOLATE && mode != MP_GUESS_PROJECT)
            return fail(c, MP_ERR_ARG, "mp_guess_mode: unknown mode");
        if ((mode == MP_GUESS_LEARNED || mode == MP_GUESS_PROJECT) && c->nprob > 1)
            return fail(c, MP_ERR_ARG,
                        "mp_guess_mode: the learned and projected guesses are for single problems");
        if (mode == MP_GUESS_PROJECT) {
            if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_guess_mode: mp_fs_create not called");
            ensure_projection(c);
        }
        if (mode == MP_GUESS_LEARNED) {
            if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_guess_mode: mp_fs_create not called");
            if (!c->net) ensure_net(c);
            if (c->guess_base == MP_GUESS_PROJECT) ensure_projection(c);
        }
        c->guess_mode = mode;
        c->guess_backoff = c->guess_skip = 0;
        return MP_OK;
    });
}

int mp_guess_order(mp_ctx* c, int order) {
    return guarded(c, "mp_guess_order", [&]() -> int {
        if (order < 2 || order > 4) return fail(c, MP_ERR_ARG, "mp_guess_order: order 2, 3 or 4");
        if (order > c->guess_order) c->hist = std::min(c->hist, c->guess_order);  // (kept so far)
        c->guess_order = order;
        return MP_OK;
    });
}

int mp_guess_base(mp_ctx* c, int base) {
    return guarded(c, "mp_guess_base", [&]() -> int {
        if (base != MP_GUESS_EXTRAPOLATE && base != MP_GUESS_PROJECT)
            return fail(c, MP_ERR_ARG, "mp_guess_base: MP_GUESS_EXTRAPOLATE or MP_GUESS_PROJECT");
        c->guess_base = base;
        c->guess_backoff = c->guess_skip = 0;
        if (c->guess_mode == MP_GUESS_LEARNED && base == MP_GUESS_PROJECT) ensure_projection(c);
        return MP_OK;
    });
}

int mp_guess_window(mp_ctx* c, int k) {
    return guarded(c, "mp_guess_window", [&]() -> int {
        if (k < 1 || k > PROJ_MAX)
            return fail(c, MP_ERR_ARG, "mp_guess_window: 1 to " + std::to_string(PROJ_MAX) + " pressures");
        c->proj_window = k;
        return MP_OK;
    });
}

int mp_guess_param_count(void) { return GNN_NPARAM; }

int mp_guess_train(mp_ctx* c, int enable, double learning_rate, unsigned long long seed) {
    return guarded(c, "mp_guess_train", [&]() -> int {
        if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_guess_train: mp_fs_create not called");
        if (c->nprob > 1) return fail(c, MP_ERR_ARG, "mp_guess_train: single problems only");
        if (!(learning_rate > 0)) return fail(c, MP_ERR_ARG, "mp_guess_train: learning rate <= 0");
        c->lr = (float)learning_rate;
        if (!c->net) ensure_net(c);
        c->net->lr = c->lr;
        if (seed) c->net->init_params(c->stream, seed);
        c->train = enable != 0;
        return MP_OK;
    });
}

int mp_guess_params(mp_ctx* c, int set, float* params) {
    return guarded(c, "mp_guess_params", [&]() -> int {
        if (!params) return fail(c, MP_ERR_ARG, "mp_guess_params: NULL");
        if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_guess_params: mp_fs_create not called");
        if (c->nprob > 1) return fail(c, MP_ERR_ARG, "mp_guess_params: single problems only");
        if (!c->net) ensure_net(c);
        if (set)
            c->net->set_params(c->stream, params);
        else
            c->net->get_params(c->stream, params);
        return MP_OK;
    });
}

int mp_guess_last_loss(mp_ctx* c, double* loss) {
    return guarded(c, "mp_guess_last_loss", [&]() -> int {
        if (!loss) return fail(c, MP_ERR_ARG, "mp_guess_last_loss: NULL");
        *loss = c->last_loss;
        return MP_OK;
    });
}

int mp_guess_eval(mp_ctx* c, const double* p1, const double* p2, const double* b,
                  const double* p_true, double* guess, float* grad, double* loss) {
    return guarded(c, "mp_guess_eval", [&]() -> int {
        if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_guess_eval: mp_fs_create not called");
        if (!p1 || !p2 || !b || !guess) return fail(c, MP_ERR_ARG, "mp_guess_eval: NULL");
        if (!c->net) ensure_net(c);
        if (c->nprob > 1) return fail(c, MP_ERR_ARG, "mp_guess_eval: single problems only");
        const PpeSolver* S = c->ppe.get();
        std::vector<double> h;
        DevBuf<double> d1, d2, db, dt, dg, dx;
        for (DevBuf<double>* q : {&d1, &d2, &db, &dt, &dg, &dx}) alloc_vec(c, *q);
        to_device_order(c, p1, h, S);
        upload(c, d1.p, h);
        to_device_order(c, p2, h, S);
        upload(c, d2.p, h);
        to_device_order(c, b, h, S);
        upload(c, db.p, h);
        // (the two-pressure form: 2 p1 - p2 on the node rows, the border
        // unknown keeps p1)
        hipLaunchKernelGGL(k_extrapolate, dim3(c->npad / 256), dim3(256), 0, c->stream, c->nv,
                           c->npad, 2, d1.p, d2.p, d2.p, d2.p, dx.p);
        const double* r = c->ppe->residual_of(c->stream, db.p, dx.p);
        c->net->features(c->stream, d1.p, d2.p, r, c->ppe->diag_inv, c->xyz.p, c->npad, c->dim);
        c->net->forward(c->stream, p_true != nullptr);
        c->net->guess(c->stream, dx.p, dg.p);
        download(c, dg.p, h);
        from_device_order(c, h, guess);
        if (p_true) {
            to_device_order(c, p_true, h, S);
            upload(c, dt.p, h);
            c->net->backward(c->stream, dx.p, dt.p);
            double ls = c->net->loss_host(c->stream);
            if (loss) *loss = ls;
            if (grad) {
                MP_HIP(hipMemcpy(grad, c->net->grad.p, sizeof(float) * GNN_NPARAM,
                                 hipMemcpyDeviceToHost));
            }
        }
        return MP_OK;
    });
}

int mp_divergence(mp_ctx* c, double* out) {
    return guarded(c, "mp_divergence", [&]() -> int {
        if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_divergence: mp_fs_create not called");
        if (!out) return fail(c, MP_ERR_ARG, "mp_divergence: NULL");
        const int g = c->npad / BLOCK;  // every row block (padding rows add zeros)
        DevBuf<double>(&U)[3] = c->U[c->cur];
        if (c->dim == 2)
            launch_divergence<2>(c->stream, c->sv, c->w[0].p, c->w[1].p, c->w[2].p, U[0].p, U[1].p,
                                 U[2].p, c->bnd.p, c->div_partial.p, g);
        else
            launch_divergence<3>(c->stream, c->sv, c->w[0].p, c->w[1].p, c->w[2].p, U[0].p, U[1].p,
                                 U[2].p, c->bnd.p, c->div_partial.p, g);
        std::vector<double> h((size_t)g * 4);
        MP_HIP(hipMemcpyAsync(h.data(), c->div_partial.p, sizeof(double) * h.size(),
                              hipMemcpyDeviceToHost, c->stream));
        MP_HIP(hipStreamSynchronize(c->stream));
        double mi = 0, si = 0, mb = 0, sb = 0;
        for (int k = 0; k < g; ++k) {
            mi = std::max(mi, h[k * 4 + 0]);
            si += h[k * 4 + 1];
            mb = std::max(mb, h[k * 4 + 2]);
            sb += h[k * 4 + 3];
        }
        int nb = c->nb, ni = c->nv - c->nb;
        out[0] = mi;
        out[1] = ni ? si / ni : 0.0;
        out[2] = mb;
        out[3] = nb ? sb / nb : 0.0;
        out[4] = std::max(mi, mb);
        out[5] = (si + sb) / c->nv;
        return MP_OK;
    });
}

int mp_divergence_noalloc(mp_ctx* c, double* out_dev) {
    return guarded(c, "mp_divergence_noalloc", [&]() -> int {
        if (!c->has_fs)
            return fail(c, MP_ERR_STATE, "mp_divergence_noalloc: mp_fs_create not called");
        if (!out_dev) return fail(c, MP_ERR_ARG, "mp_divergence_noalloc: NULL");
        const int g = c->npad / BLOCK;
        if (c->npad > MP_NOALLOC_MAX_ROWS)
            return fail(c, MP_ERR_ARG,
                        "mp_divergence_noalloc: " + std::to_string(c->npad) +
                            " device rows exceed the scratch limit of " +
                            std::to_string(MP_NOALLOC_MAX_ROWS));
        double* part = scratch_ptr[c->device];  // recorded at mp_ctx_create
        if (!part) return fail(c, MP_ERR_STATE, "mp_divergence_noalloc: no scratch address");
        DevBuf<double>(&U)[3] = c->U[c->cur];
        if (c->dim == 2)
            launch_divergence<2>(c->stream, c->sv, c->w[0].p, c->w[1].p, c->w[2].p, U[0].p, U[1].p,
                                 U[2].p, c->bnd.p, part, g);
        else
            launch_divergence<3>(c->stream, c->sv, c->w[0].p, c->w[1].p, c->w[2].p, U[0].p, U[1].p,
                                 U[2].p, c->bnd.p, part, g);
        hipLaunchKernelGGL(k_div_finish, dim3(1), dim3(64), 0, c->stream, part, g, c->nv - c->nb,
                           c->nb, out_dev);
        MP_HIP(hipGetLastError());
        return MP_OK;
    });
}

int mp_batch_params(mp_ctx* c, const double* rho, const double* mu, const double* dt) {
    return guarded(c, "mp_batch_params", [&]() -> int {
        if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_batch_params: mp_fs_create not called");
        c->release_step_graphs();
        if (!rho || !mu || !dt) return fail(c, MP_ERR_ARG, "mp_batch_params: NULL");
        std::vector<double> h((size_t)3 * c->nprob);
        for (int q = 0; q < c->nprob; ++q) {
            if (!(rho[q] > 0) || !(mu[q] >= 0) || !(dt[q] > 0))
                return fail(c, MP_ERR_ARG, "mp_batch_params: rho, dt must be > 0, mu >= 0");
            h[3 * q] = rho[q];
            h[3 * q + 1] = mu[q];
            h[3 * q + 2] = dt[q];
        }
        c->prob_prm.alloc(h.size());
        MP_HIP(hipMemcpy(c->prob_prm.p, h.data(), sizeof(double) * h.size(),
                         hipMemcpyHostToDevice));
        c->per_problem_prm = true;
        return MP_OK;
    });
}

int mp_batch_stats(mp_ctx* c, int problem, mp_step_stats* stats) {
    return guarded(c, "mp_batch_stats", [&]() -> int {
        if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_batch_stats: mp_fs_create not called");
        if (!stats || problem < 0 || problem >= c->nprob)
            return fail(c, MP_ERR_ARG, "mp_batch_stats: bad problem index");
        const SolveStats& g = c->ppe->seg_stats[problem];
        mp_step_stats o{};
        put_stats(o, g);
        o.total_iterations = c->prob_total_iters.empty() ? 0 : c->prob_total_iters[problem];
        o.total_solves = c->stats.total_solves;
        *stats = o;
        return MP_OK;
    });
}

int mp_autotune(mp_ctx* c, int* chosen, double* ms_per_candidate) {
    return guarded(c, "mp_autotune", [&]() -> int {
        if (!c->has_fs || c->stats.total_solves == 0)
            return fail(c, MP_ERR_STATE, "mp_autotune: needs a context that has stepped");
        // the pressure solve of the last step (right-hand side b, started from
        // the pressure before that step, kept in P2), repeated into scratch for
        // each poll interval; results do not depend on the interval, since a
        // system that has stopped turns the rest of its chunk into no-ops
        static const int cand[5] = {4, 8, 16, 32, 64};
        PpeSolver& S = *c->ppe;
        const int n = c->npad;
        if (S.amg) {  // multigrid layouts by timing (rounding-level change of the V-cycle)
            S.release_graphs();
            S.amg->retune(c->stream);
        }
        double best_ms = 1e300;
        int best = S.check_every;
        for (int k = 0; k < 5; ++k) {
            S.check_every = cand[k];
            double ms = 0.0;
            for (int rep = 0; rep < 2; ++rep) {  // the first run also captures the graphs
                hipLaunchKernelGGL(k_copy, dim3(n / 256), dim3(256), 0, c->stream, n, c->P2.p,
                                   c->vec_a.p);
                MP_HIP(hipGetLastError());
                MP_HIP(hipStreamSynchronize(c->stream));
                const auto t0 = std::chrono::steady_clock::now();
                S.solve(c->stream, c->b.p, c->vec_a.p);
                ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                         .count();
            }
            if (ms_per_candidate) ms_per_candidate[k] = ms;
            if (ms < best_ms) {
                best_ms = ms;
                best = cand[k];
            }
        }
        S.check_every = best;
        if (chosen) *chosen = best;
        return MP_OK;
    });
}

int mp_precond_info(mp_ctx* c, int max_levels, int* levels, int* rows, long* nnz,
                    double* setup_seconds) {
    return guarded(c, "mp_precond_info", [&]() -> int {
        if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_precond_info: mp_fs_create not called");
        if (!levels || max_levels < 0) return fail(c, MP_ERR_ARG, "mp_precond_info: bad arguments");
        const AmgDevice* a = c->ppe->amg.get();
        const int L = a ? (int)a->rows.size() : 0;
        *levels = L;
        for (int l = 0; l < L && l < max_levels; ++l) {
            if (rows) rows[l] = a->rows[l];
            if (nnz) nnz[l] = a->nnz[l];
        }
        if (setup_seconds) *setup_seconds = a ? a->setup_seconds : 0.0;
        return MP_OK;
    });
}

int mp_precond_apply(mp_ctx* c, int cycle, int n, const double* in, double* out) {
    return guarded(c, "mp_precond_apply", [&]() -> int {
        if (!c->has_fs) return fail(c, MP_ERR_STATE, "mp_precond_apply: mp_fs_create not called");
        if (!in || !out || cycle < 0 || cycle > 1)
            return fail(c, MP_ERR_ARG, "mp_precond_apply: bad arguments");
        PpeSolver& S = *c->ppe;
        AmgDevice* a = S.amg.get();
        if (!a)
            return fail(c, MP_ERR_STATE,
                        "mp_precond_apply: the pressure solve has no multigrid preconditioner");
        if (n != S.n_sys)
            return fail(c, MP_ERR_ARG, "mp_precond_apply: n is not the system size " +
                                           std::to_string(S.n_sys));
        std::vector<double> h;
        to_device_order(c, in, h, &S);
        upload(c, c->vec_a.p, h);
        a->plain_cycle = cycle == 1;
        try {
            a->apply(c->stream, c->vec_a.p, c->vec_b.p, S.nseg == 1 ? S.segs[0].border : 1,
                     nullptr);
        } catch (...) {
            a->plain_cycle = false;
            throw;
        }
        a->plain_cycle = false;
        download(c, c->vec_b.p, h);
        from_device_order(c, h, out, &S);
        return MP_OK;
    });
}

int mp_graph_mode(mp_ctx* c, int enable) {
    return guarded(c, "mp_graph_mode", [&]() -> int {
        c->use_graphs = enable != 0;
        c->release_step_graphs();
        for (PpeSolver* s : {c->solver.get(), c->ppe.get()})
            if (s) s->use_graphs = c->use_graphs;
        return MP_OK;
    });
}

int mp_kernel_timing(mp_ctx* c, int enable) {
    return guarded(c, "mp_kernel_timing", [&]() -> int {
        for (auto& t : c->timers) {
            t.enabled = enable != 0;
            t.reset();
        }
        for (PpeSolver* s : {c->solver.get(), c->ppe.get()})
            if (s) {
                s->spmv_timer.enabled = enable != 0;
                s->spmv_timer.reset();
                s->smooth_timer.enabled = enable != 0;
                s->smooth_timer.reset();
            }
        return MP_OK;
    });
}

int mp_kernel_timing_get(mp_ctx* c, int which, long* launches, double* mean_ms,
                         double* bytes_per_launch, long* all_launches) {
    return guarded(c, "mp_kernel_timing_get", [&]() -> int {
        if (which < 0 || which > 5 || !launches || !mean_ms || !bytes_per_launch)
            return fail(c, MP_ERR_ARG, "mp_kernel_timing_get: bad arguments");
        if (which == 0 || which == 5) {
            PpeSolver* s = c->ppe ? c->ppe.get() : c->solver.get();
            if (!s) return fail(c, MP_ERR_STATE, "mp_kernel_timing_get: no solver");
            long all = 0;
            (which == 0 ? s->spmv_timer : s->smooth_timer).summary(launches, mean_ms, &all);
            if (all_launches) *all_launches = all;
            *bytes_per_launch = which == 0 ? s->spmv_bytes(s->precision == MP_PREC_MIXED ? 1 : 0)
                                           : s->smooth_bytes();
        } else {
            KernelTimer& t = c->timers[which];
            long all = 0;
            t.summary(launches, mean_ms, &all);
            if (all_launches) *all_launches = all;
            *bytes_per_launch = c->bytes[which];
        }
        return MP_OK;
    });
}

}  // extern "C"
