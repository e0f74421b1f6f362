// Device side of the smoothed-aggregation preconditioner: one V(1,1)-cycle
// with damped-Jacobi smoothing, FP32 arithmetic and vectors over BF16 matrix
// values (host set-up in amg_build.hpp).
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "amg_build.hpp"
#include "mp_internal.hpp"

namespace mp {

// A sparse matrix of the hierarchy in one of two layouts, chosen per matrix
// by timing at set-up:
//   g == 0: SELL-64x4, one lane per row (short rows; ptr = slice pointers);
//   g  > 0: CSR with g lanes per row and a g-lane reduction (long rows of
//           the coarse operators and restrictions; ptr = row pointers).
struct KernelTimer;

struct AmgMat {
    int n = 0, nslices = 0, g = 0;
    long nnz = 0;
    DevBuf<int> ptr_b, col_b;
    DevBuf<uint16_t> val_b;  // BF16 values (round to nearest even)
    DevBuf<uint16_t> val2_b; // a second value array in the same slots (composed restriction)
    const int* ptr = nullptr;
    const int* col = nullptr;
    const uint16_t* val = nullptr;
    const uint16_t* col16 = nullptr;  // compressed columns of level 0 (the solver's)
    const int* cbase = nullptr;
    const uint16_t* val2 = nullptr;
};

struct AmgLevelDev {
    int n = 0;
    AmgMat A;      // operator of this level
    AmgMat P, R;   // prolongation from / restriction to the next level
    // composed steps (three or more levels): Q = R_{l-1} - A D R_{l-1} (values)
    // with R_{l-1} in the same slots (levels 2 .. L-2), T = P - D A P (levels
    // 1 .. L-2)
    AmgMat Q, T;
    DevBuf<float> dinv;  // omega / a_ii
    DevBuf<float> b, x, res, y;
};

struct AmgDevice {
    double omega = 0.7;  // Jacobi smoothing weight
    std::vector<AmgLevelDev> lv;
    int nc = 0;
    DevBuf<float> cinv;  // dense inverse(s) on the coarsest level
    bool blocked = false;            // a batch: block-diagonal coarse inverses
    DevBuf<int> cfirst, csize;       // per coarsest row: its block's first row and size
    DevBuf<long> cdata;              // per coarsest row: offset of its row of the inverse
    DevBuf<int> border_rows;
    std::vector<int> rows;
    std::vector<long> nnz;
    std::string layout;  // chosen layout per matrix, for reports
    double setup_seconds = 0.0;
    // host copies of the hierarchy (A[0] stays empty) and of the T_l, for
    // re-tuning layouts
    std::vector<mp_amg::Csr> hA, hP, hR, hT;
    // Three or more levels: the cycle below level 1 runs from the composed
    // matrices (Q_l, T_l and the coarsest solve times R_{L-2}, row-dense BF16
    // in cr_val: row r's size[r] entries at cr_data[r] act on the level-(L-2)
    // vector from row cr_first[r]).  plain_cycle runs the level-by-level
    // cycle instead (the same map, for comparisons).
    bool composed = false;
    bool plain_cycle = false;
    DevBuf<uint16_t> cr_val;
    DevBuf<int> cr_first, cr_size;
    DevBuf<long> cr_data;

    // Builds the hierarchy from the finest operator, given in SELL-64x4 slots
    // (device order) by the solver: FP64 values val64_dev in the slots of sv
    // for the host set-up; the finest level's products run on sv_apply (owned
    // by the solver, sv_apply.nv rows, possibly another order of each row's
    // slots) with a BF16 copy of val32_dev, given in sv_apply's slots.
    // blocks: (first device row, rows) of each problem; bordered[q]: problem q
    // is a full-Neumann (bordered) system, whose coarsest solve is the
    // bordered dense inverse (else the plain inverse).  A batch gets one
    // hierarchy per problem, all of one depth, stored block-diagonally, so
    // every level's products run over all problems at once.
    void setup(hipStream_t s, const SellView& sv, const double* val64_dev,
               const SellView& sv_apply, const float* val32_dev,
               const std::vector<std::pair<int, int>>& blocks, const std::vector<int>& bordered,
               const mp_amg::Options& opt);
    // device rows of a batch's border unknowns (copied through the V-cycle)
    void set_border_rows(hipStream_t s, const std::vector<int>& r);
    // Layout of every hierarchy matrix chosen by timing each candidate (SELL-64x4,
    // CSR with 4..64 lanes per row) instead of the set-up's rule of thumb.  The
    // layouts sum a row in different orders, so the V-cycle then changes at the
    // rounding level; callers holding graphs over these buffers must drop them.
    void retune(hipStream_t s);
    // out = M^-1 in for the bordered system of size nv + bordered: the V-cycle
    // on the node block, identity on the border unknown.  Every kernel returns
    // at once when done_flag (may be null) is non-zero.
    // prescaled: the level-0 smoother input w D^-1 in (FP32) is already in
    // prescaled_input() (written by the Krylov update kernels).
    // out32 (may be null): the node rows of out again in FP32 (they are FP32
    // values; the border unknown passes through only in out).
    void apply(hipStream_t s, const double* in, double* out, int bordered, const int* done_flag,
               bool prescaled = false, float* out32 = nullptr);
    // out = M^-1 in on the node block for a Krylov basis vector (FP32 or
    // FP64 values; the border unknown is not touched), written in FP32; the
    // smoother input w D^-1 in is already in prescaled_input().
    template <typename TI>
    void apply_basis(hipStream_t s, const TI* in, float* out, const int* done_flag);
    KernelTimer* post_timer = nullptr;  // times the level-0 post-smoothing product when set
    float* prescaled_input() { return lv.size() > 1 ? lv[0].x.p : nullptr; }
    const float* level0_dinv() const { return lv.empty() ? nullptr : lv[0].dinv.p; }

  private:
    template <typename TI, typename TO>
    void vcycle(hipStream_t s, const TI* in, TO* out, const double* border_in, double* border_out,
                const int* done_flag, bool prescaled, float* out32);
};

}  // namespace mp
