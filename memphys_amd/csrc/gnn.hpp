// Learned initial guess for the pressure solve: a message-passing network over
// the stencil graph (3 layers, width 64, BF16 weights, FP32 accumulation).
//
//   features  X_i = [ (p1-p2)_i / s, (D^-1 (b - A p_ext))_i / s, x_i, y_i, z_i ],
//             p_ext = the extrapolated pressure (2 p1 - p2, or through the
//             last three or four pressures),  s = rms(p1 - p2)
//   embed     h0 = relu(X W0 + b0)
//   layer l   m_l = A~ h_{l-1}  (A~ = A D^-1 restricted to the node block)
//             h_l = relu(h_{l-1} Ws_l + m_l Wn_l + b_l),   l = 1..3
//   output    d_i = h3_i . wo + bo;   guess = p_ext + s d
//
// Training: mean squared error of d against (p* - p_ext)/s, where p* is the
// converged pressure, with Adam on FP32 master weights; the forward and
// backward passes read the BF16 copy.
#pragma once
#include "mp_internal.hpp"

namespace mp {

constexpr int GNN_H = 64;
constexpr int GNN_F = 5;
constexpr int GNN_L = 3;
// parameter offsets in the packed vector
constexpr int GNN_W0 = 0;
constexpr int GNN_B0 = GNN_W0 + GNN_F * GNN_H;
constexpr int GNN_LAYER = GNN_B0 + GNN_H;                     // first hidden layer block
constexpr int GNN_LAYER_SIZE = 2 * GNN_H * GNN_H + GNN_H;    // Ws, Wn, b
constexpr int GNN_WO = GNN_LAYER + GNN_L * GNN_LAYER_SIZE;
constexpr int GNN_BO = GNN_WO + GNN_H;
constexpr int GNN_NPARAM = GNN_BO + 1;

struct GuessNet {
    int nv = 0, nblk_rows = 0;
    SellView sv{};          // node graph (device order, SELL-64)
    const float* a32 = nullptr;      // A~ values in SELL slots (FP32)
    // transposed operator (node block only), SELL-64 over columns
    DevBuf<int> tsptr, tcol, tlen;
    DevBuf<float> tval;
    SellView tsv{};
    DevBuf<float> theta, grad, adam_m, adam_v;
    DevBuf<uint16_t> theta_bf;       // BF16 copy used by the passes
    DevBuf<float> X, Z[GNN_L + 1], Hh[GNN_L + 1], Mm[GNN_L + 1], out, target, dbuf;
    DevBuf<uint16_t> Hb[GNN_L];      // BF16 copies of the hidden states the next layer aggregates
    DevBuf<float> dH, dZ, dTmp;
    DevBuf<double> scal;             // [0] s, [1] loss, [2..] scratch
    DevBuf<double> partial;
    DevBuf<float> wpart;             // weight-gradient partials
    int wpart_blocks = 0;
    long adam_t = 0;
    float lr = 1e-3f;

    void setup(hipStream_t st, const SellView& view, const float* a32_slots, int nslots_total);
    void init_params(hipStream_t st, uint64_t seed);
    void set_params(hipStream_t st, const float* host_params);
    void get_params(hipStream_t st, float* host_params);
    // features from p1, p2, b and the FP64 residual of p_ext (device order)
    void features(hipStream_t st, const double* p1, const double* p2, const double* resid,
                  const double* dinv, const double* xyz_soa, int npad, int dim);
    // keep_pre: store the pre-activations, which a following backward() needs
    void forward(hipStream_t st, bool keep_pre = true);
    // guess = p_ext + s d  (FP64 output, node block only)
    void guess(hipStream_t st, const double* p_ext, double* out);
    // loss and gradient of the MSE against p_true (sets grad); returns nothing (loss in scal[1])
    void backward(hipStream_t st, const double* p_ext, const double* p_true);
    void adam_step(hipStream_t st);
    double loss_host(hipStream_t st);
};

}  // namespace mp
