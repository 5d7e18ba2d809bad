// mpnn.h -- message-passing network on the matrix graph that predicts the
// finest-level weights of the multigrid preconditioner.
//
// Graph: one node per row (cell), one edge per off-diagonal nonzero (face).
// Node inputs (8): sum|a_ij|/a_ii, row sum/a_ii, degree/8, max and min of
// w_ij = |a_ij|/sqrt(a_ii a_jj), mean edge geometry g_ij, |r0_i|/max|r0|, 1.
// Edge inputs: w_ij and g_ij (face non-orthogonality 1 - cos(Sf, d), 0 when
// the geometry is not known).
//
//   h0 = relu(x W_in + b_in)                                  (8 -> 32)
//   3 x: m_i = sum_j w_ij h_j, g_i = sum_j g_ij h_j,
//        h <- h + relu([h, m, g] W_l + b_l)                  (96 -> 32)
//   o = h W_out + b_out;  a = 1.6 sigmoid(o_0),  t = 3.2 sigmoid(o_1)
//
// a is the smoother weight (pre = post) and t the transfer scale (restriction
// = prolongation) of every fine row, so the cycle stays symmetric; a zero
// output layer gives a = 0.8, t = 1.6 (the tuned constants).  The three
// 96 -> 32 layers and the 8 -> 32 input layer run on MFMA (FP16 storage and
// FP32 accumulation, or FP32 throughout for gradient checks), with hand-written
// backward kernels (weight gradients as MFMA reductions over the nodes).
#pragma once

#include "types.h"

#include <cstdint>
#include <vector>

namespace cfx
{

struct MpnnShape
{
    static constexpr int H = 32, FIN = 8, L = 3, NOUT = 2;
    static constexpr int K = 3 * H;  // [h, m, g]
    static constexpr int offWin = 0, offBin = FIN * H;
    static constexpr int offLayer0 = offBin + H, layerSize = K * H + H;
    static constexpr int offWout = offLayer0 + L * layerSize, offBout = offWout + H * NOUT;
    static constexpr int nParams = offBout + NOUT;
};

constexpr float kMpnnAMax = 1.6f;  // a zero output gives a = 0.8, t = 1.6 (the centre the network was trained at)
constexpr float kMpnnTMax = 3.2f;

class Mpnn
{
public:
    Mpnn(Context& ctx, std::uint64_t seed, bool fp16);
    ~Mpnn();
    bool fp16() const { return fp16_; }
    void setParams(const float* host);
    void getParams(float* host) const;
    const float* deviceParams() const { return w32_.as<float>(); }

    // Inputs of a graph: CSR values (FP64), optional per-slot geometry g and
    // optional residual r0 (FP64, device).  Computes node/edge features.
    // sameGraph: the values are those of the previous call (same arrays, same
    // contents), so only the residual feature is computed again.
    void setGraph(const CsrView& A, const double* val, const double* geo, const double* r0, bool sameGraph = false);
    // a, t per row (device FP32).  `train` keeps the activations for backward.
    void forward(float* a, float* t, bool train);
    // dL/dparams (device FP32 [nParams]) from dL/da, dL/dt (device FP32 [n])
    void backward(const float* ga, const float* gt, float* grad);
    // Adam on the FP32 master weights; refreshes the FP16 copies
    void adam(const float* grad, float lr, float beta1 = 0.9f, float beta2 = 0.999f, float eps = 1e-8f);
    int n() const { return n_; }

private:
    Context& ctx_;
    bool fp16_;
    int n_ = 0, cap_ = 0;
    CsrView A_{};
    DevBuf w32_, wT_, m_, v_;     // FP32 master weights, storage-type copy, Adam moments
    DevBuf x0_, ew_, eg_, scal_;   // node features [n][8] FP32, edge weights per slot, scalars
    DevBuf dabs_;                  // guarded |a_ii|
    DevBuf h_[2], xs_[MpnnShape::L + 1], mask_[MpnnShape::L + 1], o_;
    DevBuf dh_[2], dmg_, part_;
    int adamStep_ = 0;
    bool featuresValid_ = false;  // x0_, ew_, eg_ hold the features of (A_, lastVal_, lastGeo_)
    const double *lastVal_ = nullptr, *lastGeo_ = nullptr;
    void reserve(int n, int nnz);
    void refreshT();
};

} // namespace cfx
