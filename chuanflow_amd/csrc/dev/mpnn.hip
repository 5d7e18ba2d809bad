// mpnn.hip -- forward and backward kernels of the message-passing network
// (FP16 or FP32 storage, MFMA with FP32 accumulation) and its Adam update.
#include "mfma.h"
#include "mpnn.h"

#include <algorithm>
#include <cmath>
#include <random>

namespace cfx
{
namespace
{

using S = MpnnShape;
constexpr int B = 256;       // 4 wavefronts
constexpr int kTile = 128;   // nodes per workgroup tile (32 per wavefront)
constexpr int kMaxWg = 512;  // persistent grid of the backward kernels
constexpr int kPartStride = S::K * S::H + S::H;  // per-workgroup partial of a 96 x 32 layer

// ---------------------------------------------------------------- features
__global__ void k_absmax(int n, const double* __restrict__ v, double* __restrict__ part)
{
    __shared__ double red[B / kWave];
    double m = 0.0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) m = fmax(m, fabs(v[i]));
    m = blockMax(m, red);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// max of the G partial maxima, by one workgroup (max is exact in any order)
__global__ void k_absmax_final(int G, const double* __restrict__ part, double* __restrict__ out)
{
    __shared__ double red[B / kWave];
    double m = 0.0;
    for (int g = threadIdx.x; g < G; g += B) m = fmax(m, part[g]);
    m = blockMax(m, red);
    if (threadIdx.x == 0) *out = m;
}

// |a_ii| guarded away from zero (max(|a_ii|, 1e-300))
__global__ void k_diag_abs(CsrView A, const double* __restrict__ val, double* __restrict__ d)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= A.n) return;
    const double v = A.diagSlot[i] >= 0 ? fabs(val[A.diagSlot[i]]) : 0.0;
    d[i] = v > 1e-300 ? v : 1e-300;
}

// node features x0 [n][8] and edge features ew, eg per CSR slot (0 on the
// diagonal): 8 lanes per row, lane q taking slots q, q+8, ... of the row (a
// wavefront streams 8 consecutive rows), row reductions over the 8 lanes,
// lane c writing feature c
constexpr int kFeatLanes = 8;
static_assert(S::FIN == kFeatLanes, "one lane per node feature");
__global__ __launch_bounds__(B) void k_features(CsrView A, const double* __restrict__ val,
                                                const double* __restrict__ dabs, const double* __restrict__ geo,
                                                const double* __restrict__ r0, const double* __restrict__ rmaxPtr,
                                                float* __restrict__ x0, float* __restrict__ ew,
                                                float* __restrict__ eg)
{
    const double rmax = r0 ? *rmaxPtr : 0.0;
    const int i = blockIdx.x * (B / kFeatLanes) + static_cast<int>(threadIdx.x) / kFeatLanes;
    const int q = threadIdx.x % kFeatLanes;
    if (i >= A.n) return;  // all 8 lanes of a row leave together
    const double dd = dabs[i];
    double off = 0.0, rs = 0.0, wmax = 0.0, wmin = HUGE_VAL, gsum = 0.0;
    int deg = 0;
    for (int k = A.rowPtr[i] + q; k < A.rowPtr[i + 1]; k += kFeatLanes)
    {
        const int j = A.col[k];
        const double v = val[k];
        rs += v;
        if (j == i)
        {
            ew[k] = 0.0f;
            eg[k] = 0.0f;
            continue;
        }
        const double w = fabs(v) / sqrt(dd * dabs[j]);
        const double g = geo ? geo[k] : 0.0;
        ew[k] = static_cast<float>(w);
        eg[k] = static_cast<float>(g);
        off += fabs(v);
        wmax = fmax(wmax, w);
        wmin = fmin(wmin, w);
        gsum += g;
        ++deg;
    }
#pragma unroll
    for (int o = kFeatLanes / 2; o > 0; o >>= 1)
    {
        off += __shfl_xor(off, o, kFeatLanes);
        rs += __shfl_xor(rs, o, kFeatLanes);
        gsum += __shfl_xor(gsum, o, kFeatLanes);
        wmax = fmax(wmax, __shfl_xor(wmax, o, kFeatLanes));
        wmin = fmin(wmin, __shfl_xor(wmin, o, kFeatLanes));
        deg += __shfl_xor(deg, o, kFeatLanes);
    }
    double f;
    switch (q)
    {
    case 0: f = off / dd; break;
    case 1: f = rs / dd; break;
    case 2: f = deg / 8.0; break;
    case 3: f = wmax; break;
    case 4: f = deg > 0 ? wmin : 0.0; break;
    case 5: f = deg > 0 ? gsum / deg : 0.0; break;
    case 6: f = (r0 && rmax > 0.0) ? fabs(r0[i]) / rmax : 0.0; break;
    default: f = 1.0; break;
    }
    x0[static_cast<long long>(i) * S::FIN + q] = static_cast<float>(f);
}

// the residual feature alone (k_features' column 6, the same expression), for a graph whose other
// features are unchanged
__global__ __launch_bounds__(B) void k_feature_r0(int n, const double* __restrict__ r0,
                                                  const double* __restrict__ rmaxPtr, float* __restrict__ x0)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double rmax = r0 ? *rmaxPtr : 0.0;
    const double f = (r0 && rmax > 0.0) ? fabs(r0[i]) / rmax : 0.0;
    x0[static_cast<long long>(i) * S::FIN + 6] = static_cast<float>(f);
}

// 16 consecutive feature values (16-byte aligned) as FP32, in 16-byte loads
__device__ __forceinline__ void load16(const _Float16* p, float* v)
{
    const uint4* q = reinterpret_cast<const uint4*>(p);
    const uint4 a = q[0], b = q[1];
    const _Float16* ha = reinterpret_cast<const _Float16*>(&a);
    const _Float16* hb = reinterpret_cast<const _Float16*>(&b);
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        v[c] = static_cast<float>(ha[c]);
        v[8 + c] = static_cast<float>(hb[c]);
    }
}

__device__ __forceinline__ void load16(const float* p, float* v)
{
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int c = 0; c < 4; ++c)
    {
        const float4 f = q[c];
        v[4 * c] = f.x;
        v[4 * c + 1] = f.y;
        v[4 * c + 2] = f.z;
        v[4 * c + 3] = f.w;
    }
}

// ----------------------------------------------------------------- forward
// One layer on a tile of 128 nodes: X = [h, m, g] (AGG) or the input
// features (K = 8) staged in LDS, U = X W + b on MFMA, h_out = relu(U)
// (+ h_in for AGG).  Saves X and the sign mask of U when training.
template <typename T, int K, bool AGG>
__global__ __launch_bounds__(B) void k_layer_fwd(int n, CsrView A, const float* __restrict__ ew,
                                                 const float* __restrict__ eg, const T* __restrict__ hin,
                                                 const float* __restrict__ x0, const T* __restrict__ W,
                                                 const float* __restrict__ bias, T* __restrict__ hout,
                                                 T* __restrict__ xsave, unsigned* __restrict__ mask)
{
    constexpr int SX = K + 8;
    __shared__ T xl[kTile * SX];
    __shared__ T wl[K * S::H];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int base = blockIdx.x * kTile;
    for (int e = t; e < K * S::H; e += B) wl[e] = W[e];
    if constexpr (AGG)
    {
        // two threads per node, 16 features each
        const int node = t >> 1, f0 = (t & 1) * 16;
        const int i = base + node;
        float self[16], m[16], g[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) self[c] = m[c] = g[c] = 0.0f;
        if (i < n)
        {
            load16(hin + static_cast<long long>(i) * S::H + f0, self);
            // four neighbours per step: all four rows' loads are in flight before any is used; the sums
            // still take the neighbours in slot order
            const int kEnd = A.rowPtr[i + 1];
            for (int k = A.rowPtr[i]; k < kEnd; k += 4)
            {
                int j[4];
                float w[4], gg[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    const bool ok = k + q < kEnd;
                    j[q] = ok ? A.col[k + q] : i;
                    w[q] = ok ? ew[k + q] : 0.0f;
                    gg[q] = ok ? eg[k + q] : 0.0f;
                }
                float v[4][16];
#pragma unroll
                for (int q = 0; q < 4; ++q) load16(hin + static_cast<long long>(j[q]) * S::H + f0, v[q]);
                // fused multiply-adds (with FP16 rows they become v_fma_mix_f32, the conversion folded in)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    if (k + q < kEnd && j[q] != i)
                    {
#pragma unroll
                        for (int c = 0; c < 16; ++c)
                        {
                            m[c] = fmaf(w[q], v[q][c], m[c]);
                            g[c] = fmaf(gg[q], v[q][c], g[c]);
                        }
                    }
                }
            }
        }
        T* xr = xl + node * SX;
#pragma unroll
        for (int c = 0; c < 16; ++c)
        {
            xr[f0 + c] = static_cast<T>(self[c]);
            xr[S::H + f0 + c] = static_cast<T>(m[c]);
            xr[2 * S::H + f0 + c] = static_cast<T>(g[c]);
        }
    }
    else
    {
        for (int e = t; e < kTile * K; e += B)
        {
            const int node = e / K, c = e % K, i = base + node;
            xl[node * SX + c] = static_cast<T>(i < n ? x0[static_cast<long long>(i) * K + c] : 0.0f);
        }
    }
    __syncthreads();
    if (xsave)
        for (int e = t; e < kTile * K; e += B)
        {
            const int node = e / K, c = e % K, i = base + node;
            if (i < n) xsave[static_cast<long long>(i) * K + c] = xl[node * SX + c];
        }
    const T* xw = xl + wave * 32 * SX;
    f32x16 acc = zero16();
    acc = mfmaTile<T, K>(acc, [&](int r, int k) { return xw[r * SX + k]; },
                         [&](int k, int c) { return wl[k * S::H + c]; });
    const int col = mfmaCol(lane);
    const float bc = bias[col];
#pragma unroll
    for (int r = 0; r < 16; ++r)
    {
        const int node = wave * 32 + mfmaRow(r, lane), i = base + node;
        const float u = acc[r] + bc;
        float y = u > 0.0f ? u : 0.0f;
        if constexpr (AGG) y += static_cast<float>(xl[node * SX + col]);
        const unsigned long long bal = __ballot(u > 0.0f);
        if (i < n)
        {
            hout[static_cast<long long>(i) * S::H + col] = static_cast<T>(y);
            if (mask && lane == 0) mask[i] = static_cast<unsigned>(bal & 0xffffffffull);
            if (mask && lane == 32) mask[i] = static_cast<unsigned>(bal >> 32);
        }
    }
}

__device__ __forceinline__ float sigmoidf(float v) { return 1.0f / (1.0f + expf(-v)); }

// output heads: o = h W_out + b_out (FP32 weights), a = AMAX s(o0), t = TMAX s(o1)
template <typename T>
__global__ void k_head_fwd(int n, const T* __restrict__ h, const float* __restrict__ w, float* __restrict__ o,
                           float* __restrict__ a, float* __restrict__ t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* Wout = w + S::offWout;
    const float* bout = w + S::offBout;
    float o0 = bout[0], o1 = bout[1];
    const T* hi = h + static_cast<long long>(i) * S::H;
    for (int k = 0; k < S::H; ++k)
    {
        const float v = static_cast<float>(hi[k]);
        o0 += v * Wout[k * S::NOUT + 0];
        o1 += v * Wout[k * S::NOUT + 1];
    }
    o[2 * i] = o0;
    o[2 * i + 1] = o1;
    if (a) a[i] = kMpnnAMax * sigmoidf(o0);
    if (t) t[i] = kMpnnTMax * sigmoidf(o1);
}

// ---------------------------------------------------------------- backward
// head: go = dL/do, dh3 = go W_out^T, per-workgroup partials of dW_out, db_out
template <typename T>
__global__ __launch_bounds__(B) void k_head_bwd(int n, const T* __restrict__ h, const float* __restrict__ w,
                                                const float* __restrict__ o, const float* __restrict__ ga,
                                                const float* __restrict__ gt, float* __restrict__ dh,
                                                float* __restrict__ part)
{
    __shared__ float hs[B][S::H + 1];
    __shared__ float gs[B][2];
    const float* Wout = w + S::offWout;
    const int tid = threadIdx.x;
    float accW = 0.0f;  // thread < 64: dW_out[k = tid / 2][c = tid % 2]; 64, 65: db_out[c]
    for (int base = blockIdx.x * B; base < n; base += gridDim.x * B)
    {
        const int i = base + tid;
        float g0 = 0.0f, g1 = 0.0f;
        if (i < n)
        {
            const float s0 = sigmoidf(o[2 * i]), s1 = sigmoidf(o[2 * i + 1]);
            g0 = ga[i] * kMpnnAMax * s0 * (1.0f - s0);
            g1 = gt[i] * kMpnnTMax * s1 * (1.0f - s1);
            const T* hi = h + static_cast<long long>(i) * S::H;
            for (int k = 0; k < S::H; ++k)
            {
                hs[tid][k] = static_cast<float>(hi[k]);
                dh[static_cast<long long>(i) * S::H + k] = g0 * Wout[k * S::NOUT] + g1 * Wout[k * S::NOUT + 1];
            }
        }
        else
            for (int k = 0; k < S::H; ++k) hs[tid][k] = 0.0f;
        gs[tid][0] = g0;
        gs[tid][1] = g1;
        __syncthreads();
        if (tid < 64)
        {
            const int k = tid >> 1, c = tid & 1;
            for (int q = 0; q < B; ++q) accW += hs[q][k] * gs[q][c];
        }
        else if (tid < 66)
        {
            for (int q = 0; q < B; ++q) accW += gs[q][tid - 64];
        }
        __syncthreads();
    }
    if (tid < 66) part[blockIdx.x * 66 + tid] = accW;
}

// One layer backward on tiles of 128 nodes (persistent over tiles):
//   dU = dH_out * mask;  dX = dU W^T (AGG: dH_in = dH_out + dX_h, dM, dG);
//   dW += X^T dU (MFMA over the nodes), db += sum dU.
template <typename T, int K, bool AGG>
__global__ __launch_bounds__(B) void k_layer_bwd(int n, const float* __restrict__ dhout,
                                                 const unsigned* __restrict__ mask, const T* __restrict__ xsave,
                                                 const T* __restrict__ W, float* __restrict__ dhin,
                                                 float* __restrict__ dmg, float* __restrict__ part)
{
    constexpr int SX = K + 8, SD = S::H + 8, NB = (K + 31) / 32;
    __shared__ T xs[kTile * SX];
    __shared__ T du[kTile * SD];
    __shared__ T wl[K * S::H];
    __shared__ float red[4][32 * 33];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    for (int e = t; e < K * S::H; e += B) wl[e] = W[e];
    f32x16 accW[NB];
#pragma unroll
    for (int rb = 0; rb < NB; ++rb) accW[rb] = zero16();
    float dbAcc = 0.0f;  // column t % 32
    const int nTiles = (n + kTile - 1) / kTile;
    for (int tile = blockIdx.x; tile < nTiles; tile += gridDim.x)
    {
        const int base = tile * kTile;
        __syncthreads();
        for (int e = t; e < kTile * S::H; e += B)
        {
            const int node = e / S::H, c = e % S::H, i = base + node;
            float d = 0.0f;
            if (i < n) d = ((mask[i] >> c) & 1u) ? dhout[static_cast<long long>(i) * S::H + c] : 0.0f;
            du[node * SD + c] = static_cast<T>(d);
            dbAcc += static_cast<float>(static_cast<T>(d));
        }
        for (int e = t; e < kTile * K; e += B)
        {
            const int node = e / K, c = e % K, i = base + node;
            xs[node * SX + c] = i < n ? xsave[static_cast<long long>(i) * K + c] : static_cast<T>(0.0f);
        }
        __syncthreads();
        const T* dw = du + wave * 32 * SD;
        if constexpr (AGG)
        {
#pragma unroll
            for (int cb = 0; cb < 3; ++cb)
            {
                f32x16 acc = zero16();
                acc = mfmaTile<T, S::H>(acc, [&](int r, int k) { return dw[r * SD + k]; },
                                        [&](int k, int j) { return wl[(cb * 32 + j) * S::H + k]; });
                const int j = mfmaCol(lane);
#pragma unroll
                for (int r = 0; r < 16; ++r)
                {
                    const int i = base + wave * 32 + mfmaRow(r, lane);
                    if (i >= n) continue;
                    if (cb == 0)
                        dhin[static_cast<long long>(i) * S::H + j] = acc[r] + dhout[static_cast<long long>(i) * S::H + j];
                    else
                        dmg[static_cast<long long>(i) * 2 * S::H + (cb - 1) * S::H + j] = acc[r];
                }
            }
        }
        const T* xw = xs + wave * 32 * SX;
#pragma unroll
        for (int rb = 0; rb < NB; ++rb)
            accW[rb] = mfmaTile<T, 32>(
                accW[rb],
                [&](int r, int k) { return (rb * 32 + r < K) ? xw[k * SX + rb * 32 + r] : static_cast<T>(0.0f); },
                [&](int k, int j) { return dw[k * SD + j]; });
    }
    // sum the four wavefronts' weight-gradient tiles, then the bias columns
    float* out = part + static_cast<long long>(blockIdx.x) * kPartStride;
#pragma unroll
    for (int rb = 0; rb < NB; ++rb)
    {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; ++r) red[wave][mfmaRow(r, lane) * 33 + mfmaCol(lane)] = accW[rb][r];
        __syncthreads();
        for (int e = t; e < 32 * 32; e += B)
        {
            const int row = e >> 5, c = e & 31;
            const float v = (red[0][row * 33 + c] + red[1][row * 33 + c]) + (red[2][row * 33 + c] + red[3][row * 33 + c]);
            if (rb * 32 + row < K) out[(rb * 32 + row) * S::H + c] = v;
        }
    }
    __syncthreads();
    red[0][t] = dbAcc;
    __syncthreads();
    if (t < S::H)
    {
        float s = 0.0f;
        for (int q = 0; q < B / S::H; ++q) s += red[0][q * S::H + t];
        out[K * S::H + t] = s;
    }
}

// dH_in[i] += sum_j w_ij dM[j] + g_ij dG[j]   (w, g symmetric)
__global__ __launch_bounds__(B) void k_agg_bwd(int n, CsrView A, const float* __restrict__ ew,
                                               const float* __restrict__ eg, const float* __restrict__ dmg,
                                               float* __restrict__ dhin)
{
    const int e = blockIdx.x * B + threadIdx.x;
    const int i = e >> 1, f0 = (e & 1) * 16;
    if (i >= n) return;
    float s[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) s[c] = 0.0f;
    for (int k = A.rowPtr[i]; k < A.rowPtr[i + 1]; ++k)
    {
        const int j = A.col[k];
        if (j == i) continue;
        const float w = ew[k], g = eg[k];
        const float* d = dmg + static_cast<long long>(j) * 2 * S::H;
#pragma unroll
        for (int c = 0; c < 16; ++c) s[c] += w * d[f0 + c] + g * d[S::H + f0 + c];
    }
    float* o = dhin + static_cast<long long>(i) * S::H + f0;
#pragma unroll
    for (int c = 0; c < 16; ++c) o[c] += s[c];
}

// grad[off + p] = sum over workgroups of part[g * stride + p], p < count (fixed order)
__global__ void k_reduce_parts(int G, int stride, int count, const float* __restrict__ part, float* __restrict__ grad,
                               int off)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    double s = 0.0;
    for (int g = 0; g < G; ++g) s += part[static_cast<long long>(g) * stride + p];
    grad[off + p] = static_cast<float>(s);
}

__global__ void k_adam(int np, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.0f - b1) * gp;
    const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    w[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

template <typename T>
__global__ void k_cast_params(int np, const float* __restrict__ w, T* __restrict__ wt)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) wt[p] = static_cast<T>(w[p]);
}

} // namespace

// --------------------------------------------------------------------- host
Mpnn::Mpnn(Context& ctx, std::uint64_t seed, bool fp16) : ctx_(ctx), fp16_(fp16)
{
    std::vector<float> w(S::nParams, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fanIn, int count, float gain) {
        const float lim = gain * std::sqrt(6.0f / (fanIn + S::H));
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int k = 0; k < count; ++k) w[off + k] = U(rng);
    };
    fill(S::offWin, S::FIN, S::FIN * S::H, 1.0f);
    for (int l = 0; l < S::L; ++l) fill(S::offLayer0 + l * S::layerSize, S::K, S::K * S::H, 0.5f);
    // output layer zero: the untrained network gives the tuned constants
    w32_.alloc(sizeof(float) * S::nParams);
    wT_.alloc((fp16 ? 2 : 4) * S::nParams);
    m_.alloc(sizeof(float) * S::nParams);
    v_.alloc(sizeof(float) * S::nParams);
    scal_.alloc(sizeof(double) * 1024);
    part_.alloc(sizeof(float) * kMaxWg * kPartStride);
    CFX_HIP(hipMemsetAsync(m_.p, 0, sizeof(float) * S::nParams, ctx.stream()));
    CFX_HIP(hipMemsetAsync(v_.p, 0, sizeof(float) * S::nParams, ctx.stream()));
    setParams(w.data());
}

Mpnn::~Mpnn() = default;

void Mpnn::refreshT()
{
    hipStream_t s = ctx_.stream();
    if (fp16_)
        k_cast_params<_Float16><<<divUp(S::nParams, B), B, 0, s>>>(S::nParams, w32_.as<float>(), wT_.as<_Float16>());
    else
        k_cast_params<float><<<divUp(S::nParams, B), B, 0, s>>>(S::nParams, w32_.as<float>(), wT_.as<float>());
    CFX_CHECK_LAUNCH();
}

void Mpnn::setParams(const float* host)
{
    CFX_HIP(hipMemcpyAsync(w32_.p, host, sizeof(float) * S::nParams, hipMemcpyHostToDevice, ctx_.stream()));
    refreshT();
    ctx_.sync();
}

void Mpnn::getParams(float* host) const
{
    CFX_HIP(hipMemcpyAsync(host, w32_.p, sizeof(float) * S::nParams, hipMemcpyDeviceToHost, ctx_.stream()));
    CFX_HIP(hipStreamSynchronize(ctx_.stream()));
}

void Mpnn::reserve(int n, int nnz)
{
    if (n <= cap_ && ew_.bytes >= sizeof(float) * static_cast<std::size_t>(nnz)) return;
    const std::size_t N = static_cast<std::size_t>(n), ts = fp16_ ? 2 : 4;
    x0_.alloc(sizeof(float) * N * S::FIN);
    dabs_.alloc(sizeof(double) * N);
    ew_.alloc(sizeof(float) * nnz);
    eg_.alloc(sizeof(float) * nnz);
    for (auto& b : h_) b.alloc(ts * N * S::H);
    xs_[0].alloc(ts * N * S::FIN);
    for (int l = 1; l <= S::L; ++l) xs_[l].alloc(ts * N * S::K);
    for (auto& b : mask_) b.alloc(sizeof(unsigned) * N);
    o_.alloc(sizeof(float) * N * S::NOUT);
    for (auto& b : dh_) b.alloc(sizeof(float) * N * S::H);
    dmg_.alloc(sizeof(float) * N * 2 * S::H);
    cap_ = n;
    featuresValid_ = false;
}

void Mpnn::setGraph(const CsrView& A, const double* val, const double* geo, const double* r0, bool sameGraph)
{
    hipStream_t s = ctx_.stream();
    reserve(A.n, A.nnz);
    // the matrix-only features of the last full pass still hold: the caller says the values are those of
    // that pass, and the pattern, value and geometry arrays are the same
    const bool cached = sameGraph && featuresValid_ && A.n == n_ && A.rowPtr == A_.rowPtr && A.col == A_.col &&
                        val == lastVal_ && geo == lastGeo_;
    A_ = A;
    n_ = A.n;
    const int Gm = std::max(1, std::min(256, divUp(A.n, B)));
    double* pm = scal_.as<double>();
    if (r0)
    {
        k_absmax<<<Gm, B, 0, s>>>(A.n, r0, pm);
        k_absmax_final<<<1, B, 0, s>>>(Gm, pm, pm + 512);
    }
    if (cached)
    {
        k_feature_r0<<<divUp(A.n, B), B, 0, s>>>(A.n, r0, pm + 512, x0_.as<float>());
        CFX_CHECK_LAUNCH();
        return;
    }
    k_diag_abs<<<divUp(A.n, B), B, 0, s>>>(A, val, dabs_.as<double>());
    k_features<<<divUp(A.n, B / kFeatLanes), B, 0, s>>>(A, val, dabs_.as<double>(), geo, r0, pm + 512,
                                                       x0_.as<float>(), ew_.as<float>(), eg_.as<float>());
    CFX_CHECK_LAUNCH();
    featuresValid_ = true;
    lastVal_ = val;
    lastGeo_ = geo;
}

template <typename T>
static void forwardT(hipStream_t s, int n, const CsrView& A, const float* ew, const float* eg, const float* x0,
                     const T* wT, const float* w32, T* h0, T* h1, T* const* xs, unsigned* const* mask, float* o,
                     float* a, float* t, bool train)
{
    const int G = divUp(n, kTile);
    k_layer_fwd<T, S::FIN, false><<<G, B, 0, s>>>(n, A, ew, eg, nullptr, x0, wT + S::offWin, w32 + S::offBin, h0,
                                                  train ? xs[0] : nullptr, train ? mask[0] : nullptr);
    T* hin = h0;
    T* hout = h1;
    for (int l = 0; l < S::L; ++l)
    {
        const int off = S::offLayer0 + l * S::layerSize;
        k_layer_fwd<T, S::K, true><<<G, B, 0, s>>>(n, A, ew, eg, hin, nullptr, wT + off, w32 + off + S::K * S::H,
                                                   hout, train ? xs[l + 1] : nullptr, train ? mask[l + 1] : nullptr);
        std::swap(hin, hout);
    }
    k_head_fwd<T><<<divUp(n, B), B, 0, s>>>(n, hin, w32, o, a, t);
    CFX_CHECK_LAUNCH();
}

void Mpnn::forward(float* a, float* t, bool train)
{
    hipStream_t s = ctx_.stream();
    if (fp16_)
    {
        _Float16* xs[S::L + 1];
        unsigned* mk[S::L + 1];
        for (int l = 0; l <= S::L; ++l)
        {
            xs[l] = xs_[l].as<_Float16>();
            mk[l] = mask_[l].as<unsigned>();
        }
        forwardT<_Float16>(s, n_, A_, ew_.as<float>(), eg_.as<float>(), x0_.as<float>(), wT_.as<_Float16>(),
                           w32_.as<float>(), h_[0].as<_Float16>(), h_[1].as<_Float16>(), xs, mk, o_.as<float>(), a,
                           t, train);
    }
    else
    {
        float* xs[S::L + 1];
        unsigned* mk[S::L + 1];
        for (int l = 0; l <= S::L; ++l)
        {
            xs[l] = xs_[l].as<float>();
            mk[l] = mask_[l].as<unsigned>();
        }
        forwardT<float>(s, n_, A_, ew_.as<float>(), eg_.as<float>(), x0_.as<float>(), wT_.as<float>(),
                        w32_.as<float>(), h_[0].as<float>(), h_[1].as<float>(), xs, mk, o_.as<float>(), a, t, train);
    }
}

template <typename T>
static void backwardT(hipStream_t s, int n, const CsrView& A, const float* ew, const float* eg, const T* wT,
                      const float* w32, const T* hLast, T* const* xs, unsigned* const* mask, const float* o,
                      const float* ga, const float* gt, float* dh0, float* dh1, float* dmg, float* part,
                      float* grad)
{
    const int nTiles = divUp(n, kTile);
    const int G = std::max(1, std::min(kMaxWg, nTiles));
    const int Gh = std::max(1, std::min(kMaxWg, divUp(n, B)));
    k_head_bwd<T><<<Gh, B, 0, s>>>(n, hLast, w32, o, ga, gt, dh0, part);
    k_reduce_parts<<<1, 128, 0, s>>>(Gh, 66, S::H * S::NOUT, part, grad, S::offWout);
    k_reduce_parts<<<1, 128, 0, s>>>(Gh, 66, S::NOUT, part + S::H * S::NOUT, grad, S::offBout);
    float* dout = dh0;
    float* din = dh1;
    for (int l = S::L - 1; l >= 0; --l)
    {
        const int off = S::offLayer0 + l * S::layerSize;
        k_layer_bwd<T, S::K, true><<<G, B, 0, s>>>(n, dout, mask[l + 1], xs[l + 1], wT + off, din, dmg, part);
        k_agg_bwd<<<divUp(2 * n, B), B, 0, s>>>(n, A, ew, eg, dmg, din);
        k_reduce_parts<<<divUp(kPartStride, B), B, 0, s>>>(G, kPartStride, kPartStride, part, grad, off);
        std::swap(dout, din);
    }
    k_layer_bwd<T, S::FIN, false><<<G, B, 0, s>>>(n, dout, mask[0], xs[0], wT + S::offWin, nullptr, nullptr, part);
    k_reduce_parts<<<divUp(S::FIN * S::H, B), B, 0, s>>>(G, kPartStride, S::FIN * S::H, part, grad, S::offWin);
    k_reduce_parts<<<1, S::H, 0, s>>>(G, kPartStride, S::H, part + S::FIN * S::H, grad, S::offBin);
    CFX_CHECK_LAUNCH();
}

void Mpnn::backward(const float* ga, const float* gt, float* grad)
{
    hipStream_t s = ctx_.stream();
    // the last layer's output is in h_[1] after the input layer and 3 (odd) layers
    const int last = (S::L % 2 == 1) ? 1 : 0;
    if (fp16_)
    {
        _Float16* xs[S::L + 1];
        unsigned* mk[S::L + 1];
        for (int l = 0; l <= S::L; ++l)
        {
            xs[l] = xs_[l].as<_Float16>();
            mk[l] = mask_[l].as<unsigned>();
        }
        backwardT<_Float16>(s, n_, A_, ew_.as<float>(), eg_.as<float>(), wT_.as<_Float16>(), w32_.as<float>(),
                            h_[last].as<_Float16>(), xs, mk, o_.as<float>(), ga, gt, dh_[0].as<float>(),
                            dh_[1].as<float>(), dmg_.as<float>(), part_.as<float>(), grad);
    }
    else
    {
        float* xs[S::L + 1];
        unsigned* mk[S::L + 1];
        for (int l = 0; l <= S::L; ++l)
        {
            xs[l] = xs_[l].as<float>();
            mk[l] = mask_[l].as<unsigned>();
        }
        backwardT<float>(s, n_, A_, ew_.as<float>(), eg_.as<float>(), wT_.as<float>(), w32_.as<float>(),
                         h_[last].as<float>(), xs, mk, o_.as<float>(), ga, gt, dh_[0].as<float>(), dh_[1].as<float>(),
                         dmg_.as<float>(), part_.as<float>(), grad);
    }
}

void Mpnn::adam(const float* grad, float lr, float beta1, float beta2, float eps)
{
    ++adamStep_;
    const float c1 = 1.0f - std::pow(beta1, static_cast<float>(adamStep_));
    const float c2 = 1.0f - std::pow(beta2, static_cast<float>(adamStep_));
    k_adam<<<divUp(S::nParams, B), B, 0, ctx_.stream()>>>(S::nParams, w32_.as<float>(), grad, m_.as<float>(),
                                                          v_.as<float>(), lr, beta1, beta2, eps, c1, c2);
    CFX_CHECK_LAUNCH();
    refreshT();
}

} // namespace cfx
