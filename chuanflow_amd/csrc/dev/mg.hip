// mg.hip -- aggregation multigrid: host hierarchy build, device Galerkin
// values and the FP32 V-cycle kernels.
#include "mg.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace cfx
{
namespace
{

constexpr int B = 256;
constexpr int kCoarseThreads = 1024;
constexpr int kCoarseMax = 2048;

__device__ __forceinline__ int xcdRemap(int h, int G)
{
    constexpr int X = 8;
    const int q = G / X, r = G % X;
    const int x = h % X, idx = h / X;
    return (x < r) ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

// Greedy aggregation of a symmetric pattern: (1) a row whose whole
// neighbourhood is free becomes a root and takes it; (2) free rows join the
// aggregate of their first aggregated neighbour (phase-1 aggregates only);
// (3) leftovers form aggregates with their free neighbours.
int aggregate(const std::vector<int>& rp, const std::vector<int>& col, std::vector<int>& agg)
{
    const int n = static_cast<int>(rp.size()) - 1;
    agg.assign(n, -1);
    int na = 0;
    for (int i = 0; i < n; ++i)
    {
        if (agg[i] >= 0) continue;
        bool free = true;
        for (int k = rp[i]; k < rp[i + 1] && free; ++k) free = agg[col[k]] < 0;
        if (!free) continue;
        for (int k = rp[i]; k < rp[i + 1]; ++k) agg[col[k]] = na;
        agg[i] = na;
        ++na;
    }
    std::vector<int> a2(agg);
    for (int i = 0; i < n; ++i)
    {
        if (agg[i] >= 0) continue;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (agg[col[k]] >= 0) { a2[i] = agg[col[k]]; break; }
    }
    agg.swap(a2);
    for (int i = 0; i < n; ++i)
    {
        if (agg[i] >= 0) continue;
        agg[i] = na;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (agg[col[k]] < 0) agg[col[k]] = na;
        ++na;
    }
    return na;
}

// ------------------------------------------------------------------ kernels
__global__ void k_galerkin(int nnzC, const int* __restrict__ gStart, const int* __restrict__ gPerm,
                           const double* __restrict__ fine, double* __restrict__ coarse)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nnzC) return;
    double s = 0.0;
    for (int t = gStart[c]; t < gStart[c + 1]; ++t) s += fine[gPerm[t]];
    coarse[c] = s;
}

// per row: FP32 values, D^-1, Wpre = a D^-1, Wpost = c D^-1, prolongation scale
__global__ void k_level_setup(CsrView A, const double* __restrict__ val64, float* __restrict__ val32,
                              float* __restrict__ invD32, float* __restrict__ wPre, float* __restrict__ wPost,
                              float* __restrict__ s32, float* __restrict__ t32, const float* __restrict__ aIn,
                              float aC, const float* __restrict__ sIn, float sC, const float* __restrict__ cIn,
                              float cC, const float* __restrict__ tIn, float tC)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    for (int k = A.rowPtr[i]; k < A.rowPtr[i + 1]; ++k) val32[k] = static_cast<float>(val64[k]);
    const double d = A.diagSlot[i] >= 0 ? val64[A.diagSlot[i]] : 0.0;
    const float inv = static_cast<float>(fabs(d) > 1e-300 ? 1.0 / d : 1.0);
    invD32[i] = inv;
    wPre[i] = (aIn ? aIn[i] : aC) * inv;
    wPost[i] = (cIn ? cIn[i] : cC) * inv;
    s32[i] = sIn ? sIn[i] : sC;
    t32[i] = tIn ? tIn[i] : tC;
}

// A Wpre: column-scaled values so that r = b - A (Wpre b) is one SpMV over b
__global__ void k_scale_vals(CsrView A, const float* __restrict__ val32, const float* __restrict__ wD32,
                             float* __restrict__ valS)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    for (int k = A.rowPtr[i]; k < A.rowPtr[i + 1]; ++k) valS[k] = val32[k] * wD32[A.col[k]];
}

// CSR-stream row sums of y = M v for the rows of block blk (M = valM)
template <int LDS>
__device__ __forceinline__ bool rowSum(const CsrView& A, const float* __restrict__ valM, const float* __restrict__ v,
                                       int blk, float* lds, int& row, float& sum)
{
    const int r0 = A.rowBlocks[blk], r1 = A.rowBlocks[blk + 1];
    const int k0 = A.rowPtr[r0], nnz = A.rowPtr[r1] - k0;
    row = r0 + static_cast<int>(threadIdx.x);
    float s = 0.0f;
    if (nnz <= LDS)
    {
        for (int k = threadIdx.x; k < nnz; k += B) lds[k] = valM[k0 + k] * v[A.col[k0 + k]];
        __syncthreads();
        if (row < r1)
            for (int k = A.rowPtr[row] - k0, e = A.rowPtr[row + 1] - k0; k < e; ++k) s += lds[k];
        __syncthreads();
    }
    else if (row < r1)
    {
        for (int k = A.rowPtr[row]; k < A.rowPtr[row + 1]; ++k) s += valM[k] * v[A.col[k]];
    }
    sum = s;
    return row < r1;
}

// pre-smoothing from x = 0 and residual: x = W b, r = b - (A W) b
__global__ __launch_bounds__(B) void k_mg_down(CsrView A, const float* __restrict__ valS,
                                               const float* __restrict__ b, const float* __restrict__ wD,
                                               float* __restrict__ x, float* __restrict__ r, const int* done)
{
    __shared__ float lds[kSpmvLds];
    if (done && *done) return;
    int row;
    float s;
    if (rowSum<kSpmvLds>(A, valS, b, xcdRemap(blockIdx.x, gridDim.x), lds, row, s))
    {
        const float bi = b[row];
        x[row] = wD[row] * bi;
        r[row] = bi - s;
    }
}

// bc = P^T (T r): T = diag(ts) (null: identity)
__global__ void k_mg_restrict(int nc, const int* __restrict__ aggStart, const int* __restrict__ aggMem,
                              const float* __restrict__ ts, const float* __restrict__ r, float* __restrict__ bc,
                              const int* done)
{
    if (done && *done) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    float s = 0.0f;
    if (ts)
        for (int t = aggStart[c]; t < aggStart[c + 1]; ++t) s += ts[aggMem[t]] * r[aggMem[t]];
    else
        for (int t = aggStart[c]; t < aggStart[c + 1]; ++t) s += r[aggMem[t]];
    bc[c] = s;
}

__global__ void k_mg_prolong(int n, const int* __restrict__ agg, const float* __restrict__ sc,
                             const float* __restrict__ ec, float* __restrict__ x, const int* done)
{
    if (done && *done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] += sc[i] * ec[agg[i]];
}

// post-smoothing: z = x + W (b - A x); optional partials of z.q
__global__ __launch_bounds__(B) void k_mg_up(CsrView A, int rbPerWg, const float* __restrict__ val32,
                                             const float* __restrict__ x, const float* __restrict__ b,
                                             const float* __restrict__ wD, float* __restrict__ z,
                                             const float* __restrict__ q, double* __restrict__ partZQ,
                                             const int* done)
{
    __shared__ float lds[kSpmvLds];
    __shared__ double red[B / kWave];
    if (done && *done) return;
    const int w = xcdRemap(blockIdx.x, gridDim.x);
    double zq = 0.0;
    for (int j = 0; j < rbPerWg; ++j)
    {
        const int blk = w * rbPerWg + j;
        if (blk >= A.nRowBlocks) break;
        int row;
        float s;
        if (rowSum<kSpmvLds>(A, val32, x, blk, lds, row, s))
        {
            const float zi = x[row] + wD[row] * (b[row] - s);
            z[row] = zi;
            if (partZQ) zq += static_cast<double>(zi) * static_cast<double>(q[row]);
        }
    }
    if (partZQ)
    {
        zq = blockSum(zq, red);
        if (threadIdx.x == 0) partZQ[blockIdx.x] = zq;
    }
}

// coarsest level: damped Jacobi sweeps from zero in one workgroup (LDS vectors)
__global__ __launch_bounds__(kCoarseThreads) void k_mg_coarse(int n, const int* __restrict__ rowPtr,
                                                              const int* __restrict__ col,
                                                              const float* __restrict__ val,
                                                              const float* __restrict__ invD,
                                                              const float* __restrict__ b, float* __restrict__ x,
                                                              int sweeps, float omega, const int* done)
{
    __shared__ float xa[kCoarseMax], xb[kCoarseMax], bl[kCoarseMax];
    if (done && *done) return;
    for (int i = threadIdx.x; i < n; i += kCoarseThreads)
    {
        bl[i] = b[i];
        xa[i] = 0.0f;
    }
    __syncthreads();
    float* cur = xa;
    float* nxt = xb;
    for (int it = 0; it < sweeps; ++it)
    {
        for (int i = threadIdx.x; i < n; i += kCoarseThreads)
        {
            float s = 0.0f;
            for (int k = rowPtr[i]; k < rowPtr[i + 1]; ++k) s += val[k] * cur[col[k]];
            nxt[i] = cur[i] + omega * invD[i] * (bl[i] - s);
        }
        __syncthreads();
        float* t = cur;
        cur = nxt;
        nxt = t;
    }
    for (int i = threadIdx.x; i < n; i += kCoarseThreads) x[i] = cur[i];
}

} // namespace

Multigrid::Multigrid(Context& ctx, const DevPattern& fine, int maxLevels, int coarseMax) : ctx_(ctx), fine_(fine)
{
    hipStream_t s = ctx.stream();
    coarseMax = std::min(coarseMax, kCoarseMax);
    std::vector<int> rp = fine.hRowPtr, cl = fine.hCol;
    lv_.emplace_back();
    lv_[0].n = fine.n;
    lv_[0].nnz = fine.nnz;
    while (true)
    {
        MgLevel& L = lv_.back();
        const int n = L.n;
        if (n <= coarseMax || static_cast<int>(lv_.size()) >= maxLevels) break;
        std::vector<int> agg;
        const int na = aggregate(rp, cl, agg);
        if (na >= n) break;  // no coarsening possible
        // members of each aggregate
        std::vector<int> aStart(na + 1, 0), aMem(n);
        for (int i = 0; i < n; ++i) ++aStart[agg[i] + 1];
        for (int c = 0; c < na; ++c) aStart[c + 1] += aStart[c];
        {
            std::vector<int> fill(aStart.begin(), aStart.end() - 1);
            for (int i = 0; i < n; ++i) aMem[fill[agg[i]]++] = i;
        }
        // coarse pattern and the Galerkin map, row by coarse row
        std::vector<int> crp(na + 1, 0), ccol, gStart, gPerm;
        gPerm.reserve(rp[n]);
        std::vector<std::pair<int, int>> pairs;
        for (int c = 0; c < na; ++c)
        {
            pairs.clear();
            for (int t = aStart[c]; t < aStart[c + 1]; ++t)
            {
                const int i = aMem[t];
                for (int k = rp[i]; k < rp[i + 1]; ++k) pairs.emplace_back(agg[cl[k]], k);
            }
            std::sort(pairs.begin(), pairs.end());
            // one coarse nonzero per distinct coarse column, summing its fine
            // nonzeros in ascending slot order
            for (std::size_t t = 0; t < pairs.size(); ++t)
            {
                if (t == 0 || pairs[t].first != pairs[t - 1].first)
                {
                    gStart.push_back(static_cast<int>(gPerm.size()));
                    ccol.push_back(pairs[t].first);
                }
                gPerm.push_back(pairs[t].second);
            }
            crp[c + 1] = static_cast<int>(ccol.size());
        }
        gStart.push_back(static_cast<int>(gPerm.size()));
        L.agg = upload(agg, s);
        L.aggStart = upload(aStart, s);
        L.aggMem = upload(aMem, s);
        L.gStart = upload(gStart, s);
        L.gPerm = upload(gPerm, s);
        L.hAgg = agg;
        MgLevel C;
        C.n = na;
        C.nnz = crp[na];
        lv_.push_back(std::move(C));
        lv_.back().pat.build(crp, ccol, s);
        rp.swap(crp);
        cl.swap(ccol);
    }
    if (lv_.size() > 1 && lv_.back().n > kCoarseMax)
        throw std::invalid_argument("multigrid: coarsest level exceeds 2048 rows (raise maxLevels)");
    for (std::size_t l = 0; l < lv_.size(); ++l)
    {
        MgLevel& L = lv_[l];
        const std::size_t n = L.n, nnz = L.nnz;
        if (l > 0) L.val64.alloc(sizeof(double) * nnz);
        L.val32.alloc(sizeof(float) * nnz);
        L.valS32.alloc(sizeof(float) * nnz);
        for (DevBuf* b : {&L.invD32, &L.wPre32, &L.wPost32, &L.s32, &L.t32, &L.b, &L.x, &L.r})
            b->alloc(sizeof(float) * n);
    }
    ctx.sync();
}

void Multigrid::setMatrix(const double* fineVal64, const float* pre, const float* scale, const float* post,
                          const float* restr, const MgParams& p)
{
    hipStream_t s = ctx_.stream();
    p_ = p;
    for (std::size_t l = 0; l < lv_.size(); ++l)
    {
        MgLevel& L = lv_[l];
        const CsrView A = l == 0 ? fine_.view() : L.pat.view();
        const double* v64 = l == 0 ? fineVal64 : L.val64.as<double>();
        if (l > 0)
        {
            const MgLevel& F = lv_[l - 1];
            const double* fv = l == 1 ? fineVal64 : F.val64.as<double>();
            k_galerkin<<<divUp(L.nnz, B), B, 0, s>>>(L.nnz, F.gStart.as<int>(), F.gPerm.as<int>(), fv,
                                                     L.val64.as<double>());
        }
        const bool fineLevel = l == 0;
        const float aC = fineLevel ? p.omegaPre : p.coarseOmega;
        const float cC = fineLevel ? p.omegaPost : p.coarseOmega;
        const float sC = fineLevel ? p.scale : p.coarseScale;
        k_level_setup<<<divUp(L.n, B), B, 0, s>>>(A, v64, L.val32.as<float>(), L.invD32.as<float>(),
                                                  L.wPre32.as<float>(), L.wPost32.as<float>(), L.s32.as<float>(),
                                                  L.t32.as<float>(), fineLevel ? pre : nullptr, aC,
                                                  fineLevel ? scale : nullptr, sC, fineLevel ? post : nullptr, cC,
                                                  fineLevel ? restr : nullptr, fineLevel ? p.restrictScale : 1.0f);
        k_scale_vals<<<divUp(L.n, B), B, 0, s>>>(A, L.val32.as<float>(), L.wPre32.as<float>(), L.valS32.as<float>());
    }
    CFX_CHECK_LAUNCH();
}

void Multigrid::recurse(int l, const int* done)
{
    hipStream_t s = ctx_.stream();
    MgLevel& L = lv_[l];
    const int last = levels() - 1;
    if (l == last)
    {
        const CsrView A = L.pat.view();
        k_mg_coarse<<<1, kCoarseThreads, 0, s>>>(L.n, A.rowPtr, A.col, L.val32.as<float>(), L.invD32.as<float>(),
                                                 L.b.as<float>(), L.x.as<float>(), p_.coarseSweeps, p_.coarsestOmega, done);
        return;
    }
    // levels > 0: b in L.b, result left in L.r
    const CsrView A = L.pat.view();
    MgLevel& C = lv_[l + 1];
    k_mg_down<<<A.nRowBlocks, B, 0, s>>>(A, L.valS32.as<float>(), L.b.as<float>(), L.wPre32.as<float>(),
                                         L.x.as<float>(), L.r.as<float>(), done);
    k_mg_restrict<<<divUp(C.n, B), B, 0, s>>>(C.n, L.aggStart.as<int>(), L.aggMem.as<int>(), nullptr,
                                              L.r.as<float>(), C.b.as<float>(), done);
    recurse(l + 1, done);
    const float* ec = (l + 1 == last) ? C.x.as<float>() : C.r.as<float>();
    k_mg_prolong<<<divUp(L.n, B), B, 0, s>>>(L.n, L.agg.as<int>(), L.s32.as<float>(), ec, L.x.as<float>(), done);
    k_mg_up<<<A.nRowBlocks, B, 0, s>>>(A, 1, L.val32.as<float>(), L.x.as<float>(), L.b.as<float>(),
                                       L.wPost32.as<float>(), L.r.as<float>(), nullptr, nullptr, done);
}

void Multigrid::apply(const float* r, float* z, const float* q, double* partZQ, int Gs, int rbPerWg, const int* done)
{
    hipStream_t s = ctx_.stream();
    MgLevel& L = lv_[0];
    const CsrView A = fine_.view();
    if (levels() == 1)
    {
        // no coarse level: two Jacobi sweeps
        k_mg_down<<<A.nRowBlocks, B, 0, s>>>(A, L.valS32.as<float>(), r, L.wPre32.as<float>(), L.x.as<float>(),
                                             L.r.as<float>(), done);
        k_mg_up<<<Gs, B, 0, s>>>(A, rbPerWg, L.val32.as<float>(), L.x.as<float>(), r, L.wPost32.as<float>(), z, q,
                                 partZQ, done);
        CFX_CHECK_LAUNCH();
        return;
    }
    MgLevel& C = lv_[1];
    k_mg_down<<<A.nRowBlocks, B, 0, s>>>(A, L.valS32.as<float>(), r, L.wPre32.as<float>(), L.x.as<float>(),
                                         L.r.as<float>(), done);
    k_mg_restrict<<<divUp(C.n, B), B, 0, s>>>(C.n, L.aggStart.as<int>(), L.aggMem.as<int>(), L.t32.as<float>(),
                                              L.r.as<float>(), C.b.as<float>(), done);
    recurse(1, done);
    const float* ec = (levels() == 2) ? C.x.as<float>() : C.r.as<float>();
    k_mg_prolong<<<divUp(L.n, B), B, 0, s>>>(L.n, L.agg.as<int>(), L.s32.as<float>(), ec, L.x.as<float>(), done);
    k_mg_up<<<Gs, B, 0, s>>>(A, rbPerWg, L.val32.as<float>(), L.x.as<float>(), r, L.wPost32.as<float>(), z, q,
                             partZQ, done);
    CFX_CHECK_LAUNCH();
}

} // namespace cfx
