// mg.hip -- aggregation multigrid: host hierarchy build, device Galerkin
// values and the FP32 V-cycle kernels.
#include "mg.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace cfx
{
namespace
{

constexpr int B = 256;
constexpr int kCoarseThreads = 1024;
constexpr int kCoarseMax = 2048;
constexpr int kTailThreads = 1024;
constexpr int kFuseRowsMax = 65536;  // largest level whose prolongation is fused into post-smoothing
constexpr int kDeferGroupsMax = 256; // largest K-cycle grid whose combination may move into the prolongation

// the global address of array k of tail level T (null where the level has none)
__host__ __device__ inline const void* tailGlobal(const TailLevel& T, int k)
{
    switch (k)
    {
    case kTRowPtr: return T.rowPtr;
    case kTCol: return T.col;
    case kTValS: return T.valS;
    case kTVal: return T.val;
    case kTWPre: return T.wPre;
    case kTWPost: return T.wPost;
    case kTSc: return T.sc;
    case kTInvD: return T.invD;
    case kTAgg: return T.agg;
    case kTRRowPtr: return T.rRowPtr;
    case kTRCol: return T.rCol;
    case kTRval: return T.rval;
    default: return T.dense;
    }
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

// The coarse pattern of aggregation `agg` (na aggregates) of the pattern (rp, cl): coarse rows in
// aggregate order, each row's columns ascending.  With gStart/gPerm also the Galerkin map: coarse
// nonzero c sums the fine nonzeros gPerm[gStart[c] .. gStart[c+1]), in ascending slot order (the order
// in which the members, ascending, and their slots are visited, so the slots need no sort).
void coarsePattern(const std::vector<int>& rp, const std::vector<int>& cl, const std::vector<int>& agg, int na,
                   std::vector<int>& crp, std::vector<int>& ccol, std::vector<int>* gStart, std::vector<int>* gPerm)
{
    const int n = static_cast<int>(rp.size()) - 1;
    std::vector<int> aStart(na + 1, 0), aMem(n);
    for (int i = 0; i < n; ++i) ++aStart[agg[i] + 1];
    for (int c = 0; c < na; ++c) aStart[c + 1] += aStart[c];
    {
        std::vector<int> fill(aStart.begin(), aStart.end() - 1);
        for (int i = 0; i < n; ++i) aMem[fill[agg[i]]++] = i;
    }
    crp.assign(na + 1, 0);
    ccol.clear();
    ccol.reserve(static_cast<std::size_t>(na) * 16);
    if (gPerm)
    {
        gStart->clear();
        gStart->reserve(static_cast<std::size_t>(na) * 16 + 1);
        gPerm->assign(rp[n], 0);
    }
    // per coarse column: the coarse row that last saw it, its fine-nonzero count and next output slot
    std::vector<int> mark(na, -1), cnt(na), pos(na), row;
    int out = 0;
    for (int c = 0; c < na; ++c)
    {
        row.clear();
        for (int t = aStart[c]; t < aStart[c + 1]; ++t)
        {
            const int i = aMem[t];
            for (int k = rp[i]; k < rp[i + 1]; ++k)
            {
                const int cc = agg[cl[k]];
                if (mark[cc] != c)
                {
                    mark[cc] = c;
                    cnt[cc] = 0;
                    row.push_back(cc);
                }
                ++cnt[cc];
            }
        }
        std::sort(row.begin(), row.end());
        for (int cc : row)
        {
            if (gPerm)
            {
                gStart->push_back(out);
                pos[cc] = out;
                out += cnt[cc];
            }
            ccol.push_back(cc);
        }
        crp[c + 1] = static_cast<int>(ccol.size());
        if (gPerm)
            for (int t = aStart[c]; t < aStart[c + 1]; ++t)
            {
                const int i = aMem[t];
                for (int k = rp[i]; k < rp[i + 1]; ++k) (*gPerm)[pos[agg[cl[k]]]++] = k;
            }
    }
    if (gPerm) gStart->push_back(out);
}

// ------------------------------------------------------------------ kernels
// coarse nonzero c = sum of its fine nonzeros gPerm[gStart[c] .. gStart[c+1]): L lanes per coarse
// nonzero (lane q takes every L-th fine slot), then a fixed butterfly over the L lanes, so the result
// does not depend on the launch
template <int L>
__global__ void k_galerkin(int nnzC, const int* __restrict__ gStart, const int* __restrict__ gPerm,
                           const double* __restrict__ fine, double* __restrict__ coarse)
{
    const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int c = static_cast<int>(g / L), q = static_cast<int>(g % L);
    if (c >= nnzC) return;  // the L lanes of one coarse nonzero leave together
    double s = 0.0;
    for (int t = gStart[c] + q; t < gStart[c + 1]; t += L) s += fine[gPerm[t]];
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, L);
    if (q == 0) coarse[c] = s;
}

// per row: D^-1, Wpre = a D^-1, Wpost = c D^-1, prolongation and restriction scale
__global__ void k_level_setup(CsrView A, const double* __restrict__ val64,
                              float* __restrict__ invD32, float* __restrict__ wPre, float* __restrict__ wPost,
                              float* __restrict__ s32, float* __restrict__ t32, const float* __restrict__ aIn,
                              float aC, const float* __restrict__ sIn, float sC, const float* __restrict__ cIn,
                              float cC, const float* __restrict__ tIn, float tC)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const double d = A.diagSlot[i] >= 0 ? val64[A.diagSlot[i]] : 0.0;
    const float inv = static_cast<float>(fabs(d) > 1e-300 ? 1.0 / d : 1.0);
    invD32[i] = inv;
    wPre[i] = (aIn ? aIn[i] : aC) * inv;
    wPost[i] = (cIn ? cIn[i] : cC) * inv;
    s32[i] = sIn ? sIn[i] : sC;
    t32[i] = tIn ? tIn[i] : tC;
}

// per nonzero (flat over the slots): the FP32 copy and A Wpre, the column-scaled values with which
// r = b - A (Wpre b) is one SpMV over b
__global__ void k_scale_vals(CsrView A, const double* __restrict__ val64, float* __restrict__ val32,
                             const float* __restrict__ wD32, float* __restrict__ valS)
{
    for (long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; k < A.nnz;
         k += static_cast<long long>(gridDim.x) * blockDim.x)
    {
        const float v = static_cast<float>(val64[k]);
        if (val32) val32[k] = v;
        if (valS) valS[k] = v * wD32[A.col[k]];
    }
}

// the finest FP32 values in slot order: out[t] = val32[k] for the CSR slot k of slot t, 0 in padding
__global__ void k_slice_vals(long long slots, const int* __restrict__ idx, const float* __restrict__ val32,
                             float* __restrict__ out)
{
    for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < slots;
         t += static_cast<long long>(gridDim.x) * blockDim.x)
    {
        const int k = idx[t];
        out[t] = k >= 0 ? val32[k] : 0.0f;
    }
}

// the column-scaled values in slot order from the slot-order FP32 values: out[t] = val[t] * wD[col of t], 0
// in padding (the same FP32 products as k_scale_vals'); one workgroup per row block, one wave per slice
__global__ __launch_bounds__(B) void k_slice_scale(CsrView A, SliceView S, const float* __restrict__ val,
                                                   const float* __restrict__ wD, float* __restrict__ out)
{
    const int blk = blockIdx.x;
    const int r0 = A.rowBlocks[blk], r1 = A.rowBlocks[blk + 1];
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    if (r0 + wave * kWave >= r1) return;
    const int sl = S.blockSlice[blk] + wave;
    const int p1 = S.slicePtr[sl + 1];
    for (int t = S.slicePtr[sl] + static_cast<int>(threadIdx.x & (kWave - 1)); t < p1; t += kWave)
    {
        const int c = S.col[t];
        out[t] = c == kSlicePad ? 0.0f : val[t] * wD[r0 + c];
    }
}

__global__ void k_slice_vals64(long long slots, const int* __restrict__ idx, const double* __restrict__ val64,
                               double* __restrict__ out)
{
    for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < slots;
         t += static_cast<long long>(gridDim.x) * blockDim.x)
    {
        const int k = idx[t];
        out[t] = k >= 0 ? val64[k] : 0.0;
    }
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
        stageProducts(lds, valM, A, v, k0, r0, nnz);
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

template <bool kSliced>
__device__ __forceinline__ bool levelRowSum(const CsrView& A, const SliceView& S, const float* __restrict__ valM,
                                            const float* __restrict__ v, int blk, float* lds, int& row, float& sum)
{
    if constexpr (kSliced)
        return sliceRowSum<float>(A, S, valM, v, blk, row, sum);
    else
        return rowSum<kSpmvLds>(A, valM, v, blk, lds, row, sum);
}

// The restriction of one row block of a level numbered aggregate by
// aggregate (the block holds whole aggregates): bc[c] = sum of r over the
// rows of aggregate c, in row order from zero -- k_mg_restrict_csr's sum with
// unit weights, bit for bit.  The aggregate range and member starts are read
// at kernel entry (blockAggregates), before the block's products.
struct BlockAggregates
{
    int r0 = 0, c0 = 0, cEnd = -1, m0 = 0, m1 = 0;  // first row, aggregates [c0, cEnd], this thread's members
};
__device__ __forceinline__ BlockAggregates blockAggregates(const CsrView& A, int blk, const int* __restrict__ agg,
                                                           const int* __restrict__ aStart)
{
    BlockAggregates g;
    const int r0 = A.rowBlocks[blk], r1 = A.rowBlocks[blk + 1];
    g.r0 = r0;
    if (r1 > r0)
    {
        g.c0 = agg[r0];
        g.cEnd = agg[r1 - 1];
        const int c = g.c0 + static_cast<int>(threadIdx.x);
        if (c <= g.cEnd)
        {
            g.m0 = aStart[c];
            g.m1 = aStart[c + 1];
        }
    }
    return g;
}
__device__ __forceinline__ void restrictBlock(const BlockAggregates& g, float ri, bool mine, float* rl,
                                              float* __restrict__ bc)
{
    if (mine) rl[threadIdx.x] = ri;
    __syncthreads();
    const int c = g.c0 + static_cast<int>(threadIdx.x);
    if (c <= g.cEnd)
    {
        float sum = 0.0f;
        for (int i = g.m0; i < g.m1; ++i) sum += rl[i - g.r0];
        bc[c] = sum;
    }
}

// pre-smoothing from x = 0 and residual: x = W b, r = b - (A W) b; with bc
// (a level numbered aggregate by aggregate, unit restriction weights) also
// the coarse right-hand side bc = P^T r of the block's aggregates
template <bool kSliced>
__global__ __launch_bounds__(B) void k_mg_down(CsrView A, SliceView S, const float* __restrict__ valS,
                                               const float* __restrict__ b, const float* __restrict__ wD,
                                               float* __restrict__ x, float* __restrict__ r,
                                               const int* __restrict__ agg, const int* __restrict__ aStart,
                                               float* __restrict__ bc, const int* done)
{
    __shared__ float lds[kSliced ? 1 : kSpmvLds];
    __shared__ float rl[B];
    if (done && *done) return;
    const int blk = xcdRemap(blockIdx.x, gridDim.x);
    BlockAggregates g;
    if (bc) g = blockAggregates(A, blk, agg, aStart);
    int row;
    float s, ri = 0.0f;
    // the row's own operands are loaded before its sum (in flight with the slot loads)
    const int own = A.rowBlocks[blk] + static_cast<int>(threadIdx.x);
    const bool ownRow = own < A.rowBlocks[blk + 1];
    const float bi = ownRow ? b[own] : 0.0f, wi = ownRow ? wD[own] : 0.0f;
    const bool mine = levelRowSum<kSliced>(A, S, valS, b, blk, lds, row, s);
    if (mine)
    {
        x[row] = wi * bi;
        ri = bi - s;
        r[row] = ri;
    }
    if (bc) restrictBlock(g, ri, mine, rl, bc);
}

// a further pre-smoothing sweep: x += W r, r' = r - (A W) r (= b - A x of the
// new x); with bc as in k_mg_down, the coarse right-hand side from r'
template <bool kSliced>
__global__ __launch_bounds__(B) void k_mg_down_more(CsrView A, SliceView S, const float* __restrict__ valS,
                                                    const float* __restrict__ rIn, const float* __restrict__ wD,
                                                    float* __restrict__ x, float* __restrict__ rOut,
                                                    const int* __restrict__ agg, const int* __restrict__ aStart,
                                                    float* __restrict__ bc, const int* done)
{
    __shared__ float lds[kSliced ? 1 : kSpmvLds];
    __shared__ float rl[B];
    if (done && *done) return;
    const int blk = xcdRemap(blockIdx.x, gridDim.x);
    BlockAggregates g;
    if (bc) g = blockAggregates(A, blk, agg, aStart);
    int row;
    float s, ri = 0.0f;
    const int own = A.rowBlocks[blk] + static_cast<int>(threadIdx.x);
    const bool ownRow = own < A.rowBlocks[blk + 1];
    const float r0 = ownRow ? rIn[own] : 0.0f, wi = ownRow ? wD[own] : 0.0f, xi = ownRow ? x[own] : 0.0f;
    const bool mine = levelRowSum<kSliced>(A, S, valS, rIn, blk, lds, row, s);
    if (mine)
    {
        x[row] = xi + wi * r0;
        ri = r0 - s;
        rOut[row] = ri;
    }
    if (bc) restrictBlock(g, ri, mine, rl, bc);
}

// bc = P^T (T r) as a CSR-stream product over the transfer pattern
__global__ __launch_bounds__(B) void k_mg_restrict_csr(CsrView R, const float* __restrict__ rval,
                                                       const float* __restrict__ r, float* __restrict__ bc,
                                                       const int* done)
{
    __shared__ float lds[kSpmvLds];
    if (done && *done) return;
    int row;
    float s;
    if (rowSum<kSpmvLds>(R, rval, r, xcdRemap(blockIdx.x, gridDim.x), lds, row, s)) bc[row] = s;
}

// transfer values: restriction scale of each member (null: 1)
__global__ void k_restrict_vals(int n, const int* __restrict__ mem, const float* __restrict__ ts,
                                float* __restrict__ rval)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) rval[k] = ts ? ts[mem[k]] : 1.0f;
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
template <bool kSliced>
__global__ __launch_bounds__(B) void k_mg_up(CsrView A, SliceView S, int rbPerWg, const float* __restrict__ val32,
                                             const float* __restrict__ x, const float* __restrict__ b,
                                             const float* __restrict__ wD, float* __restrict__ z,
                                             const float* __restrict__ q, double* __restrict__ partZQ,
                                             const int* done)
{
    __shared__ float lds[kSliced ? 1 : kSpmvLds];
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
        const int own = A.rowBlocks[blk] + static_cast<int>(threadIdx.x);
        const bool ownRow = own < A.rowBlocks[blk + 1];
        const float xi = ownRow ? x[own] : 0.0f, bi = ownRow ? b[own] : 0.0f, wi = ownRow ? wD[own] : 0.0f;
        if (levelRowSum<kSliced>(A, S, val32, x, blk, lds, row, s))
        {
            const float zi = xi + wi * (bi - s);
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

// sum over k in [k0, k1) of val[k] * x[col[k]], added in slot order.  The
// loads go out 8 at a time: the x gathers are issued unconditionally
// (padding lanes read x[0]) and only the accumulation is predicated, by a
// select, so the compiler cannot sink each gather behind its own wait; the
// sum is the plain loop's, bit for bit.
__device__ __forceinline__ float rowDot8(const float* __restrict__ val, const int* __restrict__ col,
                                         const float* x, int k0, int k1)
{
    float s = 0.0f;
    for (int k = k0; k < k1; k += 8)
    {
        float v[8], xv[8];
        int c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            const bool ok = k + u < k1;
            v[u] = ok ? val[k + u] : 0.0f;
            c[u] = ok ? col[k + u] : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            const float t = s + v[u] * xv[u];
            s = k + u < k1 ? t : s;
        }
    }
    return s;
}

// sum over j < n of m[j * n + i] * x[j] in column order, 8 loads at a time
__device__ __forceinline__ float colDot8(const float* __restrict__ m, const float* x, int n, int i)
{
    float s = 0.0f;
    for (int j = 0; j < n; j += 8)
    {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = j + u < n ? m[(j + u) * n + i] : 0.0f;
        float xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = x[j + u < n ? j + u : 0];
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            const float t = s + v[u] * xv[u];
            s = j + u < n ? t : s;
        }
    }
    return s;
}

// rowDot8 over arrays in LDS: the value and column loads are unconditional
// (each staged array is followed by kLdsPad readable words) and only their
// use is masked, so the loads of a row go out together without a branch per
// load; the sum is rowDot8's, bit for bit.
constexpr int kLdsPad = 8;
__device__ __forceinline__ float rowDotLds8(const float* val, const int* col, const float* x, int k0, int k1)
{
    float s = 0.0f;
    for (int k = k0; k < k1; k += 8)
    {
        float v[8], xv[8];
        int c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            const float vv = val[k + u];
            const int cc = col[k + u];
            const bool ok = k + u < k1;
            v[u] = ok ? vv : 0.0f;
            c[u] = ok ? cc : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            const float t = s + v[u] * xv[u];
            s = k + u < k1 ? t : s;
        }
    }
    return s;
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
            const float s = rowDot8(val, col, cur, rowPtr[i], rowPtr[i + 1]);
            nxt[i] = cur[i] + omega * invD[i] * (bl[i] - s);
        }
        __syncthreads();
        float* t = cur;
        cur = nxt;
        nxt = t;
    }
    for (int i = threadIdx.x; i < n; i += kCoarseThreads) x[i] = cur[i];
}

// Dense inverse of the coarsest matrix (n <= 128), FP64 Gauss-Jordan in LDS
// by one workgroup, stored as FP32.  A pure Neumann operator (every row sum
// below 1e-9 max|diag|) is inverted as A + gamma 1 1^T with gamma =
// mean(diag) / n; a non-positive pivot is replaced by max|diag| (the result
// is then an approximate inverse, still a valid preconditioner).
constexpr int kDense = Multigrid::kDenseMax;
__global__ __launch_bounds__(1024) void k_mg_dense_inverse(CsrView A, const double* __restrict__ val,
                                                           float* __restrict__ inv)
{
    __shared__ double a[kDense * kDense];
    __shared__ double f[kDense];
    __shared__ double red[1024 / kWave];
    const int n = A.n, t = threadIdx.x;
    for (int i = t; i < n * n; i += 1024) a[i] = 0.0;
    __syncthreads();
    double dmax = 0.0, dsum = 0.0, rmax = 0.0;
    for (int i = t; i < n; i += 1024)
    {
        double rs = 0.0;
        for (int k = A.rowPtr[i]; k < A.rowPtr[i + 1]; ++k)
        {
            a[i * n + A.col[k]] = val[k];
            rs += val[k];
        }
        const double d = a[i * n + i];
        dmax = fmax(dmax, fabs(d));
        dsum += d;
        rmax = fmax(rmax, fabs(rs));
    }
    __syncthreads();
    dmax = blockMax(dmax, red);
    if (t == 0) f[0] = dmax;
    __syncthreads();
    dmax = f[0];
    __syncthreads();
    rmax = blockMax(rmax, red);
    if (t == 0) f[0] = rmax;
    __syncthreads();
    rmax = f[0];
    __syncthreads();
    dsum = blockSum(dsum, red);
    if (t == 0) f[0] = dsum;
    __syncthreads();
    dsum = f[0];
    __syncthreads();
    if (rmax <= 1e-9 * dmax)
    {
        const double gamma = dsum / n / n;
        for (int i = t; i < n * n; i += 1024) a[i] += gamma;
        __syncthreads();
    }
    for (int k = 0; k < n; ++k)
    {
        if (t < n) f[t] = a[t * n + k];
        __syncthreads();
        double piv = f[k];
        if (!(piv > 1e-14 * dmax)) piv = dmax > 0 ? dmax : 1.0;
        if (t < n) a[k * n + t] = (t == k) ? 1.0 / piv : a[k * n + t] / piv;
        __syncthreads();
        for (int e = t; e < n * n; e += 1024)
        {
            const int i = e / n, j = e % n;
            if (i == k) continue;
            a[e] = (j == k) ? -f[i] / piv : a[e] - f[i] * a[k * n + j];
        }
        __syncthreads();
    }
    for (int e = t; e < n * n; e += 1024) inv[e] = static_cast<float>(a[e]);
}

// x = inv b on the coarsest level: thread i forms row i (inv is symmetric,
// so column reads inv[j][i] are coalesced across threads)
__global__ __launch_bounds__(kDense) void k_mg_dense_apply(int n, const float* __restrict__ inv,
                                                           const float* __restrict__ b, float* __restrict__ x,
                                                           const int* done)
{
    __shared__ float bl[kDense];
    if (done && *done) return;
    const int i = threadIdx.x;
    if (i < n) bl[i] = b[i];
    __syncthreads();
    if (i >= n) return;
    x[i] = colDot8(inv, bl, n, i);
}

// ------------------------------------------------------------ coarse tail
// The cycle of the tail levels (0 = the tail's first level) in ONE workgroup,
// vectors in LDS: the same operations in the same order as k_mg_down,
// k_mg_restrict_csr, the coarsest solve, k_mg_prolong and k_mg_up (row sums
// in slot order, member sums in ascending member order, dense products in
// column order), so the result is bit for bit that of the separate kernels.
// The level arrays come from the accessor (acc.get<T>(level, array): global
// memory, or LDS at a known offset so the compiler issues LDS reads); v
// holds b of tail level 0 at entry; the correction of level 0 is written to
// eOut.
template <typename Acc>
__device__ __forceinline__ void tailCycle(const Acc& acc, float* v, float* eOut)
{
    const int t = threadIdx.x, nt = blockDim.x;
    const int count = acc.count(), last = count - 1;
    // level m: b at off[m], x at +n, r (later z) at +2n; coarsest: b, e (and the Jacobi ping-pong buffer)
    int off[kTailLevels];
    {
        int o = 0;
        for (int m = 0; m < count; ++m)
        {
            off[m] = o;
            o += 3 * acc.n(m);
        }
    }
    for (int m = 0; m < last; ++m)
    {
        const int n = acc.n(m), nc = acc.n(m + 1);
        const int* rowPtr = acc.template get<int>(m, kTRowPtr);
        const int* col = acc.template get<int>(m, kTCol);
        const float* valS = acc.template get<float>(m, kTValS);
        const float* wPre = acc.template get<float>(m, kTWPre);
        float *b = v + off[m], *x = b + n, *r = x + n;
        for (int i = t; i < n; i += nt)
        {
            const float s = acc.dot(valS, col, b, rowPtr[i], rowPtr[i + 1]);
            const float bi = b[i];
            x[i] = wPre[i] * bi;
            r[i] = bi - s;
        }
        __syncthreads();
        const int* rRowPtr = acc.template get<int>(m, kTRRowPtr);
        const int* rCol = acc.template get<int>(m, kTRCol);
        const float* rval = acc.template get<float>(m, kTRval);
        float* bc = v + off[m + 1];
        for (int c = t; c < nc; c += nt) bc[c] = acc.dot(rval, rCol, r, rRowPtr[c], rRowPtr[c + 1]);
        __syncthreads();
    }
    const float* ec;
    {
        const int n = acc.n(last);
        float *b = v + off[last], *e = b + n, *w = e + n;
        const float* dense = acc.template get<float>(last, kTDense);
        if (dense)
        {
            // each row's column-order sum (colDot8's order)
            for (int i = t; i < n; i += nt)
            {
                float sd = 0.0f;
                for (int j = 0; j < n; ++j) sd += dense[j * n + i] * b[j];
                e[i] = sd;
            }
            __syncthreads();
            ec = e;
        }
        else
        {
            // damped Jacobi sweeps from zero (k_mg_coarse)
            const int* rowPtr = acc.template get<int>(last, kTRowPtr);
            const int* col = acc.template get<int>(last, kTCol);
            const float* val = acc.template get<float>(last, kTVal);
            const float* invD = acc.template get<float>(last, kTInvD);
            for (int i = t; i < n; i += nt) e[i] = 0.0f;
            __syncthreads();
            float *cur = e, *nxt = w;
            const int sweeps = acc.sweeps();
            const float omega = acc.omega();
            for (int it = 0; it < sweeps; ++it)
            {
                for (int i = t; i < n; i += nt)
                {
                    const float s = acc.dot(val, col, cur, rowPtr[i], rowPtr[i + 1]);
                    nxt[i] = cur[i] + omega * invD[i] * (b[i] - s);
                }
                __syncthreads();
                float* tmp = cur;
                cur = nxt;
                nxt = tmp;
            }
            ec = cur;
        }
    }
    for (int m = last - 1; m >= 0; --m)
    {
        const int n = acc.n(m);
        const int* rowPtr = acc.template get<int>(m, kTRowPtr);
        const int* col = acc.template get<int>(m, kTCol);
        const float* val = acc.template get<float>(m, kTVal);
        const float* wPost = acc.template get<float>(m, kTWPost);
        const float* sc = acc.template get<float>(m, kTSc);
        const int* agg = acc.template get<int>(m, kTAgg);
        float *b = v + off[m], *x = b + n;
        float* z = m == 0 ? eOut : x + n;
        for (int i = t; i < n; i += nt) x[i] += sc[i] * ec[agg[i]];
        __syncthreads();
        for (int i = t; i < n; i += nt)
        {
            const float s = acc.dot(val, col, x, rowPtr[i], rowPtr[i + 1]);
            z[i] = x[i] + wPost[i] * (b[i] - s);
        }
        __syncthreads();
        ec = z;
    }
}

// tail arrays in global memory, the coarsest inverse in LDS
struct TailGlobalAcc
{
    const TailArgs& a;
    const float* dense;
    __device__ __forceinline__ int count() const { return a.count; }
    __device__ __forceinline__ int n(int m) const { return a.L[m].n; }
    __device__ __forceinline__ int sweeps() const { return a.sweeps; }
    __device__ __forceinline__ float omega() const { return a.omega; }
    __device__ __forceinline__ float dot(const float* v, const int* c, const float* x, int k0, int k1) const
    {
        return rowDot8(v, c, x, k0, k1);
    }
    template <typename T>
    __device__ __forceinline__ const T* get(int m, int k) const
    {
        if (k == kTDense) return a.L[m].dense ? reinterpret_cast<const T*>(dense) : nullptr;
        return static_cast<const T*>(tailGlobal(a.L[m], k));
    }
};

// The tail's sizes and staging slots, copied from the kernel arguments into
// LDS in one round at kernel entry (read per level from the arguments, each
// level's first use waits on a scalar load).
struct TailMeta
{
    int count, sweeps;
    float omega;
    int n[kTailLevels];
    int slot[kTailLevels * kTailArrays];
};
__device__ __forceinline__ void loadTailMeta(const TailArgs& a, const TailStage& st, TailMeta* mt)
{
    const int t = threadIdx.x;
    if (t < kTailLevels * kTailArrays) mt->slot[t] = st.slot[t / kTailArrays][t % kTailArrays];
    if (t < kTailLevels) mt->n[t] = a.L[t].n;
    if (t == 0)
    {
        mt->count = a.count;
        mt->sweeps = a.sweeps;
        mt->omega = a.omega;
    }
}

// tail arrays staged in LDS at the word offsets of the slot table
struct TailLdsAcc
{
    const float* arena;
    const TailMeta& mt;
    __device__ __forceinline__ int count() const { return mt.count; }
    __device__ __forceinline__ int n(int m) const { return mt.n[m]; }
    __device__ __forceinline__ int sweeps() const { return mt.sweeps; }
    __device__ __forceinline__ float omega() const { return mt.omega; }
    __device__ __forceinline__ float dot(const float* v, const int* c, const float* x, int k0, int k1) const
    {
        return rowDotLds8(v, c, x, k0, k1);
    }
    template <typename T>
    __device__ __forceinline__ const T* get(int m, int k) const
    {
        // every array the cycle reads is staged; the coarsest inverse exists only where it was formed
        const int w = mt.slot[m * kTailArrays + k];
        return w >= 0 ? reinterpret_cast<const T*>(arena + w) : nullptr;
    }
};

__global__ __launch_bounds__(kTailThreads) void k_mg_tail(TailArgs a, const float* __restrict__ bIn,
                                                          float* __restrict__ eOut, const int* done)
{
    __shared__ float v[kTailLds];
    __shared__ float dl[kDense * kDense];
    if (done && *done) return;
    const int t = threadIdx.x;
    const int last = a.count - 1;
    for (int i = t; i < a.L[0].n; i += kTailThreads) v[i] = bIn[i];
    // the coarsest inverse in LDS in one coalesced pass
    if (a.L[last].dense)
    {
        const int nn = a.L[last].n * a.L[last].n;
        for (int k = t; k < nn; k += kTailThreads) dl[k] = a.L[last].dense[k];
    }
    __syncthreads();
    tailCycle(TailGlobalAcc{a, dl}, v, eOut);
}

// Copies the staged tail arrays into LDS.  The chunk table is read into LDS
// first (one round trip); then the float4 slots of all chunks form one index
// space, eight slots in flight per lane (a chunk's last slot may be partial
// and is read word by word).
__device__ __forceinline__ void stageTail(const TailStage& st, float* arena, int* cs4, int* cw, int* cd,
                                          const float** csrc)
{
    const int t = threadIdx.x, nt = blockDim.x;
    if (t <= st.chunks) cs4[t] = st.start4[t];
    if (t < st.chunks)
    {
        cw[t] = st.words[t];
        cd[t] = st.dst[t];
        csrc[t] = static_cast<const float*>(st.src[t]);
    }
    __syncthreads();
    int c = 0;
    for (int base = t; base < st.total4; base += 8 * nt)
    {
        float4 val[8];
        int dst[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            const int s = base + u * nt;
            dst[u] = -1;
            if (s < st.total4)
            {
                while (s >= cs4[c + 1]) ++c;
                const int j = s - cs4[c];
                const int w = cw[c];
                const float* src = csrc[c];
                if (4 * j + 3 < w)
                    val[u] = reinterpret_cast<const float4*>(src)[j];
                else
                {
                    val[u].x = src[4 * j];
                    val[u].y = 4 * j + 1 < w ? src[4 * j + 1] : 0.0f;
                    val[u].z = 4 * j + 2 < w ? src[4 * j + 2] : 0.0f;
                    val[u].w = 0.0f;
                }
                dst[u] = cd[c] + 4 * j;
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (dst[u] >= 0) *reinterpret_cast<float4*>(arena + dst[u]) = val[u];
    }
}

// sum over k in [k0, k1) of val[k] * x[col[k]] in slot order (rowDot8's
// sums), the loads of 32 nonzeros in flight together; unit: val is all ones
// (the product is then x itself, exactly)
template <bool UNIT>
__device__ __forceinline__ float rowDot32(const float* __restrict__ val, const int* __restrict__ col,
                                          const float* __restrict__ x, int k0, int k1)
{
    float s = 0.0f;
    for (int k = k0; k < k1; k += 32)
    {
        float v[32], xv[32];
        int c[32];
#pragma unroll
        for (int u = 0; u < 32; ++u)
        {
            const bool ok = k + u < k1;
            c[u] = ok ? col[k + u] : 0;
            v[u] = UNIT ? 1.0f : (ok ? val[k + u] : 0.0f);
        }
#pragma unroll
        for (int u = 0; u < 32; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < 32; ++u)
        {
            const float t = s + (UNIT ? xv[u] : v[u] * xv[u]);
            s = k + u < k1 ? t : s;
        }
    }
    return s;
}

// Copies two arrays (wa and wb words) into LDS as one index space of float4
// slots, eight in flight per lane; a last partial slot is read word by word.
__device__ __forceinline__ void copyTwo(float* da, const float* __restrict__ sa, int wa, float* db,
                                        const float* __restrict__ sb, int wb)
{
    const int nt = blockDim.x;
    const int a4 = (wa + 3) / 4, total = a4 + (wb + 3) / 4;
    for (int base = threadIdx.x; base < total; base += 8 * nt)
    {
        float4 v[8];
        float* d[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            const int s = base + u * nt;
            d[u] = nullptr;
            if (s < total)
            {
                const bool first = s < a4;
                const int j = first ? s : s - a4;
                const int w = first ? wa : wb;
                const float* src = first ? sa : sb;
                if (4 * j + 3 < w)
                    v[u] = reinterpret_cast<const float4*>(src)[j];
                else
                {
                    v[u].x = src[4 * j];
                    v[u].y = 4 * j + 1 < w ? src[4 * j + 1] : 0.0f;
                    v[u].z = 4 * j + 2 < w ? src[4 * j + 2] : 0.0f;
                    v[u].w = 0.0f;
                }
                d[u] = (first ? da : db) + 4 * j;
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (d[u]) *reinterpret_cast<float4*>(d[u]) = v[u];
    }
}

// The tail fused into the post-smoothing of the level above it (level l):
// every workgroup copies l's residual and the transfer's member list into
// LDS, restricts into the tail, stages the tail's arrays over them, runs the
// tail cycle on its own from LDS, forms the prolongated iterate
// xp = x + S P e of all of l's rows in LDS, then post-smooths four row blocks
// (one per 256 threads) the CSR-stream way: products staged in LDS from
// coalesced value and column loads, each row summed in slot order.  Every
// value is the float expression of k_mg_restrict_csr, the tail,
// k_mg_prolong and k_mg_up in their summation orders, so z is bit for bit
// theirs; the launches and the tail's chains of dependent global loads go.
constexpr int kTailUpThreads = 1024;
constexpr int kTailUpBlocks = kTailUpThreads / B;  // row blocks per workgroup
constexpr int kXpAhead = 16;                        // prolongation rows per thread loaded at kernel entry
__global__ __launch_bounds__(kTailUpThreads) void k_mg_tail_up(
    CsrView A, const float* __restrict__ val32, const float* __restrict__ x, const float* __restrict__ b,
    const float* __restrict__ wD, const int* __restrict__ agg, const float* __restrict__ sc,
    const float* __restrict__ r, const int* __restrict__ rRowPtr, const int* __restrict__ rCol,
    const float* __restrict__ rval, int unitRestrict, int restricted, TailArgs ta, TailStage st, int rOff, int cOff,
    int xpOff, int prodOff, float* __restrict__ z, const int* done)
{
    __shared__ float arena[kTailUpArena];
    __shared__ int cs4[kTailChunks + 1], cw[kTailChunks], cd[kTailChunks];
    __shared__ const float* csrc[kTailChunks];
    __shared__ TailMeta mt;
    if (done && *done) return;
    const int t = threadIdx.x;
    const int n = A.n, n0 = ta.L[0].n;
    // loads that depend on nothing the tail computes, issued first so that they arrive while the tail's
    // arrays are staged and its cycle runs: the prolongation's x, S and aggregate of the first
    // kXpAhead * 1024 rows, and this workgroup's post-smoothing row block
    float xv[kXpAhead], sv[kXpAhead];
    int av[kXpAhead];
#pragma unroll
    for (int u = 0; u < kXpAhead; ++u)
    {
        const int j = t + u * kTailUpThreads;
        const bool ok = j < n;
        xv[u] = ok ? x[j] : 0.0f;
        sv[u] = ok ? sc[j] : 0.0f;
        av[u] = ok ? agg[j] : 0;
    }
    const int q = t / B, lt = t % B;
    const int blk = blockIdx.x * kTailUpBlocks + q;
    int r0 = 0, r1 = 0, k0 = 0, nnz = 0;
    if (blk < A.nRowBlocks)
    {
        r0 = A.rowBlocks[blk];
        r1 = A.rowBlocks[blk + 1];
        k0 = A.rowPtr[r0];
        nnz = A.rowPtr[r1] - k0;
    }
    loadTailMeta(ta, st, &mt);
    // restricted: the tail's right-hand side was formed by l's pre-smoothing and is a staged chunk
    if (!restricted)
    {
        // l's residual and the members of every coarse row
        float* rl = arena + rOff;
        int* cl = reinterpret_cast<int*>(arena + cOff);
        copyTwo(rl, r, n, reinterpret_cast<float*>(cl), reinterpret_cast<const float*>(rCol), n);
        __syncthreads();
        // b of the tail's first level = P^T (T r), every row, members in ascending order (k_mg_restrict_csr's sums)
        for (int c = t; c < n0; c += kTailUpThreads)
        {
            const int k0 = rRowPtr[c], k1 = rRowPtr[c + 1];
            float sum = 0.0f;
            for (int k = k0; k < k1; k += 8)
            {
                int m[8];
                float v[8], w[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                {
                    m[u] = k + u < k1 ? cl[k + u] : 0;
                    w[u] = unitRestrict || k + u >= k1 ? 1.0f : rval[k + u];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = rl[m[u]];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                {
                    const float t2 = sum + (unitRestrict ? v[u] : w[u] * v[u]);
                    sum = k + u < k1 ? t2 : sum;
                }
            }
            arena[c] = sum;
        }
        __syncthreads();
    }
    stageTail(st, arena, cs4, cw, cd, csrc);
    __syncthreads();
    float* e = arena + 2 * n0;  // the tail's level-0 residual slot, free once restricted
    tailCycle(TailLdsAcc{arena, mt}, arena, e);
    // xp of every row of level l, over the staged arrays (no longer read)
    float* xp = arena + xpOff;
#pragma unroll
    for (int u = 0; u < kXpAhead; ++u)
    {
        const int j = t + u * kTailUpThreads;
        if (j < n) xp[j] = xv[u] + sv[u] * e[av[u]];
    }
    for (int j0 = t + kXpAhead * kTailUpThreads; j0 < n; j0 += 16 * kTailUpThreads)
    {
        float xv[16], sv[16];
        int av[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
        {
            const int j = j0 + u * kTailUpThreads;
            const bool ok = j < n;
            xv[u] = ok ? x[j] : 0.0f;
            sv[u] = ok ? sc[j] : 0.0f;
            av[u] = ok ? agg[j] : 0;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u)
        {
            const int j = j0 + u * kTailUpThreads;
            if (j < n) xp[j] = xv[u] + sv[u] * e[av[u]];
        }
    }
    __syncthreads();
    // post-smoothing of this workgroup's row blocks, one per 256 threads
    float* lq = arena + prodOff + q * kSpmvLds;
    const bool staged = nnz <= kSpmvLds;
    if (staged)
    {
        int k = lt;
        for (; k + 3 * B < nnz; k += 4 * B)
        {
            float v[4];
            int c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
            {
                v[u] = val32[k0 + k + u * B];
                c[u] = A.col[k0 + k + u * B];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) lq[k + u * B] = v[u] * xp[c[u]];
        }
        for (; k < nnz; k += B) lq[k] = val32[k0 + k] * xp[A.col[k0 + k]];
    }
    __syncthreads();
    const int row = r0 + lt;
    if (row < r1)
    {
        float s = 0.0f;
        if (staged)
            for (int k = A.rowPtr[row] - k0, ke = A.rowPtr[row + 1] - k0; k < ke; ++k) s += lq[k];
        else
            for (int k = A.rowPtr[row]; k < A.rowPtr[row + 1]; ++k) s += val32[k] * xp[A.col[k]];
        z[row] = xp[row] + wD[row] * (b[row] - s);
    }
}

// ----------------------------------------------------------------- K-cycle
// deterministic all-workgroup sum of G <= 2048 partials (blockDim B)
__device__ __forceinline__ double redAll(const double* __restrict__ p, int G, double* lds)
{
    double t[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
        const int i = threadIdx.x + j * B;
        t[j] = i < G ? p[i] : 0.0;
    }
    double v = ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
    v = blockSum(v, lds);
    if (threadIdx.x == 0) lds[B / kWave] = v;
    __syncthreads();
    v = lds[B / kWave];
    __syncthreads();
    return v;
}

// v1 = A c1; partials c1.v1 (P0), c1.b (P1)
__global__ __launch_bounds__(B) void k_kc_spmv1(CsrView A, int rbPerWg, const float* __restrict__ val,
                                                const float* __restrict__ c, const float* __restrict__ b,
                                                float* __restrict__ v, double* __restrict__ part, const int* done)
{
    __shared__ float lds[kSpmvLds];
    __shared__ double red[B / kWave];
    if (done && *done) return;
    const int w = xcdRemap(blockIdx.x, gridDim.x);
    double p0 = 0.0, p1 = 0.0;
    for (int j = 0; j < rbPerWg; ++j)
    {
        const int blk = w * rbPerWg + j;
        if (blk >= A.nRowBlocks) break;
        int row;
        float sum;
        if (rowSum<kSpmvLds>(A, val, c, blk, lds, row, sum))
        {
            v[row] = sum;
            const double ci = c[row];
            p0 += ci * static_cast<double>(sum);
            p1 += ci * static_cast<double>(b[row]);
        }
    }
    p0 = blockSum(p0, red);
    p1 = blockSum(p1, red);
    if (threadIdx.x == 0)
    {
        part[blockIdx.x] = p0;
        part[Multigrid::kPartMax + blockIdx.x] = p1;
    }
}

// r2 = b - (c1.b / c1.v1) v1
__global__ __launch_bounds__(B) void k_kc_axpy(int n, const float* __restrict__ b, const float* __restrict__ v,
                                               float* __restrict__ r2, const double* __restrict__ part, int G,
                                               const int* done)
{
    __shared__ double red[B / kWave + 1];
    if (done && *done) return;
    const double rho = redAll(part, G, red);
    const double a1 = redAll(part + Multigrid::kPartMax, G, red);
    const float coef = static_cast<float>(rho > 0.0 ? a1 / rho : 0.0);
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) r2[i] = b[i] - coef * v[i];
}

// v2 = A c2 (not stored); partials c2.v1 (P2), c2.v2 (P3), c2.r2 (P4)
__global__ __launch_bounds__(B) void k_kc_spmv2(CsrView A, int rbPerWg, const float* __restrict__ val,
                                                const float* __restrict__ c, const float* __restrict__ v1,
                                                const float* __restrict__ r2, double* __restrict__ part,
                                                const int* done)
{
    __shared__ float lds[kSpmvLds];
    __shared__ double red[B / kWave];
    if (done && *done) return;
    const int w = xcdRemap(blockIdx.x, gridDim.x);
    double p2 = 0.0, p3 = 0.0, p4 = 0.0;
    for (int j = 0; j < rbPerWg; ++j)
    {
        const int blk = w * rbPerWg + j;
        if (blk >= A.nRowBlocks) break;
        int row;
        float sum;
        if (rowSum<kSpmvLds>(A, val, c, blk, lds, row, sum))
        {
            const double ci = c[row];
            p2 += ci * static_cast<double>(v1[row]);
            p3 += ci * static_cast<double>(sum);
            p4 += ci * static_cast<double>(r2[row]);
        }
    }
    p2 = blockSum(p2, red);
    p3 = blockSum(p3, red);
    p4 = blockSum(p4, red);
    if (threadIdx.x == 0)
    {
        part[2 * Multigrid::kPartMax + blockIdx.x] = p2;
        part[3 * Multigrid::kPartMax + blockIdx.x] = p3;
        part[4 * Multigrid::kPartMax + blockIdx.x] = p4;
    }
}

// e = k1 c1 + k2 c2 with the two-step FCG coefficients
//   rho1 = c1.v1, a1 = c1.b, g = c2.v1, beta = c2.v2, a2 = c2.r2,
//   rho2 = beta - g^2 / rho1, k1 = a1/rho1 - g a2 / (rho1 rho2), k2 = a2 / rho2
// (one step, k2 = 0, if rho2 <= 0; e = c1 if rho1 <= 0)
__device__ __forceinline__ void kcCoefficients(const double* part, int G, double* red, float& k1, float& k2)
{
    const double rho1 = redAll(part, G, red);
    const double a1 = redAll(part + Multigrid::kPartMax, G, red);
    const double g = redAll(part + 2 * Multigrid::kPartMax, G, red);
    const double be = redAll(part + 3 * Multigrid::kPartMax, G, red);
    const double a2 = redAll(part + 4 * Multigrid::kPartMax, G, red);
    double c1 = 1.0, c2 = 0.0;
    if (rho1 > 0.0)
    {
        const double rho2 = be - g * g / rho1;
        if (rho2 > 0.0)
        {
            c1 = a1 / rho1 - g * a2 / (rho1 * rho2);
            c2 = a2 / rho2;
        }
        else
            c1 = a1 / rho1;
    }
    k1 = static_cast<float>(c1);
    k2 = static_cast<float>(c2);
}

__global__ __launch_bounds__(B) void k_kc_combine(int n, const float* __restrict__ c1, const float* __restrict__ c2,
                                                  float* __restrict__ e, const double* __restrict__ part, int G,
                                                  const int* done)
{
    __shared__ double red[B / kWave + 1];
    if (done && *done) return;
    float k1, k2;
    kcCoefficients(part, G, red, k1, k2);
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) e[i] = k1 * c1[i] + k2 * c2[i];
}

// Pre-smoothing of a K-cycle's second visit with its right-hand side formed
// where it is read: r2 = b - coef v1, coef = c1.b / c1.v1 from the partials
// (every workgroup reduces them in k_kc_axpy's order), so x = W r2 and
// r = r2 - (A W) r2 are k_kc_axpy + k_mg_down bit for bit; r2 of the
// workgroup's rows is stored for the rest of the visit.
__global__ __launch_bounds__(B) void k_mg_down_kc(CsrView A, const float* __restrict__ valS,
                                                  const float* __restrict__ b, const float* __restrict__ v1,
                                                  const double* __restrict__ part, int G,
                                                  const float* __restrict__ wD, float* __restrict__ r2,
                                                  float* __restrict__ x, float* __restrict__ r,
                                                  const int* __restrict__ agg, const int* __restrict__ aStart,
                                                  float* __restrict__ bc, const int* done)
{
    __shared__ float lds[kSpmvLds];
    __shared__ double red[B / kWave + 1];
    __shared__ float rl[B];
    if (done && *done) return;
    BlockAggregates g;
    if (bc) g = blockAggregates(A, xcdRemap(blockIdx.x, gridDim.x), agg, aStart);
    const double rho = redAll(part, G, red);
    const double a1 = redAll(part + Multigrid::kPartMax, G, red);
    const float coef = static_cast<float>(rho > 0.0 ? a1 / rho : 0.0);
    const int blk = xcdRemap(blockIdx.x, gridDim.x);
    const int r0 = A.rowBlocks[blk], r1 = A.rowBlocks[blk + 1];
    const int k0 = A.rowPtr[r0], nnz = A.rowPtr[r1] - k0;
    const int row = r0 + static_cast<int>(threadIdx.x);
    float s = 0.0f;
    if (nnz <= kSpmvLds)
    {
        int k = threadIdx.x;
        for (; k + 3 * B < nnz; k += 4 * B)
        {
            float v[4], bb[4], vv[4];
            int c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
            {
                v[u] = valS[k0 + k + u * B];
                c[u] = A.col[k0 + k + u * B];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
            {
                bb[u] = b[c[u]];
                vv[u] = v1[c[u]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) lds[k + u * B] = v[u] * (bb[u] - coef * vv[u]);
        }
        for (; k < nnz; k += B)
        {
            const int c = A.col[k0 + k];
            lds[k] = valS[k0 + k] * (b[c] - coef * v1[c]);
        }
        __syncthreads();
        if (row < r1)
            for (int k = A.rowPtr[row] - k0, e = A.rowPtr[row + 1] - k0; k < e; ++k) s += lds[k];
    }
    else if (row < r1)
    {
        for (int k = A.rowPtr[row]; k < A.rowPtr[row + 1]; ++k)
        {
            const int c = A.col[k];
            s += valS[k] * (b[c] - coef * v1[c]);
        }
    }
    float ri = 0.0f;
    if (row < r1)
    {
        const float bi = b[row] - coef * v1[row];
        r2[row] = bi;
        x[row] = wD[row] * bi;
        ri = bi - s;
        r[row] = ri;
    }
    if (bc) restrictBlock(g, ri, row < r1, rl, bc);
}

// Prolongation of a deferred K-cycle correction: x += S P (k1 c1 + k2 c2),
// the coefficients reduced by every workgroup (k_kc_combine + k_mg_prolong
// bit for bit).
__global__ __launch_bounds__(B) void k_mg_prolong_kc(int n, const int* __restrict__ agg,
                                                     const float* __restrict__ sc, const float* __restrict__ c1,
                                                     const float* __restrict__ c2, const double* __restrict__ kcPart,
                                                     int kcG, float* __restrict__ x, const int* done)
{
    __shared__ double red[B / kWave + 1];
    if (done && *done) return;
    float k1, k2;
    kcCoefficients(kcPart, kcG, red, k1, k2);
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B)
    {
        const int a = agg[i];
        const float e = k1 * c1[a] + k2 * c2[a];
        x[i] += sc[i] * e;
    }
}

// Prolongation fused into the post-smoothing of a coarse level: with
// xp = x + S P e, z = xp + W (b - A xp), xp formed where it is read.  When the
// coarse correction was a K-cycle its combination e = k1 c1 + k2 c2 is formed
// here too (coefficients from its partials, as k_kc_combine).  Each value is
// the same float expression as in k_mg_prolong / k_kc_combine, so z is bit
// for bit that of the separate kernels.
__global__ __launch_bounds__(B) void k_mg_prolong_up(CsrView A, int rbPerWg, const float* __restrict__ val32,
                                                     const float* __restrict__ x, const float* __restrict__ b,
                                                     const float* __restrict__ wD, const int* __restrict__ agg,
                                                     const float* __restrict__ sc, const float* __restrict__ ec,
                                                     const float* __restrict__ c1, const float* __restrict__ c2,
                                                     const double* __restrict__ kcPart, int kcG,
                                                     float* __restrict__ z, const int* done)
{
    __shared__ float lds[kSpmvLds];
    __shared__ double red[B / kWave + 1];
    if (done && *done) return;
    float k1 = 0.0f, k2 = 0.0f;
    if (kcPart) kcCoefficients(kcPart, kcG, red, k1, k2);
    auto xp = [&](int j) {
        const int a = agg[j];
        const float e = kcPart ? k1 * c1[a] + k2 * c2[a] : ec[a];
        return x[j] + sc[j] * e;
    };
    const int w = xcdRemap(blockIdx.x, gridDim.x);
    for (int jb = 0; jb < rbPerWg; ++jb)
    {
        const int blk = w * rbPerWg + jb;
        if (blk >= A.nRowBlocks) break;
        const int r0 = A.rowBlocks[blk], r1 = A.rowBlocks[blk + 1];
        const int k0 = A.rowPtr[r0], nnz = A.rowPtr[r1] - k0;
        const int row = r0 + static_cast<int>(threadIdx.x);
        float s = 0.0f;
        if (nnz <= kSpmvLds)
        {
            for (int k = threadIdx.x; k < nnz; k += B) lds[k] = val32[k0 + k] * xp(A.col[k0 + k]);
            __syncthreads();
            if (row < r1)
                for (int k = A.rowPtr[row] - k0, e = A.rowPtr[row + 1] - k0; k < e; ++k) s += lds[k];
            __syncthreads();
        }
        else if (row < r1)
        {
            for (int k = A.rowPtr[row]; k < A.rowPtr[row + 1]; ++k) s += val32[k] * xp(A.col[k]);
        }
        if (row < r1) z[row] = xp(row) + wD[row] * (b[row] - s);
    }
}


// ------------------------------------------------------ training objective
__global__ __launch_bounds__(B) void k_spmv_f(CsrView A, const float* __restrict__ val, const float* __restrict__ x,
                                              float* __restrict__ y)
{
    __shared__ float lds[kSpmvLds];
    int row;
    float s;
    if (rowSum<kSpmvLds>(A, val, x, xcdRemap(blockIdx.x, gridDim.x), lds, row, s)) y[row] = s;
}

__global__ void k_dot_part(int n, const float* __restrict__ x, const float* __restrict__ y, double* __restrict__ part)
{
    __shared__ double red[B / kWave];
    double v = 0.0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B)
        v += static_cast<double>(x[i]) * static_cast<double>(y[i]);
    v = blockSum(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void k_absmax_part(int n, const float* __restrict__ x, double* __restrict__ part)
{
    __shared__ double red[B / kWave];
    double v = 0.0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) v = fmax(v, fabs(static_cast<double>(x[i])));
    v = blockMax(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// out[0] = sum (mode 0) or max (mode 1) of G <= 2048 partials
__global__ void k_final(int G, const double* __restrict__ part, double* __restrict__ out, int mode)
{
    __shared__ double red[B / kWave + 1];
    if (mode == 0)
    {
        const double v = redAll(part, G, red);
        if (threadIdx.x == 0) *out = v;
    }
    else
    {
        double v = 0.0;
        for (int i = threadIdx.x; i < G; i += B) v = fmax(v, part[i]);
        v = blockMax(v, red);
        if (threadIdx.x == 0) *out = v;
    }
}

__global__ void k_lg_start(int n, const float* __restrict__ invD, const float* __restrict__ a,
                           const float* __restrict__ b, float* __restrict__ db, float* __restrict__ x1)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const float d = invD[i] * b[i];
    db[i] = d;
    x1[i] = a[i] * d;
}

__global__ void k_lg_sub(int n, const float* __restrict__ b, const float* __restrict__ ax, float* __restrict__ r)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) r[i] = b[i] - ax[i];
}

__global__ void k_lg_gather(int n, const int* __restrict__ agg, const float* __restrict__ ec, float* __restrict__ u)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) u[i] = ec[agg[i]];
}

__global__ void k_lg_x2(int n, const float* __restrict__ x1, const float* __restrict__ t, const float* __restrict__ u,
                        float* __restrict__ x2)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) x2[i] = x1[i] + t[i] * u[i];
}

// pd = D^-1 (b - A x2), z = x2 + a pd, eps = e - z
__global__ void k_lg_z(int n, const float* __restrict__ b, const float* __restrict__ ax2,
                       const float* __restrict__ invD, const float* __restrict__ a, const float* __restrict__ x2,
                       const float* __restrict__ e, float* __restrict__ pd, float* __restrict__ eps)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const float p = invD[i] * (b[i] - ax2[i]);
    pd[i] = p;
    eps[i] = e[i] - (x2[i] + a[i] * p);
}

// L = 1 - p^2 / (c q) with z = e - eps, c = e.b, p = z.b = c - d, q = z.Az =
// c - 2 d + f (d = eps.b, f = eps.A eps, A symmetric):
// gz = dL/dz = (2 p / (c q)) (-b + (p / q) A z), A z = b - A eps;
// then ga += gz pd; w1 = D^-1 (a gz)
__global__ void k_lg_g(int n, const float* __restrict__ b, const float* __restrict__ aeps,
                       const double* __restrict__ sc, const float* __restrict__ a, const float* __restrict__ invD,
                       const float* __restrict__ pd, float* __restrict__ gz, float* __restrict__ w1,
                       float* __restrict__ ga)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double c = sc[0], d = sc[1], f = sc[2];
    const double p = c - d, q = c - 2.0 * d + f;
    const double k1 = 2.0 * p / (c * q), k2 = p / q;
    const double bi = b[i];
    const float g = static_cast<float>(k1 * (-bi + k2 * (bi - static_cast<double>(aeps[i]))));
    gz[i] = g;
    ga[i] += g * pd[i];
    w1[i] = invD[i] * (a[i] * g);
}

// gx2 = gz - A w1; gt += gx2 u
__global__ void k_lg_gx2(int n, const float* __restrict__ gz, const float* __restrict__ aw1,
                         const float* __restrict__ u, float* __restrict__ gx2, float* __restrict__ gt)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const float g = gz[i] - aw1[i];
    gx2[i] = g;
    gt[i] += g * u[i];
}

// gt += r1 py; tp = t py
__global__ void k_lg_tail(int n, const float* __restrict__ r1, const float* __restrict__ py,
                          const float* __restrict__ t, float* __restrict__ gt, float* __restrict__ tp)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    gt[i] += r1[i] * py[i];
    tp[i] = t[i] * py[i];
}

// ga += (gx2 - A tp) D^-1 b
__global__ void k_lg_end(int n, const float* __restrict__ gx2, const float* __restrict__ atp,
                         const float* __restrict__ db, float* __restrict__ ga)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) ga[i] += (gx2[i] - atp[i]) * db[i];
}

// one smoothing sweep of the objective's cycle: p = D^-1 (b - A x), x' = x + a p (x' may be x); p is kept
// for the adjoint
__global__ void k_lg_sweep(int n, const float* __restrict__ b, const float* __restrict__ ax,
                           const float* __restrict__ invD, const float* __restrict__ a, const float* x, float* xo,
                           float* __restrict__ pj)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const float p = invD[i] * (b[i] - ax[i]);
    pj[i] = p;
    xo[i] = x[i] + a[i] * p;
}

// the adjoint of that sweep, first half: ga += g p, w = D^-1 (a g) (the gradient passed back is g - A w)
__global__ void k_lg_back(int n, const float* __restrict__ g, const float* __restrict__ pj,
                          const float* __restrict__ invD, const float* __restrict__ a, float* __restrict__ ga,
                          float* __restrict__ w)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    ga[i] += g[i] * pj[i];
    w[i] = invD[i] * (a[i] * g[i]);
}

__global__ void k_lg_isub(int n, float* __restrict__ e, const float* __restrict__ z)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) e[i] -= z[i];
}

__global__ void k_scale_by(int n, float* __restrict__ x, const double* __restrict__ m)
{
    const int i = blockIdx.x * B + threadIdx.x;
    const double s = *m;
    if (i < n && s > 0.0) x[i] = static_cast<float>(x[i] / s);
}

} // namespace

Multigrid::Multigrid(Context& ctx, const DevPattern& fine, int maxLevels, int coarseMax) : ctx_(ctx), fine_(fine)
{
    static std::atomic<std::uintptr_t> counter{0};
    id_ = ++counter;
    hipStream_t s = ctx.stream();
    coarseMax = std::min(coarseMax, kCoarseMax);
    // 1. the aggregation of every level, each coarse level numbered in the order its aggregates were made
    //    (RP[l], CL[l]: the pattern of level l + 1; level 0 is the fine pattern itself)
    std::vector<std::vector<int>> RP, CL, AG;
    auto rpOf = [&](int l) -> const std::vector<int>& { return l == 0 ? fine.hRowPtr : RP[l - 1]; };
    auto clOf = [&](int l) -> const std::vector<int>& { return l == 0 ? fine.hCol : CL[l - 1]; };
    while (true)
    {
        const int l = static_cast<int>(RP.size());
        const int n = static_cast<int>(rpOf(l).size()) - 1;
        if (n <= coarseMax || l + 1 >= maxLevels) break;
        std::vector<int> agg;
        const int na = aggregate(rpOf(l), clOf(l), agg);
        if (na >= n) break;  // no coarsening possible
        std::vector<int> crp, ccol;
        coarsePattern(rpOf(l), clOf(l), agg, na, crp, ccol, nullptr, nullptr);
        AG.push_back(std::move(agg));
        RP.push_back(std::move(crp));
        CL.push_back(std::move(ccol));
    }
    const int last = static_cast<int>(RP.size());
    // 2. every coarse level above the coarsest renumbered aggregate by aggregate (rows sorted by their
    //    aggregate, then by their first number), from the top down so that each level is sorted by the
    //    final numbers of the level above it: the same hierarchy, whose restriction from such a level sums
    //    contiguous rows and whose aggregates follow its rows in order
    for (int l = last - 1; l >= 1; --l)
    {
        const int n = static_cast<int>(AG[l].size()), na = static_cast<int>(RP[l].size()) - 1;
        std::vector<int> start(na + 1, 0), pi(n);
        for (int i = 0; i < n; ++i) ++start[AG[l][i] + 1];
        for (int c = 0; c < na; ++c) start[c + 1] += start[c];
        for (int i = 0; i < n; ++i) pi[i] = start[AG[l][i]]++;
        for (int& a : AG[l - 1]) a = pi[a];
        std::vector<int> moved(n);
        for (int i = 0; i < n; ++i) moved[pi[i]] = AG[l][i];
        AG[l].swap(moved);
    }
    // 3. patterns, transfers and Galerkin maps in the final numbering (rp, cl: level l's, after level 0)
    std::vector<int> rp, cl;
    lv_.emplace_back();
    lv_[0].n = fine.n;
    lv_[0].nnz = fine.nnz;
    for (int l = 0; l < last; ++l)
    {
        MgLevel& L = lv_.back();
        const int n = L.n;
        const std::vector<int>& agg = AG[l];
        const int na = static_cast<int>(RP[l].size()) - 1;
        const std::vector<int>& lrp = l == 0 ? fine.hRowPtr : rp;
        const std::vector<int>& lcl = l == 0 ? fine.hCol : cl;
        // members of each aggregate
        std::vector<int> aStart(na + 1, 0), aMem(n);
        for (int i = 0; i < n; ++i) ++aStart[agg[i] + 1];
        for (int c = 0; c < na; ++c) aStart[c + 1] += aStart[c];
        {
            std::vector<int> fill(aStart.begin(), aStart.end() - 1);
            for (int i = 0; i < n; ++i) aMem[fill[agg[i]]++] = i;
        }
        std::vector<int> crp, ccol, gStart, gPerm;
        coarsePattern(lrp, lcl, agg, na, crp, ccol, &gStart, &gPerm);
        L.agg = upload(agg, s);
        L.aggStart = upload(aStart, s);
        L.aggMem = upload(aMem, s);
        L.rpat.build(aStart, aMem, s);
        L.rval.alloc(sizeof(float) * n);
        L.gStart = upload(gStart, s);
        L.gPerm = upload(gPerm, s);
        L.hAgg = agg;
        if (l > 0)
        {
            // this coarse level's rows and row blocks: numbered aggregate by aggregate, blocks of whole
            // aggregates (its restriction then runs inside its pre-smoothing)
            std::vector<int> blocks = buildRowBlocksGrouped(rp, agg);
            L.groupedRows = !blocks.empty();
            L.pat.build(rp, cl, s, std::move(blocks));
        }
        MgLevel C;
        C.n = na;
        C.nnz = crp[na];
        lv_.push_back(std::move(C));
        rp.swap(crp);
        cl.swap(ccol);
    }
    if (last > 0) lv_.back().pat.build(rp, cl, s);
    if (lv_.size() > 1 && lv_.back().n > kCoarseMax)
        throw std::invalid_argument("multigrid: coarsest level exceeds 2048 rows (raise maxLevels)");
    // the coarse tail: the first level from which every level fits the tail's LDS
    for (int l = 1; l < last && tailFrom_ < 0; ++l)
    {
        long long rows = 0;
        for (int m = l; m <= last; ++m) rows += lv_[m].n;
        if (last - l + 1 <= kTailLevels && 3 * rows <= kTailLds) tailFrom_ = l;
    }
    for (std::size_t l = 0; l < lv_.size(); ++l)
    {
        MgLevel& L = lv_[l];
        const std::size_t n = L.n, nnz = L.nnz;
        if (l > 0) L.val64.alloc(sizeof(double) * nnz);
        L.val32.alloc(sizeof(float) * nnz);
        L.valS32.alloc(sizeof(float) * nnz);
        for (DevBuf* b : {&L.invD32, &L.wPre32, &L.wPost32, &L.s32, &L.t32, &L.b, &L.x, &L.r, &L.t2, &L.e})
            b->alloc(sizeof(float) * n);
        const int nrb = l == 0 ? fine_.nRowBlocks : L.pat.nRowBlocks;
        L.G = std::max(1, std::min(nrb, kPartMax));
        L.rbPerWg = divUp(nrb, L.G);
        L.G = std::max(1, divUp(nrb, L.rbPerWg));
        if (l > 0 && l + 1 < lv_.size())
        {
            // K-cycle buffers on every intermediate coarse level (the depth is a parameter)
            for (DevBuf* b : {&L.c1, &L.c2, &L.v1, &L.r2}) b->alloc(sizeof(float) * n);
            L.part.alloc(sizeof(double) * 5 * kPartMax);
        }
        if (l > 0 && l + 1 == lv_.size() && L.n <= kDenseMax) L.dense.alloc(sizeof(float) * n * n);
    }
    buildSlices();
    ctx.sync();
}

void Multigrid::buildSlices()
{
    const std::vector<int>& rp = fine_.hRowPtr;
    const std::vector<int>& cl = fine_.hCol;
    const int nrb = fine_.nRowBlocks;
    if (nrb < 1) return;
    std::vector<int> blocks(static_cast<std::size_t>(nrb) + 1);
    CFX_HIP(hipMemcpy(blocks.data(), fine_.rowBlocks.p, sizeof(int) * blocks.size(), hipMemcpyDeviceToHost));
    // slice widths: the longest row of each 64-row slice of each row block
    std::vector<int> slicePtr{0}, blockSlice{0};
    long long slots = 0;
    for (int b = 0; b < nrb; ++b)
    {
        for (int q = blocks[b]; q < blocks[b + 1]; q += kWave)
        {
            int w = 0;
            for (int r = q; r < std::min(q + kWave, blocks[b + 1]); ++r) w = std::max(w, rp[r + 1] - rp[r]);
            slots += static_cast<long long>(kWave) * w;
            if (slots > std::numeric_limits<int>::max()) return;
            slicePtr.push_back(static_cast<int>(slots));
        }
        blockSlice.push_back(static_cast<int>(slicePtr.size()) - 1);
    }
    if (4 * slots > 5LL * std::max(1, fine_.nnz)) return;  // padding above a quarter of the nonzeros
    std::vector<short> col(static_cast<std::size_t>(slots), kSlicePad);
    std::vector<int> idx(static_cast<std::size_t>(slots), -1);
    int sl = 0;
    for (int b = 0; b < nrb; ++b)
    {
        const int r0 = blocks[b];
        for (int q = r0; q < blocks[b + 1]; q += kWave, ++sl)
        {
            const int p0 = slicePtr[sl];
            for (int r = q; r < std::min(q + kWave, blocks[b + 1]); ++r)
                for (int k = rp[r]; k < rp[r + 1]; ++k)
                {
                    const int o = cl[k] - r0;
                    if (o < -32767 || o > 32767) return;  // kSlicePad is reserved
                    const std::size_t t = static_cast<std::size_t>(p0) + static_cast<std::size_t>(k - rp[r]) * kWave +
                                          static_cast<std::size_t>(r - q);
                    col[t] = static_cast<short>(o);
                    idx[t] = k;
                }
        }
    }
    // readable slots past the last slice for the row sums' unconditional loads (padding, value 0)
    col.resize(col.size() + kSliceTail, kSlicePad);
    hipStream_t s = ctx_.stream();
    sl_.slicePtr = upload(slicePtr, s);
    sl_.blockSlice = upload(blockSlice, s);
    sl_.col = upload(col, s);
    sl_.idx = upload(idx, s);
    for (DevBuf* v : {&sl_.val, &sl_.valS})
    {
        v->alloc(sizeof(float) * (slots + kSliceTail));
        CFX_HIP(hipMemsetAsync(v->as<float>() + slots, 0, sizeof(float) * kSliceTail, s));
    }
    sl_.val64.alloc(sizeof(double) * (slots + kSliceTail));
    CFX_HIP(hipMemsetAsync(sl_.val64.as<double>() + slots, 0, sizeof(double) * kSliceTail, s));
    sl_.slots = slots;
    sl_.ok = true;
    CFX_HIP(hipStreamSynchronize(s));
}

void Multigrid::setMatrix(const double* fineVal64, const float* pre, const float* scale, const float* post,
                          const float* restr, const MgParams& p)
{
    setValues(fineVal64, p);
    setWeights(pre, scale, post, restr, p);
}

void Multigrid::setValues(const double* fineVal64, const MgParams& p)
{
    hipStream_t s = ctx_.stream();
    if (p.kcycleDepth != p_.kcycleDepth || p.fusion != p_.fusion || p.coarseSweeps != p_.coarseSweeps ||
        p.coarsestOmega != p_.coarsestOmega || p.smoothFine != p_.smoothFine || p.smoothCoarse != p_.smoothCoarse ||
        p.smoothDeep != p_.smoothDeep)
        ++structVersion_;
    p_ = p;
    fineVal64_ = fineVal64;
    for (std::size_t l = 0; l < lv_.size(); ++l)
    {
        MgLevel& L = lv_[l];
        const CsrView A = l == 0 ? fine_.view() : L.pat.view();
        const double* v64 = l == 0 ? fineVal64 : L.val64.as<double>();
        if (l == 0)
        {
            // the FP32 copy of the finest values (its weight-dependent arrays are setWeights')
            k_scale_vals<<<std::max(1, std::min(8192, divUp(L.nnz, B))), B, 0, s>>>(A, v64, L.val32.as<float>(),
                                                                                   nullptr, nullptr);
            if (sliced(0))
            {
                const int g = static_cast<int>(std::min<long long>(8192, divUp(sl_.slots, B)));
                k_slice_vals<<<g, B, 0, s>>>(sl_.slots, sl_.idx.as<int>(), L.val32.as<float>(), sl_.val.as<float>());
                // the FP64 values too, for the flexible CG's product on this level
                k_slice_vals64<<<g, B, 0, s>>>(sl_.slots, sl_.idx.as<int>(), v64, sl_.val64.as<double>());
            }
            continue;
        }
        const MgLevel& F = lv_[l - 1];
        const double* fv = l == 1 ? fineVal64 : F.val64.as<double>();
        // lanes per coarse nonzero from the mean number of fine nonzeros it sums
        const double per = static_cast<double>(F.nnz) / std::max(1, L.nnz);
        if (per >= 48.0)
            k_galerkin<64><<<divUp(64LL * L.nnz, B), B, 0, s>>>(L.nnz, F.gStart.as<int>(), F.gPerm.as<int>(), fv,
                                                               L.val64.as<double>());
        else if (per >= 3.0)
            k_galerkin<8><<<divUp(8LL * L.nnz, B), B, 0, s>>>(L.nnz, F.gStart.as<int>(), F.gPerm.as<int>(), fv,
                                                             L.val64.as<double>());
        else
            k_galerkin<1><<<divUp(L.nnz, B), B, 0, s>>>(L.nnz, F.gStart.as<int>(), F.gPerm.as<int>(), fv,
                                                       L.val64.as<double>());
        // coarse levels take the constants
        k_level_setup<<<divUp(L.n, B), B, 0, s>>>(A, v64, L.invD32.as<float>(), L.wPre32.as<float>(),
                                                  L.wPost32.as<float>(), L.s32.as<float>(), L.t32.as<float>(), nullptr,
                                                  p.coarseOmega, nullptr, p.coarseScale, nullptr, p.coarseOmega,
                                                  nullptr, 1.0f);
        k_scale_vals<<<std::max(1, std::min(8192, divUp(L.nnz, B))), B, 0, s>>>(A, v64, L.val32.as<float>(),
                                                                               L.wPre32.as<float>(),
                                                                               L.valS32.as<float>());
        if (L.dense.p) k_mg_dense_inverse<<<1, 1024, 0, s>>>(A, L.val64.as<double>(), L.dense.as<float>());
        if (L.rval.p)
            k_restrict_vals<<<divUp(L.n, B), B, 0, s>>>(L.n, L.aggMem.as<int>(), nullptr, L.rval.as<float>());
    }
    CFX_CHECK_LAUNCH();
}

void Multigrid::setWeights(const float* pre, const float* scale, const float* post, const float* restr,
                           const MgParams& p)
{
    if (!fineVal64_) throw std::logic_error("multigrid: weights set before the matrix values");
    hipStream_t s = ctx_.stream();
    MgLevel& L = lv_[0];
    const CsrView A = fine_.view();
    k_level_setup<<<divUp(L.n, B), B, 0, s>>>(A, fineVal64_, L.invD32.as<float>(), L.wPre32.as<float>(),
                                              L.wPost32.as<float>(), L.s32.as<float>(), L.t32.as<float>(), pre,
                                              p.omegaPre, scale, p.scale, post, p.omegaPost, restr, p.restrictScale);
    if (sliced(0))
        // the column-scaled values in slot order only (the finest pre-smoothing is their one reader)
        k_slice_scale<<<A.nRowBlocks, B, 0, s>>>(A, sl_.view(), sl_.val.as<float>(), L.wPre32.as<float>(),
                                                 sl_.valS.as<float>());
    else
        k_scale_vals<<<std::max(1, std::min(8192, divUp(L.nnz, B))), B, 0, s>>>(A, fineVal64_, nullptr,
                                                                               L.wPre32.as<float>(),
                                                                               L.valS32.as<float>());
    if (L.rval.p)
        k_restrict_vals<<<divUp(L.n, B), B, 0, s>>>(L.n, L.aggMem.as<int>(), L.t32.as<float>(), L.rval.as<float>());
    CFX_CHECK_LAUNCH();
}

bool Multigrid::tailAt(int l) const
{
    return tailFrom_ > 0 && l >= tailFrom_ && l > p_.kcycleDepth && l < levels() - 1;
}

bool Multigrid::useTail(int l) const { return (p_.fusion & 1) && tailAt(l); }

TailArgs Multigrid::tailArgs(int l) const
{
    const int last = levels() - 1;
    TailArgs a;
    a.count = last - l + 1;
    a.sweeps = p_.coarseSweeps;
    a.omega = p_.coarsestOmega;
    for (int m = 0; m < a.count; ++m)
    {
        const MgLevel& L = lv_[l + m];
        const CsrView A = L.pat.view();
        TailLevel& T = a.L[m];
        T.n = L.n;
        T.rowPtr = A.rowPtr;
        T.col = A.col;
        T.valS = L.valS32.as<float>();
        T.val = L.val32.as<float>();
        T.wPre = L.wPre32.as<float>();
        T.wPost = L.wPost32.as<float>();
        T.sc = L.s32.as<float>();
        T.invD = L.invD32.as<float>();
        if (l + m < last)
        {
            const CsrView R = L.rpat.view();
            T.agg = L.agg.as<int>();
            T.rRowPtr = R.rowPtr;
            T.rCol = R.col;
            T.rval = L.rval.as<float>();
        }
        else
            T.dense = L.dense.p ? L.dense.as<float>() : nullptr;
    }
    return a;
}

void Multigrid::launchTail(int l, const int* done)
{
    k_mg_tail<<<1, kTailThreads, 0, ctx_.stream()>>>(tailArgs(l), lv_[l].b.as<float>(), lv_[l].e.as<float>(), done);
}

bool Multigrid::launchTailUp(int l, const float* in, float* out, const float* res, const double* partZQ,
                             const int* done, bool restricted)
{
    // z.q partials follow the grid of the separate post-smoothing: with them the separate kernels run
    if (!(p_.fusion & 4) || partZQ || !tailAt(l + 1)) return false;
    const MgLevel& L = lv_[l];
    const CsrView A = l == 0 ? fine_.view() : L.pat.view();
    // every workgroup repeats the tail: only where the level takes a single wave of workgroups
    const int G = divUp(A.nRowBlocks, kTailUpBlocks);
    if (G > 256) return false;
    const TailArgs ta = tailArgs(l + 1);
    const int last = ta.count - 1;
    int vec = 0;
    for (int m = 0; m < ta.count; ++m) vec += 3 * ta.L[m].n;
    auto up4 = [](int w) { return (w + 3) / 4 * 4; };
    // LDS: the tail's vectors, then l's residual and member list (restriction), later the tail's staged
    // arrays, later l's prolongated iterate and the products of four row blocks
    const int rOff = up4(vec), cOff = rOff + up4(L.n);
    const int xpOff = up4(vec), prodOff = xpOff + up4(L.n);
    if ((!restricted && cOff + up4(L.n) > kTailUpArena) || prodOff + kTailUpBlocks * kSpmvLds > kTailUpArena)
        return false;
    TailStage st;
    for (auto& row : st.slot)
        for (int& v : row) v = -1;
    int off = up4(vec);
    st.start4[0] = 0;
    if (restricted)
    {
        // the tail's right-hand side, formed by l's pre-smoothing, staged with its arrays: the tail's
        // level-0 b at word 0 (the rest of the vector slots is written by the cycle before it is read)
        const int words4 = (ta.L[0].n + 3) / 4;
        st.src[0] = lv_[l + 1].b.as<float>();
        st.words[0] = ta.L[0].n;
        st.dst[0] = 0;
        st.start4[1] = words4;
        st.chunks = 1;
    }
    for (int m = 0; m < ta.count; ++m)
    {
        const TailLevel& T = ta.L[m];
        const int n = T.n, nnz = lv_[l + 1 + m].nnz, nc = m < last ? ta.L[m + 1].n : 0;
        struct Want
        {
            int k, words;
        };
        std::vector<Want> want;
        if (m < last)
            want = {{kTRowPtr, n + 1}, {kTCol, nnz}, {kTValS, nnz}, {kTWPre, n}, {kTRRowPtr, nc + 1}, {kTRCol, n},
                    {kTRval, n}, {kTSc, n}, {kTAgg, n}, {kTVal, nnz}, {kTWPost, n}};
        else if (T.dense)
            want = {{kTDense, n * n}};
        else
            want = {{kTRowPtr, n + 1}, {kTCol, nnz}, {kTVal, nnz}, {kTInvD, n}};
        for (const Want& w : want)
        {
            // each array is followed by kLdsPad words that the cycle's unconditional LDS loads may read
            const int words4 = (w.words + 3) / 4;
            if (off + 4 * words4 + kLdsPad > kTailUpArena) return false;  // the tail does not fit in LDS
            const int c = st.chunks++;
            st.src[c] = tailGlobal(T, w.k);
            st.words[c] = w.words;
            st.dst[c] = off;
            st.start4[c + 1] = st.start4[c] + words4;
            st.slot[m][w.k] = off;
            off += 4 * words4 + kLdsPad;
        }
    }
    st.total4 = st.start4[st.chunks];
    const CsrView R = L.rpat.view();
    // off the finest level the restriction scale is 1 (k_restrict_vals): its products are exact
    k_mg_tail_up<<<G, kTailUpThreads, 0, ctx_.stream()>>>(
        A, L.val32.as<float>(), L.x.as<float>(), in, L.wPost32.as<float>(), L.agg.as<int>(), L.s32.as<float>(),
        res, R.rowPtr, R.col, L.rval.as<float>(), l > 0 ? 1 : 0, restricted ? 1 : 0, ta, st, rOff, cOff,
        xpOff, prodOff, out, done);
    return true;
}

bool Multigrid::solveLevel(int l, const int* done, bool deferCombine)
{
    hipStream_t s = ctx_.stream();
    MgLevel& L = lv_[l];
    const int last = levels() - 1;
    if (useTail(l))
    {
        launchTail(l, done);
        return false;
    }
    if (l == last)
    {
        if (L.dense.p)
            k_mg_dense_apply<<<1, kDense, 0, s>>>(L.n, L.dense.as<float>(), L.b.as<float>(), L.e.as<float>(), done);
        else
        {
            const CsrView A = L.pat.view();
            k_mg_coarse<<<1, kCoarseThreads, 0, s>>>(L.n, A.rowPtr, A.col, L.val32.as<float>(), L.invD32.as<float>(),
                                                     L.b.as<float>(), L.e.as<float>(), p_.coarseSweeps,
                                                     p_.coarsestOmega, done);
        }
        return false;
    }
    if (l > p_.kcycleDepth)
    {
        visit(l, L.b.as<float>(), L.e.as<float>(), nullptr, nullptr, L.G, L.rbPerWg, done);
        return false;
    }
    const CsrView A = L.pat.view();
    // vector kernels of the K-cycle: about four elements per thread (each workgroup also reduces the
    // level's partials, which costs little next to a thread walking a long stride of elements)
    const int Gv = std::max(1, std::min(1024, divUp(L.n, 4 * B)));
    double* part = L.part.as<double>();
    visit(l, L.b.as<float>(), L.c1.as<float>(), nullptr, nullptr, L.G, L.rbPerWg, done);
    k_kc_spmv1<<<L.G, B, 0, s>>>(A, L.rbPerWg, L.val32.as<float>(), L.c1.as<float>(), L.b.as<float>(),
                                 L.v1.as<float>(), part, done);
    if (p_.fusion & 8)
    {
        // the second visit's pre-smoothing forms r2 = b - coef v1 itself
        visit(l, L.r2.as<float>(), L.c2.as<float>(), nullptr, nullptr, L.G, L.rbPerWg, done, L.b.as<float>(),
              L.v1.as<float>());
    }
    else
    {
        k_kc_axpy<<<Gv, B, 0, s>>>(L.n, L.b.as<float>(), L.v1.as<float>(), L.r2.as<float>(), part, L.G, done);
        visit(l, L.r2.as<float>(), L.c2.as<float>(), nullptr, nullptr, L.G, L.rbPerWg, done);
    }
    k_kc_spmv2<<<L.G, B, 0, s>>>(A, L.rbPerWg, L.val32.as<float>(), L.c2.as<float>(), L.v1.as<float>(),
                                 L.r2.as<float>(), part, done);
    if (deferCombine) return true;  // the caller forms e = k1 c1 + k2 c2 where it reads it
    k_kc_combine<<<Gv, B, 0, s>>>(L.n, L.c1.as<float>(), L.c2.as<float>(), L.e.as<float>(), part, L.G, done);
    return false;
}

void Multigrid::visit(int l, const float* in, float* out, const float* q, double* partZQ, int Gs, int rbPerWg,
                      const int* done, const float* kcB, const float* kcV)
{
    hipStream_t s = ctx_.stream();
    MgLevel& L = lv_[l];
    const CsrView A = l == 0 ? fine_.view() : L.pat.view();
    // smoothing sweeps before and after the coarse correction (the same number): smoothFine on the finest
    // level, smoothCoarse on level 1 and smoothDeep on the further K-cycle levels, one below them
    const int sweeps =
        std::max(1, l == 0 ? p_.smoothFine : (l > p_.kcycleDepth ? 1 : (l == 1 ? p_.smoothCoarse : p_.smoothDeep)));
    // fusion bit 5: a coarse level numbered aggregate by aggregate restricts inside its last pre-smoothing
    // sweep (its restriction weights are 1)
    const bool fuseR = (p_.fusion & 32) && l >= 1 && L.groupedRows && l + 1 < levels();
    float* bc = fuseR ? lv_[l + 1].b.as<float>() : nullptr;
    float* rb[2] = {L.r.as<float>(), L.t2.as<float>()};  // residuals of successive sweeps
    if (kcB)
        k_mg_down_kc<<<A.nRowBlocks, B, 0, s>>>(A, L.valS32.as<float>(), kcB, kcV, L.part.as<double>(), L.G,
                                                L.wPre32.as<float>(), const_cast<float*>(in), L.x.as<float>(), rb[0],
                                                L.agg.as<int>(), L.aggStart.as<int>(), sweeps == 1 ? bc : nullptr,
                                                done);
    else if (sliced(l))
        k_mg_down<true><<<A.nRowBlocks, B, 0, s>>>(A, sl_.view(), sl_.valS.as<float>(), in, L.wPre32.as<float>(),
                                                   L.x.as<float>(), rb[0], L.agg.as<int>(), L.aggStart.as<int>(),
                                                   sweeps == 1 ? bc : nullptr, done);
    else
        k_mg_down<false><<<A.nRowBlocks, B, 0, s>>>(A, SliceView{}, L.valS32.as<float>(), in, L.wPre32.as<float>(),
                                                    L.x.as<float>(), rb[0], L.agg.as<int>(), L.aggStart.as<int>(),
                                                    sweeps == 1 ? bc : nullptr, done);
    int cur = 0;
    for (int k = 1; k < sweeps; ++k, cur ^= 1)
    {
        float* bck = k + 1 == sweeps ? bc : nullptr;
        if (sliced(l))
            k_mg_down_more<true><<<A.nRowBlocks, B, 0, s>>>(A, sl_.view(), sl_.valS.as<float>(), rb[cur],
                                                            L.wPre32.as<float>(), L.x.as<float>(), rb[cur ^ 1],
                                                            L.agg.as<int>(), L.aggStart.as<int>(), bck, done);
        else
            k_mg_down_more<false><<<A.nRowBlocks, B, 0, s>>>(A, SliceView{}, L.valS32.as<float>(), rb[cur],
                                                             L.wPre32.as<float>(), L.x.as<float>(), rb[cur ^ 1],
                                                             L.agg.as<int>(), L.aggStart.as<int>(), bck, done);
    }
    float* res = rb[cur];
    // the first post-smoothing sweep writes z1: out itself, or (further sweeps follow) the residual buffer
    // that the restriction has consumed
    float* z1 = sweeps == 1 ? out : rb[cur ^ 1];
    // the z.q partials belong to the last sweep
    const float* q1 = sweeps == 1 ? q : nullptr;
    double* part1 = sweeps == 1 ? partZQ : nullptr;
    bool first = false;  // the first post-smoothing sweep already ran (fused)
    if (l + 1 < levels())
    {
        if (launchTailUp(l, in, z1, res, part1, done, fuseR))
            first = true;
        else
        {
            MgLevel& C = lv_[l + 1];
            if (!fuseR)
                k_mg_restrict_csr<<<L.rpat.nRowBlocks, B, 0, s>>>(L.rpat.view(), L.rval.as<float>(), res,
                                                                  C.b.as<float>(), done);
            // fused only where the level is small: on a large level forming S P e per nonzero costs more
            // gathers than the separate prolongation saves
            const bool fuse = (p_.fusion & 2) && l >= 1 && L.n <= kFuseRowsMax;
            // a K-cycle below may leave its combination to the prolongation (its coefficients are reduced by
            // every workgroup there, so only for a level of few workgroups)
            const bool deferKc = fuse || ((p_.fusion & 16) && C.G <= kDeferGroupsMax);
            const bool deferred = solveLevel(l + 1, done, deferKc);
            if (fuse)
            {
                k_mg_prolong_up<<<Gs, B, 0, s>>>(A, rbPerWg, L.val32.as<float>(), L.x.as<float>(), in,
                                                 L.wPost32.as<float>(), L.agg.as<int>(), L.s32.as<float>(),
                                                 C.e.as<float>(), C.c1.as<float>(), C.c2.as<float>(),
                                                 deferred ? C.part.as<double>() : nullptr, C.G, z1, done);
                first = true;
            }
            else if (deferred)
                k_mg_prolong_kc<<<std::max(1, std::min(1024, divUp(L.n, B))), B, 0, s>>>(
                    L.n, L.agg.as<int>(), L.s32.as<float>(), C.c1.as<float>(), C.c2.as<float>(),
                    C.part.as<double>(), C.G, L.x.as<float>(), done);
            else
                k_mg_prolong<<<divUp(L.n, B), B, 0, s>>>(L.n, L.agg.as<int>(), L.s32.as<float>(), C.e.as<float>(),
                                                          L.x.as<float>(), done);
        }
    }
    if (!first) launchUp(l, A, Gs, rbPerWg, L.x.as<float>(), in, z1, q1, part1, done);
    // further sweeps z = z' + W (b - A z'), alternating between the two residual buffers, the last into out
    float* zc = z1;
    for (int k = 1; k < sweeps; ++k)
    {
        const bool last = k + 1 == sweeps;
        float* zn = last ? out : (zc == rb[0] ? rb[1] : rb[0]);
        launchUp(l, A, Gs, rbPerWg, zc, in, zn, last ? q : nullptr, last ? partZQ : nullptr, done);
        zc = zn;
    }
}

void Multigrid::launchUp(int l, const CsrView& A, int Gs, int rbPerWg, const float* x, const float* b, float* z,
                         const float* q, double* partZQ, const int* done)
{
    hipStream_t s = ctx_.stream();
    const MgLevel& L = lv_[l];
    if (probeOn_ && l == 0)
    {
        // the command processor records the events at the kernel's own start and end on this stream
        while (static_cast<int>(probeEv_.size()) < 2 * (probeUsed_ + 1))
        {
            hipEvent_t e;
            CFX_HIP(hipEventCreate(&e));
            probeEv_.push_back(e);
        }
        if (sliced(l))
            hipExtLaunchKernelGGL(k_mg_up<true>, dim3(Gs), dim3(B), 0, s, probeEv_[2 * probeUsed_],
                                  probeEv_[2 * probeUsed_ + 1], 0, A, sl_.view(), rbPerWg, sl_.val.as<float>(), x, b,
                                  L.wPost32.as<float>(), z, q, partZQ, done);
        else
            hipExtLaunchKernelGGL(k_mg_up<false>, dim3(Gs), dim3(B), 0, s, probeEv_[2 * probeUsed_],
                                  probeEv_[2 * probeUsed_ + 1], 0, A, SliceView{}, rbPerWg, L.val32.as<float>(), x, b,
                                  L.wPost32.as<float>(), z, q, partZQ, done);
        ++probeUsed_;
        return;
    }
    if (sliced(l))
        k_mg_up<true><<<Gs, B, 0, s>>>(A, sl_.view(), rbPerWg, sl_.val.as<float>(), x, b, L.wPost32.as<float>(), z, q,
                                       partZQ, done);
    else
        k_mg_up<false><<<Gs, B, 0, s>>>(A, SliceView{}, rbPerWg, L.val32.as<float>(), x, b, L.wPost32.as<float>(), z,
                                        q, partZQ, done);
}

void Multigrid::setProbe(bool on)
{
    probeOn_ = on;
    probeUsed_ = 0;
}

bool Multigrid::probeResult(double& avgMs, int& count)
{
    if (probeUsed_ == 0) return false;
    CFX_HIP(hipStreamSynchronize(ctx_.stream()));
    // launches after the stop decision return at once: average over those that ran (> 20 % of the longest)
    std::vector<float> t(probeUsed_);
    float mx = 0;
    for (int i = 0; i < probeUsed_; ++i)
    {
        CFX_HIP(hipEventElapsedTime(&t[i], probeEv_[2 * i], probeEv_[2 * i + 1]));
        mx = std::max(mx, t[i]);
    }
    double sum = 0;
    int cnt = 0;
    for (float v : t)
        if (v > 0.2f * mx)
        {
            sum += v;
            ++cnt;
        }
    avgMs = cnt ? sum / cnt : 0.0;
    count = cnt;
    return cnt > 0;
}

double Multigrid::timeFineSmooth(int reps)
{
    hipStream_t s = ctx_.stream();
    MgLevel& L = lv_[0];
    const CsrView A = fine_.view();
    hipEvent_t e0, e1;
    CFX_HIP(hipEventCreate(&e0));
    CFX_HIP(hipEventCreate(&e1));
    // z = x + W (b - A x) on the level's own vectors (their contents do not matter for the time), in the
    // layout the cycle uses
    for (int i = 0; i <= reps; ++i)
    {
        if (i == 1) CFX_HIP(hipEventRecord(e0, s));
        launchUp(0, A, A.nRowBlocks, 1, L.x.as<float>(), L.b.as<float>(), L.e.as<float>(), nullptr, nullptr, nullptr);
    }
    CFX_HIP(hipEventRecord(e1, s));
    CFX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    CFX_HIP(hipEventElapsedTime(&ms, e0, e1));
    CFX_HIP(hipEventDestroy(e0));
    CFX_HIP(hipEventDestroy(e1));
    return ms / std::max(1, reps);
}

void Multigrid::apply(const float* r, float* z, const float* q, double* partZQ, int Gs, int rbPerWg, const int* done)
{
    visit(0, r, z, q, partZQ, Gs, rbPerWg, done);
    CFX_CHECK_LAUNCH();
}

double Multigrid::lossGrad(const float* e, const float* a, const float* t, float* ga, float* gt)
{
    if (levels() < 2) throw std::invalid_argument("lossGrad: the hierarchy has no coarse level");
    hipStream_t s = ctx_.stream();
    MgLevel& L = lv_[0];
    MgLevel& C = lv_[1];
    const int n = L.n;
    const int k = std::max(1, p_.smoothFine);  // smoothing sweeps on each side
    if (!tv_[0].p)
    {
        for (DevBuf& v : tv_) v.alloc(sizeof(float) * n);
        tpart_.alloc(sizeof(double) * (2 * kPartMax + 16));
    }
    // D^-1 (b - A x) of the pre-smoothing sweeps 2..k and of the post-smoothing sweeps 1..k-1
    while (static_cast<int>(tk_.size()) < 2 * (k - 1))
    {
        tk_.emplace_back();
        tk_.back().alloc(sizeof(float) * n);
    }
    auto pre = [&](int j) { return tk_[j].as<float>(); };           // j = 0 .. k-2
    auto post = [&](int j) { return tk_[k - 1 + j].as<float>(); };  // j = 0 .. k-2
    float *b = tv_[0].as<float>(), *db = tv_[1].as<float>(), *x1 = tv_[2].as<float>(), *ax = tv_[3].as<float>(),
          *r1 = tv_[4].as<float>(), *u = tv_[5].as<float>(), *x2 = tv_[6].as<float>(), *pd = tv_[7].as<float>(),
          *eps = tv_[8].as<float>(), *aeps = tv_[9].as<float>(), *gz = tv_[10].as<float>(),
          *w1 = tv_[11].as<float>(), *gx2 = tv_[12].as<float>(), *py = tv_[13].as<float>();
    double* part = tpart_.as<double>();
    double* sc = part + 2 * kPartMax;  // [0] e.A e, [1] eps.A e, [2] eps.A eps
    const CsrView A = fine_.view();
    const float* v32 = L.val32.as<float>();
    const float* invD = L.invD32.as<float>();
    const int G = divUp(n, B), Gd = std::max(1, std::min(kPartMax, divUp(n, 4 * B)));
    // forward: the cycle applied to b = A e
    k_spmv_f<<<A.nRowBlocks, B, 0, s>>>(A, v32, e, b);
    k_lg_start<<<G, B, 0, s>>>(n, invD, a, b, db, x1);
    for (int j = 0; j + 1 < k; ++j)
    {
        k_spmv_f<<<A.nRowBlocks, B, 0, s>>>(A, v32, x1, ax);
        k_lg_sweep<<<G, B, 0, s>>>(n, b, ax, invD, a, x1, x1, pre(j));
    }
    k_spmv_f<<<A.nRowBlocks, B, 0, s>>>(A, v32, x1, ax);
    k_lg_sub<<<G, B, 0, s>>>(n, b, ax, r1);
    k_mg_restrict_csr<<<L.rpat.nRowBlocks, B, 0, s>>>(L.rpat.view(), L.rval.as<float>(), r1, C.b.as<float>(), nullptr);
    solveLevel(1, nullptr);
    k_lg_gather<<<G, B, 0, s>>>(n, L.agg.as<int>(), C.e.as<float>(), u);
    k_lg_x2<<<G, B, 0, s>>>(n, x1, t, u, x2);
    for (int j = 0; j + 1 < k; ++j)
    {
        k_spmv_f<<<A.nRowBlocks, B, 0, s>>>(A, v32, x2, ax);
        k_lg_sweep<<<G, B, 0, s>>>(n, b, ax, invD, a, x2, x2, post(j));
    }
    k_spmv_f<<<A.nRowBlocks, B, 0, s>>>(A, v32, x2, ax);
    k_lg_z<<<G, B, 0, s>>>(n, b, ax, invD, a, x2, e, pd, eps);
    k_spmv_f<<<A.nRowBlocks, B, 0, s>>>(A, v32, eps, aeps);
    k_dot_part<<<Gd, B, 0, s>>>(n, e, b, part);
    k_final<<<1, B, 0, s>>>(Gd, part, sc, 0);
    k_dot_part<<<Gd, B, 0, s>>>(n, eps, b, part + kPartMax);
    k_final<<<1, B, 0, s>>>(Gd, part + kPartMax, sc + 1, 0);
    k_dot_part<<<Gd, B, 0, s>>>(n, eps, aeps, part);
    k_final<<<1, B, 0, s>>>(Gd, part, sc + 2, 0);
    // adjoint: back through the post-smoothing sweeps (the last in k_lg_g), each passing g - A D^-1 (a g)
    k_lg_g<<<G, B, 0, s>>>(n, b, aeps, sc, a, invD, pd, gz, w1, ga);
    for (int j = k - 2; j >= 0; --j)
    {
        k_spmv_f<<<A.nRowBlocks, B, 0, s>>>(A, v32, w1, ax);
        k_lg_isub<<<G, B, 0, s>>>(n, gz, ax);
        k_lg_back<<<G, B, 0, s>>>(n, gz, post(j), invD, a, ga, w1);
    }
    k_spmv_f<<<A.nRowBlocks, B, 0, s>>>(A, v32, w1, ax);
    k_lg_gx2<<<G, B, 0, s>>>(n, gz, ax, u, gx2, gt);
    // the coarse correction (symmetric) and the restriction
    k_mg_restrict_csr<<<L.rpat.nRowBlocks, B, 0, s>>>(L.rpat.view(), L.rval.as<float>(), gx2, C.b.as<float>(), nullptr);
    solveLevel(1, nullptr);
    k_lg_gather<<<G, B, 0, s>>>(n, L.agg.as<int>(), C.e.as<float>(), py);
    k_lg_tail<<<G, B, 0, s>>>(n, r1, py, t, gt, x1);
    k_spmv_f<<<A.nRowBlocks, B, 0, s>>>(A, v32, x1, ax);
    // back through the pre-smoothing sweeps k..2, then x1 = a D^-1 b
    if (k > 1) k_lg_isub<<<G, B, 0, s>>>(n, gx2, ax);
    for (int j = k - 2; j >= 0; --j)
    {
        k_lg_back<<<G, B, 0, s>>>(n, gx2, pre(j), invD, a, ga, w1);
        k_spmv_f<<<A.nRowBlocks, B, 0, s>>>(A, v32, w1, ax);
        if (j > 0) k_lg_isub<<<G, B, 0, s>>>(n, gx2, ax);
    }
    k_lg_end<<<G, B, 0, s>>>(n, gx2, ax, db, ga);
    CFX_CHECK_LAUNCH();
    double h[3];
    CFX_HIP(hipMemcpyAsync(h, sc, sizeof(h), hipMemcpyDeviceToHost, s));
    CFX_HIP(hipStreamSynchronize(s));
    const double pz = h[0] - h[1], qz = h[0] - 2.0 * h[1] + h[2];
    return 1.0 - pz * pz / (h[0] * qz);
}

void Multigrid::smoothSample(float* e, int sweeps)
{
    hipStream_t s = ctx_.stream();
    MgLevel& L = lv_[0];
    const int n = L.n;
    if (!tv_[0].p)
    {
        for (DevBuf& v : tv_) v.alloc(sizeof(float) * n);
        tpart_.alloc(sizeof(double) * (2 * kPartMax + 16));
    }
    const CsrView A = fine_.view();
    float *b = tv_[0].as<float>(), *z = tv_[1].as<float>();
    double* part = tpart_.as<double>();
    const int G = divUp(n, B), Gd = std::max(1, std::min(kPartMax, divUp(n, 4 * B)));
    for (int k = 0; k < sweeps; ++k)
    {
        k_spmv_f<<<A.nRowBlocks, B, 0, s>>>(A, L.val32.as<float>(), e, b);
        visit(0, b, z, nullptr, nullptr, L.G, L.rbPerWg, nullptr);
        k_lg_isub<<<G, B, 0, s>>>(n, e, z);
    }
    k_absmax_part<<<Gd, B, 0, s>>>(n, e, part);
    k_final<<<1, B, 0, s>>>(Gd, part, part + 2 * kPartMax + 2, 1);
    k_scale_by<<<G, B, 0, s>>>(n, e, part + 2 * kPartMax + 2);
    CFX_CHECK_LAUNCH();
}

} // namespace cfx
