// pcg.hip -- Jacobi-PCG with three kernels per iteration and no atomics.
//
// Iteration k runs
//   A(k): q = A p (CSR-stream: each row block's products staged in LDS, each
//         row summed in column order), partial p.q per workgroup
//   B(k): reduce the p.q partials -- every workgroup redundantly, in one
//         fixed order, so all agree bit for bit -- alpha = rz / p.q,
//         x += alpha p, r -= alpha q, partials r.(invD r) and max|r|
//   C(k): reduce those partials the same way, take the stop decision,
//         beta = rz_k / rz_{k-1}, p = invD r + beta p
// Workgroup 0 records the scalars for the host, which only polls `done`
// between batches of iterations.  Scalars that a kernel reads while its own
// workgroup 0 writes the next value are double-buffered by iteration parity.
//
// Mixed precision (reliable updates): the Krylov iteration runs in FP32 on
// an FP32 copy of the matrix and accumulates a correction x32; whenever the
// FP32 recurrence residual has dropped by `delta` below the last true
// residual (or claims convergence) the iteration pauses, the host folds x32
// into the FP64 solution, the true residual r = b - A x is recomputed in
// FP64 and replaces the FP32 residual, and the iteration resumes with the
// same search direction.  Convergence is only ever declared on a true FP64
// residual.
#include "batch.h"
#include "krylov.h"
#include "mg.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>

namespace cfx
{
namespace
{

constexpr int kRedMax = 2048;  // partial count cap (grids are capped to this)
constexpr int kB = 256;

struct CgState
{
    double rz[2];        // r.z after iteration k (slot k & 1)
    double bnorm, initialResidual, rnorm, alpha, pAp, updAbs, beta;
    double updFast;      // flexible CG, automatic factor: the early reliable-update threshold
    int iter, done, converged, breakdown, paused, nan;
    int kcur;            // flexible CG: iteration index of the running step (set by B, read by C)
    int kRef;            // flexible CG: iteration of the last true residual
};

// Deterministic all-workgroup reduction of G <= 2048 partials (blockDim 256):
// 8 independent loads per thread, a fixed tree, broadcast through LDS.
__device__ __forceinline__ double allSum(const double* __restrict__ p, int G, double* lds)
{
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
        const int i = threadIdx.x + j * kB;
        v[j] = i < G ? p[i] : 0.0;
    }
    double s = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
    s = blockSum(s, lds);
    if (threadIdx.x == 0) lds[0] = s;
    __syncthreads();
    s = lds[0];
    __syncthreads();
    return s;
}

// max over partials, where a partial of -1 marks a NaN in that workgroup
__device__ __forceinline__ double allMaxNan(const double* __restrict__ p, int G, double* lds)
{
    double m = 0.0, mn = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
        const int i = threadIdx.x + j * kB;
        const double v = i < G ? p[i] : 0.0;
        m = fmax(m, v);
        mn = fmin(mn, v);
    }
    m = blockMax(m, lds);
    const double anyNan = blockMax(-mn, lds);  // 1 if some partial was -1
    if (threadIdx.x == 0) lds[0] = anyNan > 0 ? NAN : m;
    __syncthreads();
    m = lds[0];
    __syncthreads();
    return m;
}

// ------------------------------------------------------------------ kernels
// r = b - A x (FP64); rT = (T) r; partials: max|r| (NaN -> -1), max|b|,
// rT.(invD rT).  Also used after a reliable update.
template <typename T>
__device__ __forceinline__ void body_pcg_residual(CsrView A, int rbPerWg, const double* __restrict__ val,
                                                     const double* __restrict__ x, const double* __restrict__ b,
                                                     double* __restrict__ r64, T* __restrict__ rT,
                                                     const T* __restrict__ invD, double* __restrict__ partRZ,
                                                     double* __restrict__ partRM, double* __restrict__ partBM, int bid, int G)
{
    __shared__ double lds[kSpmvLds];
    __shared__ double red[kB / kWave];
    const int w = xcdRemap(bid, G);
    double rm = 0, bm = 0, rz = 0;
    int nan = 0;
    for (int j = 0; j < rbPerWg; ++j)
    {
        const int blk = w * rbPerWg + j;
        if (blk >= A.nRowBlocks) break;
        const int r0 = A.rowBlocks[blk], r1 = A.rowBlocks[blk + 1];
        const int k0 = A.rowPtr[r0], nnz = A.rowPtr[r1] - k0;
        const int row = r0 + static_cast<int>(threadIdx.x);
        double s = 0;
        if (nnz <= kSpmvLds)
        {
            stageProducts(lds, val, A, x, k0, r0, nnz);
            __syncthreads();
            if (row < r1)
                for (int k = A.rowPtr[row] - k0, e = A.rowPtr[row + 1] - k0; k < e; ++k) s += lds[k];
            __syncthreads();
        }
        else if (row < r1)
        {
            for (int k = A.rowPtr[row]; k < A.rowPtr[row + 1]; ++k) s += val[k] * x[A.col[k]];
        }
        if (row < r1)
        {
            const double bb = b[row];
            const double rr = bb - s;
            if (r64) r64[row] = rr;
            const T rt = static_cast<T>(rr);
            rT[row] = rt;
            const T z = invD[row] * rt;
            rz += static_cast<double>(rt) * static_cast<double>(z);
            rm = fmax(rm, fabs(rr));
            bm = fmax(bm, fabs(bb));
            nan |= (rr != rr) | ((bb != bb) << 1);
        }
    }
    rz = blockSum(rz, red);
    rm = blockMax(rm, red);
    bm = blockMax(bm, red);
    __shared__ int ired[kB / kWave];
    nan = blockOr(nan, ired);
    if (threadIdx.x == 0)
    {
        partRZ[bid] = rz;
        partRM[bid] = (nan & 1) ? -1.0 : rm;
        partBM[bid] = (nan & 2) ? -1.0 : bm;
    }
}
template <typename T>
__global__ __launch_bounds__(kB) void k_pcg_residual(CsrView A, int rbPerWg, const double* __restrict__ val,
                                                     const double* __restrict__ x, const double* __restrict__ b,
                                                     double* __restrict__ r64, T* __restrict__ rT,
                                                     const T* __restrict__ invD, double* __restrict__ partRZ,
                                                     double* __restrict__ partRM, double* __restrict__ partBM)
{
    body_pcg_residual<T>(A, rbPerWg, val, x, b, r64, rT, invD, partRZ, partRM, partBM, blockIdx.x, gridDim.x);
}

// A(k): q = A p (CSR-stream, LDS-staged), partial p.q.
template <typename T>
__device__ __forceinline__ void body_pcg_A(CsrView A, int rbPerWg, const T* __restrict__ val,
                                              const T* __restrict__ p, T* __restrict__ q, const CgState* st,
                                              double* __restrict__ partPQ, int bid, int G)
{
    __shared__ T lds[kSpmvLds];
    __shared__ double red[kB / kWave];
    if (st->done) return;
    const int w = xcdRemap(bid, G);
    double pq = 0;
    for (int j = 0; j < rbPerWg; ++j)
    {
        const int blk = w * rbPerWg + j;
        if (blk >= A.nRowBlocks) break;
        const int r0 = A.rowBlocks[blk], r1 = A.rowBlocks[blk + 1];
        const int k0 = A.rowPtr[r0], nnz = A.rowPtr[r1] - k0;
        const int row = r0 + static_cast<int>(threadIdx.x);
        T s = T(0);
        if (nnz <= kSpmvLds)
        {
            stageProducts(lds, val, A, p, k0, r0, nnz);
            __syncthreads();
            if (row < r1)
                for (int kk = A.rowPtr[row] - k0, e = A.rowPtr[row + 1] - k0; kk < e; ++kk) s += lds[kk];
            __syncthreads();
        }
        else if (row < r1)
        {
            for (int kk = A.rowPtr[row]; kk < A.rowPtr[row + 1]; ++kk) s += val[kk] * p[A.col[kk]];
        }
        if (row < r1)
        {
            q[row] = s;
            pq += static_cast<double>(p[row]) * static_cast<double>(s);
        }
    }
    pq = blockSum(pq, red);
    if (threadIdx.x == 0) partPQ[bid] = pq;
}
template <typename T>
__global__ __launch_bounds__(kB) void k_pcg_A(CsrView A, int rbPerWg, const T* __restrict__ val,
                                              const T* __restrict__ p, T* __restrict__ q, const CgState* st,
                                              double* __restrict__ partPQ)
{
    body_pcg_A<T>(A, rbPerWg, val, p, q, st, partPQ, blockIdx.x, gridDim.x);
}

// B(k): reduce p.q, alpha = rz / p.q; x += alpha p, r -= alpha q;
// partials r.(invD r), max|r|   (solveCG :189-211)
template <typename T>
__device__ __forceinline__ void body_pcg_B(int n, const T* __restrict__ invD, T* __restrict__ x,
                                              T* __restrict__ r, const T* __restrict__ p,
                                              const T* __restrict__ q, CgState* st,
                                              const double* __restrict__ partPQ, int Gs, double* __restrict__ partRZ,
                                              double* __restrict__ partRM, int k, int bid, int G)
{
    __shared__ double red[kB / kWave];
    __shared__ int ired[kB / kWave];
    if (st->done) return;
    const double PAP = allSum(partPQ, Gs, red);
    const double rz = st->rz[(k - 1) & 1];
    if (PAP <= 0 || fabs(rz) < 1e-300)
    {
        // matrix not sufficiently SPD (solveCG breaks here, Matrix.cpp:191-194)
        if (bid == 0 && threadIdx.x == 0) { st->breakdown = 1; st->done = 1; st->pAp = PAP; }
        return;
    }
    const double alphaD = rz / PAP;
    if (bid == 0 && threadIdx.x == 0) { st->alpha = alphaD; st->pAp = PAP; }
    const T alpha = static_cast<T>(alphaD);
    double rzp = 0, rm = 0;
    int nan = 0;
    for (int i = bid * kB + threadIdx.x; i < n; i += G * kB)
    {
        x[i] += alpha * p[i];
        const T ri = r[i] - alpha * q[i];
        r[i] = ri;
        const T z = invD[i] * ri;
        rzp += static_cast<double>(ri) * static_cast<double>(z);
        rm = fmax(rm, fabs(static_cast<double>(ri)));
        nan |= (ri != ri);
    }
    rzp = blockSum(rzp, red);
    rm = blockMax(rm, red);
    nan = blockOr(nan, ired);
    if (threadIdx.x == 0)
    {
        partRZ[bid] = rzp;
        partRM[bid] = nan ? -1.0 : rm;
    }
}
template <typename T>
__global__ __launch_bounds__(kB) void k_pcg_B(int n, const T* __restrict__ invD, T* __restrict__ x,
                                              T* __restrict__ r, const T* __restrict__ p,
                                              const T* __restrict__ q, CgState* st,
                                              const double* __restrict__ partPQ, int Gs, double* __restrict__ partRZ,
                                              double* __restrict__ partRM, int k)
{
    body_pcg_B<T>(n, invD, x, r, p, q, st, partPQ, Gs, partRZ, partRM, k, blockIdx.x, gridDim.x);
}

// C(k): stop decision on the residual of iteration k (k = 0: the initial
// residual), beta = rz_k / rz_{k-1}, p = invD r + beta p  (solveCG :202-217).
// flags bit0: the partials hold a TRUE residual (after a residual kernel);
//       bit1: mixed precision (pause instead of converging on sloppy residuals)
template <typename T>
__device__ __forceinline__ void body_pcg_C(int n, const T* __restrict__ invD, const T* __restrict__ r,
                                              T* __restrict__ p, CgState* st, const double* __restrict__ partRZ,
                                              const double* __restrict__ partRM, const double* __restrict__ partBM,
                                              int nPart, int k, double tol, int maxIter, double delta, int flags,
                                              double* resHist, int bid, int G)
{
    __shared__ double red[kB / kWave];
    if (st->done) return;
    const bool lead = bid == 0 && threadIdx.x == 0;
    const double RZ = allSum(partRZ, nPart, red);
    const double RM = allMaxNan(partRM, nPart, red);
    const bool trueRes = flags & 1, mixed = flags & 2;
    if (k == 0)
    {
        const double BM = allMaxNan(partBM, nPart, red);
        const double bnorm = (BM < 1e-300) ? 1e-300 : BM;  // std::max(||b||inf, 1e-300)
        if (lead)
        {
            st->bnorm = bnorm;
            st->initialResidual = RM;
            st->rnorm = RM;
            st->iter = 0;
            if (resHist) resHist[0] = RM;
        }
        if (RM / bnorm < tol)
        {
            if (lead) { st->converged = 1; st->done = 1; }
            return;
        }
        if (maxIter <= 0)
        {
            if (lead) st->done = 1;
            return;
        }
    }
    else
    {
        const double bnorm = st->bnorm;
        if (lead)
        {
            st->iter = k;
            st->rnorm = RM;
            if (resHist) resHist[k] = RM;
        }
        const bool conv = RM / bnorm < tol;
        if (conv && (!mixed || trueRes))
        {
            if (lead) { st->converged = 1; st->done = 1; }
            return;
        }
        if (k >= maxIter)
        {
            if (lead) st->done = 1;
            return;
        }
        if (mixed && !trueRes && (conv || RM < st->updAbs))
        {
            if (lead) { st->paused = 1; st->done = 1; }
            return;
        }
    }
    if (mixed && trueRes && lead) st->updAbs = delta * RM;
    const double beta = k == 0 ? 0.0 : RZ / st->rz[(k - 1) & 1];
    if (lead)
    {
        st->rz[k & 1] = RZ;
        st->beta = beta;
    }
    const T tb = static_cast<T>(beta);
    for (int i = bid * kB + threadIdx.x; i < n; i += G * kB)
    {
        const T z = invD[i] * r[i];
        p[i] = k == 0 ? z : z + tb * p[i];
    }
}
template <typename T>
__global__ __launch_bounds__(kB) void k_pcg_C(int n, const T* __restrict__ invD, const T* __restrict__ r,
                                              T* __restrict__ p, CgState* st, const double* __restrict__ partRZ,
                                              const double* __restrict__ partRM, const double* __restrict__ partBM,
                                              int nPart, int k, double tol, int maxIter, double delta, int flags,
                                              double* resHist)
{
    body_pcg_C<T>(n, invD, r, p, st, partRZ, partRM, partBM, nPart, k, tol, maxIter, delta, flags, resHist, blockIdx.x, gridDim.x);
}

// reliable update: x64 += x32, x32 = 0
__global__ void k_fold(int n, double* __restrict__ x, float* __restrict__ x32)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        x[i] += static_cast<double>(x32[i]);
        x32[i] = 0.0f;
    }
}

__global__ void k_resume(CgState* st)
{
    st->done = 0;
    st->paused = 0;
}

__global__ void k_inv_diag32(CsrView A, const double* __restrict__ val, float* __restrict__ invD32,
                             double* __restrict__ invD64)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const int k = A.diagSlot[i];
    const double d = k >= 0 ? val[k] : 0.0;
    const double inv = (fabs(d) > 1e-300) ? 1.0 / d : 1.0;
    invD64[i] = inv;
    invD32[i] = static_cast<float>(inv);
}


// ------------------------------------------------------- flexible CG (FP32)
// FCG A(k): q = A p with the FP64 matrix values (FP32 values would perturb
// the near-null modes of a (nearly) singular pressure matrix by more than
// their eigenvalues), q rounded to FP32; partials p.q and p.r.  kSliced: the
// values in the matrix's 64-row slices (S), the same sums bit for bit
template <bool kSliced>
__global__ __launch_bounds__(kB) void k_fcg_A(CsrView A, SliceView S, int rbPerWg, const double* __restrict__ val,
                                              const float* __restrict__ p, const float* __restrict__ r,
                                              float* __restrict__ q, const CgState* st, double* __restrict__ partPQ,
                                              double* __restrict__ partPR)
{
    __shared__ double lds[kSliced ? 1 : kSpmvLds];
    __shared__ double red[kB / kWave];
    if (st->done) return;
    const int w = xcdRemap(blockIdx.x, gridDim.x);
    double pq = 0, pr = 0;
    for (int j = 0; j < rbPerWg; ++j)
    {
        const int blk = w * rbPerWg + j;
        if (blk >= A.nRowBlocks) break;
        const int r0 = A.rowBlocks[blk], r1 = A.rowBlocks[blk + 1];
        int row = r0 + static_cast<int>(threadIdx.x);
        double s = 0.0;
        // the row's own operands, loaded before its sum (in flight with the value loads)
        const float pOwn = row < r1 ? p[row] : 0.0f, rOwn = row < r1 ? r[row] : 0.0f;
        if constexpr (kSliced)
            sliceRowSum<double>(A, S, val, p, blk, row, s);
        else
        {
            const int k0 = A.rowPtr[r0], nnz = A.rowPtr[r1] - k0;
            if (nnz <= kSpmvLds)
            {
                stageProducts(lds, val, A, p, k0, r0, nnz);
                __syncthreads();
                if (row < r1)
                    for (int kk = A.rowPtr[row] - k0, e = A.rowPtr[row + 1] - k0; kk < e; ++kk) s += lds[kk];
                __syncthreads();
            }
            else if (row < r1)
            {
                for (int kk = A.rowPtr[row]; kk < A.rowPtr[row + 1]; ++kk)
                    s += val[kk] * static_cast<double>(p[A.col[kk]]);
            }
        }
        if (row < r1)
        {
            const float qs = static_cast<float>(s);
            q[row] = qs;
            const double pi = pOwn;
            pq += pi * static_cast<double>(qs);
            pr += pi * static_cast<double>(rOwn);
        }
    }
    pq = blockSum(pq, red);
    pr = blockSum(pr, red);
    if (threadIdx.x == 0)
    {
        partPQ[blockIdx.x] = pq;
        partPR[blockIdx.x] = pr;
    }
}

// FCG B(k): alpha = p.r / p.q; x += alpha p, r -= alpha q; partial max|r|
__global__ __launch_bounds__(kB) void k_fcg_B(int n, float* __restrict__ x, float* __restrict__ r,
                                              const float* __restrict__ p, const float* __restrict__ q, CgState* st,
                                              const double* __restrict__ partPQ, const double* __restrict__ partPR,
                                              int Gs, double* __restrict__ partRM)
{
    __shared__ double red[kB / kWave];
    __shared__ int ired[kB / kWave];
    if (st->done) return;
    const double PQ = allSum(partPQ, Gs, red);
    const double PR = allSum(partPR, Gs, red);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (lead) st->kcur = st->iter + 1;
    if (!(PQ > 0) || !(fabs(PR) < INFINITY))
    {
        // A p.p <= 0 (not SPD) or a non-finite value from the preconditioner
        if (lead)
        {
            st->breakdown = PQ <= 0;
            st->nan = !(PQ <= 0);
            st->done = 1;
            st->pAp = PQ;
        }
        return;
    }
    const double alphaD = PR / PQ;
    if (lead)
    {
        st->alpha = alphaD;
        st->pAp = PQ;
    }
    const float alpha = static_cast<float>(alphaD);
    double rm = 0;
    int nan = 0;
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB)
    {
        x[i] += alpha * p[i];
        const float ri = r[i] - alpha * q[i];
        r[i] = ri;
        rm = fmax(rm, fabs(static_cast<double>(ri)));
        nan |= (ri != ri);
    }
    rm = blockMax(rm, red);
    nan = blockOr(nan, ired);
    if (threadIdx.x == 0) partRM[blockIdx.x] = nan ? -1.0 : rm;
}

// FCG C(k): stop decision on max|r_k| (k = 0: initial residual),
// beta = -z.q / p.q, p = z + beta p.  flags as in k_pcg_C (bit0 true residual,
// bit1 mixed).  kArg < 0 takes k from the state (set by B), so that one
// captured iteration can be replayed.
__global__ __launch_bounds__(kB) void k_fcg_C(int n, const float* __restrict__ z, float* __restrict__ p, CgState* st,
                                              const double* __restrict__ partZQ, int nZQ,
                                              const double* __restrict__ partRM, const double* __restrict__ partBM,
                                              int nRM, int kArg, double tol, int maxIter, double delta, int flags,
                                              double* resHist)
{
    __shared__ double red[kB / kWave];
    if (st->done) return;
    const int k = kArg >= 0 ? kArg : st->kcur;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const double RM = allMaxNan(partRM, nRM, red);
    const double ZQ = k == 0 ? 0.0 : allSum(partZQ, nZQ, red);
    const bool trueRes = flags & 1;
    if (k == 0)
    {
        const double BM = allMaxNan(partBM, nRM, red);
        const double bnorm = (BM < 1e-300) ? 1e-300 : BM;
        if (lead)
        {
            st->bnorm = bnorm;
            st->initialResidual = RM;
            st->rnorm = RM;
            st->iter = 0;
            if (resHist) resHist[0] = RM;
        }
        if (RM != RM || BM != BM)
        {
            if (lead) { st->nan = 1; st->done = 1; }
            return;
        }
        if (RM / bnorm < tol)
        {
            if (lead) { st->converged = 1; st->done = 1; }
            return;
        }
        if (maxIter <= 0)
        {
            if (lead) st->done = 1;
            return;
        }
    }
    else
    {
        if (lead)
        {
            st->iter = k;
            st->rnorm = RM;
            if (resHist) resHist[k] = RM;
        }
        if (RM != RM || ZQ != ZQ || !(fabs(ZQ) < INFINITY))
        {
            if (lead) { st->nan = 1; st->done = 1; }
            return;
        }
        const bool conv = RM / st->bnorm < tol;
        if (conv && trueRes)
        {
            if (lead) { st->converged = 1; st->done = 1; }
            return;
        }
        if (k >= maxIter)
        {
            if (lead) st->done = 1;
            return;
        }
        // automatic factor (delta < 0): the update comes after a drop of kFcgDelta, or after one of
        // kPcgDelta once kFcgSegment iterations have passed, whichever is first
        const bool due = delta > 0 ? RM < st->updAbs
                                   : (RM < st->updFast || (RM < st->updAbs && k - st->kRef >= kFcgSegment));
        if (!trueRes && (conv || due))
        {
            if (lead) { st->paused = 1; st->done = 1; }
            return;
        }
    }
    if (trueRes && lead)
    {
        st->updAbs = (delta > 0 ? delta : kPcgDelta) * RM;
        st->updFast = kFcgDelta * RM;
        st->kRef = k;
    }
    const double beta = k == 0 ? 0.0 : -ZQ / st->pAp;
    if (lead) st->beta = beta;
    const float tb = static_cast<float>(beta);
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB)
        p[i] = k == 0 ? z[i] : z[i] + tb * p[i];
}


// ------------------------------------------------------ batched Jacobi-PCG
// Per system: its matrix, vectors, state and partial arrays; workgroup
// tables map a launch's workgroup to (system, local index) for the SpMV
// grid (Gs per system) and for the vector grid (Gv per system).
struct BatchSysDev
{
    CsrView A;
    const double* val;
    const double* invD;
    const double* b;
    double *x, *r, *p, *q;
    CgState* st;
    double* part;  // RZ, RM, BM, PQ: kRedMax each
    int n, Gs, rbPerWg, Gv;
};

__global__ __launch_bounds__(kB) void k_bpcg_residual(const BatchSysDev* __restrict__ S, const int2* __restrict__ tab)
{
    const int2 t = tab[blockIdx.x];
    const BatchSysDev s = S[t.x];
    body_pcg_residual<double>(s.A, s.rbPerWg, s.val, s.x, s.b, nullptr, s.r, s.invD, s.part, s.part + kRedMax,
                              s.part + 2 * kRedMax, t.y, s.Gs);
}

__global__ __launch_bounds__(kB) void k_bpcg_C0(const BatchSysDev* __restrict__ S, const int2* __restrict__ tab,
                                                double tol, int maxIter)
{
    const int2 t = tab[blockIdx.x];
    const BatchSysDev s = S[t.x];
    body_pcg_C<double>(s.n, s.invD, s.r, s.p, s.st, s.part, s.part + kRedMax, s.part + 2 * kRedMax, s.Gs, 0, tol,
                       maxIter, 0.0, 0, nullptr, t.y, s.Gv);
}

__global__ __launch_bounds__(kB) void k_bpcg_A(const BatchSysDev* __restrict__ S, const int2* __restrict__ tab)
{
    const int2 t = tab[blockIdx.x];
    const BatchSysDev s = S[t.x];
    body_pcg_A<double>(s.A, s.rbPerWg, s.val, s.p, s.q, s.st, s.part + 3 * kRedMax, t.y, s.Gs);
}

// iteration index from the system's state: C(k-1) recorded iter = k-1
__global__ __launch_bounds__(kB) void k_bpcg_B(const BatchSysDev* __restrict__ S, const int2* __restrict__ tab)
{
    const int2 t = tab[blockIdx.x];
    const BatchSysDev s = S[t.x];
    if (s.st->done) return;
    const int k = s.st->iter + 1;
    if (t.y == 0 && threadIdx.x == 0) s.st->kcur = k;
    body_pcg_B<double>(s.n, s.invD, s.x, s.r, s.p, s.q, s.st, s.part + 3 * kRedMax, s.Gs, s.part, s.part + kRedMax, k,
                       t.y, s.Gv);
}

__global__ __launch_bounds__(kB) void k_bpcg_C(const BatchSysDev* __restrict__ S, const int2* __restrict__ tab,
                                               double tol, int maxIter)
{
    const int2 t = tab[blockIdx.x];
    const BatchSysDev s = S[t.x];
    if (s.st->done) return;
    body_pcg_C<double>(s.n, s.invD, s.r, s.p, s.st, s.part, s.part + kRedMax, s.part + 2 * kRedMax, s.Gv,
                       s.st->kcur, tol, maxIter, 0.0, 0, nullptr, t.y, s.Gv);
}

// ------------------------------------------- persistent batched Jacobi-PCG
// The whole batch in one launch.  Each system walks the steps residual, C(0),
// then A, B, C per iteration; step `st` of system s consists of the chunks
// that the launches above would give their workgroups (Gs for residual and
// A, Gv for C(0), B and C), and chunk c runs exactly the body workgroup c
// runs there, so every partial and every iterate is the same.  Workgroups
// are not tied to systems: wave 0 of each workgroup looks at every system's
// control word (step << 32 | chunks claimed), claims a run of chunks (the
// system's claim size) of a step that has some left with one atomic add,
// runs them, and counts them finished; the workgroup that finishes a step's last chunk moves the system
// to its next step (or retires it once C has set `done`).  Release/acquire
// fences at agent scope order each step's results before the next step's
// claims across the XCDs' L2s.  No workgroup ever waits for another one, so
// the launch needs no co-residency; one that finds nothing to claim sleeps
// briefly and looks again.  It exits once no system is active, or -- the
// bound on every wait -- once no step anywhere has finished for `stallTicks`
// of the wall clock, which sets the abort word for the host to report.
constexpr unsigned int kStepDone = 0xffffffffu;
// each system's claim word and finish counter sit on cache lines of their own
constexpr int kCtlStride = 256;

struct BatchCtl
{
    char* word;                  // [K], stride kCtlStride: u64 step << 32 | chunks claimed
    char* fin;                   // [K], stride kCtlStride: u32 chunks finished of the current step
    long long* finish;           // [K] wall-clock tick at which the system stopped
    unsigned int* glob;          // active systems, abort, steps finished
    long long* t0;               // start tick
    const int4* chunks;          // [K] (Gs, Gv, chunks per claim of residual/A steps, of C/B steps)
};
__device__ __forceinline__ unsigned long long* ctlWord(const BatchCtl& c, int s)
{
    return reinterpret_cast<unsigned long long*>(c.word + static_cast<std::size_t>(s) * kCtlStride);
}
__device__ __forceinline__ unsigned int* ctlFin(const BatchCtl& c, int s)
{
    return reinterpret_cast<unsigned int*>(c.fin + static_cast<std::size_t>(s) * kCtlStride);
}

// step kinds: 0 residual, 1 C(0), then A, B, C; residual and A have Gs chunks, the others Gv
__device__ __forceinline__ int stepKind(unsigned int step) { return step < 2 ? static_cast<int>(step) : 2 + static_cast<int>((step - 2) % 3); }
__device__ __forceinline__ bool spmvStep(unsigned int step)
{
    const int kind = stepKind(step);
    return kind == 0 || kind == 2;
}
__device__ __forceinline__ unsigned int stepChunks(int4 g, unsigned int step)
{
    return static_cast<unsigned int>(spmvStep(step) ? g.x : g.y);
}
__device__ __forceinline__ unsigned int stepClaim(int4 g, unsigned int step)
{
    return static_cast<unsigned int>(spmvStep(step) ? g.z : g.w);
}

__global__ void k_bpcg_ctl_init(BatchCtl c, CgState* __restrict__ st, int K)
{
    for (int s = threadIdx.x; s < K; s += blockDim.x)
    {
        *ctlWord(c, s) = 0ull;
        *ctlFin(c, s) = 0u;
        c.finish[s] = -1;
        unsigned int* w = reinterpret_cast<unsigned int*>(st + s);
        for (int i = 0; i < static_cast<int>(sizeof(CgState) / 4); ++i) w[i] = 0u;
    }
    if (threadIdx.x == 0)
    {
        c.glob[0] = static_cast<unsigned int>(K);
        c.glob[1] = 0u;
        c.glob[2] = 0u;
        *c.t0 = wall_clock64();
    }
}

__global__ __launch_bounds__(kB) void k_bpcg_persistent(const BatchSysDev* __restrict__ S, int K, BatchCtl c,
                                                        const int* __restrict__ home, double tol, int maxIter,
                                                        long long stallTicks)
{
    __shared__ int sh[4];
    int cursor = home[blockIdx.x];
    bool idle = false;
    unsigned int idleProgress = 0;
    long long idleSince = 0;
    int backoff = 1;  // sleeps of about 0.2 us between looks while nothing can be claimed (doubles to 32)
    for (;;)
    {
        // ---- wave 0 looks for chunks to claim
        if (threadIdx.x < kWave)
        {
            const int lane = static_cast<int>(threadIdx.x);
            int got = -2;  // >= 0: system, -1: leave, -2: nothing to claim now
            unsigned int gStep = 0, gChunk = 0, gCount = 0;
            const bool over = __hip_atomic_load(c.glob, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u ||
                              __hip_atomic_load(c.glob + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
            if (over) got = -1;
            for (int base = 0; got == -2 && base < K; base += kWave)
            {
                const int s = (cursor + base + lane) % K;
                bool avail = false;
                unsigned int seen = 0;  // the step this lane saw
                if (base + lane < K)
                {
                    const unsigned long long w = __hip_atomic_load(ctlWord(c, s), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    seen = static_cast<unsigned int>(w >> 32);
                    avail = seen != kStepDone && static_cast<unsigned int>(w) < stepChunks(c.chunks[s], seen);
                }
                unsigned long long mask = __ballot(avail);
                while (mask && got == -2)
                {
                    const int first = __ffsll(static_cast<long long>(mask)) - 1;
                    mask &= mask - 1;
                    int res = -2;
                    unsigned int rs = 0, rc = 0, rn = 0;
                    if (lane == first)
                    {
                        // the claim size is that of the step this lane saw; the add returns the step it
                        // claims in, whose chunk count bounds what is taken
                        const unsigned int take = stepClaim(c.chunks[s], seen);
                        const unsigned long long w2 = __hip_atomic_fetch_add(
                            ctlWord(c, s), static_cast<unsigned long long>(take), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        rs = static_cast<unsigned int>(w2 >> 32);
                        rc = static_cast<unsigned int>(w2);
                        if (rs != kStepDone)
                        {
                            const unsigned int n = stepChunks(c.chunks[s], rs);
                            if (rc < n)
                            {
                                res = s;
                                rn = min(take, n - rc);
                            }
                        }
                    }
                    got = __shfl(res, first);
                    gStep = __shfl(rs, first);
                    gChunk = __shfl(rc, first);
                    gCount = __shfl(rn, first);
                }
            }
            if (lane == 0)
            {
                if (got >= 0)
                {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    idle = false;
                    backoff = 1;
                    cursor = got;
                }
                else if (got == -2)
                {
                    // bounded wait: give up only when no step anywhere has finished for stallTicks
                    const unsigned int prog = __hip_atomic_load(c.glob + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const long long now = wall_clock64();
                    if (!idle || prog != idleProgress)
                    {
                        idle = true;
                        idleProgress = prog;
                        idleSince = now;
                    }
                    else if (now - idleSince > stallTicks)
                    {
                        __hip_atomic_store(c.glob + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        got = -1;
                    }
                    for (int z = 0; z < backoff; ++z) __builtin_amdgcn_s_sleep(8);
                    backoff = min(2 * backoff, 32);
                }
                sh[0] = got;
                sh[1] = static_cast<int>(gStep);
                sh[2] = static_cast<int>(gChunk);
                sh[3] = static_cast<int>(gCount);
            }
        }
        __syncthreads();
        const int s = sh[0];
        const unsigned int step = static_cast<unsigned int>(sh[1]);
        const int c0 = sh[2], cn = sh[3];
        __syncthreads();
        if (s == -1) return;
        if (s == -2) continue;
        // ---- run chunks c0 .. c0 + cn - 1 of the step, as workgroups c of the launch
        const BatchSysDev d = S[s];
        const int kind = stepKind(step);
        for (int ch = c0; ch < c0 + cn; ++ch)
        {
            if (kind == 0)
                body_pcg_residual<double>(d.A, d.rbPerWg, d.val, d.x, d.b, nullptr, d.r, d.invD, d.part,
                                          d.part + kRedMax, d.part + 2 * kRedMax, ch, d.Gs);
            else if (kind == 1)
                body_pcg_C<double>(d.n, d.invD, d.r, d.p, d.st, d.part, d.part + kRedMax, d.part + 2 * kRedMax, d.Gs, 0,
                                   tol, maxIter, 0.0, 0, nullptr, ch, d.Gv);
            else if (kind == 2)
                body_pcg_A<double>(d.A, d.rbPerWg, d.val, d.p, d.q, d.st, d.part + 3 * kRedMax, ch, d.Gs);
            else if (kind == 3)
            {
                if (!d.st->done)
                {
                    const int k = d.st->iter + 1;
                    if (ch == 0 && threadIdx.x == 0) d.st->kcur = k;
                    body_pcg_B<double>(d.n, d.invD, d.x, d.r, d.p, d.q, d.st, d.part + 3 * kRedMax, d.Gs, d.part,
                                       d.part + kRedMax, k, ch, d.Gv);
                }
            }
            else if (!d.st->done)
                body_pcg_C<double>(d.n, d.invD, d.r, d.p, d.st, d.part, d.part + kRedMax, d.part + 2 * kRedMax, d.Gv,
                                   d.st->kcur, tol, maxIter, 0.0, 0, nullptr, ch, d.Gv);
            __syncthreads();
        }
        // ---- count them finished; the last one moves the system on.  Every wave
        // releases its own stores (a workgroup barrier does not wait for them).
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (threadIdx.x == 0)
        {
            const unsigned int n = stepChunks(c.chunks[s], step);
            const unsigned int f = __hip_atomic_fetch_add(ctlFin(c, s), static_cast<unsigned int>(cn), __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT);
            if (f + static_cast<unsigned int>(cn) == n)
            {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                __hip_atomic_store(ctlFin(c, s), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const bool stop = (kind == 1 || kind == 4) &&
                                  __hip_atomic_load(&d.st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
                if (stop) c.finish[s] = wall_clock64();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_store(ctlWord(c, s),
                                   static_cast<unsigned long long>(stop ? kStepDone : step + 1) << 32,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (stop) __hip_atomic_fetch_sub(c.glob, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(c.glob + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

} // namespace

namespace
{
__global__ void k_clear_words(unsigned int* __restrict__ p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}
} // namespace

static_assert(sizeof(CgState) <= kPcgStateBytes, "the PCG state outgrows its workspace slot");

// ------------------------------------------------------------------- host
struct PcgLaunch
{
    int Gs, rbPerWg, Gv;
};

static PcgLaunch pcgGrid(const CsrView& A, int n)
{
    PcgLaunch g;
    g.Gs = std::min(A.nRowBlocks, kRedMax);
    g.rbPerWg = divUp(A.nRowBlocks, g.Gs);
    g.Gs = divUp(A.nRowBlocks, g.rbPerWg);
    g.Gv = std::max(1, std::min(kRedMax, divUp(n, 4 * kB)));
    return g;
}

// The Jacobi-PCG driver on an explicit workspace (no allocation here): the
// state is cleared by a kernel, the host polls pinned state between batches.
template <typename T>
static SolveReport pcgCore(Context& ctx, int n_, const PcgWork& w, int minBatch, int maxBatch, const CsrView& A,
                           const double* val64, const T* valT, const T* invD, const double* b, double* x64, T* xT,
                           T* r, double tol, int maxIter, double* resHist, bool mixed, double delta)
{
    hipStream_t s = ctx.stream();
    const PcgLaunch g = pcgGrid(A, n_);
    if (g.Gs > kRedMax || g.Gv > kRedMax) throw std::invalid_argument("pcg: grid exceeds the partial buffer");
    double* part = w.part;
    double *partRZ = part, *partRM = part + kRedMax, *partBM = part + 2 * kRedMax, *partPQ = part + 3 * kRedMax;
    CgState* st = static_cast<CgState*>(w.state);
    k_clear_words<<<1, kB, 0, s>>>(reinterpret_cast<unsigned int*>(st), sizeof(CgState) / 4);
    T* p = static_cast<T*>(w.p);
    T* q = static_cast<T*>(w.q);
    double* r64 = w.r64;

    // initial residual (FP64); for FP64 runs r is the residual itself
    k_pcg_residual<T><<<g.Gs, kB, 0, s>>>(A, g.rbPerWg, val64, x64, b, r64, r, invD, partRZ, partRM, partBM);
    CFX_CHECK_LAUNCH();
    CgState* hs = reinterpret_cast<CgState*>(ctx.hostState() + 2);
    auto pollState = [&] {
        CFX_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
        CFX_HIP(hipStreamSynchronize(s));
    };
    int k = 0, outer = 1;
    int flags = mixed ? 3 : 0;  // the C(k) after a residual kernel sees a true residual
    double r0 = -1;
    // C(0) decides on the initial residual and forms p; then A,B,C per iteration
    k_pcg_C<T><<<g.Gv, kB, 0, s>>>(n_, invD, r, p, st, partRZ, partRM, partBM, g.Gs, 0, tol, maxIter, delta, flags,
                                   resHist);
    flags &= ~1;
    for (;;)
    {
        int batch = minBatch;
        if (r0 > 0 && hs->iter > 0 && hs->rnorm > 0 && hs->rnorm < r0)
        {
            const double rate = std::log(hs->rnorm / r0) / hs->iter;
            const double target = mixed ? std::max(tol * hs->bnorm, hs->updAbs) : tol * hs->bnorm;
            const double need = std::log(target / hs->rnorm) / rate;
            if (std::isfinite(need)) batch = static_cast<int>(0.9 * need);
        }
        batch = std::max(minBatch, std::min(maxBatch, batch));
        batch = std::min(batch, maxIter - k);
        if (batch <= 0) break;  // C(maxIter) has been launched and stops
        for (int j = 0; j < batch; ++j)
        {
            ++k;
            k_pcg_A<T><<<g.Gs, kB, 0, s>>>(A, g.rbPerWg, valT, p, q, st, partPQ);
            k_pcg_B<T><<<g.Gv, kB, 0, s>>>(n_, invD, xT, r, p, q, st, partPQ, g.Gs, partRZ, partRM, k);
            k_pcg_C<T><<<g.Gv, kB, 0, s>>>(n_, invD, r, p, st, partRZ, partRM, partBM, g.Gv, k, tol, maxIter, delta,
                                           flags, resHist);
        }
        CFX_CHECK_LAUNCH();
        pollState();
        if (r0 < 0) r0 = hs->initialResidual;
        if (!hs->done) continue;
        if (mixed && hs->paused && !hs->converged)
        {
            // reliable update at iteration hs->iter: fold, true residual, decide again there
            k = hs->iter;
            k_fold<<<g.Gv, kB, 0, s>>>(n_, x64, reinterpret_cast<float*>(xT));
            k_pcg_residual<T><<<g.Gs, kB, 0, s>>>(A, g.rbPerWg, val64, x64, b, r64, r, invD, partRZ, partRM, partBM);
            k_resume<<<1, 1, 0, s>>>(st);
            k_pcg_C<T><<<g.Gv, kB, 0, s>>>(n_, invD, r, p, st, partRZ, partRM, partBM, g.Gs, k, tol, maxIter, delta,
                                           3, resHist);
            CFX_CHECK_LAUNCH();
            ++outer;
            continue;
        }
        break;
    }
    if (mixed && !hs->converged)
    {
        k_fold<<<g.Gv, kB, 0, s>>>(n_, x64, reinterpret_cast<float*>(xT));
        CFX_CHECK_LAUNCH();
    }
    SolveReport rep;
    rep.iterations = hs->iter;
    rep.initialResidual = hs->initialResidual;
    rep.finalResidual = hs->iter > 0 ? hs->rnorm : hs->initialResidual;
    rep.converged = hs->converged;
    rep.breakdown = hs->breakdown;
    rep.outerIterations = mixed ? outer : 0;
    if (mixed) rep.trueResidual = hs->converged ? hs->rnorm / hs->bnorm : -1;
    return rep;
}

template <typename T>
SolveReport Krylov::pcgRun(const CsrView& A, const double* val64, const T* valT, const T* invD, const double* b,
                           double* x64, T* xT, T* r, double tol, int maxIter, double* resHist, bool mixed,
                           double delta)
{
    const PcgWork w{partials_.as<double>(), state_.as<char>() + 2 * sizeof(SolveState), mixed ? p32_.p : p_.p,
                    mixed ? q32_.p : q_.p, mixed ? r0_.as<double>() : nullptr};
    return pcgCore<T>(ctx_, n_, w, minBatch, maxBatch, A, val64, valT, invD, b, x64, xT, r, tol, maxIter, resHist,
                      mixed, delta);
}

SolveReport pcgSolveWith(Context& ctx, const CsrView& A, const double* val, const double* invD, const double* b,
                         double* x, double tol, int maxIter, const PcgWork& w, double* r)
{
    return pcgCore<double>(ctx, A.n, w, 8, 256, A, val, val, invD, b, x, x, r, tol, maxIter, nullptr, false, 0.0);
}

namespace
{

// The flexible CG's product q = A p: on the preconditioner's sliced copy of the values where it holds one
// for `val`, else on the CSR values; with start/stop events (measurement) when given.
struct FcgProduct
{
    CsrView A;
    SliceView S;
    const double* val = nullptr;
    bool sliced = false;
    void launch(const PcgLaunch& g, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const float* p, const float* r,
                float* q, const CgState* st, double* partPQ, double* partPR) const
    {
        if (e0 && sliced)
            hipExtLaunchKernelGGL(k_fcg_A<true>, dim3(g.Gs), dim3(kB), 0, s, e0, e1, 0, A, S, g.rbPerWg, val, p, r, q,
                                  st, partPQ, partPR);
        else if (e0)
            hipExtLaunchKernelGGL(k_fcg_A<false>, dim3(g.Gs), dim3(kB), 0, s, e0, e1, 0, A, S, g.rbPerWg, val, p, r,
                                  q, st, partPQ, partPR);
        else if (sliced)
            k_fcg_A<true><<<g.Gs, kB, 0, s>>>(A, S, g.rbPerWg, val, p, r, q, st, partPQ, partPR);
        else
            k_fcg_A<false><<<g.Gs, kB, 0, s>>>(A, S, g.rbPerWg, val, p, r, q, st, partPQ, partPR);
    }
};
FcgProduct fcgProduct(const CsrView& A, const double* val, const Preconditioner* pre)
{
    FcgProduct f;
    f.A = A;
    f.val = val;
    const double* sv = nullptr;
    if (pre && pre->fineSlices(val, f.S, sv))
    {
        f.val = sv;
        f.sliced = true;
    }
    return f;
}

} // namespace

double Krylov::timeFcgSpmv(const CsrView& A, const double* val, int reps, const Preconditioner* pre)
{
    hipStream_t s = ctx_.stream();
    const PcgLaunch g = pcgGrid(A, n_);
    if (A.n != n_ || reps < 1) throw std::invalid_argument("timeFcgSpmv: bad matrix or repetition count");
    if (!fcgPart_.p)
    {
        fcgPart_.alloc(sizeof(double) * 5 * kRedMax);
        z32_.alloc(sizeof(float) * static_cast<std::size_t>(n_));
    }
    double* part = fcgPart_.as<double>();
    CgState* st = reinterpret_cast<CgState*>(state_.as<char>() + 2 * sizeof(SolveState));
    k_clear_words<<<1, kB, 0, s>>>(reinterpret_cast<unsigned int*>(st), sizeof(CgState) / 4);  // not done
    FcgProduct fa = fcgProduct(A, val, pre);
    for (int i = 0; i <= reps; ++i)  // the first launch warms
    {
        if (i == 1) CFX_HIP(hipEventRecord(ev0_, s));
        fa.launch(g, s, nullptr, nullptr, p32_.as<float>(), r32_.as<float>(), q32_.as<float>(), st, part,
                  part + kRedMax);
    }
    CFX_HIP(hipEventRecord(ev1_, s));
    CFX_HIP(hipEventSynchronize(ev1_));
    CFX_CHECK_LAUNCH();
    float ms = 0;
    CFX_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    return ms / reps;
}

SolveReport Krylov::cg(const CsrView& A, const double* val, const double* b, double* x, double tol, int maxIter,
                       double* resHist)
{
    hipStream_t s = ctx_.stream();
    CFX_HIP(hipEventRecord(ev0_, s));
    computeInvD(ctx_, A, val, invD_.as<double>(), false);
    SolveReport rep = pcgRun<double>(A, val, val, invD_.as<double>(), b, x, x, r_.as<double>(), tol, maxIter,
                                     resHist, false, 0.0);
    CFX_HIP(hipEventRecord(ev1_, s));
    CFX_HIP(hipEventSynchronize(ev1_));
    float ms = 0;
    CFX_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    lastMs_ = rep.ms = ms;
    return rep;
}

SolveReport Krylov::cgMixed(const CsrView& A, const double* val, const float* val32, const double* b, double* x,
                            double tol, int maxIter, double delta, bool /*project*/)
{
    if (delta <= 0) delta = kPcgDelta;
    hipStream_t s = ctx_.stream();
    CFX_HIP(hipEventRecord(ev0_, s));
    k_inv_diag32<<<divUp(A.n, 256), 256, 0, s>>>(A, val, invD32_.as<float>(), invD_.as<double>());
    CFX_HIP(hipMemsetAsync(d32_.p, 0, sizeof(float) * n_, s));
    CFX_CHECK_LAUNCH();
    SolveReport rep = pcgRun<float>(A, val, val32, invD32_.as<float>(), b, x, d32_.as<float>(), r32_.as<float>(),
                                    tol, maxIter, nullptr, true, delta);
    if (rep.converged)
    {
        double inf = 0, l2 = 0;
        trueResidual(A, val, b, x, inf, l2);
        rep.trueResidual = inf;
        rep.l2Residual = l2;
    }
    CFX_HIP(hipEventRecord(ev1_, s));
    CFX_HIP(hipEventSynchronize(ev1_));
    float ms = 0;
    CFX_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    lastMs_ = rep.ms = ms;
    return rep;
}


SolveReport Krylov::fcgMixed(const CsrView& A, const double* val, Preconditioner& mg, const double* b, double* x,
                             double tol, int maxIter, double deltaArg, double* resHist)
{
    const double delta = deltaArg > 0 ? deltaArg : -1.0;  // negative: the automatic rule of k_fcg_C
    hipStream_t s = ctx_.stream();
    CFX_HIP(hipEventRecord(ev0_, s));
    const PcgLaunch g = pcgGrid(A, n_);
    if (g.Gs > kRedMax || g.Gv > kRedMax) throw std::invalid_argument("fcg: grid exceeds the partial buffer");
    if (mg.rows() != A.n) throw std::invalid_argument("fcg: preconditioner does not match the matrix");
    if (!fcgPart_.p)
    {
        fcgPart_.alloc(sizeof(double) * 5 * kRedMax);
        z32_.alloc(sizeof(float) * static_cast<std::size_t>(n_));
    }
    double* part = fcgPart_.as<double>();
    double *partPQ = part, *partPR = part + kRedMax, *partRM = part + 2 * kRedMax, *partBM = part + 3 * kRedMax,
           *partZQ = part + 4 * kRedMax;
    double* partRZ = partials_.as<double>();  // written by the residual kernel, unused here
    CgState* st = reinterpret_cast<CgState*>(state_.as<char>() + 2 * sizeof(SolveState));
    probeUsed_ = 0;
    mg.setProbe(probeFcgA);
    CFX_HIP(hipMemsetAsync(st, 0, sizeof(CgState), s));
    CFX_HIP(hipMemsetAsync(d32_.p, 0, sizeof(float) * n_, s));
    float *p = p32_.as<float>(), *q = q32_.as<float>(), *r = r32_.as<float>(), *z = z32_.as<float>(),
          *x32 = d32_.as<float>();
    const float* val32 = mg.val32();
    const float* invD32 = mg.invD32();
    const int* done = &st->done;
    const FcgProduct fa = fcgProduct(A, val, &mg);

    k_pcg_residual<float><<<g.Gs, kB, 0, s>>>(A, g.rbPerWg, val, x, b, nullptr, r, invD32, partRZ, partRM, partBM);
    mg.apply(r, z, nullptr, nullptr, g.Gs, g.rbPerWg, done);
    k_fcg_C<<<g.Gv, kB, 0, s>>>(n_, z, p, st, partZQ, g.Gs, partRM, partBM, g.Gs, 0, tol, maxIter, delta, 3, resHist);
    CFX_CHECK_LAUNCH();
    CgState* hs = reinterpret_cast<CgState*>(ctx_.hostState() + 2);
    auto pollState = [&] {
        CFX_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
        CFX_HIP(hipStreamSynchronize(s));
    };
    int k = 0, outer = 1;
    int k0Ref = 0;  // iteration of the last true residual
    double r0 = -1;
    // the device's reliable-update thresholds as of the last true residual, and the last poll's state
    double updAbs = 0, updFast = 0;
    int kRef = 0, lastIter = -1;
    double lastRnorm = -1;
    for (;;)
    {
        // iterations to the next stop the device will make: convergence of the recurrence residual or a
        // reliable update (after a drop to updFast, or to updAbs once kFcgSegment iterations have passed;
        // a fixed factor: updAbs), at the rate of the last batch (since the start while that ran fewer than
        // two iterations).  An iteration launched past the stop costs its launches (about 0.1 ms at 2 M rows),
        // a batch that ends before it one more poll, which is cheaper: the batch ends at the predicted stop.
        int batch = 2;
        if (r0 > 0 && hs->iter > 0 && hs->rnorm > 0 && hs->rnorm < r0)
        {
            double rate = std::log(hs->rnorm / r0) / hs->iter;
            if (lastIter >= 0 && hs->iter - lastIter >= 2 && lastRnorm > hs->rnorm)
                rate = std::log(hs->rnorm / lastRnorm) / (hs->iter - lastIter);
            auto itersTo = [&](double level) {
                return std::log(std::max(level, tol * hs->bnorm) / hs->rnorm) / rate;
            };
            const double need = delta > 0 ? itersTo(updAbs)
                                          : std::min(itersTo(updFast), std::max(itersTo(updAbs),
                                                                                 double(kRef + kFcgSegment - k)));
            if (std::isfinite(need)) batch = static_cast<int>(std::ceil(need - 0.25));
        }
        batch = std::max(1, std::min(64, batch));
        batch = std::min(batch, maxIter - k);
        if (batch <= 0) break;
        if (!useGraphs || probeFcgA)
        {
            for (int j = 0; j < batch; ++j)
            {
                if (probeFcgA)
                {
                    // measurement mode: the launch carries HIP start and stop events, which the command
                    // processor records at the kernel's own start and end on the solver's stream
                    while (static_cast<int>(probeEv_.size()) < 2 * (probeUsed_ + 1))
                    {
                        hipEvent_t e;
                        CFX_HIP(hipEventCreate(&e));
                        probeEv_.push_back(e);
                    }
                    fa.launch(g, s, probeEv_[2 * probeUsed_], probeEv_[2 * probeUsed_ + 1], p, r, q, st, partPQ,
                              partPR);
                    ++probeUsed_;
                }
                else
                    fa.launch(g, s, nullptr, nullptr, p, r, q, st, partPQ, partPR);
                k_fcg_B<<<g.Gv, kB, 0, s>>>(n_, x32, r, p, q, st, partPQ, partPR, g.Gs, partRM);
                mg.apply(r, z, q, partZQ, g.Gs, g.rbPerWg, done);
                k_fcg_C<<<g.Gv, kB, 0, s>>>(n_, z, p, st, partZQ, g.Gs, partRM, partBM, g.Gv, -1, tol, maxIter,
                                            delta, 2, resHist);
            }
        }
        else
        {
            // one iteration (A, B, the cycle, C) as a captured graph, replayed per step
            // every address and size the captured kernels take: the matrix pattern and values, the
            // preconditioner (its id changes with its buffers and launch structure), this solver's vectors,
            // partials and state, the history buffer and the grids
            const std::vector<const void*> key{&mg,
                                               reinterpret_cast<const void*>(mg.id()),
                                               val,
                                               fa.val,
                                               resHist,
                                               A.rowPtr,
                                               A.col,
                                               A.rowBlocks,
                                               reinterpret_cast<const void*>(static_cast<std::uintptr_t>(A.n)),
                                               reinterpret_cast<const void*>(static_cast<std::uintptr_t>(A.nRowBlocks)),
                                               p,
                                               q,
                                               r,
                                               z,
                                               x32,
                                               part,
                                               st,
                                               reinterpret_cast<const void*>(static_cast<std::uintptr_t>(g.Gs)),
                                               reinterpret_cast<const void*>(static_cast<std::uintptr_t>(g.Gv))};
            if (!fcgExec_ || key != fcgKey_ || tol != fcgTol_ || maxIter != fcgMaxIter_ || delta != fcgDelta_)
            {
                if (fcgExec_) CFX_HIP(hipGraphExecDestroy(fcgExec_));
                fcgExec_ = nullptr;
                hipGraph_t graph;
                CFX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                fa.launch(g, s, nullptr, nullptr, p, r, q, st, partPQ, partPR);
                k_fcg_B<<<g.Gv, kB, 0, s>>>(n_, x32, r, p, q, st, partPQ, partPR, g.Gs, partRM);
                mg.apply(r, z, q, partZQ, g.Gs, g.rbPerWg, done);
                k_fcg_C<<<g.Gv, kB, 0, s>>>(n_, z, p, st, partZQ, g.Gs, partRM, partBM, g.Gv, -1, tol, maxIter, delta, 2,
                                            resHist);
                CFX_HIP(hipStreamEndCapture(s, &graph));
                CFX_HIP(hipGraphInstantiate(&fcgExec_, graph, nullptr, nullptr, 0));
                CFX_HIP(hipGraphDestroy(graph));
                fcgKey_ = key;
                fcgTol_ = tol;
                fcgMaxIter_ = maxIter;
                fcgDelta_ = delta;
            }
            for (int j = 0; j < batch; ++j) CFX_HIP(hipGraphLaunch(fcgExec_, s));
        }
        k += batch;
        CFX_CHECK_LAUNCH();
        pollState();
        if (r0 < 0)
        {
            r0 = hs->initialResidual;
            updAbs = hs->updAbs;
            updFast = hs->updFast;
        }
        if (!hs->done)
        {
            lastIter = hs->iter;
            lastRnorm = hs->rnorm;
            continue;
        }
        if (hs->paused && !hs->converged)
        {
            // reliable update at iteration hs->iter: the direction p_{k-1} and q = A p_{k-1} are kept; r
            // becomes the true residual, from which C(k) decides and B takes the next step length.  After a
            // short run of FP32 iterations (at most kFcgSegment) the recurrence residual is still close to
            // the true one, so z = M r of the paused iteration and its z.q partials are kept too (flexible
            // CG takes any preconditioned vector for z; applying M again would cost one more cycle); after
            // a longer run z is recomputed from the true residual
            const int kPrev = k0Ref;
            k = hs->iter;
            k0Ref = k;
            k_fold<<<g.Gv, kB, 0, s>>>(n_, x, x32);
            k_pcg_residual<float><<<g.Gs, kB, 0, s>>>(A, g.rbPerWg, val, x, b, nullptr, r, invD32, partRZ, partRM,
                                                      partBM);
            k_resume<<<1, 1, 0, s>>>(st);
            if (k - kPrev > kFcgSegment) mg.apply(r, z, q, partZQ, g.Gs, g.rbPerWg, done);
            k_fcg_C<<<g.Gv, kB, 0, s>>>(n_, z, p, st, partZQ, g.Gs, partRM, partBM, g.Gs, k, tol, maxIter, delta, 3,
                                        resHist);
            CFX_CHECK_LAUNCH();
            ++outer;
            // the thresholds the device takes from this true residual (close to the recurrence residual)
            updAbs = (delta > 0 ? delta : kPcgDelta) * hs->rnorm;
            updFast = kFcgDelta * hs->rnorm;
            kRef = k;
            lastIter = hs->iter;
            lastRnorm = hs->rnorm;
            if (hs->rnorm < tol * hs->bnorm)
            {
                // the pause came from the recurrence residual's convergence: the true residual most likely
                // converges too, and then any iteration launched now would only return
                pollState();
                if (hs->done) break;
            }
            continue;
        }
        break;
    }
    if (probeFcgA)
    {
        probeSmoothMs = 0;
        probeSmoothCount = 0;
        mg.probeResult(probeSmoothMs, probeSmoothCount);
        mg.setProbe(false);
    }
    if (probeFcgA && probeUsed_ > 0)
    {
        // launches after the stop decision return at once: average over those that ran (> 20 % of the longest)
        CFX_HIP(hipStreamSynchronize(s));
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
        probeMs = cnt ? sum / cnt : 0.0;
        probeCount = cnt;
    }
    const bool failed = hs->nan || hs->breakdown;
    if (!hs->converged && !failed)
    {
        k_fold<<<g.Gv, kB, 0, s>>>(n_, x, x32);
        CFX_CHECK_LAUNCH();
    }
    SolveReport rep;
    rep.iterations = hs->iter;
    rep.initialResidual = hs->initialResidual;
    rep.finalResidual = hs->iter > 0 ? hs->rnorm : hs->initialResidual;
    rep.converged = hs->converged;
    rep.breakdown = hs->breakdown;
    rep.outerIterations = outer;
    rep.trueResidual = hs->converged ? hs->rnorm / hs->bnorm : -1;
    if (failed)
    {
        // the preconditioner produced a non-finite value or lost positivity:
        // continue from the last FP64 iterate with Jacobi-PCG (FP32 + reliable updates)
        const int used = rep.iterations;
        const double init = rep.initialResidual;
        k_inv_diag32<<<divUp(A.n, 256), 256, 0, s>>>(A, val, invD32_.as<float>(), invD_.as<double>());
        CFX_HIP(hipMemsetAsync(d32_.p, 0, sizeof(float) * n_, s));
        CFX_CHECK_LAUNCH();
        rep = pcgRun<float>(A, val, val32, invD32_.as<float>(), b, x, d32_.as<float>(), r32_.as<float>(), tol,
                            std::max(0, maxIter - used), nullptr, true, deltaArg > 0 ? deltaArg : kPcgDelta);
        rep.iterations += used;
        rep.initialResidual = init;
        rep.precondFallbacks = 1;
    }
    if (rep.converged)
    {
        double inf = 0, l2 = 0;
        trueResidual(A, val, b, x, inf, l2);
        rep.trueResidual = inf;
        rep.l2Residual = l2;
    }
    CFX_HIP(hipEventRecord(ev1_, s));
    CFX_HIP(hipEventSynchronize(ev1_));
    float ms = 0;
    CFX_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    lastMs_ = rep.ms = ms;
    return rep;
}

// ------------------------------------------------------ batched Jacobi-PCG
BatchedPcg::BatchedPcg(Context& ctx) : ctx_(ctx)
{
    CFX_HIP(hipEventCreate(&e0_));
    CFX_HIP(hipEventCreate(&e1_));
}

BatchedPcg::~BatchedPcg()
{
    if (e0_) (void)hipEventDestroy(e0_);
    if (e1_) (void)hipEventDestroy(e1_);
}

void BatchedPcg::reserve(std::size_t n, std::size_t K)
{
    if (n > capN_)
    {
        for (DevBuf* b : {&r_, &p_, &q_, &invD_, &xs_}) b->alloc(sizeof(double) * n);
        capN_ = n;
    }
    if (K > capK_)
    {
        part_.alloc(sizeof(double) * 4 * kRedMax * K);
        state_.alloc(sizeof(CgState) * K);
        sys_.alloc(sizeof(BatchSysDev) * K);
        capK_ = K;
    }
}

BatchResult BatchedPcg::solve(const std::vector<BatchSystem>& sys, double tol, int maxIter, const BatchSchedule& sc)
{
    hipStream_t s = ctx_.stream();
    const std::size_t K = sys.size();
    if (K == 0) throw std::invalid_argument("batch: no systems");
    if (sc.vecElems < 1 || sc.wgCap < 1 || sc.wgCap > kRedMax || sc.pollEvery < 1 || sc.claim < 1 || sc.wgs < 0)
        throw std::invalid_argument("batch: bad schedule");
    std::size_t N = 0;
    for (const BatchSystem& b : sys) N += static_cast<std::size_t>(b.A.n);
    reserve(N, K);
    std::vector<BatchSysDev> hs(K);
    std::vector<int2> tabS, tabV;
    std::size_t off = 0;
    CgState* st = state_.as<CgState>();
    for (std::size_t k = 0; k < K; ++k)
    {
        const BatchSystem& b = sys[k];
        BatchSysDev& d = hs[k];
        d.A = b.A;
        d.val = b.val;
        d.b = b.b;
        d.x = b.x;
        d.invD = invD_.as<double>() + off;
        d.r = r_.as<double>() + off;
        d.p = p_.as<double>() + off;
        d.q = q_.as<double>() + off;
        d.st = st + k;
        d.part = part_.as<double>() + 4 * kRedMax * k;
        d.n = b.A.n;
        int Gs = std::max(1, std::min(b.A.nRowBlocks, sc.wgCap));
        d.rbPerWg = divUp(b.A.nRowBlocks, Gs);
        d.Gs = std::max(1, divUp(b.A.nRowBlocks, d.rbPerWg));
        d.Gv = std::max(1, std::min(kRedMax, divUp(b.A.n, sc.vecElems * kB)));
        for (int j = 0; j < d.Gs; ++j) tabS.push_back(make_int2(static_cast<int>(k), j));
        for (int j = 0; j < d.Gv; ++j) tabV.push_back(make_int2(static_cast<int>(k), j));
        computeInvD(ctx_, b.A, b.val, const_cast<double*>(d.invD), false);
        off += static_cast<std::size_t>(b.A.n);
    }
    CFX_HIP(hipMemcpyAsync(sys_.p, hs.data(), sizeof(BatchSysDev) * K, hipMemcpyHostToDevice, s));
    BatchResult res;
    if (sc.persistent)
    {
        std::vector<int2> ch(K);
        for (std::size_t q = 0; q < K; ++q) ch[q] = make_int2(hs[q].Gs, hs[q].Gv);
        runPersistent(sys, ch, sc, tol, maxIter, N, res, res.finishMs);
        return res;
    }
    tabS_ = upload(tabS, s);
    tabV_ = upload(tabV, s);
    CFX_HIP(hipMemsetAsync(st, 0, sizeof(CgState) * K, s));
    const BatchSysDev* S = sys_.as<BatchSysDev>();
    const int GS = static_cast<int>(tabS.size()), GV = static_cast<int>(tabV.size());
    const int2* TS = tabS_.as<int2>();
    const int2* TV = tabV_.as<int2>();
    std::vector<CgState> h(K);
    res.finishMs.assign(K, -1.0);
    // capture (and instantiate) the iteration before the timed region
    hipGraphExec_t exec = nullptr;
    if (sc.graph)
    {
        hipGraph_t g;
        CFX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        k_bpcg_A<<<GS, kB, 0, s>>>(S, TS);
        k_bpcg_B<<<GV, kB, 0, s>>>(S, TV);
        k_bpcg_C<<<GV, kB, 0, s>>>(S, TV, tol, maxIter);
        CFX_HIP(hipStreamEndCapture(s, &g));
        CFX_HIP(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
        CFX_HIP(hipGraphDestroy(g));
    }
    CFX_HIP(hipEventRecord(e0_, s));
    k_bpcg_residual<<<GS, kB, 0, s>>>(S, TS);
    k_bpcg_C0<<<GV, kB, 0, s>>>(S, TV, tol, maxIter);
    CFX_CHECK_LAUNCH();
    int k = 0;
    std::size_t left = K;
    while (left > 0 && k < maxIter)
    {
        const int batch = std::min(sc.pollEvery, maxIter - k);
        for (int j = 0; j < batch; ++j)
        {
            if (exec)
                CFX_HIP(hipGraphLaunch(exec, s));
            else
            {
                k_bpcg_A<<<GS, kB, 0, s>>>(S, TS);
                k_bpcg_B<<<GV, kB, 0, s>>>(S, TV);
                k_bpcg_C<<<GV, kB, 0, s>>>(S, TV, tol, maxIter);
            }
        }
        k += batch;
        CFX_CHECK_LAUNCH();
        CFX_HIP(hipMemcpyAsync(h.data(), st, sizeof(CgState) * K, hipMemcpyDeviceToHost, s));
        CFX_HIP(hipEventRecord(e1_, s));
        CFX_HIP(hipEventSynchronize(e1_));
        float ms = 0;
        CFX_HIP(hipEventElapsedTime(&ms, e0_, e1_));
        left = 0;
        for (std::size_t q = 0; q < K; ++q)
        {
            if (h[q].done && res.finishMs[q] < 0) res.finishMs[q] = ms;
            if (!h[q].done) ++left;
        }
    }
    if (exec) CFX_HIP(hipGraphExecDestroy(exec));
    CFX_HIP(hipMemcpyAsync(h.data(), st, sizeof(CgState) * K, hipMemcpyDeviceToHost, s));
    CFX_HIP(hipEventRecord(e1_, s));
    CFX_HIP(hipEventSynchronize(e1_));
    float ms = 0;
    CFX_HIP(hipEventElapsedTime(&ms, e0_, e1_));
    res.ms = ms;
    res.reports.resize(K);
    for (std::size_t q = 0; q < K; ++q)
    {
        SolveReport& r = res.reports[q];
        r.iterations = h[q].iter;
        r.initialResidual = h[q].initialResidual;
        r.finalResidual = h[q].iter > 0 ? h[q].rnorm : h[q].initialResidual;
        r.converged = h[q].converged;
        r.breakdown = h[q].breakdown;
        if (res.finishMs[q] < 0) res.finishMs[q] = ms;
        r.ms = res.finishMs[q];
        res.maxIterations = std::max(res.maxIterations, r.iterations);
    }
    return res;
}

void BatchedPcg::runPersistent(const std::vector<BatchSystem>& sys, const std::vector<int2>& grids,
                               const BatchSchedule& sc, double tol, int maxIter, std::size_t N, BatchResult& res,
                               std::vector<double>& finishMs)
{
    hipStream_t s = ctx_.stream();
    const int K = static_cast<int>(sys.size());
    int dev = 0, cus = 0, perCu = 0, rateKHz = 0;
    CFX_HIP(hipGetDevice(&dev));
    CFX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    CFX_HIP(hipDeviceGetAttribute(&rateKHz, hipDeviceAttributeWallClockRate, dev));
    CFX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, reinterpret_cast<const void*>(k_bpcg_persistent), kB, 0));
    const int G = sc.wgs > 0 ? sc.wgs : std::max(1, perCu) * std::max(1, cus);
    if (rateKHz <= 0) throw std::runtime_error("batch: no device wall clock rate");
    // first system each workgroup looks at: the one holding its share of the rows
    std::vector<int> home(G);
    {
        std::size_t q = 0, end = static_cast<std::size_t>(sys[0].A.n);
        for (int g = 0; g < G; ++g)
        {
            const std::size_t row = N * static_cast<std::size_t>(g) / static_cast<std::size_t>(G);
            while (row >= end && q + 1 < sys.size()) end += static_cast<std::size_t>(sys[++q].A.n);
            home[g] = static_cast<int>(q);
        }
    }
    home_ = upload(home, s);
    // chunks per claim: about `claim` claims per workgroup and step, the workgroups dealt to systems by rows
    std::vector<int4> ch(sys.size());
    for (std::size_t q = 0; q < sys.size(); ++q)
    {
        const double share = static_cast<double>(G) * sys[q].A.n / static_cast<double>(N);
        const int wgs = std::max(1, static_cast<int>(share));
        ch[q] = make_int4(grids[q].x, grids[q].y, std::max(1, divUp(grids[q].x, static_cast<long long>(sc.claim) * wgs)),
                          std::max(1, divUp(grids[q].y, static_cast<long long>(sc.claim) * wgs)));
    }
    chunks_ = upload(ch, s);
    // control block: claim words and finish counters (a cache line each), finish ticks (8 B) per system,
    // then the start tick and 4 words
    const std::size_t Kz = static_cast<std::size_t>(K);
    const std::size_t bytes = 2 * kCtlStride * Kz + 8 * Kz + 16 + 16;
    if (!ctl_.p || ctl_.bytes < bytes) ctl_.alloc(bytes);
    char* base = ctl_.as<char>();
    BatchCtl c;
    c.word = base;
    c.fin = base + kCtlStride * Kz;
    c.finish = reinterpret_cast<long long*>(base + 2 * kCtlStride * Kz);
    c.t0 = reinterpret_cast<long long*>(base + 2 * kCtlStride * Kz + 8 * Kz);
    c.glob = reinterpret_cast<unsigned int*>(base + 2 * kCtlStride * Kz + 8 * Kz + 16);
    c.chunks = chunks_.as<int4>();
    CgState* st = state_.as<CgState>();
    const BatchSysDev* S = sys_.as<BatchSysDev>();
    // a wait ends after one second in which no step anywhere finished
    const long long stallTicks = static_cast<long long>(rateKHz) * 1000;
    auto launch = [&] {
        k_bpcg_ctl_init<<<1, kB, 0, s>>>(c, st, K);
        k_bpcg_persistent<<<G, kB, 0, s>>>(S, K, c, home_.as<int>(), tol, maxIter, stallTicks);
    };
    hipGraphExec_t exec = nullptr;
    if (sc.graph)
    {
        hipGraph_t g;
        CFX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        launch();
        CFX_HIP(hipStreamEndCapture(s, &g));
        CFX_HIP(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
        CFX_HIP(hipGraphDestroy(g));
    }
    CFX_HIP(hipEventRecord(e0_, s));
    if (exec)
        CFX_HIP(hipGraphLaunch(exec, s));
    else
        launch();
    CFX_CHECK_LAUNCH();
    CFX_HIP(hipEventRecord(e1_, s));
    CFX_HIP(hipEventSynchronize(e1_));
    if (exec) CFX_HIP(hipGraphExecDestroy(exec));
    float ms = 0;
    CFX_HIP(hipEventElapsedTime(&ms, e0_, e1_));
    std::vector<CgState> h(Kz);
    std::vector<long long> fin(Kz + 2);
    unsigned int glob[4];
    CFX_HIP(hipMemcpyAsync(h.data(), st, sizeof(CgState) * Kz, hipMemcpyDeviceToHost, s));
    CFX_HIP(hipMemcpyAsync(fin.data(), c.finish, sizeof(long long) * (Kz + 1), hipMemcpyDeviceToHost, s));
    CFX_HIP(hipMemcpyAsync(glob, c.glob, sizeof(glob), hipMemcpyDeviceToHost, s));
    CFX_HIP(hipStreamSynchronize(s));
    if (glob[1])
        throw std::runtime_error("batch: the persistent solve stopped making progress (" + std::to_string(glob[0]) +
                                 " systems unfinished)");
    const long long t0 = fin[Kz];
    res.ms = ms;
    res.reports.resize(Kz);
    finishMs.assign(Kz, ms);
    for (std::size_t q = 0; q < Kz; ++q)
    {
        SolveReport& r = res.reports[q];
        r.iterations = h[q].iter;
        r.initialResidual = h[q].initialResidual;
        r.finalResidual = h[q].iter > 0 ? h[q].rnorm : h[q].initialResidual;
        r.converged = h[q].converged;
        r.breakdown = h[q].breakdown;
        if (fin[q] >= 0) finishMs[q] = static_cast<double>(fin[q] - t0) / rateKHz;
        r.ms = finishMs[q];
        res.maxIterations = std::max(res.maxIterations, r.iterations);
    }
}

BatchSchedule BatchedPcg::autotune(const std::vector<BatchSystem>& sys, int iters,
                                   std::vector<std::pair<BatchSchedule, double>>* log)
{
    std::vector<int> key;
    for (const BatchSystem& b : sys) key.push_back(b.A.n);
    key.push_back(iters);
    auto it = tuned_.find(key);
    if (it != tuned_.end() && !log) return it->second;
    hipStream_t s = ctx_.stream();
    std::size_t N = 0;
    for (const BatchSystem& b : sys) N += static_cast<std::size_t>(b.A.n);
    reserve(N, sys.size());
    // time on scratch copies of x so the caller's initial guesses are kept
    std::vector<BatchSystem> tmp = sys;
    std::size_t off = 0;
    for (BatchSystem& b : tmp)
    {
        b.x = xs_.as<double>() + off;
        off += static_cast<std::size_t>(b.A.n);
    }
    std::vector<BatchSchedule> cands;
    for (int ve : {2, 4, 8})
        for (int cap : {512, 1024, 2048})
            for (int claim : {1, 2, 4})
            {
                BatchSchedule c;
                c.persistent = 1;
                c.vecElems = ve;
                c.wgCap = cap;
                c.claim = claim;
                cands.push_back(c);
            }
    for (int graph : {1, 0})
        for (int ve : {2, 4, 8})
            for (int cap : {512, 1024, 2048})
            {
                BatchSchedule c;
                c.persistent = 0;
                c.vecElems = ve;
                c.wgCap = cap;
                c.graph = graph;
                c.pollEvery = iters;
                cands.push_back(c);
            }
    BatchSchedule best;
    double bestMs = 1e300;
    for (const BatchSchedule& c : cands)
    {
        for (std::size_t q = 0; q < sys.size(); ++q)
            CFX_HIP(hipMemcpyAsync(tmp[q].x, sys[q].x, sizeof(double) * sys[q].A.n, hipMemcpyDeviceToDevice, s));
        // tol 0: no system stops early, every candidate runs `iters` iterations
        const BatchResult r = solve(tmp, 0.0, iters, c);
        if (log) log->push_back({c, r.ms});
        if (r.ms < bestMs)
        {
            bestMs = r.ms;
            best = c;
        }
    }
    best.pollEvery = 16;
    tuned_[key] = best;
    return best;
}

template SolveReport Krylov::pcgRun<double>(const CsrView&, const double*, const double*, const double*,
                                            const double*, double*, double*, double*, double, int, double*, bool,
                                            double);
template SolveReport Krylov::pcgRun<float>(const CsrView&, const double*, const float*, const float*,
                                           const double*, double*, float*, float*, double, int, double*, bool,
                                           double);

} // namespace cfx
