// krylov.hip -- CSR-stream SpMV, Jacobi-PCG, Jacobi-BiCGStab and the
// FP32-inner / FP64-refined PCG for gfx950.
//
// Per CG iteration three kernels run back to back on one stream:
//   spmv_dot : q = A p  (LDS-staged CSR-stream)        + partial p.q
//   update   : x += a p, r -= a q, z = invD r          + partial r.z, max|r|
//   direction: p = invD r + b p
// Reductions leave per-workgroup partials (grids capped at 2048 workgroups);
// a one-workgroup kernel, or the consumer kernel itself, sums them in a fixed
// order and writes the new scalars into the device-resident SolveState (no
// atomics, deterministic); the host only polls `done` between batches of
// iterations, sized from the observed convergence rate.
#include "krylov.h"

#include <algorithm>
#include <atomic>
#include <cmath>

namespace cfx
{

std::vector<int> buildRowBlocks(const std::vector<int>& rowPtr)
{
    const int n = static_cast<int>(rowPtr.size()) - 1;
    std::vector<int> b{0};
    int r = 0;
    while (r < n)
    {
        int rows = 0;
        long long nnz = 0;
        while (r < n && rows < kSpmvBlock)
        {
            const int len = rowPtr[r + 1] - rowPtr[r];
            if (rows > 0 && nnz + len > kSpmvLds) break;
            nnz += len;
            ++rows;
            ++r;
            if (nnz > kSpmvLds) break;  // one long row alone in its block
        }
        b.push_back(r);
    }
    return b;
}

std::vector<int> buildRowBlocksGrouped(const std::vector<int>& rowPtr, const std::vector<int>& group)
{
    const int n = static_cast<int>(rowPtr.size()) - 1;
    std::vector<int> b{0};
    int rows = 0;
    long long nnz = 0;
    for (int g0 = 0; g0 < n;)
    {
        int g1 = g0 + 1;
        while (g1 < n && group[g1] == group[g0]) ++g1;
        const int len = g1 - g0;
        const long long gnz = rowPtr[g1] - rowPtr[g0];
        if (len > kSpmvBlock) return {};
        if (rows > 0 && (rows + len > kSpmvBlock || nnz + gnz > kSpmvLds))
        {
            b.push_back(g0);
            rows = 0;
            nnz = 0;
        }
        rows += len;
        nnz += gnz;
        g0 = g1;
    }
    if (n > 0) b.push_back(n);
    return b;
}

void DevPattern::build(std::vector<int> rowPtrH, std::vector<int> colH, hipStream_t s, std::vector<int> blocks)
{
    static std::atomic<std::uint64_t> counter{0};
    uid = ++counter;
    n = static_cast<int>(rowPtrH.size()) - 1;
    nnz = rowPtrH.back();
    std::vector<int> diag(static_cast<std::size_t>(n), -1);
    for (int r = 0; r < n; ++r)
        for (int k = rowPtrH[r]; k < rowPtrH[r + 1]; ++k)
            if (colH[k] == r) { diag[r] = k; break; }
    if (blocks.empty()) blocks = buildRowBlocks(rowPtrH);
    nRowBlocks = static_cast<int>(blocks.size()) - 1;
    // the columns as int16 offsets from each block's first row, where all fit
    {
        std::vector<short> c16(static_cast<std::size_t>(nnz));
        bool fits = true;
        for (int b = 0; b < nRowBlocks && fits; ++b)
        {
            const int r0 = blocks[b];
            for (int k = rowPtrH[r0]; k < rowPtrH[blocks[b + 1]]; ++k)
            {
                const int o = colH[k] - r0;
                if (o < -32768 || o > 32767)
                {
                    fits = false;
                    break;
                }
                c16[k] = static_cast<short>(o);
            }
        }
        if (fits && nnz > 0) col16 = upload(c16, s);
        else col16 = DevBuf();
    }
    rowPtr = upload(rowPtrH, s);
    col = upload(colH, s);
    diagSlot = upload(diag, s);
    rowBlocks = upload(blocks, s);
    hRowPtr = std::move(rowPtrH);
    hCol = std::move(colH);
    CFX_HIP(hipStreamSynchronize(s));
}

namespace
{

constexpr int kVecGridMax = 2048;


// Row sum of the CSR-stream block `blk` for the row this thread owns.
// Products are formed once (coalesced val/col reads, gathered x) and staged
// in LDS; each row is then summed sequentially in column order from zero,
// which is exactly SparseMatrix::product's arithmetic.
template <typename T>
__device__ __forceinline__ bool spmvRowSum(const CsrView& A, const T* __restrict__ val,
                                           const T* __restrict__ x, int blk, T* lds, int& row,
                                           T& sum)
{
    const int r0 = A.rowBlocks[blk], r1 = A.rowBlocks[blk + 1];
    const int k0 = A.rowPtr[r0];
    const int nnz = A.rowPtr[r1] - k0;
    row = r0 + static_cast<int>(threadIdx.x);
    if (nnz <= kSpmvLds)
    {
        stageProducts(lds, val, A, x, k0, r0, nnz);
        __syncthreads();
        if (row < r1)
        {
            T s = T(0);
            const int e = A.rowPtr[row + 1] - k0;
            for (int k = A.rowPtr[row] - k0; k < e; ++k) s += lds[k];
            sum = s;
            return true;
        }
        return false;
    }
    if (row < r1)
    {
        T s = T(0);
        for (int k = A.rowPtr[row]; k < A.rowPtr[row + 1]; ++k) s += val[k] * x[A.col[k]];
        sum = s;
        return true;
    }
    return false;
}

template <typename T>
__global__ __launch_bounds__(kSpmvBlock) void k_spmv(CsrView A, const T* __restrict__ val,
                                                     const T* __restrict__ x, T* __restrict__ y)
{
    __shared__ T lds[kSpmvLds];
    int row;
    T s;
    if (spmvRowSum(A, val, x, xcdRemap(blockIdx.x, gridDim.x), lds, row, s)) y[row] = s;
}

// True residual of a finished solve, two kernels and no atomics: per-workgroup
// partials of max|r|, max|b|, sum r^2, sum b^2 and NaN flags over rbPerWg
// row blocks, then one workgroup reduces them into st->sum0 (max|r|/max|b|)
// and st->rho (|r|2 / |b|2).
__global__ __launch_bounds__(kSpmvBlock) void k_true_resid_part(CsrView A, int rbPerWg, const double* __restrict__ val,
                                                                const double* __restrict__ x,
                                                                const double* __restrict__ b, double* part,
                                                                int* ipart)
{
    __shared__ double lds[kSpmvLds];
    __shared__ double red[kSpmvBlock / kWave];
    __shared__ int ired[kSpmvBlock / kWave];
    const int G = gridDim.x;
    const int w = xcdRemap(blockIdx.x, G);
    double rm = 0, bm = 0, r2 = 0, b2 = 0;
    int nan = 0;
    for (int j = 0; j < rbPerWg; ++j)
    {
        const int blk = w * rbPerWg + j;
        if (blk >= A.nRowBlocks) break;
        int row;
        double s;
        if (spmvRowSum(A, val, x, blk, lds, row, s))
        {
            const double bd = b[row], rd = bd - s;
            rm = fmax(rm, fabs(rd));
            bm = fmax(bm, fabs(bd));
            r2 += rd * rd;
            b2 += bd * bd;
            nan |= (rd != rd ? 1 : 0) | (bd != bd ? 2 : 0);
        }
    }
    rm = blockMax(rm, red);
    bm = blockMax(bm, red);
    r2 = blockSum(r2, red);
    b2 = blockSum(b2, red);
    nan = blockOr(nan, ired);
    if (threadIdx.x == 0)
    {
        part[blockIdx.x] = rm;
        part[G + blockIdx.x] = bm;
        part[2 * G + blockIdx.x] = r2;
        part[3 * G + blockIdx.x] = b2;
        ipart[blockIdx.x] = nan;
    }
}

__global__ __launch_bounds__(kSpmvBlock) void k_true_resid_final(int G, const double* part, const int* ipart,
                                                                 SolveState* st)
{
    __shared__ double red[kSpmvBlock / kWave];
    __shared__ int ired[kSpmvBlock / kWave];
    const double RM = maxPartials(part, G, red);
    const double BM = maxPartials(part + G, G, red);
    const double R2 = sumPartials(part + 2 * G, G, red);
    const double B2 = sumPartials(part + 3 * G, G, red);
    const int NAN_ = orPartials(ipart, G, ired);
    if (threadIdx.x == 0)
    {
        const double binf = (NAN_ & 2) ? NAN : BM;
        st->sum0 = ((NAN_ & 1) ? NAN : RM) / (binf < 1e-300 ? 1e-300 : binf);
        st->rho = sqrt(R2) / (sqrt(B2) > 1e-300 ? sqrt(B2) : 1e-300);
    }
}

// r = b - A x over rbPerWg row blocks per workgroup, with per-workgroup
// partials of max|r|, max|b| (NaN-flagged), sum r^2, sum b^2; k_residual_final
// (one workgroup) reduces them in a fixed order and sets the state.
// mode 0: initial residual of a solve (sets initialResidual, bnorm and the
//         early-convergence exit, solveCG :167-180);
// mode 1: refinement residual (sets rnorm; converged if rnorm/bnorm < tol);
// mode 2: reductions only (sum0 = rnorm, rho = ||r||2 / ||b||2).
template <typename T>
__global__ __launch_bounds__(kSpmvBlock) void k_residual(CsrView A, int rbPerWg, const T* __restrict__ val,
                                                         const T* __restrict__ x, const T* __restrict__ b,
                                                         T* __restrict__ r, double* part, int* ipart)
{
    __shared__ T lds[kSpmvLds];
    __shared__ double red[kSpmvBlock / kWave];
    __shared__ int ired[kSpmvBlock / kWave];
    const int G = gridDim.x;
    const int w = xcdRemap(blockIdx.x, G);
    double rm = 0, bm = 0, r2 = 0, b2 = 0;
    int nan = 0;
    for (int j = 0; j < rbPerWg; ++j)
    {
        const int blk = w * rbPerWg + j;
        if (blk >= A.nRowBlocks) break;
        int row;
        T s;
        if (spmvRowSum(A, val, x, blk, lds, row, s))
        {
            const T bb = b[row];
            const T rr = bb - s;
            r[row] = rr;
            const double rd = rr, bd = bb;
            rm = fmax(rm, fabs(rd));
            bm = fmax(bm, fabs(bd));
            r2 += rd * rd;
            b2 += bd * bd;
            nan |= (rd != rd ? 1 : 0) | (bd != bd ? 2 : 0);
        }
    }
    rm = blockMax(rm, red);
    bm = blockMax(bm, red);
    r2 = blockSum(r2, red);
    b2 = blockSum(b2, red);
    nan = blockOr(nan, ired);
    if (threadIdx.x == 0)
    {
        part[w] = rm;
        part[G + w] = bm;
        part[2 * G + w] = r2;
        part[3 * G + w] = b2;
        ipart[w] = nan;
    }
}

__global__ __launch_bounds__(kSpmvBlock) void k_residual_final(int G, const double* part, const int* ipart,
                                                               SolveState* st, int mode)
{
    __shared__ double red[kSpmvBlock / kWave];
    __shared__ int ired[kSpmvBlock / kWave];
    const double RM = maxPartials(part, G, red);
    const double BM = maxPartials(part + G, G, red);
    const double R2 = sumPartials(part + 2 * G, G, red);
    const double B2 = sumPartials(part + 3 * G, G, red);
    const int NAN_ = orPartials(ipart, G, ired);
    if (threadIdx.x == 0)
    {
        const double rnorm = (NAN_ & 1) ? NAN : RM;
        const double binf = (NAN_ & 2) ? NAN : BM;
        const double l2 = sqrt(R2) / (sqrt(B2) > 1e-300 ? sqrt(B2) : 1e-300);
        if (mode == 0)
        {
            const double bnorm = (binf < 1e-300) ? 1e-300 : binf;  // std::max(binf, 1e-300)
            st->initialResidual = rnorm;
            st->rnorm = rnorm;
            st->bnorm = bnorm;
            st->iter = 0;
            st->rho = l2;
            if (rnorm / bnorm < st->tol)
            {
                st->converged = 1;
                st->done = 1;
            }
        }
        else if (mode == 1)
        {
            st->rnorm = rnorm;
            st->rho = l2;
            if (rnorm / st->bnorm < st->tol)
            {
                st->converged = 1;
                st->done = 1;
            }
        }
        else
        {
            st->sum0 = rnorm / ((binf < 1e-300) ? 1e-300 : binf);
            st->rho = l2;
        }
    }
}

// ---------------------------------------------------------------- BiCGStab
// r0 = r, p = v = 0, rho/beta of iteration 1 (solveBiCGStab :253-266)
template <typename T>
__global__ __launch_bounds__(kVecBlock) void k_bicg_init(int n, const T* __restrict__ r, T* __restrict__ r0,
                                                         T* __restrict__ p, T* __restrict__ v, const SolveState* st,
                                                         double* part)
{
    __shared__ double red[kVecBlock / kWave];
    if (st->done) return;
    double s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const T ri = r[i];
        r0[i] = ri;
        p[i] = T(0);
        v[i] = T(0);
        s += static_cast<double>(ri) * static_cast<double>(ri);
    }
    s = blockSum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kVecBlock) void k_bicg_init_final(int G, const double* part, SolveState* st)
{
    __shared__ double red[kVecBlock / kWave];
    if (st->done) return;
    const double rhoNew = sumPartials(part, G, red);
    if (threadIdx.x == 0)
    {
        st->rho = 1.0;
        st->alpha = 1.0;
        st->omega = 1.0;
        if (fabs(rhoNew) < 1e-300)
        {
            st->done = 1;
        }
        else
        {
            st->beta = (rhoNew / st->rho) * (st->alpha / st->omega);
            st->rho = rhoNew;
        }
    }
}

// The iteration kernels below solve blockIdx.y's system: component c uses the
// vectors at offset c*vs, the matrix values at c*ms, the partials at c*ps
// (doubles) and c*ips (ints) and state st[c]; a single solve is c = 0 with
// zero strides.  Converged components return at once.

// p = r + beta (p - omega v); z = p invD   (:267-272)
template <typename T>
__global__ __launch_bounds__(kVecBlock) void k_bicg_p(int n, long long vs, const T* __restrict__ invD,
                                                      const T* __restrict__ r, T* __restrict__ p,
                                                      const T* __restrict__ v, T* __restrict__ z,
                                                      const SolveState* st)
{
    const long long o = blockIdx.y * vs;
    st += blockIdx.y;
    if (st->done) return;
    invD += o;
    r += o;
    p += o;
    v += o;
    z += o;
    const T beta = static_cast<T>(st->beta), omega = static_cast<T>(st->omega);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const T pi = r[i] + beta * (p[i] - omega * v[i]);
        p[i] = pi;
        z[i] = pi * invD[i];
    }
}

// y = A x, then a dot with `w` (r0.v) or two dots (t.t, t.s) in the epilogue.
// which == 0: v = A z, denom = r0.v, alpha = rho/denom   (:273-277)
// which == 1: t = A sh, tt = t.t, ts = t.s, omega = ts/tt (:290-300)
// (per-workgroup partials over rbPerWg row blocks; the next vector kernel reduces them)
template <typename T>
__global__ __launch_bounds__(kSpmvBlock) void k_bicg_spmv(CsrView A, int rbPerWg, long long ms, long long vs, int ps,
                                                          const T* __restrict__ val, const T* __restrict__ x,
                                                          T* __restrict__ y, const T* __restrict__ w,
                                                          const SolveState* st, double* part, int which)
{
    __shared__ T lds[kSpmvLds];
    __shared__ double red[kSpmvBlock / kWave];
    st += blockIdx.y;
    if (st->done) return;
    const long long o = blockIdx.y * vs;
    val += blockIdx.y * ms;
    x += o;
    y += o;
    w += o;
    part += blockIdx.y * ps;
    const int G = gridDim.x;
    const int wg = xcdRemap(blockIdx.x, G);
    double d0 = 0, d1 = 0;
    for (int j = 0; j < rbPerWg; ++j)
    {
        const int blk = wg * rbPerWg + j;
        if (blk >= A.nRowBlocks) break;
        int row;
        T s;
        if (spmvRowSum(A, val, x, blk, lds, row, s))
        {
            y[row] = s;
            const double sd = s, wd = w[row];
            if (which == 0) d0 += wd * sd;
            else { d0 += sd * sd; d1 += sd * wd; }
        }
    }
    d0 = blockSum(d0, red);
    d1 = blockSum(d1, red);
    if (threadIdx.x == 0) { part[wg] = d0; part[G + wg] = d1; }
}

// the sum of partials[0..n) (sumPartials' fixed order) in every thread
__device__ __forceinline__ double sumPartialsAll(const double* partials, int n, double* lds, double* slot)
{
    const double v = sumPartials(partials, n, lds);
    if (threadIdx.x == 0) *slot = v;
    __syncthreads();
    return *slot;
}

// partials of k_bicg_r live past those of the SpMVs, which k_bicg_r reads
constexpr int kRPartOff = 2 * kVecGridMax;

// alpha = rho / r0.v from the first SpMV's Gs partials, reduced by every
// workgroup in the same fixed order (workgroup 0 records it; r0.v ~ 0 ends
// the solve, :273-277); x += alpha z; s = r - alpha v; sh = s invD (:278-289)
template <typename T>
__global__ __launch_bounds__(kVecBlock) void k_bicg_s(int n, long long vs, int ps, int Gs,
                                                      const T* __restrict__ invD, T* __restrict__ x,
                                                      const T* __restrict__ z, const T* __restrict__ r,
                                                      const T* __restrict__ v, T* __restrict__ s,
                                                      T* __restrict__ sh, SolveState* st, const double* part)
{
    __shared__ double red[kVecBlock / kWave];
    __shared__ double slot;
    const long long o = blockIdx.y * vs;
    st += blockIdx.y;
    if (st->done) return;
    const double D0 = sumPartialsAll(part + blockIdx.y * ps, Gs, red, &slot);
    const bool brk = fabs(D0) < 1e-300;
    const double alphaD = brk ? 0.0 : st->rho / D0;
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        st->denom = D0;
        if (brk) st->done = 1;
        else st->alpha = alphaD;
    }
    if (brk) return;
    invD += o;
    x += o;
    z += o;
    r += o;
    v += o;
    s += o;
    sh += o;
    const T alpha = static_cast<T>(alphaD);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        x[i] += alpha * z[i];
        const T si = r[i] - alpha * v[i];
        s[i] = si;
        sh[i] = si * invD[i];
    }
}

// omega = t.s / t.t from the second SpMV's partials, reduced by every
// workgroup (workgroup 0 records it; t.t ~ 0 is the breakdown that takes
// r = s and stops, :290-300); x += omega sh; r = s - omega t; ||r||inf;
// next rho = r0.r (:301-314, :263-266), partials at kRPartOff
template <typename T>
__global__ __launch_bounds__(kVecBlock) void k_bicg_r(int n, long long vs, int ps, int ips, int Gs, T* __restrict__ x,
                                                      const T* __restrict__ sh, const T* __restrict__ s,
                                                      const T* __restrict__ t, const T* __restrict__ r0,
                                                      T* __restrict__ r, SolveState* st, double* part,
                                                      int* ipart)
{
    __shared__ double red[kVecBlock / kWave];
    __shared__ int ired[kVecBlock / kWave];
    __shared__ double slot[2];
    const long long o = blockIdx.y * vs;
    st += blockIdx.y;
    if (st->done) return;
    x += o;
    sh += o;
    s += o;
    t += o;
    r0 += o;
    r += o;
    part += blockIdx.y * ps;
    ipart += blockIdx.y * ips;
    const double TT = sumPartialsAll(part, Gs, red, &slot[0]);
    const double TS = sumPartialsAll(part + Gs, Gs, red, &slot[1]);
    const bool tsmall = TT < 1e-300;
    const double omegaD = tsmall ? 1.0 : TS / TT;
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        st->tt = TT;
        st->ts = TS;
        if (tsmall) st->breakdown = 2;
        else st->omega = omegaD;
    }
    part += kRPartOff;
    const T omega = static_cast<T>(omegaD);
    double rm = 0, rr0 = 0;
    int nan = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        T ri;
        if (tsmall)
        {
            ri = s[i];
        }
        else
        {
            x[i] += omega * sh[i];
            ri = s[i] - omega * t[i];
        }
        r[i] = ri;
        const double rd = ri;
        rm = fmax(rm, fabs(rd));
        nan |= (rd != rd);
        rr0 += static_cast<double>(r0[i]) * rd;
    }
    const int G = gridDim.x;
    rm = blockMax(rm, red);
    rr0 = blockSum(rr0, red);
    nan = blockOr(nan, ired);
    if (threadIdx.x == 0) { part[blockIdx.x] = rm; part[G + blockIdx.x] = rr0; ipart[blockIdx.x] = nan; }
}

__global__ __launch_bounds__(kVecBlock) void k_bicg_r_final(int G, int ps, int ips, const double* part,
                                                            const int* ipart, SolveState* st)
{
    __shared__ double red[kVecBlock / kWave];
    __shared__ int ired[kVecBlock / kWave];
    st += blockIdx.y;
    if (st->done) return;
    part += blockIdx.y * ps + kRPartOff;
    ipart += blockIdx.y * ips;
    const bool tsmall = st->breakdown == 2;
    const double RM = maxPartials(part, G, red);
    const double RR0 = sumPartials(part + G, G, red);
    const int NAN_ = orPartials(ipart, G, ired);
    if (threadIdx.x == 0)
    {
        const int it = st->iter + 1;
        const double rnorm = NAN_ ? NAN : RM;
        st->iter = it;
        st->rnorm = rnorm;
        if (tsmall)
        {
            st->converged = rnorm / st->bnorm < st->tol;
            st->done = 1;
        }
        else if (rnorm / st->bnorm < st->tol)
        {
            st->converged = 1;
            st->done = 1;
        }
        else if (it >= st->maxIter)
        {
            st->done = 1;
        }
        else if (fabs(RR0) < 1e-300)
        {
            st->done = 1;  // rho breakdown at the top of the next iteration
        }
        else
        {
            st->beta = (RR0 / st->rho) * (st->alpha / st->omega);
            st->rho = RR0;
        }
    }
}

// ---------------------------------------------------------------- helpers
__global__ void k_invD(CsrView A, const double* __restrict__ val, double* __restrict__ invD,
                       int bicgGuard)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const int k = A.diagSlot[i];
    const double d = k >= 0 ? val[k] : 0.0;
    invD[i] = bicgGuard ? ((d > 1e-300) ? 1.0 / d : 1.0) : ((fabs(d) > 1e-300) ? 1.0 / d : 1.0);
}

__global__ void k_cast(const double* __restrict__ in, float* __restrict__ out, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = static_cast<float>(in[i]);
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// x += d (FP32 correction folded into the FP64 iterate)
__global__ void k_add_float(int n, const float* __restrict__ d, double* __restrict__ x)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x[i] += static_cast<double>(d[i]);
}

int vecGridFor(long long n) { return std::max(1, std::min(kVecGridMax, divUp(n, 4LL * kVecBlock))); }

// One BiCGStab iteration (solveBiCGStab :267-314) of `nc` systems on the
// pattern A: vectors w = {r, r0, p, v, z, s, sh, t} and x at stride vs,
// values at stride ms, partials at strides ps / ips, states st[0..nc).
template <typename T>
void launchBicgIteration(hipStream_t s, const CsrView& A, int Gv, int Gs, int rbPerWg, int nc, long long vs,
                         long long ms, int ps, int ips, const T* val, const T* invD, T* x, T* const* w,
                         SolveState* st, double* part, int* ipart)
{
    T *r = w[0], *r0 = w[1], *p = w[2], *v = w[3], *z = w[4], *sv = w[5], *sh = w[6], *t = w[7];
    const int n = A.n;
    const dim3 gv(Gv, nc), gs(Gs, nc), g1(1, nc);
    k_bicg_p<T><<<gv, kVecBlock, 0, s>>>(n, vs, invD, r, p, v, z, st);
    k_bicg_spmv<T><<<gs, kSpmvBlock, 0, s>>>(A, rbPerWg, ms, vs, ps, val, z, v, r0, st, part, 0);
    k_bicg_s<T><<<gv, kVecBlock, 0, s>>>(n, vs, ps, Gs, invD, x, z, r, v, sv, sh, st, part);
    k_bicg_spmv<T><<<gs, kSpmvBlock, 0, s>>>(A, rbPerWg, ms, vs, ps, val, sh, t, sv, st, part, 1);
    k_bicg_r<T><<<gv, kVecBlock, 0, s>>>(n, vs, ps, ips, Gs, x, sh, sv, t, r0, r, st, part, ipart);
    k_bicg_r_final<<<g1, kVecBlock, 0, s>>>(Gv, ps, ips, part, ipart, st);
}

} // namespace

void computeInvD(Context& ctx, const CsrView& A, const double* val, double* invD, bool bicgGuard)
{
    k_invD<<<divUp(A.n, 256), 256, 0, ctx.stream()>>>(A, val, invD, bicgGuard ? 1 : 0);
    CFX_CHECK_LAUNCH();
}

void castToFloat(Context& ctx, const double* in, float* out, long long n)
{
    k_cast<<<std::max(1, std::min(4096, divUp(n, 256))), 256, 0, ctx.stream()>>>(in, out, n);
    CFX_CHECK_LAUNCH();
}

Krylov::Krylov(Context& ctx, int n, int maxRowBlocks) : ctx_(ctx), n_(n)
{
    maxBlocks_ = std::max(maxRowBlocks, kVecGridMax);
    const std::size_t d = sizeof(double) * static_cast<std::size_t>(n), f = sizeof(float) * static_cast<std::size_t>(n);
    state_.alloc(3 * sizeof(SolveState));  // slots: refinement, Krylov, PCG state
    CFX_HIP(hipMemsetAsync(state_.p, 0, 3 * sizeof(SolveState), ctx.stream()));
    partials_.alloc(sizeof(double) * 4 * maxBlocks_ + sizeof(int) * maxBlocks_);
    for (DevBuf* b : {&r_, &p_, &q_, &r0_, &v_, &s_, &sh_, &t_, &z_, &invD_, &invDb_}) b->alloc(d);
    for (DevBuf* b : {&r32_, &p32_, &q32_, &d32_, &b32_, &invD32_}) b->alloc(f);
    CFX_HIP(hipEventCreate(&ev0_));
    CFX_HIP(hipEventCreate(&ev1_));
    ctx.sync();
}

Krylov::~Krylov()
{
    for (hipEvent_t e : probeEv_) (void)hipEventDestroy(e);
    if (hostC_) (void)hipHostFree(hostC_);
    if (fcgExec_) (void)hipGraphExecDestroy(fcgExec_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
}

void Krylov::poll()
{
    CFX_HIP(hipMemcpyAsync(ctx_.hostState(), state_.p, 2 * sizeof(SolveState), hipMemcpyDeviceToHost,
                           ctx_.stream()));
    CFX_HIP(hipStreamSynchronize(ctx_.stream()));
}

// Iterations to launch before the next poll: extrapolate the observed
// geometric residual decay to the tolerance, undershooting by 10%.
int Krylov::nextBatch(int done, double r0, double rNow, double tolAbs, int maxIter)
{
    int b = minBatch;
    if (done > 0 && rNow > 0 && r0 > 0 && rNow < r0 && tolAbs > 0)
    {
        const double rate = std::log(rNow / r0) / done;
        const double need = std::log(tolAbs / rNow) / rate;
        if (std::isfinite(need)) b = static_cast<int>(0.9 * need);
    }
    b = std::max(minBatch, std::min(maxBatch, b));
    return std::max(1, std::min(b, maxIter - done));
}

void Krylov::spmv(const CsrView& A, const double* val, const double* x, double* y)
{
    k_spmv<double><<<A.nRowBlocks, kSpmvBlock, 0, ctx_.stream()>>>(A, val, x, y);
    CFX_CHECK_LAUNCH();
}

void Krylov::spmv32(const CsrView& A, const float* val, const float* x, float* y)
{
    k_spmv<float><<<A.nRowBlocks, kSpmvBlock, 0, ctx_.stream()>>>(A, val, x, y);
    CFX_CHECK_LAUNCH();
}

void Krylov::trueResidual(const CsrView& A, const double* val, const double* b, const double* x,
                          double& infRel, double& l2Rel)
{
    hipStream_t s = ctx_.stream();
    SolveState* st = state_.as<SolveState>();
    double* part = partials_.as<double>();
    int* ipart = reinterpret_cast<int*>(part + 4 * maxBlocks_);
    int G = std::min(A.nRowBlocks, kVecGridMax);
    const int rbPerWg = divUp(A.nRowBlocks, G);
    G = std::max(1, divUp(A.nRowBlocks, rbPerWg));
    k_true_resid_part<<<G, kSpmvBlock, 0, s>>>(A, rbPerWg, val, x, b, part, ipart);
    k_true_resid_final<<<1, kSpmvBlock, 0, s>>>(G, part, ipart, st);
    CFX_CHECK_LAUNCH();
    poll();
    infRel = ctx_.hostState()[0].sum0;
    l2Rel = ctx_.hostState()[0].rho;
}

template <typename T>
SolveReport Krylov::bicgRun(const CsrView& A, const T* val, const T* invD, const T* b, T* x, double tol, int maxIter,
                            T* const* w)
{
    // w: r, r0, p, v, z, s, sh, t
    hipStream_t s = ctx_.stream();
    SolveState h{};
    h.tol = tol;
    h.maxIter = maxIter;
    SolveState* st = state_.as<SolveState>() + 1;
    CFX_HIP(hipMemcpyAsync(st, &h, offsetof(SolveState, ticket), hipMemcpyHostToDevice, s));
    double* part = partials_.as<double>();
    int* ipart = reinterpret_cast<int*>(part + 4 * maxBlocks_);
    const int Gv = vecGridFor(n_);
    int Gs = std::min(A.nRowBlocks, kVecGridMax);
    const int rbPerWg = divUp(A.nRowBlocks, Gs);
    Gs = std::max(1, divUp(A.nRowBlocks, rbPerWg));
    T *r = w[0], *r0 = w[1], *p = w[2], *v = w[3], *z = w[4], *sv = w[5], *sh = w[6], *t = w[7];
    k_residual<T><<<Gs, kSpmvBlock, 0, s>>>(A, rbPerWg, val, x, b, r, part, ipart);
    k_residual_final<<<1, kSpmvBlock, 0, s>>>(Gs, part, ipart, st, 0);
    k_bicg_init<T><<<Gv, kVecBlock, 0, s>>>(n_, r, r0, p, v, st, part);
    k_bicg_init_final<<<1, kVecBlock, 0, s>>>(Gv, part, st);
    CFX_CHECK_LAUNCH();
    poll();
    SolveState* hs = ctx_.hostState() + 1;
    const double rInit = hs->initialResidual, bnorm = hs->bnorm;
    int launched = 0;
    while (!hs->done && launched < maxIter)
    {
        const int batch = nextBatch(hs->iter, rInit, hs->rnorm, tol * bnorm, maxIter);
        for (int k = 0; k < batch; ++k)
        {
            launchBicgIteration<T>(s, A, Gv, Gs, rbPerWg, 1, 0, 0, 0, 0, val, invD, x, w, st, part, ipart);
        }
        CFX_CHECK_LAUNCH();
        launched += batch;
        poll();
    }
    SolveReport rep;
    rep.iterations = hs->iter;
    rep.initialResidual = hs->initialResidual;
    rep.finalResidual = hs->iter > 0 ? hs->rnorm : hs->initialResidual;
    rep.converged = hs->converged;
    return rep;
}

SolveReport Krylov::bicgstab(const CsrView& A, const double* val, const double* b, double* x, double tol,
                             int maxIter)
{
    hipStream_t s = ctx_.stream();
    CFX_HIP(hipEventRecord(ev0_, s));
    computeInvD(ctx_, A, val, invDb_.as<double>(), true);
    double* w[8] = {r_.as<double>(), r0_.as<double>(), p_.as<double>(), v_.as<double>(),
                    z_.as<double>(), s_.as<double>(), sh_.as<double>(), t_.as<double>()};
    SolveReport rep = bicgRun<double>(A, val, invDb_.as<double>(), b, x, tol, maxIter, w);
    CFX_HIP(hipEventRecord(ev1_, s));
    CFX_HIP(hipEventSynchronize(ev1_));
    float ms = 0;
    CFX_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    lastMs_ = rep.ms = ms;
    return rep;
}

void Krylov::reserveComponents()
{
    if (wc_.p) return;
    const std::size_t nn = static_cast<std::size_t>(n_);
    const int ps = 4 * maxBlocks_, ips = maxBlocks_;
    wc_.alloc(sizeof(double) * 9 * kMaxComponents * nn);
    // states: the component solves' [0, 3), the mixed solve's refinement states [3, 6)
    stC_.alloc(sizeof(SolveState) * 2 * kMaxComponents);
    partC_.alloc((sizeof(double) * ps + sizeof(int) * ips) * kMaxComponents);
    CFX_HIP(hipHostMalloc(reinterpret_cast<void**>(&hostC_), sizeof(SolveState) * 2 * kMaxComponents,
                          hipHostMallocDefault));
}

// BiCGStab (bicgRun's arithmetic) of the components c < ncomp with run[c]:
// values at val + c valStride, invD, b, x and the work vectors w[j] as [c][n];
// tolerance and iteration budget per component.  The others are left as they
// are (their state starts done, so their workgroups return at once).
template <typename T>
void Krylov::bicgRunComponents(const CsrView& A, const T* val, long long valStride, const T* invD, const T* b, T* x,
                               int ncomp, const double* tol, const int* maxIter, const bool* run, T* const* w,
                               SolveReport* reps)
{
    hipStream_t s = ctx_.stream();
    const long long nn = n_;
    const int ps = 4 * maxBlocks_, ips = maxBlocks_;
    SolveState* st = stC_.as<SolveState>();
    double* part = partC_.as<double>();
    int* ipart = reinterpret_cast<int*>(part + static_cast<std::size_t>(ps) * kMaxComponents);
    int Gs = std::min(A.nRowBlocks, kVecGridMax);
    const int rbPerWg = divUp(A.nRowBlocks, Gs);
    Gs = std::max(1, divUp(A.nRowBlocks, rbPerWg));
    const int Gv = vecGridFor(n_);
    int budget = 0;
    for (int c = 0; c < ncomp; ++c)
    {
        SolveState h{};
        h.tol = tol[c];
        h.maxIter = maxIter[c];
        h.done = run[c] ? 0 : 1;
        CFX_HIP(hipMemcpyAsync(st + c, &h, offsetof(SolveState, ticket), hipMemcpyHostToDevice, s));
        if (!run[c]) continue;
        budget = std::max(budget, maxIter[c]);
        // each component starts as bicgRun starts a single solve (:244-266)
        const long long o = c * nn;
        double* pc = part + static_cast<std::size_t>(c) * ps;
        int* ic = ipart + static_cast<std::size_t>(c) * ips;
        k_residual<T><<<Gs, kSpmvBlock, 0, s>>>(A, rbPerWg, val + c * valStride, x + o, b + o, w[0] + o, pc, ic);
        k_residual_final<<<1, kSpmvBlock, 0, s>>>(Gs, pc, ic, st + c, 0);
        k_bicg_init<T><<<Gv, kVecBlock, 0, s>>>(n_, w[0] + o, w[1] + o, w[2] + o, w[3] + o, st + c, pc);
        k_bicg_init_final<<<1, kVecBlock, 0, s>>>(Gv, pc, st + c);
    }
    CFX_CHECK_LAUNCH();
    auto pollC = [&] {
        CFX_HIP(hipMemcpyAsync(hostC_, st, sizeof(SolveState) * ncomp, hipMemcpyDeviceToHost, s));
        CFX_HIP(hipStreamSynchronize(s));
    };
    pollC();
    double rInit[kMaxComponents], bnorm[kMaxComponents];
    for (int c = 0; c < ncomp; ++c)
    {
        rInit[c] = hostC_[c].initialResidual;
        bnorm[c] = hostC_[c].bnorm;
    }
    int launched = 0;
    for (;;)
    {
        // launch enough iterations for the slowest component still running
        int batch = 0;
        for (int c = 0; c < ncomp; ++c)
            if (!hostC_[c].done)
                batch = std::max(batch, nextBatch(hostC_[c].iter, rInit[c], hostC_[c].rnorm, tol[c] * bnorm[c],
                                                  maxIter[c]));
        batch = std::min(batch, budget - launched);
        if (batch <= 0) break;
        for (int k = 0; k < batch; ++k)
            launchBicgIteration<T>(s, A, Gv, Gs, rbPerWg, ncomp, nn, valStride, ps, ips, val, invD, x, w, st, part,
                                   ipart);
        CFX_CHECK_LAUNCH();
        launched += batch;
        pollC();
    }
    for (int c = 0; c < ncomp; ++c)
    {
        if (!run[c]) continue;
        const SolveState& hs = hostC_[c];
        SolveReport& rep = reps[c];
        rep = SolveReport{};
        rep.iterations = hs.iter;
        rep.initialResidual = hs.initialResidual;
        rep.finalResidual = hs.iter > 0 ? hs.rnorm : hs.initialResidual;
        rep.converged = hs.converged;
    }
}

void Krylov::bicgstabComponents(const CsrView& A, const double* val, long long valStride, const double* b, double* x,
                                int ncomp, double tol, int maxIter, SolveReport* reps)
{
    if (ncomp < 1 || ncomp > kMaxComponents)
        throw std::invalid_argument("bicgstabComponents: 1 to 3 components");
    if (A.n != n_) throw std::invalid_argument("bicgstabComponents: pattern size differs from the solver's");
    hipStream_t s = ctx_.stream();
    CFX_HIP(hipEventRecord(ev0_, s));
    reserveComponents();
    const long long nn = n_;
    // work vector j of component c at wc + (j * ncomp + c) n: each vector kind is [c][n] with stride n
    double* wc = wc_.as<double>();
    double* w[8];
    for (int j = 0; j < 8; ++j) w[j] = wc + j * ncomp * nn;
    double* invD = wc + 8 * ncomp * nn;
    double tols[kMaxComponents];
    int its[kMaxComponents];
    bool run[kMaxComponents];
    for (int c = 0; c < ncomp; ++c)
    {
        computeInvD(ctx_, A, val + c * valStride, invD + c * nn, true);
        tols[c] = tol;
        its[c] = maxIter;
        run[c] = true;
    }
    bicgRunComponents<double>(A, val, valStride, invD, b, x, ncomp, tols, its, run, w, reps);
    CFX_HIP(hipEventRecord(ev1_, s));
    CFX_HIP(hipEventSynchronize(ev1_));
    float ms = 0;
    CFX_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    lastMs_ = ms;
    for (int c = 0; c < ncomp; ++c) reps[c].ms = ms;
}

void Krylov::bicgstabMixedComponents(const CsrView& A, const double* val, long long valStride, const double* b,
                                     double* x, int ncomp, double tol, int maxIter, SolveReport* reps,
                                     double innerTol)
{
    if (ncomp < 1 || ncomp > kMaxComponents)
        throw std::invalid_argument("bicgstabMixedComponents: 1 to 3 components");
    if (A.n != n_) throw std::invalid_argument("bicgstabMixedComponents: pattern size differs from the solver's");
    hipStream_t s = ctx_.stream();
    CFX_HIP(hipEventRecord(ev0_, s));
    reserveComponents();
    const long long nn = n_;
    const std::size_t nnz = static_cast<std::size_t>(A.nnz);
    if (!wc32_.p)
    {
        wc32_.alloc(sizeof(float) * 11 * kMaxComponents * static_cast<std::size_t>(nn));
        // FP64 residuals [c][n], then each component's x before its last correction [c][n]
        r64c_.alloc(sizeof(double) * 2 * kMaxComponents * static_cast<std::size_t>(nn));
    }
    if (!val32c_.p || val32c_.bytes < sizeof(float) * kMaxComponents * nnz)
        val32c_.alloc(sizeof(float) * kMaxComponents * nnz);
    // FP32: 8 Krylov vectors, invD, the correction d and its right-hand side, each [c][n]
    float* wf = wc32_.as<float>();
    float* w32[8];
    for (int j = 0; j < 8; ++j) w32[j] = wf + j * ncomp * nn;
    float *invD32 = wf + 8 * ncomp * nn, *d = wf + 9 * ncomp * nn, *b32 = wf + 10 * ncomp * nn;
    float* val32 = val32c_.as<float>();
    double* wc = wc_.as<double>();
    double* invD64 = wc + 8 * ncomp * nn;
    double* r64 = r64c_.as<double>();
    double* xPrev = r64 + kMaxComponents * nn;
    double* part = partC_.as<double>();
    const int ps = 4 * maxBlocks_, ips = maxBlocks_;
    int* ipart = reinterpret_cast<int*>(part + static_cast<std::size_t>(ps) * kMaxComponents);
    SolveState* st = stC_.as<SolveState>() + kMaxComponents;  // refinement states
    SolveState* hostR = hostC_ + kMaxComponents;
    int Gs = std::min(A.nRowBlocks, kVecGridMax);
    const int rbPerWg = divUp(A.nRowBlocks, Gs);
    Gs = std::max(1, divUp(A.nRowBlocks, rbPerWg));
    const int Gv = vecGridFor(n_);
    for (int c = 0; c < ncomp; ++c)
    {
        // as bicgstabMixed: invD (BiCGStab guard), its FP32 copy and the FP32 values
        computeInvD(ctx_, A, val + c * valStride, invD64 + c * nn, true);
        castToFloat(ctx_, invD64 + c * nn, invD32 + c * nn, n_);
        castToFloat(ctx_, val + c * valStride, val32 + c * nnz, A.nnz);
    }
    int its[kMaxComponents] = {0, 0, 0}, outer[kMaxComponents] = {0, 0, 0};
    double last[kMaxComponents] = {INFINITY, INFINITY, INFINITY};
    bool active[kMaxComponents], fallback[kMaxComponents] = {false, false, false};
    bool restore[kMaxComponents] = {false, false, false};
    for (int c = 0; c < ncomp; ++c)
    {
        reps[c] = SolveReport{};
        active[c] = true;
    }
    for (;;)
    {
        // true FP64 residual r = b - A x of every running component and the reference's stop test on it
        for (int c = 0; c < ncomp; ++c)
        {
            if (!active[c]) continue;
            SolveState h{};
            h.tol = tol;
            if (outer[c] == 0)
                CFX_HIP(hipMemcpyAsync(st + c, &h, offsetof(SolveState, ticket), hipMemcpyHostToDevice, s));
            double* pc = part + static_cast<std::size_t>(c) * ps;
            int* ic = ipart + static_cast<std::size_t>(c) * ips;
            k_residual<double><<<Gs, kSpmvBlock, 0, s>>>(A, rbPerWg, val + c * valStride, x + c * nn, b + c * nn,
                                                         r64 + c * nn, pc, ic);
            k_residual_final<<<1, kSpmvBlock, 0, s>>>(Gs, pc, ic, st + c, outer[c] == 0 ? 0 : 1);
        }
        CFX_CHECK_LAUNCH();
        CFX_HIP(hipMemcpyAsync(hostR, st, sizeof(SolveState) * ncomp, hipMemcpyDeviceToHost, s));
        CFX_HIP(hipStreamSynchronize(s));
        SolveState o[kMaxComponents];
        for (int c = 0; c < ncomp; ++c) o[c] = hostR[c];
        double inner[kMaxComponents];
        int budget[kMaxComponents];
        bool solve[kMaxComponents];
        bool any = false;
        for (int c = 0; c < ncomp; ++c)
        {
            solve[c] = false;
            inner[c] = tol;
            budget[c] = 0;
            if (!active[c]) continue;
            if (outer[c] == 0) reps[c].initialResidual = o[c].initialResidual;
            reps[c].finalResidual = o[c].rnorm;
            reps[c].converged = o[c].converged;
            if (o[c].converged || its[c] >= maxIter || (outer[c] == 0 && !std::isfinite(o[c].rnorm)))
            {
                active[c] = false;
                continue;
            }
            if (outer[c] > 0 && !(o[c].rnorm < 0.5 * last[c]))
            {
                // the FP32 correction no longer halves the FP64 residual: finish in FP64, from the x before
                // it when it made the residual worse or not finite
                fallback[c] = true;
                restore[c] = !(o[c].rnorm < last[c]);
                active[c] = false;
                continue;
            }
            last[c] = o[c].rnorm;
            // FP32 BiCGStab on A d = r from d = 0, to innerTol relative (or to the overall tolerance)
            castToFloat(ctx_, r64 + c * nn, b32 + c * nn, n_);
            k_fill<float><<<Gv, kVecBlock, 0, s>>>(d + c * nn, 0.0f, n_);
            inner[c] = std::max(innerTol, 0.5 * tol * o[c].bnorm / o[c].rnorm);
            budget[c] = maxIter - its[c];
            solve[c] = any = true;
        }
        if (!any) break;
        SolveReport in[kMaxComponents];
        bicgRunComponents<float>(A, val32, static_cast<long long>(nnz), invD32, b32, d, ncomp, inner, budget, solve,
                                 w32, in);
        for (int c = 0; c < ncomp; ++c)
        {
            if (!solve[c]) continue;
            its[c] += in[c].iterations;
            if (!std::isfinite(in[c].finalResidual))
            {
                fallback[c] = true;  // the FP32 solve itself failed: x is left as it is and finished in FP64
                active[c] = false;
                continue;
            }
            CFX_HIP(hipMemcpyAsync(xPrev + c * nn, x + c * nn, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
            k_add_float<<<Gv, kVecBlock, 0, s>>>(n_, d + c * nn, x + c * nn);
            ++outer[c];
            if (in[c].iterations == 0) active[c] = false;  // nothing more the FP32 solve can do
        }
        CFX_CHECK_LAUNCH();
    }
    bool anyFallback = false;
    int fbBudget[kMaxComponents];
    double tols[kMaxComponents];
    for (int c = 0; c < ncomp; ++c)
    {
        fallback[c] = fallback[c] && its[c] < maxIter;
        fbBudget[c] = maxIter - its[c];
        tols[c] = tol;
        anyFallback |= fallback[c];
        if (fallback[c] && restore[c])
            CFX_HIP(hipMemcpyAsync(x + c * nn, xPrev + c * nn, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
    }
    if (anyFallback)
    {
        double* w64[8];
        for (int j = 0; j < 8; ++j) w64[j] = wc + j * ncomp * nn;
        SolveReport f64[kMaxComponents];
        bicgRunComponents<double>(A, val, valStride, invD64, b, x, ncomp, tols, fbBudget, fallback, w64, f64);
        for (int c = 0; c < ncomp; ++c)
            if (fallback[c])
            {
                its[c] += f64[c].iterations;
                reps[c].finalResidual = f64[c].finalResidual;
                reps[c].converged = f64[c].converged;
                reps[c].precondFallbacks = 1;
            }
    }
    CFX_HIP(hipEventRecord(ev1_, s));
    CFX_HIP(hipEventSynchronize(ev1_));
    float ms = 0;
    CFX_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    lastMs_ = ms;
    for (int c = 0; c < ncomp; ++c)
    {
        reps[c].iterations = its[c];
        reps[c].outerIterations = outer[c];
        reps[c].ms = ms;
    }
}

SolveReport Krylov::bicgstabMixed(const CsrView& A, const double* val, const double* b, double* x, double tol,
                                  int maxIter, double innerTol)
{
    hipStream_t s = ctx_.stream();
    CFX_HIP(hipEventRecord(ev0_, s));
    const std::size_t f = sizeof(float) * static_cast<std::size_t>(n_);
    if (!bf_[0].p)
        for (DevBuf& v : bf_) v.alloc(f);
    if (!val32b_.p || val32b_.bytes < sizeof(float) * static_cast<std::size_t>(A.nnz))
        val32b_.alloc(sizeof(float) * static_cast<std::size_t>(A.nnz));
    computeInvD(ctx_, A, val, invDb_.as<double>(), true);
    castToFloat(ctx_, invDb_.as<double>(), invD32_.as<float>(), n_);
    castToFloat(ctx_, val, val32b_.as<float>(), A.nnz);
    float* w[8] = {r32_.as<float>(), bf_[0].as<float>(), p32_.as<float>(), bf_[1].as<float>(),
                   bf_[2].as<float>(), bf_[3].as<float>(), bf_[4].as<float>(), bf_[5].as<float>()};
    float *d = d32_.as<float>(), *b32 = b32_.as<float>();
    double* r64 = r_.as<double>();
    double* part = partials_.as<double>();
    int* ipart = reinterpret_cast<int*>(part + 4 * maxBlocks_);
    int Gs = std::min(A.nRowBlocks, kVecGridMax);
    const int rbPerWg = divUp(A.nRowBlocks, Gs);
    Gs = std::max(1, divUp(A.nRowBlocks, rbPerWg));
    const int Gv = vecGridFor(n_);
    SolveState* st0 = state_.as<SolveState>();
    SolveState h{};
    h.tol = tol;
    CFX_HIP(hipMemcpyAsync(st0, &h, offsetof(SolveState, ticket), hipMemcpyHostToDevice, s));
    SolveReport rep;
    int its = 0, outer = 0;
    double last = INFINITY;
    double* xPrev = q_.as<double>();  // x before the last correction
    bool fallback = false, restore = false;
    for (;;)
    {
        // true FP64 residual r = b - A x and the reference's stop test on it
        k_residual<double><<<Gs, kSpmvBlock, 0, s>>>(A, rbPerWg, val, x, b, r64, part, ipart);
        k_residual_final<<<1, kSpmvBlock, 0, s>>>(Gs, part, ipart, st0, outer == 0 ? 0 : 1);
        CFX_CHECK_LAUNCH();
        poll();
        const SolveState& o = ctx_.hostState()[0];
        if (outer == 0) rep.initialResidual = o.initialResidual;
        rep.finalResidual = o.rnorm;
        rep.converged = o.converged;
        if (o.converged || its >= maxIter || (outer == 0 && !std::isfinite(o.rnorm))) break;
        if (outer > 0 && !(o.rnorm < 0.5 * last))
        {
            // the FP32 correction no longer halves the FP64 residual: finish in FP64, from the x before it
            // when it made the residual worse or not finite
            fallback = true;
            restore = !(o.rnorm < last);
            break;
        }
        last = o.rnorm;
        // FP32 BiCGStab on A d = r from d = 0, to innerTol relative (or to the overall tolerance)
        castToFloat(ctx_, r64, b32, n_);
        k_fill<float><<<Gv, kVecBlock, 0, s>>>(d, 0.0f, n_);
        const double inner = std::max(innerTol, 0.5 * tol * o.bnorm / o.rnorm);
        const SolveReport in = bicgRun<float>(A, val32b_.as<float>(), invD32_.as<float>(), b32, d, inner,
                                              maxIter - its, w);
        its += in.iterations;
        if (!std::isfinite(in.finalResidual))
        {
            fallback = true;  // the FP32 solve itself failed: x is left as it is and finished in FP64
            break;
        }
        CFX_HIP(hipMemcpyAsync(xPrev, x, sizeof(double) * n_, hipMemcpyDeviceToDevice, s));
        k_add_float<<<Gv, kVecBlock, 0, s>>>(n_, d, x);
        ++outer;
        if (in.iterations == 0) break;  // nothing more the FP32 solve can do (breakdown at its start)
    }
    if (fallback && its < maxIter)
    {
        if (restore) CFX_HIP(hipMemcpyAsync(x, xPrev, sizeof(double) * n_, hipMemcpyDeviceToDevice, s));
        double* w64[8] = {r_.as<double>(), r0_.as<double>(), p_.as<double>(), v_.as<double>(),
                          z_.as<double>(), s_.as<double>(), sh_.as<double>(), t_.as<double>()};
        const SolveReport f64 = bicgRun<double>(A, val, invDb_.as<double>(), b, x, tol, maxIter - its, w64);
        its += f64.iterations;
        rep.finalResidual = f64.finalResidual;
        rep.converged = f64.converged;
        rep.precondFallbacks = 1;
    }
    rep.iterations = its;
    rep.outerIterations = outer;
    CFX_HIP(hipEventRecord(ev1_, s));
    CFX_HIP(hipEventSynchronize(ev1_));
    float ms = 0;
    CFX_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    lastMs_ = rep.ms = ms;
    return rep;
}

} // namespace cfx
