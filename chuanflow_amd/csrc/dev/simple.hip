// simple.hip -- fvMatrix pass loop, SIMPLE and PISO on the device.
//
// The host sequences device kernels exactly as the reference drivers order
// their field operations; every field stays in HBM for the whole run.  Per
// outer iteration the host reads back one small block of scalars (the
// continuity errors and residual norms the stop logic needs); the linear
// solvers poll their own convergence flags.
#include "simple.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace cfx
{
namespace
{

constexpr int B = 256;
inline int grid(long long n) { return std::max(1, divUp(n, B)); }
inline int gridCap(long long n) { return std::max(1, std::min(1024, divUp(n, B))); }

// std::max / std::min: a NaN second argument is dropped
__device__ __forceinline__ double smax(double a, double b) { return a < b ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return b < a ? b : a; }

// slots of the per-iteration scalar block (each reduction result is
// followed by its ticket)
enum Slot
{
    kContErr = 0,      // +1 ticket
    kErrAfter = 2,     // +3
    kPredCont = 4,     // +5
    kPStep = 6,
    kPScale = 7,
    kAP = 8,
    kPDelta = 9,
    kUMax = 10,
    kDUMax = 11,
    kNonFinite = 12,
    kDiagBad = 13,
    kTicketA = 16,     // ticket words for the kernels below
    kTicketB = 17,
    kTicketC = 18,
    kNumSlots = 32
};

__device__ __forceinline__ unsigned int* ticketOf(double* sc, int slot)
{
    return reinterpret_cast<unsigned int*>(sc + slot);
}

// volScalarField / volVectorField::linearise
__global__ void k_linearise(int nI, int nB, int nF, int ncomp, const double* __restrict__ Sf,
                            const double* __restrict__ d, const int* __restrict__ type,
                            const double* __restrict__ value, double* __restrict__ linc, double* __restrict__ link,
                            int* err)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nB) return;
    const int f = nI + b;
    const double sx = Sf[f], sy = Sf[nF + f], sz = Sf[2 * nF + f];
    for (int c = 0; c < ncomp; ++c)
    {
        const double v = value[b * ncomp + c];
        double cc = 0, kk = 0;
        switch (type[b])
        {
        case 0: cc = 0; kk = v; break;
        case 1:
        {
            // normalDistance (volFields.cpp:14-21): d . (Sf / |Sf|)
            const double a = sqrt(sx * sx + sy * sy + sz * sz);
            const double nx = sx / a, ny = sy / a, nz = sz / a;
            cc = 1;
            kk = v * (d[f] * nx + d[nF + f] * ny + d[2 * nF + f] * nz);
            break;
        }
        case 2: cc = 1; kk = 0; break;
        case 3:
            if (ncomp == 1) { cc = 1; kk = 0; }
            else
            {
                // dominantAxis (volFields.cpp:23-38) of Sf / |Sf|
                const double a = sqrt(sx * sx + sy * sy + sz * sz);
                const double ax = fabs(sx / a), ay = fabs(sy / a), az = fabs(sz / a);
                int k;
                if (ax >= ay && ax >= az) k = ax < 1.0 - 1e-6 ? -1 : 0;
                else if (ay >= az) k = ay < 1.0 - 1e-6 ? -1 : 1;
                else k = az < 1.0 - 1e-6 ? -1 : 2;
                if (k < 0) *err = 1;
                cc = (c == k) ? 0.0 : 1.0;
                kk = 0;
            }
            break;
        default: *err = 2;
        }
        linc[b * ncomp + c] = cc;
        link[b * ncomp + c] = kk;
    }
}

// rAU = V / max(aDiag, 1e-300); flags a component diagonal that differs
// (simpleSolver.cpp:123-144, pisoSolver.cpp:78-101)
__global__ void k_rau(int n, const double* __restrict__ diag, const double* __restrict__ V, double* __restrict__ rAU,
                      double* sc)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double d0 = diag[c];
    for (int k = 1; k < 3; ++k)
    {
        const double dk = diag[k * n + c];
        if (fabs(d0 - dk) > 1e-10 * (fabs(d0) + fabs(dk) + 1.0)) sc[kDiagBad] = 1.0;
    }
    rAU[c] = V[c] / smax(d0, 1e-300);
}

// Hfree = (1 - relaxU) UOld + relaxU x   (simpleSolver.cpp:158-172)
__global__ void k_hfree(int n, double relaxU, const double* __restrict__ UOld, const double* __restrict__ x,
                        double* __restrict__ H)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    H[i] = (1.0 - relaxU) * UOld[i] + relaxU * x[i];
}

// adjustPhi (simpleSolver.cpp:211-247), in two parallel passes.  Each patch's
// flux sum is one workgroup's: lane t adds faces t, t + B, ... in order, then a
// fixed tree over the lanes (a sum of the same terms as the reference's face-order
// loop, grouped differently).
__global__ __launch_bounds__(B) void k_patch_sums(const long long* __restrict__ pStart,
                                                  const long long* __restrict__ pSize, const double* __restrict__ phi,
                                                  double* __restrict__ psum)
{
    __shared__ double red[B / kWave];
    const int pi = blockIdx.x;
    const long long st = pStart[pi], sz = pSize[pi];
    double s = 0.0;
    for (long long f = threadIdx.x; f < sz; f += B) s += phi[st + f];
    s = blockSum(s, red);
    if (threadIdx.x == 0) psum[pi] = s;
}
// ... then the net and free sums in patch order (the reference's), and every
// face of a patch whose velocity is not fixed scaled by (free - net) / free
__global__ void k_adjust_phi(int nPatch, const long long* __restrict__ pStart, const long long* __restrict__ pSize,
                             const int* __restrict__ fixedU, const double* __restrict__ psum, double* __restrict__ phi)
{
    double netSum = 0, freeSum = 0;
    for (int pi = 0; pi < nPatch; ++pi)
    {
        netSum += psum[pi];
        if (!fixedU[pi]) freeSum += psum[pi];
    }
    if (!(fabs(freeSum) > 1e-30)) return;
    const double scale = (freeSum - netSum) / freeSum;
    for (int pi = 0; pi < nPatch; ++pi)
    {
        if (fixedU[pi]) continue;
        const long long st = pStart[pi], sz = pSize[pi];
        for (long long f = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; f < sz;
             f += static_cast<long long>(gridDim.x) * blockDim.x)
            phi[st + f] *= scale;
    }
}

// pStep = max|pNew - p|, pScale = max|p|, then aP (simpleSolver.cpp:303-313)
__global__ __launch_bounds__(B) void k_pstep(int n, const double* __restrict__ pNew, const double* __restrict__ p,
                                             double* sc, double* part, double relaxP, double pRelaxStart)
{
    __shared__ double red[B / kWave];
    __shared__ int flag;
    double st = 0, scl = 0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B)
    {
        st = fmax(st, fabs(pNew[i] - p[i]));
        scl = fmax(scl, fabs(p[i]));
    }
    st = blockMax(st, red);
    scl = blockMax(scl, red);
    if (threadIdx.x == 0) { part[blockIdx.x] = st; part[gridDim.x + blockIdx.x] = scl; }
    if (!lastBlockDone(ticketOf(sc, kTicketA), &flag)) return;
    const double ST = maxPartials(part, gridDim.x, red);
    const double SC = maxPartials(part + gridDim.x, gridDim.x, red);
    if (threadIdx.x == 0)
    {
        sc[kPStep] = ST;
        sc[kPScale] = SC;
        sc[kAP] = pRelaxStart > 0 ? smin(relaxP, smax(pRelaxStart, 0.15 * SC / smax(ST, 1e-300))) : relaxP;
    }
}

// p += aP (pNew - p); U = Hfree - aP rAU (grad pNew - grad p^n);
// pDeltaMax, max|U|, max|U - UOld|, finiteness (simpleSolver.cpp:315-347)
__global__ __launch_bounds__(B) void k_simple_update(int n, const double* __restrict__ pNew, double* __restrict__ p,
                                                     const double* __restrict__ H, const double* __restrict__ gF,
                                                     const double* __restrict__ gN, const double* __restrict__ rAU,
                                                     double* __restrict__ U, const double* __restrict__ UOld,
                                                     double* sc, double* part)
{
    __shared__ double red[B / kWave];
    __shared__ int ired[B / kWave];
    __shared__ int flag;
    const double aP = sc[kAP];
    double pd = 0, um = 0, dum = 0;
    int bad = 0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B)
    {
        const double dp = aP * (pNew[i] - p[i]);
        pd = fmax(pd, fabs(dp));
        const double pi = p[i] + dp;
        p[i] = pi;
        const double s = aP * rAU[i];
        double u[3], du[3];
        for (int k = 0; k < 3; ++k)
        {
            u[k] = H[k * n + i] - (gF[k * n + i] - gN[k * n + i]) * s;
            U[k * n + i] = u[k];
            du[k] = u[k] - UOld[k * n + i];
        }
        um = fmax(um, sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
        dum = fmax(dum, sqrt(du[0] * du[0] + du[1] * du[1] + du[2] * du[2]));
        bad |= !(isfinite(u[0]) && isfinite(u[1]) && isfinite(u[2]) && isfinite(pi));
    }
    pd = blockMax(pd, red);
    um = blockMax(um, red);
    dum = blockMax(dum, red);
    bad = blockOr(bad, ired);
    if (threadIdx.x == 0)
    {
        part[blockIdx.x] = pd;
        part[gridDim.x + blockIdx.x] = um;
        part[2 * gridDim.x + blockIdx.x] = dum;
        reinterpret_cast<int*>(part + 3 * gridDim.x)[blockIdx.x] = bad;
    }
    if (!lastBlockDone(ticketOf(sc, kTicketB), &flag)) return;
    const int G = gridDim.x;
    const double PD = maxPartials(part, G, red);
    const double UM = maxPartials(part + G, G, red);
    const double DUM = maxPartials(part + 2 * G, G, red);
    const int BAD = orPartials(reinterpret_cast<const int*>(part + 3 * G), G, ired);
    if (threadIdx.x == 0)
    {
        sc[kPDelta] = PD;
        sc[kUMax] = UM;
        sc[kDUMax] = DUM;
        sc[kNonFinite] = BAD ? 1.0 : 0.0;
    }
}

// max |a - b| into sc[slot] (PISO pressureCorrectionMax)
__global__ __launch_bounds__(B) void k_maxdiff(int n, const double* __restrict__ a, const double* __restrict__ b,
                                               double* sc, int slot, double* part)
{
    __shared__ double red[B / kWave];
    __shared__ int flag;
    double m = 0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) m = fmax(m, fabs(a[i] - b[i]));
    m = blockMax(m, red);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
    if (!lastBlockDone(ticketOf(sc, kTicketC), &flag)) return;
    const double M = maxPartials(part, gridDim.x, red);
    if (threadIdx.x == 0) sc[slot] = M;
}

// HbyA_k = rAU (h + V gradPn_k) / V   (pisoSolver.cpp:133-138)
__global__ void k_hbya(int n, const double* __restrict__ rAU, const double* __restrict__ h,
                       const double* __restrict__ V, const double* __restrict__ g, double* __restrict__ H)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    H[i] = rAU[i] * (h[i] + V[i] * g[i]) / V[i];
}

// U = HbyA - rAU gradP   (pisoSolver.cpp:206-210)
__global__ void k_u_piso(int n, const double* __restrict__ H, const double* __restrict__ g,
                         const double* __restrict__ rAU, double* __restrict__ U)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; ++k) U[k * n + i] = H[k * n + i] - g[k * n + i] * rAU[i];
}

__global__ void k_explicit_h(int n, const double* __restrict__ rhs, const double* __restrict__ ax,
                             const double* __restrict__ diag, const double* __restrict__ x, double* __restrict__ h)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) h[i] = rhs[i] - (ax[i] - diag[i] * x[i]);
}

__global__ void k_zero(double* p, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// Buffers of one driver run
struct Work
{
    DevBuf sc;
    DevBuf gradA, gradB, xU, UOld, H, rAU, rAUf, divPhi, pNew, gammaU, faceTmp, gradTmp, cellTmp, hTmp;
    DevBuf Udiag, Urhs, Ubase, Uup, Ulo, Uval, pdiag, prhs, pbase, pup, plo, pval, val32;
    DevBuf Ulc, Ulk, plc, plk, Hlc, Hlk, types, values, patchStart, patchSize, fixedU;
    int nB = 0;

    void alloc(DevMesh& m)
    {
        const std::size_t n = m.n, nF = m.nF, nI = std::max(1, m.nI), nnz = m.pattern.nnz, nB3 = 3 * std::max(1, m.nB);
        const std::size_t d = sizeof(double);
        sc.alloc(d * kNumSlots);
        CFX_HIP(hipMemsetAsync(sc.p, 0, d * kNumSlots, m.ctx->stream()));
        for (DevBuf* b : {&gradA, &gradB, &xU, &UOld, &H, &gradTmp, &Udiag, &Urhs, &Ubase}) b->alloc(d * 3 * n);
        for (DevBuf* b : {&rAU, &divPhi, &pNew, &cellTmp, &hTmp, &pdiag, &prhs, &pbase}) b->alloc(d * n);
        for (DevBuf* b : {&rAUf, &gammaU, &faceTmp}) b->alloc(d * nF);
        for (DevBuf* b : {&Uup, &Ulo, &pup, &plo}) b->alloc(d * nI);
        Uval.alloc(d * 3 * nnz);
        pval.alloc(d * nnz);
        for (DevBuf* b : {&Ulc, &Ulk, &Hlc, &Hlk, &values}) b->alloc(d * nB3);
        for (DevBuf* b : {&plc, &plk}) b->alloc(d * std::max(1, m.nB));
        types.alloc(sizeof(int) * std::max(1, m.nB));
        patchStart.alloc(sizeof(long long) * std::max(1, m.nPatches));
        patchSize.alloc(sizeof(long long) * std::max(1, m.nPatches));
        fixedU.alloc(sizeof(int) * std::max(1, m.nPatches));
        nB = m.nB;
    }
    Ldu U() const { return {3, Udiag.as<double>(), Urhs.as<double>(), Uup.as<double>(), Ulo.as<double>()}; }
    Ldu P() const { return {1, pdiag.as<double>(), prhs.as<double>(), pup.as<double>(), plo.as<double>()}; }
};

// device time of a driver call on its stream (events freed on every exit)
class StepTimer
{
public:
    explicit StepTimer(hipStream_t s) : s_(s)
    {
        CFX_HIP(hipEventCreate(&e0_));
        CFX_HIP(hipEventCreate(&e1_));
        CFX_HIP(hipEventRecord(e0_, s_));
    }
    ~StepTimer()
    {
        (void)hipEventDestroy(e0_);
        (void)hipEventDestroy(e1_);
    }
    StepTimer(const StepTimer&) = delete;
    StepTimer& operator=(const StepTimer&) = delete;
    double stop()
    {
        CFX_HIP(hipEventRecord(e1_, s_));
        CFX_HIP(hipEventSynchronize(e1_));
        float ms = 0;
        CFX_HIP(hipEventElapsedTime(&ms, e0_, e1_));
        return ms;
    }

private:
    hipStream_t s_;
    hipEvent_t e0_ = nullptr, e1_ = nullptr;
};

// The mesh's driver work buffers: allocated by the first driver call on the
// mesh and kept (a time step of a 6M-cell mesh otherwise allocates and frees
// ~3 GB, and every free waits for the device); the state slots start at zero
// in every call, as a fresh allocation did.
static Work& driverWork(DevMesh& m)
{
    if (!m.driverWork)
    {
        auto w = std::make_shared<Work>();
        w->alloc(m);
        m.driverWork = w;
    }
    Work& w = *static_cast<Work*>(m.driverWork.get());
    CFX_HIP(hipMemsetAsync(w.sc.p, 0, sizeof(double) * kNumSlots, m.ctx->stream()));
    return w;
}

// BCs: U, p from the caller; HbyA = fixedValue (U's face values) where U is
// fixedValue, zeroGradient elsewhere (simpleSolver.cpp:175-187)
void setupBcs(DevMesh& m, Work& w, const int* Utype, const double* Uvalue, const int* ptype, const double* pvalue,
              bool& needReference)
{
    hipStream_t s = m.ctx->stream();
    const int nB = m.nB;
    if (nB == 0) { needReference = true; return; }
    std::vector<int> ut(Utype, Utype + nB), pt(ptype, ptype + nB), ht(nB);
    for (int b = 0; b < nB; ++b) ht[b] = ut[b] == 0 ? 0 : 2;
    auto lin = [&](const std::vector<int>& t, const double* v, int nc, double* lc, double* lk) {
        CFX_HIP(hipMemcpyAsync(w.types.p, t.data(), sizeof(int) * nB, hipMemcpyHostToDevice, s));
        CFX_HIP(hipMemcpyAsync(w.values.p, v, sizeof(double) * nB * nc, hipMemcpyHostToDevice, s));
        if (!lineariseBc(m, nc, w.types.as<int>(), w.values.as<double>(), lc, lk))
            throw std::invalid_argument("fields: symmetry normal not axis aligned, or unknown bc type");
    };
    lin(ut, Uvalue, 3, w.Ulc.as<double>(), w.Ulk.as<double>());
    lin(ht, Uvalue, 3, w.Hlc.as<double>(), w.Hlk.as<double>());
    lin(pt, pvalue, 1, w.plc.as<double>(), w.plk.as<double>());
    needReference = true;
    for (int b = 0; b < nB; ++b)
        if (pt[b] == 0) needReference = false;
    std::vector<int> fixedU(std::max(1, m.nPatches), 0);
    for (int pi = 0; pi < m.nPatches; ++pi)
        if (m.patchSize[pi] > 0) fixedU[pi] = ut[m.patchStart[pi] - m.nI] == 0;
    CFX_HIP(hipMemcpyAsync(w.patchStart.p, m.patchStart.data(), sizeof(long long) * m.nPatches, hipMemcpyHostToDevice, s));
    CFX_HIP(hipMemcpyAsync(w.patchSize.p, m.patchSize.data(), sizeof(long long) * m.nPatches, hipMemcpyHostToDevice, s));
    CFX_HIP(hipMemcpyAsync(w.fixedU.p, fixedU.data(), sizeof(int) * fixedU.size(), hipMemcpyHostToDevice, s));
    m.ctx->sync();
}

} // namespace

bool lineariseBc(DevMesh& m, int ncomp, const int* dType, const double* dValue, double* linc, double* link)
{
    if (m.nB == 0) return true;
    DevBuf err(sizeof(int));
    hipStream_t s = m.ctx->stream();
    CFX_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
    k_linearise<<<grid(m.nB), B, 0, s>>>(m.nI, m.nB, m.nF, ncomp, m.Sf.as<double>(), m.d.as<double>(), dType, dValue,
                                         linc, link, err.as<int>());
    CFX_CHECK_LAUNCH();
    int h = 0;
    CFX_HIP(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    CFX_HIP(hipStreamSynchronize(s));
    return h == 0;
}

// rhs -= mean(rhs): per-workgroup partial sums, then every workgroup reduces
// them (fixed order) and subtracts the mean from its elements
__global__ void k_rhs_sum(int n, const double* __restrict__ v, double* __restrict__ part)
{
    __shared__ double red[B / kWave];
    double s = 0.0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) s += v[i];
    s = blockSum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_rhs_sub_mean(int n, double* __restrict__ v, const double* __restrict__ part, int G)
{
    __shared__ double red[B / kWave];
    __shared__ double meanS;
    double s = 0.0;
    for (int g = threadIdx.x; g < G; g += B) s += part[g];
    s = blockSum(s, red);
    if (threadIdx.x == 0) meanS = s / n;
    __syncthreads();
    const double mean = meanS;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) v[i] -= mean;
}

static void projectZeroMean(DevMesh& m, double* v)
{
    hipStream_t s = m.ctx->stream();
    const int G = std::max(1, std::min(256, divUp(m.n, 4 * B)));
    double* part = m.ctx->partials(3);
    k_rhs_sum<<<G, B, 0, s>>>(m.n, v, part);
    k_rhs_sub_mean<<<G, B, 0, s>>>(m.n, v, part, G);
    CFX_CHECK_LAUNCH();
}

// pressure-solve options of the drivers
static void setPressureSolver(FvSolveArgs& pa, int precision, const PressureSolver* ps)
{
    pa.precision = precision == kFP32MG ? kFP32IR : precision;
    if (precision != kFP32MG) return;
    if (!ps || !ps->mg) throw std::invalid_argument("multigrid pressure solve without a multigrid state");
    pa.mg = ps->mg;
    pa.net = ps->net;
    pa.geo = ps->geo;
    pa.wa = ps->wa;
    pa.wt = ps->wt;
}

static void dumpPressure(DevMesh& m, const Ldu& PE, const double* pval, PressureDump* d)
{
    if (!d || d->every <= 0) return;
    const int k = d->seen++;
    if (k % d->every != 0 || d->count >= d->max) return;
    hipStream_t s = m.ctx->stream();
    const std::size_t nnz = m.pattern.nnz, n = m.n;
    CFX_HIP(hipMemcpyAsync(d->val + d->count * nnz, pval, sizeof(double) * nnz, hipMemcpyDeviceToDevice, s));
    CFX_HIP(hipMemcpyAsync(d->rhs + d->count * n, PE.rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    ++d->count;
}

namespace
{
// x = xOld + rl (x - xOld), the reference's expression (fvMatrix.cpp:116)
__global__ void k_relax(long long n, const double* __restrict__ xOld, double* __restrict__ x, double rl)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] = xOld[i] + rl * (x[i] - xOld[i]);
}
} // namespace

SolveReport fvSolve(DevMesh& m, Krylov& K, const Ldu& A, double* val, double* rhsBase, DevBuf& val32, double* x,
                    const FvSolveArgs& a, double* gradTmp, double* faceTmp)
{
    hipStream_t s = m.ctx->stream();
    const int n = m.n, nnz = m.pattern.nnz;
    for (int k = 0; k < A.ncomp; ++k) lduToCsr(m, A, k, val + static_cast<long long>(k) * nnz);
    const std::size_t bytes = sizeof(double) * A.ncomp * n;
    CFX_HIP(hipMemcpyAsync(rhsBase, A.rhs, bytes, hipMemcpyDeviceToDevice, s));
    const int passes = a.gammaFace ? std::max(1, a.nonOrth + 1) : 1;
    const CsrView view = m.pattern.view();
    SolveReport last;
    const bool relax = a.relaxation < 1.0;
    if (relax && !a.relaxOld) throw std::invalid_argument("fvSolve: relaxation needs a buffer for the old solution");
    for (int pass = 0; pass < passes; ++pass)
    {
        const bool finalPass = pass + 1 == passes;
        // the solution before the final pass's solves (each component's solve leaves the others untouched)
        if (finalPass && relax) CFX_HIP(hipMemcpyAsync(a.relaxOld, x, bytes, hipMemcpyDeviceToDevice, s));
        CFX_HIP(hipMemcpyAsync(A.rhs, rhsBase, bytes, hipMemcpyDeviceToDevice, s));
        if (a.gammaFace)
            for (int k = 0; k < A.ncomp; ++k)
                nonOrthCorrect(m, a.gammaFace, x + static_cast<long long>(k) * n, a.linc, a.link, A.ncomp, k,
                               a.gradCorrected, a.sign, A.rhs + static_cast<long long>(k) * n, gradTmp, faceTmp);
        if (a.projectMean)
            for (int k = 0; k < A.ncomp; ++k) projectZeroMean(m, A.rhs + static_cast<long long>(k) * n);
        if (a.solver == 1 && A.ncomp > 1 && K.batchComponents)
        {
            // the components' BiCGStab solves in shared launches, equal to the loop below bit for bit
            SolveReport reps[Krylov::kMaxComponents];
            if (a.precision == kFP32IR)
                K.bicgstabMixedComponents(view, val, nnz, A.rhs, x, A.ncomp, a.tol, a.maxIter, reps);
            else
                K.bicgstabComponents(view, val, nnz, A.rhs, x, A.ncomp, a.tol, a.maxIter, reps);
            last = reps[A.ncomp - 1];
            continue;
        }
        for (int k = 0; k < A.ncomp; ++k)
        {
            const double* v = val + static_cast<long long>(k) * nnz;
            double* b = A.rhs + static_cast<long long>(k) * n;
            double* xk = x + static_cast<long long>(k) * n;
            if (a.solver == 1)
            {
                last = a.precision == kFP32IR ? K.bicgstabMixed(view, v, b, xk, a.tol, a.maxIter)
                                              : K.bicgstab(view, v, b, xk, a.tol, a.maxIter);
            }
            else if (a.mg)
            {
                // the values are those of the hierarchy's last setup in the later passes of this solve and for
                // a repeated matrix (one component: a multigrid solve is scalar)
                const bool sameValues = (pass > 0 || a.sameMatrix) && A.ncomp == 1;
                if (a.net)
                {
                    // the network's weights follow the right-hand side: new weights on every pass; its
                    // matrix-only features and the hierarchy's values stay while the values do
                    a.net->setGraph(view, v, a.geo, b, sameValues);
                    a.net->forward(a.wa, a.wt, false);
                    a.mg->useWeights(a.wa, a.wt, a.wa, a.wt);
                    a.mg->weightsValid = false;
                }
                if (!sameValues) a.mg->valid = false;
                Multigrid& M = a.mg->ready(*m.ctx, m.pattern, v);
                last = K.fcgMixed(view, v, M, b, xk, a.tol, a.maxIter, a.delta, nullptr);
            }
            else if (a.precision == kFP64)
            {
                last = K.cg(view, v, b, xk, a.tol, a.maxIter);
            }
            else
            {
                if (!val32.p || val32.bytes < sizeof(float) * nnz) val32.alloc(sizeof(float) * nnz);
                castToFloat(*m.ctx, v, val32.as<float>(), nnz);
                last = K.cgMixed(view, v, val32.as<float>(), b, xk, a.tol, a.maxIter, a.delta, false);
            }
        }
    }
    if (relax)
    {
        const long long total = static_cast<long long>(A.ncomp) * n;
        k_relax<<<std::max(1, std::min(4096, divUp(total, 256))), 256, 0, s>>>(total, a.relaxOld, x, a.relaxation);
        CFX_CHECK_LAUNCH();
    }
    return last;
}

SimpleResult simpleSolve(DevMesh& m, Krylov& K, double nu, const Schemes& sch, const SimpleControls& c,
                         const int* Utype, const double* Uvalue, const int* ptype, const double* pvalue, double* U,
                         double* p, double* phi, double* hist)
{
    Context& ctx = *m.ctx;
    hipStream_t s = ctx.stream();
    const int n = m.n;
    Work& w = driverWork(m);
    bool needReference = true;
    setupBcs(m, w, Utype, Uvalue, ptype, pvalue, needReference);
    double* sc = w.sc.as<double>();
    double* part = ctx.partials(0);
    harmonicGammaUniform(m, nu, w.gammaU.as<double>());
    StepTimer timer(s);

    SimpleResult res;
    double contLag1 = INFINITY, contLag2 = INFINITY;
    double* gradPn = w.gradA.as<double>();
    double* gradPFull = w.gradB.as<double>();
    for (int iter = 1; iter <= c.maxIterations; ++iter)
    {
        gradGauss(m, p, w.plc.as<double>(), w.plk.as<double>(), 1, 0, sch.gradCorrected, true, gradPn,
                  w.faceTmp.as<double>());
        // momentum
        const Ldu UE = w.U();
        lduReset(m, UE);
        assembleDiv(m, UE, phi, sch.div, sch.blendCoeff, w.Ulc.as<double>(), w.Ulk.as<double>(), 1);
        assembleLaplacian(m, UE, w.gammaU.as<double>(), w.Ulc.as<double>(), w.Ulk.as<double>(), -1);
        addVolumeSource(m, UE, c.bodyForce, gradPn);
        k_zero<<<gridCap(3LL * n), B, 0, s>>>(w.xU.as<double>(), 3LL * n);
        FvSolveArgs ua;
        ua.solver = 1;
        ua.precision = c.uPrecision;
        ua.tol = c.linearTol;
        ua.maxIter = c.linearMaxIter;
        ua.nonOrth = c.nonOrthCorrectors;
        ua.gammaFace = sch.laplacianCorrected ? w.gammaU.as<double>() : nullptr;
        ua.linc = w.Ulc.as<double>();
        ua.link = w.Ulk.as<double>();
        ua.sign = -1;
        ua.gradCorrected = sch.gradCorrected;
        fvSolve(m, K, UE, w.Uval.as<double>(), w.Ubase.as<double>(), w.val32, w.xU.as<double>(), ua,
                w.gradTmp.as<double>(), w.faceTmp.as<double>());
        CFX_HIP(hipMemcpyAsync(w.UOld.p, U, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, s));
        k_rau<<<grid(n), B, 0, s>>>(n, w.Udiag.as<double>(), m.V.as<double>(), w.rAU.as<double>(), sc);
        // HbyA (pressure-free blend) and its predictor flux
        k_hfree<<<grid(3LL * n), B, 0, s>>>(n, c.relaxU, w.UOld.as<double>(), w.xU.as<double>(), w.H.as<double>());
        updatePhiInternal(m, w.H.as<double>(), phi);
        correctBoundaryFlux(m, w.H.as<double>(), w.Hlc.as<double>(), w.Hlk.as<double>(), phi);
        if (needReference && m.nPatches > 0)
        {
            k_patch_sums<<<m.nPatches, B, 0, s>>>(w.patchStart.as<long long>(), w.patchSize.as<long long>(), phi,
                                                  part);
            k_adjust_phi<<<gridCap(m.nB), B, 0, s>>>(m.nPatches, w.patchStart.as<long long>(),
                                                     w.patchSize.as<long long>(), w.fixedU.as<int>(), part, phi);
        }
        continuityError(m, phi, w.divPhi.as<double>(), sc + kContErr, part);
        // pressure
        interpolateToFaces(m, w.rAU.as<double>(), w.rAUf.as<double>());
        const Ldu PE = w.P();
        lduReset(m, PE);
        assembleLaplacian(m, PE, w.rAUf.as<double>(), w.plc.as<double>(), w.plk.as<double>(), -1);
        addNegRhs(m, PE, w.divPhi.as<double>());
        k_zero<<<gridCap(n), B, 0, s>>>(w.pNew.as<double>(), n);
        FvSolveArgs pa;
        pa.solver = 0;
        pa.tol = c.linearTol;
        pa.maxIter = c.linearMaxIter;
        pa.nonOrth = c.nonOrthCorrectors;
        pa.gammaFace = sch.laplacianCorrected ? w.rAUf.as<double>() : nullptr;
        pa.linc = w.plc.as<double>();
        pa.link = w.plk.as<double>();
        pa.sign = -1;
        pa.gradCorrected = sch.gradCorrected;
        setPressureSolver(pa, c.pPrecision, c.pSolver);
        fvSolve(m, K, PE, w.pval.as<double>(), w.pbase.as<double>(), w.val32, w.pNew.as<double>(), pa,
                w.gradTmp.as<double>(), w.faceTmp.as<double>());
        dumpPressure(m, PE, w.pval.as<double>(), c.dump);
        // correct
        gradGauss(m, w.pNew.as<double>(), w.plc.as<double>(), w.plk.as<double>(), 1, 0, sch.gradCorrected, true,
                  gradPFull, w.faceTmp.as<double>());
        correctFluxes(m, w.pNew.as<double>(), gradPFull, w.rAUf.as<double>(), w.plc.as<double>(),
                      w.plk.as<double>(), sch.laplacianCorrected, phi);
        k_pstep<<<gridCap(n), B, 0, s>>>(n, w.pNew.as<double>(), p, sc, part, c.relaxP, c.pRelaxStart);
        k_simple_update<<<gridCap(n), B, 0, s>>>(n, w.pNew.as<double>(), p, w.H.as<double>(), gradPFull, gradPn,
                                                 w.rAU.as<double>(), U, w.UOld.as<double>(), sc, part);
        continuityError(m, phi, w.cellTmp.as<double>(), sc + kErrAfter, part);
        CFX_CHECK_LAUNCH();
        double h[kNumSlots];
        CFX_HIP(hipMemcpyAsync(h, sc, sizeof(h), hipMemcpyDeviceToHost, s));
        CFX_HIP(hipStreamSynchronize(s));
        if (h[kDiagBad] != 0) throw std::invalid_argument("SIMPLE: momentum diagonal differs between components");
        const double contErr = h[kContErr], pDeltaMax = h[kPDelta];
        const double uRes = h[kDUMax] / (h[kUMax] < 1e-300 ? 1e-300 : h[kUMax]);
        if (hist)
        {
            hist[4 * (iter - 1)] = contErr;
            hist[4 * (iter - 1) + 1] = h[kErrAfter];
            hist[4 * (iter - 1) + 2] = pDeltaMax;
            hist[4 * (iter - 1) + 3] = uRes;
        }
        res.iterations = iter;
        res.finalContinuityError = contErr;
        res.finalMomentumResidual = uRes;
        if (!std::isfinite(contErr) || !std::isfinite(pDeltaMax) || h[kNonFinite] != 0)
        {
            res.diverged = 1;
            break;
        }
        const bool settled = contErr <= contLag2 * 1.01 + 1e-12;
        const bool momentumSettled = !(c.momentumTol > 0) || uRes <= c.momentumTol;
        if (momentumSettled &&
            (contErr < c.continuityTol || (iter >= 5 && c.pcorrTol > 0 && pDeltaMax < c.pcorrTol && settled)))
        {
            res.converged = 1;
            break;
        }
        contLag2 = contLag1;
        contLag1 = contErr;
    }
    res.ms = timer.stop();
    return res;
}

PisoResult pisoStep(DevMesh& m, Krylov& K, double nu, const Schemes& sch, const PisoControls& c, const int* Utype,
                    const double* Uvalue, const int* ptype, const double* pvalue, double* U, double* p, double* phi)
{
    if (!(c.dt > 0)) throw std::invalid_argument("PISO: non-positive time step");
    Context& ctx = *m.ctx;
    hipStream_t s = ctx.stream();
    const int n = m.n;
    Work& w = driverWork(m);
    bool needReference = true;
    setupBcs(m, w, Utype, Uvalue, ptype, pvalue, needReference);
    double* sc = w.sc.as<double>();
    double* part = ctx.partials(0);
    harmonicGammaUniform(m, nu, w.gammaU.as<double>());
    StepTimer timer(s);

    PisoResult res;
    double* gradP = w.gradB.as<double>();
    double* gradPn = w.gradA.as<double>();
    gradGauss(m, p, w.plc.as<double>(), w.plk.as<double>(), 1, 0, sch.gradCorrected, true, gradPn,
              w.faceTmp.as<double>());
    CFX_HIP(hipMemcpyAsync(w.UOld.p, U, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, s));
    const Ldu UE = w.U();
    lduReset(m, UE);
    assembleDdt(m, UE, w.UOld.as<double>(), c.dt, 1);
    assembleDiv(m, UE, phi, sch.div, sch.blendCoeff, w.Ulc.as<double>(), w.Ulk.as<double>(), 1);
    assembleLaplacian(m, UE, w.gammaU.as<double>(), w.Ulc.as<double>(), w.Ulk.as<double>(), -1);
    addVolumeSource(m, UE, c.bodyForce, gradPn);
    k_zero<<<gridCap(3LL * n), B, 0, s>>>(w.xU.as<double>(), 3LL * n);
    FvSolveArgs ua;
    ua.solver = 1;
    ua.precision = c.uPrecision;
    ua.tol = c.linearTol;
    ua.maxIter = c.linearMaxIter;
    ua.nonOrth = c.nonOrthCorrectors;
    ua.gammaFace = sch.laplacianCorrected ? w.gammaU.as<double>() : nullptr;
    ua.linc = w.Ulc.as<double>();
    ua.link = w.Ulk.as<double>();
    ua.gradCorrected = sch.gradCorrected;
    fvSolve(m, K, UE, w.Uval.as<double>(), w.Ubase.as<double>(), w.val32, w.xU.as<double>(), ua,
            w.gradTmp.as<double>(), w.faceTmp.as<double>());
    CFX_HIP(hipMemcpyAsync(U, w.xU.p, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, s));
    k_rau<<<grid(n), B, 0, s>>>(n, w.Udiag.as<double>(), m.V.as<double>(), w.rAU.as<double>(), sc);

    const CsrView view = m.pattern.view();
    auto rebuildHbyA = [&] {
        for (int k = 0; k < 3; ++k)
        {
            const double* xk = U + static_cast<long long>(k) * n;
            K.spmv(view, w.Uval.as<double>() + static_cast<long long>(k) * m.pattern.nnz, xk, w.cellTmp.as<double>());
            k_explicit_h<<<grid(n), B, 0, s>>>(n, w.Urhs.as<double>() + static_cast<long long>(k) * n,
                                               w.cellTmp.as<double>(), w.Udiag.as<double>() + static_cast<long long>(k) * n,
                                               xk, w.hTmp.as<double>());
            k_hbya<<<grid(n), B, 0, s>>>(n, w.rAU.as<double>(), w.hTmp.as<double>(), m.V.as<double>(),
                                         gradPn + static_cast<long long>(k) * n, w.H.as<double>() + static_cast<long long>(k) * n);
        }
    };
    rebuildHbyA();
    interpolateToFaces(m, w.rAU.as<double>(), w.rAUf.as<double>());
    for (int corr = 1; corr <= c.correctors; ++corr)
    {
        if (corr > 1) rebuildHbyA();
        updatePhiInternal(m, w.H.as<double>(), phi);
        correctBoundaryFlux(m, w.H.as<double>(), w.Hlc.as<double>(), w.Hlk.as<double>(), phi);
        continuityError(m, phi, w.divPhi.as<double>(), sc + kPredCont, part);
        const Ldu PE = w.P();
        lduReset(m, PE);
        assembleLaplacian(m, PE, w.rAUf.as<double>(), w.plc.as<double>(), w.plk.as<double>(), -1);
        addNegRhs(m, PE, w.divPhi.as<double>());
        CFX_HIP(hipMemcpyAsync(w.pNew.p, p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));  // pBefore
        FvSolveArgs pa;
        pa.solver = 0;
        pa.tol = c.linearTol;
        pa.maxIter = c.linearMaxIter;
        pa.nonOrth = c.nonOrthCorrectors;
        pa.gammaFace = sch.laplacianCorrected ? w.rAUf.as<double>() : nullptr;
        pa.linc = w.plc.as<double>();
        pa.link = w.plk.as<double>();
        pa.gradCorrected = sch.gradCorrected;
        pa.projectMean = c.projectNeumann && needReference;
        pa.sameMatrix = corr > 1;  // rAUf, and so the pressure matrix, is the same for every corrector
        setPressureSolver(pa, c.pPrecision, c.pSolver);
        fvSolve(m, K, PE, w.pval.as<double>(), w.pbase.as<double>(), w.val32, p, pa, w.gradTmp.as<double>(),
                w.faceTmp.as<double>());
        dumpPressure(m, PE, w.pval.as<double>(), c.dump);
        k_maxdiff<<<gridCap(n), B, 0, s>>>(n, p, w.pNew.as<double>(), sc, kPDelta, part);
        gradGauss(m, p, w.plc.as<double>(), w.plk.as<double>(), 1, 0, sch.gradCorrected, true, gradP,
                  w.faceTmp.as<double>());
        k_u_piso<<<grid(n), B, 0, s>>>(n, w.H.as<double>(), gradP, w.rAU.as<double>(), U);
        correctFluxes(m, p, gradP, w.rAUf.as<double>(), w.plc.as<double>(), w.plk.as<double>(), sch.laplacianCorrected,
                      phi);
        continuityError(m, phi, w.cellTmp.as<double>(), sc + kErrAfter, part);
        CFX_CHECK_LAUNCH();
        double h[kNumSlots];
        CFX_HIP(hipMemcpyAsync(h, sc, sizeof(h), hipMemcpyDeviceToHost, s));
        CFX_HIP(hipStreamSynchronize(s));
        if (h[kDiagBad] != 0) throw std::invalid_argument("PISO: momentum diagonal differs between components");
        res.predictedContinuityError = h[kPredCont];
        res.pressureCorrectionMax = h[kPDelta];
        res.correctorsUsed = corr;
        res.continuityError = h[kErrAfter];
        if (!std::isfinite(res.continuityError) || !std::isfinite(h[kPDelta])) break;
        if (res.continuityError < c.continuityTol)
        {
            res.converged = 1;
            break;
        }
    }
    res.ms = timer.stop();
    return res;
}

} // namespace cfx
