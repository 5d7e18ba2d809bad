// krylov.h -- SpMV and Krylov solvers on the device.
//
// solveCG / solveBiCGStab follow the reference algorithms step for step
// (src/common/src/Matrix.cpp:153-222, :224-319): Jacobi preconditioning with
// the same guards, the same infinity-norm stop ||r||inf / max(||b||inf,
// 1e-300) < tol on the recurrence residual, the same breakdown exits.  The
// SpMV sums each row in column order, so A*x is bit-identical to the
// reference SparseMatrix::product; only the dot products differ (tree
// instead of serial summation), which moves iterates by rounding only.
#pragma once

#include "types.h"

#include <vector>

namespace cfx
{


struct SolveReport
{
    int iterations = 0;
    double initialResidual = 0, finalResidual = 0;
    int converged = 0;
    int breakdown = 0;
    // extensions
    double l2Residual = 0;        // ||b - A x||2 / ||b||2 (true FP64 residual)
    double trueResidual = 0;      // ||b - A x||inf / ||b||inf (true FP64 residual)
    int outerIterations = 0;      // refinement sweeps (mixed precision)
    int precondFallbacks = 0;     // learned preconditioner rejected -> Jacobi
    double ms = 0;                // device time of the solve
};

enum Precond { kJacobi = 0, kDIC = 1, kLearned = 2 };
enum Precision { kFP64 = 0, kFP32IR = 1 };

class Krylov
{
public:
    Krylov(Context& ctx, int n, int maxRowBlocks);
    ~Krylov();
    int n() const { return n_; }
    int maxBlocks() const { return maxBlocks_; }

    // y = A x (bit-identical to the reference product)
    void spmv(const CsrView& A, const double* val, const double* x, double* y);
    void spmv32(const CsrView& A, const float* val, const float* x, float* y);

    // Jacobi-PCG in FP64 (solveCG).  x is the initial guess and the result.
    // resHist (device, optional, [maxIter+1]) receives ||r||inf per iteration.
    SolveReport cg(const CsrView& A, const double* val, const double* b, double* x, double tol,
                   int maxIter, double* resHist = nullptr);

    // Jacobi-BiCGStab in FP64 (solveBiCGStab).
    SolveReport bicgstab(const CsrView& A, const double* val, const double* b, double* x, double tol,
                         int maxIter);

    // Jacobi-BiCGStab in FP64 of the ncomp (1..3) systems of one pattern: the
    // components of a vector equation (fvMatrix::solve, fvMatrix.cpp:93-122),
    // values at val + c*valStride, b and x as [c][n].  Each component runs
    // exactly the arithmetic of bicgstab() on its own system (same kernels,
    // same reductions), so x and reps[c] equal ncomp separate solves bit for
    // bit; the components share each launch, a third of the launches.
    void bicgstabComponents(const CsrView& A, const double* val, long long valStride, const double* b, double* x,
                            int ncomp, double tol, int maxIter, SolveReport* reps);
    static constexpr int kMaxComponents = 3;

    // bicgstabMixed for the ncomp components in shared launches: per
    // component the same true residuals, FP32 inner solves, corrections and
    // FP64 fallback as bicgstabMixed on its system alone, bit for bit.
    void bicgstabMixedComponents(const CsrView& A, const double* val, long long valStride, const double* b,
                                 double* x, int ncomp, double tol, int maxIter, SolveReport* reps,
                                 double innerTol = 1e-4);

    // Jacobi-BiCGStab in FP32 with FP64 iterative refinement: r = b - A x in
    // FP64, the FP32 solve of A d = r from d = 0 to `innerTol` relative (or to
    // the overall tolerance when that is closer), x += d; the reference's stop
    // test is applied to the true FP64 residual only.  When a correction no
    // longer halves the FP64 residual the solve finishes in FP64
    // (precondFallbacks = 1), from the x before that correction if it made the
    // residual worse or not finite; an FP32 solve that fails (non-finite) is
    // not applied.  iterations counts the FP32 (and FP64) BiCGStab
    // iterations, outerIterations the corrections.
    SolveReport bicgstabMixed(const CsrView& A, const double* val, const double* b, double* x, double tol,
                              int maxIter, double innerTol = 1e-4);

    // FP32-inner Jacobi-PCG with FP64 iterative refinement by reliable
    // updates: the Krylov recurrence runs in FP32 (matrix `val32`) and
    // accumulates a correction; each time its residual has fallen by `delta`
    // below the last true residual, the correction is folded into the FP64 x
    // and r = b - A x is recomputed in FP64 and replaces the FP32 residual.
    // Convergence is declared on true FP64 residuals only.  delta <= 0
    // takes kPcgDelta.
    SolveReport cgMixed(const CsrView& A, const double* val, const float* val32, const double* b,
                        double* x, double tol, int maxIter, double delta, bool project);

    // Flexible CG (FCG(1): alpha = p.r / p.q, beta = -z.q / p.q) with one
    // multigrid V-cycle as the preconditioner, FP32 Krylov vectors and the
    // same reliable FP64 updates as cgMixed.  `mg` must hold the hierarchy
    // of `val` (Multigrid::setMatrix).  On a breakdown or a non-finite value
    // the solve continues from the last FP64 iterate with cgMixed (Jacobi)
    // and reports precondFallbacks = 1.  delta <= 0: the update comes after
    // a drop of kFcgDelta, or after one of kPcgDelta once kFcgSegment
    // iterations have passed (the Jacobi fallback takes kPcgDelta).
    SolveReport fcgMixed(const CsrView& A, const double* val, Preconditioner& mg, const double* b, double* x,
                         double tol, int maxIter, double delta, double* resHist = nullptr);

    // Average device time (ms) of the flexible CG's fine-level kernel (q = A p with FP64 values on FP32
    // vectors, partials p.q and p.r) over `reps` launches on this solver's vectors, HIP events on the
    // context's stream: the roofline yardstick of the multigrid solves (on the preconditioner's sliced
    // values of `val` where it holds them, as the solve does).
    double timeFcgSpmv(const CsrView& A, const double* val, int reps, const Preconditioner* pre = nullptr);

    // True residual ||b - A x|| in FP64: returns {inf-norm rel, 2-norm rel}.
    void trueResidual(const CsrView& A, const double* val, const double* b, const double* x,
                      double& infRel, double& l2Rel);

    // Device timing of the last solve (ms), from HIP events on the stream.
    double lastMs() const { return lastMs_; }

    // Batch control: iterations launched between host polls.
    int minBatch = 8, maxBatch = 256;
    // flexible CG: replay one captured iteration (true) or launch its kernels one by one
    bool useGraphs = false;
    // measurement: flexible CG launches eagerly and times every fine-level k_fcg_A launch with HIP events;
    // probeMs / probeCount: average device time and count of those that ran (not the ones after the stop)
    bool probeFcgA = false;
    double probeMs = 0;
    int probeCount = 0;
    // the same probe on the preconditioner's finest-level post-smoothing launches (multigrid)
    double probeSmoothMs = 0;
    int probeSmoothCount = 0;
    // the drivers' FP64 vector solves: all components in shared launches (true) or one by one
    bool batchComponents = true;

private:
    template <typename T>
    SolveReport bicgRun(const CsrView& A, const T* val, const T* invD, const T* b, T* x, double tol, int maxIter,
                        T* const* work);
    template <typename T>
    void bicgRunComponents(const CsrView& A, const T* val, long long valStride, const T* invD, const T* b, T* x,
                           int ncomp, const double* tol, const int* maxIter, const bool* run, T* const* w,
                           SolveReport* reps);
    void reserveComponents();
    template <typename T>
    SolveReport pcgRun(const CsrView& A, const double* val64, const T* valT, const T* invD,
                       const double* b, double* x64, T* xT, T* r, double tol, int maxIter,
                       double* resHist, bool mixed, double delta);
    void poll();
    int nextBatch(int done, double r0, double rNow, double tolAbs, int maxIter);

    Context& ctx_;
    int n_, maxBlocks_;
    DevBuf state_, partials_;
    DevBuf r_, p_, q_, r0_, v_, s_, sh_, t_, z_, invD_, invDb_;
    DevBuf r32_, p32_, q32_, d32_, b32_, invD32_;
    DevBuf z32_, fcgPart_;  // allocated by the first fcgMixed
    DevBuf bf_[6], val32b_;  // FP32 BiCGStab vectors and matrix, allocated by the first bicgstabMixed
    DevBuf wc_, stC_, partC_;  // component BiCGStab vectors, states, partials (first component solve)
    DevBuf wc32_, r64c_, val32c_;  // FP32 vectors, FP64 residuals and FP32 values of the mixed component solve
    SolveState* hostC_ = nullptr;  // pinned
    // captured flexible-CG iteration, valid for the key (multigrid, buffers) and the stop parameters
    hipGraphExec_t fcgExec_ = nullptr;
    std::vector<hipEvent_t> probeEv_;  // event pairs of the k_fcg_A probe
    int probeUsed_ = 0;
    std::vector<const void*> fcgKey_;
    double fcgTol_ = -1, fcgDelta_ = -1;
    int fcgMaxIter_ = -1;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    double lastMs_ = 0;
};

// Default reliable-update factors.  An update of the flexible CG after a
// long run re-applies the preconditioner to the new true residual (about
// one iteration's cost).  Over a few iterations its FP32 recurrence stays
// accurate through a drop of 1e-5 (multigrid on the 128^3 pressure system:
// 21 iterations at 1e-2 .. 1e-5, 22 at 1e-6), but over long runs it drifts
// as Jacobi-PCG's does (711 iterations at 1e-2, 795 at 1e-4, 2,615 at 1e-6;
// Jacobi-preconditioned flexible CG 295 at 1e-2, 309 at 1e-4).  So the
// automatic rule updates after a drop of kFcgDelta, or of kPcgDelta once
// kFcgSegment iterations have passed since the last true residual.
constexpr double kFcgDelta = 1e-4;
constexpr double kPcgDelta = 1e-2;
constexpr int kFcgSegment = 16;

// Workspace of one Jacobi-PCG run: 4 x 2048 partials, the device state
// (kPcgStateBytes), the direction and product vectors (n entries of the
// Krylov type each) and, for FP32 runs with FP64 refinement, the FP64
// residual.
constexpr std::size_t kPcgStateBytes = 256;
constexpr int kPcgPartials = 4 * 2048;
struct PcgWork
{
    double* part;
    void* state;
    void* p;
    void* q;
    double* r64;
};
// FP64 Jacobi-PCG (solveCG, Matrix.cpp:153-222) on a caller's workspace: the
// kernels and the host loop of Krylov::cg, nothing allocated.  invD holds the
// CG-guarded inverse diagonal, r is the residual vector (n doubles).
SolveReport pcgSolveWith(Context& ctx, const CsrView& A, const double* val, const double* invD, const double* b,
                         double* x, double tol, int maxIter, const PcgWork& w, double* r);

// invD from a CSR diagonal: CG guard |d| > 1e-300 (Matrix.cpp:160-165),
// BiCGStab guard d > 1e-300 (Matrix.cpp:231-236).
void computeInvD(Context& ctx, const CsrView& A, const double* val, double* invD, bool bicgGuard);
void castToFloat(Context& ctx, const double* in, float* out, long long n);

} // namespace cfx
