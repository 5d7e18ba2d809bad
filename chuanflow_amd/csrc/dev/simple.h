// simple.h -- the pressure-velocity drivers on the device.
#pragma once

#include "fvops.h"
#include "krylov.h"
#include "mg.h"
#include "mpnn.h"

namespace cfx
{

struct Schemes
{
    int div = 0;              // 0 upwind, 1 linear, 2 blend
    int laplacianCorrected = 1;
    int gradCorrected = 1;
    double blendCoeff = 0.9;
};

// fvMatrix::solve (fvMatrix.cpp:73-129) on the fixed pattern: CSR values
// of every component from the LDU, base rhs kept, then per pass the rhs is
// restored, the laplacian corrector (if gammaFace) adds its terms from the
// current solution, and every component is solved (CG or BiCGStab) from x.
// On return A.rhs holds the last pass' rhs (what explicitOperator reads).
struct FvSolveArgs
{
    int solver = 0;  // 0 CG, 1 BiCGStab
    double tol = 1e-7;
    int maxIter = 2000;
    int nonOrth = 2;
    const double* gammaFace = nullptr;  // corrector registered when non-null
    const double* linc = nullptr;
    const double* link = nullptr;
    int sign = -1;
    bool gradCorrected = true;
    int precision = kFP64;
    double delta = 0.0;  // reliable-update factor of FP32-refined solves (0: the solver's default)
    MgCache* mg = nullptr;  // non-null: flexible CG with the multigrid preconditioner (FP32 + FP64 refinement)
    Mpnn* net = nullptr;    // non-null: finest-level weights from the network (per-slot geometry geo)
    const double* geo = nullptr;
    float *wa = nullptr, *wt = nullptr;
    bool projectMean = false;  // subtract the mean of every pass' rhs (singular Neumann systems)
    // the matrix values equal those of the previous solve on this mesh (a later PISO corrector): the
    // multigrid hierarchy's values are kept instead of set up again (without a network only, whose
    // weights depend on the right-hand side)
    bool sameMatrix = false;
    // Schemes::relaxation (FvMatrix.h:30-38): after the final pass x = xOld + relaxation (x - xOld) with
    // xOld the solution before that pass (fvMatrix.cpp:110-118); relaxOld holds ncomp * n doubles
    double relaxation = 1.0;
    double* relaxOld = nullptr;
};
SolveReport fvSolve(DevMesh& m, Krylov& K, const Ldu& A, double* val, double* rhsBase, DevBuf& val32,
                    double* x, const FvSolveArgs& a, double* gradTmp, double* faceTmp);

// volScalarField/volVectorField::linearise (volFields.cpp:117-147, :246-283)
// on the device: types/values per boundary face (value [nB][ncomp]) ->
// lin_c/lin_k [nB][ncomp].  Returns false if a symmetry/wall normal is not
// axis aligned (the reference throws there).
bool lineariseBc(DevMesh& m, int ncomp, const int* dType, const double* dValue, double* linc, double* link);

// Copies of the assembled pressure systems (CSR values and right-hand side
// of the last corrector pass) into caller buffers [max][nnz] / [max][n],
// every `every` pressure solves (SIMPLE iterations, PISO correctors).
struct PressureDump
{
    double* val = nullptr;
    double* rhs = nullptr;
    int every = 0, max = 0, count = 0, seen = 0;
};

// pPrecision == kFP32MG: flexible CG with the multigrid preconditioner and,
// when `net` is set, the network's finest-level weights (geometry per slot).
struct PressureSolver
{
    MgCache* mg = nullptr;
    Mpnn* net = nullptr;
    const double* geo = nullptr;
    float *wa = nullptr, *wt = nullptr;
};
constexpr int kFP32MG = 2;

struct SimpleControls
{
    int maxIterations = 200;
    double continuityTol = 1e-6, pcorrTol = 0, momentumTol = 0;
    double linearTol = 1e-7;
    int linearMaxIter = 2000, nonOrthCorrectors = 2;
    double relaxU = 0.7, relaxP = 0.3, pRelaxStart = 0;
    double bodyForce[3] = {0, 0, 0};
    int pPrecision = kFP64;  // pressure solve: FP64, FP32 inner + FP64 refinement, or kFP32MG
    int uPrecision = kFP64;  // momentum solve: FP64 or FP32 BiCGStab with FP64 refinement
    PressureSolver* pSolver = nullptr;
    PressureDump* dump = nullptr;
};

struct SimpleResult
{
    int iterations = 0;
    double finalContinuityError = 0, finalMomentumResidual = 0;
    int converged = 0, diverged = 0;
    double ms = 0;
};

// solveIncompressibleSimple (simpleSolver.cpp:33-394), laminar, steady.
// U [3][n], p [n], phi [nF] device, in/out.  Boundary conditions per
// boundary face (host): type codes 0 fixedValue, 1 fixedGradient,
// 2 zeroGradient/outlet, 3 symmetry (vectors: symmetry or wall).
// hist (host, optional) gets {contErr, errAfter, pDeltaMax, uRes} per iteration.
SimpleResult simpleSolve(DevMesh& m, Krylov& K, double nu, const Schemes& s, const SimpleControls& c,
                         const int* Utype, const double* Uvalue, const int* ptype, const double* pvalue, double* U,
                         double* p, double* phi, double* hist);

struct PisoControls
{
    double dt = 0.01;
    int correctors = 2;
    double continuityTol = 1e-10, linearTol = 1e-8;
    int linearMaxIter = 2000, nonOrthCorrectors = 2;
    double bodyForce[3] = {0, 0, 0};
    int pPrecision = kFP64;
    int uPrecision = kFP64;
    PressureSolver* pSolver = nullptr;
    PressureDump* dump = nullptr;
    bool projectNeumann = false;  // zero-sum pressure rhs when no pressure face is fixed
};

struct PisoResult
{
    int correctorsUsed = 0;
    double predictedContinuityError = 0, continuityError = 0, pressureCorrectionMax = 0;
    int converged = 0;
    double ms = 0;
};

// solveTransientIncompressiblePiso (pisoSolver.cpp:13-235), laminar, one step.
PisoResult pisoStep(DevMesh& m, Krylov& K, double nu, const Schemes& s, const PisoControls& c, const int* Utype,
                    const double* Uvalue, const int* ptype, const double* pvalue, double* U, double* p, double* phi);

} // namespace cfx
