/*
 * cfo.h -- CPU oracle for the SIMPLE pressure-velocity hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, single-threaded
 * restatement of the reference algorithms, used as the checker for the HIP
 * path and as the timed CPU baseline.  Nothing in the shipped library links,
 * loads or calls it; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may.
 *
 * Every function follows a reference function in operation order (so that
 * results are bit-identical to the reference build on the same inputs); the
 * reference file:line each one restates is given beside it.
 *
 * Conventions: indices are int64 (reference `label`), scalars are double,
 * vectors are AoS [n][3] like the reference's Field<Vector>.
 */
#ifndef CFO_H
#define CFO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- RNG */
/* MT19937-64 (Matsumoto & Nishimura 2004), the engine std::mt19937_64 names. */
typedef struct { uint64_t mt[312]; int mti; } cfo_mt64;
void cfo_mt64_seed(cfo_mt64* g, uint64_t seed);
uint64_t cfo_mt64_next(cfo_mt64* g);
/* n-th output (1-based) of a freshly seeded engine. */
uint64_t cfo_mt64_nth(uint64_t seed, int64_t n);
/* uniform double in [0,1): top 53 bits. */
double cfo_mt64_uniform(cfo_mt64* g);

/* ---------------------------------------------------------------- mesh */
typedef struct cfo_mesh {
    int64_t nPoints, nFaces, nInternal, nCells, nPatches;
    const double*  points;      /* [nPoints][3] */
    const int64_t* faceStart;   /* [nFaces+1]   */
    const int64_t* faceVerts;   /* [faceStart[nFaces]] */
    const int64_t* owner;       /* [nFaces]     */
    const int64_t* neighbour;   /* [nInternal]  */
    const int64_t* patchStart;  /* [nPatches]   */
    const int64_t* patchSize;   /* [nPatches]   */
    /* geometry, caller-allocated, filled by cfo_geometry */
    double* Sf;      /* [nFaces][3] */
    double* Cf;      /* [nFaces][3] */
    double* magSf;   /* [nFaces]    */
    double* C;       /* [nCells][3] */
    double* V;       /* [nCells]    */
    double* dvec;    /* [nFaces][3] */
    double* invDist; /* [nFaces]    */
} cfo_mesh;

/* BlockMesh::create (src/3d/src/mesh/blockMesh.cpp:36-189): faces in
 * first-encounter order, internal first, then the six sides
 * xmin xmax ymin ymax zmin zmax.  4 vertices per face.
 * Sizes: points (nx+1)(ny+1)(nz+1); faces 3nxnynz+nxny+nynz+nznx. */
int cfo_blockmesh(int64_t nx, int64_t ny, int64_t nz,
                  const double mn[3], const double mx[3],
                  double* points, int64_t* faceVerts,
                  int64_t* owner, int64_t* neighbour,
                  int64_t* patchStart, int64_t* patchSize,
                  int64_t* nInternalOut);

/* Seeded uniform jitter of strictly interior vertices of a structured
 * (nx+1)(ny+1)(nz+1) point lattice: x += (2u-1)*amp*h per component, u from
 * MT19937-64, points visited in lattice order (i fastest).  No reference
 * counterpart (config generator). */
void cfo_jitter(int64_t nx, int64_t ny, int64_t nz,
                const double mn[3], const double mx[3],
                double amp, uint64_t seed, double* points);

/* polyMesh constructor checks + calcGeometry
 * (src/3d/src/mesh/polyMesh.cpp:12-64,83-288).  Returns 0, or -1 on a
 * degenerate cell / zero distance (the reference throws there). */
int cfo_geometry(const cfo_mesh* m);

/* -------------------------------------------------------------- fields */
enum { CFO_BC_FIXED_VALUE = 0, CFO_BC_FIXED_GRADIENT = 1,
       CFO_BC_ZERO_GRADIENT = 2, CFO_BC_SYMMETRY = 3 };

/* Boundary linearisation Tf = c*Tp + k per boundary face and component
 * (volFields.cpp:117-147 scalar, :246-283 vector).  type/value are given per
 * boundary face (face f at index f-nInternal); value is [nB][ncomp].
 * For vectors SYMMETRY covers the reference's "symmetry" and "wall".
 * Output lin_c/lin_k are [nB][ncomp].  Returns -1 if a symmetry normal is not
 * axis aligned. */
int cfo_linearise(const cfo_mesh* m, int ncomp, const int32_t* type,
                  const double* value, double* lin_c, double* lin_k);

/* makeInternalOwnerWeights (interpolate.cpp:15-35): w[F]. */
void cfo_owner_weights(const cfo_mesh* m, double* w);

/* Gauss gradient (gradScheme.cpp:15-96).  corrected!=0 -> bounded face
 * interpolation + 2 skew passes; boundary faces use lin (ncomp stride and
 * component index select the column of lin_c/lin_k).
 * quirkLinear!=0 reproduces volScalarField::faceValue on internal faces
 * (volFields.cpp:101-111) for the uncorrected scheme. */
void cfo_grad(const cfo_mesh* m, const double* data, const double* lin_c,
              const double* lin_k, int ncomp, int comp, int corrected,
              int quirkLinear, double* gradOut /*[n][3]*/);

/* harmonicFaceGamma for a uniform cell gamma (laplacianScheme.cpp:37-62). */
void cfo_harmonic_gamma_uniform(const cfo_mesh* m, double gamma, double* gf);

/* interpolateToFaces (pressureCorrection.cpp:27-35). */
void cfo_interpolate_to_faces(const cfo_mesh* m, const double* cell, double* face);

/* ----------------------------------------------------------- matrices */
/* Assembled scalar-block matrix in the reference's triplet order:
 * per component diag[n], rhs[n], and the off-diagonal coefficients per
 * internal face: upper = entry (owner row, neighbour col), lower = entry
 * (neighbour row, owner col).  Duplicates (div + laplacian on the same face)
 * are summed as the reference CSR merge does (Matrix.cpp:102-105). */
typedef struct cfo_ldu {
    int64_t n, F;
    int ncomp;
    double* diag;   /* [ncomp][n] */
    double* rhs;    /* [ncomp][n] */
    double* upper;  /* [F] (shared by components) */
    double* lower;  /* [F] */
} cfo_ldu;

/* Reset an LDU: diag/rhs to +0.0 (the reference zero-initialises them),
 * upper/lower to -0.0 (the additive identity: the first contribution is
 * stored unchanged, as the reference stores the first triplet). */
void cfo_ldu_reset(cfo_ldu* A);

/* fvm::laplacian/laplacianFace implicit part (laplacianScheme.cpp:67-104).
 * gammaFace[nFaces]; adds into an LDU that may already hold terms. */
void cfo_assemble_laplacian(const cfo_mesh* m, const double* gammaFace,
                            const double* lin_c, const double* lin_k,
                            int sign, cfo_ldu* A);

/* fvm::div (divScheme.cpp:11-62); scheme 0 upwind, 1 linear, 2 blend. */
void cfo_assemble_div(const cfo_mesh* m, const double* phi, int scheme,
                      double blendCoeff, const double* lin_c,
                      const double* lin_k, int sign, cfo_ldu* A);

/* fvm::ddt for a vector/scalar old field [n][ncomp] (divScheme.cpp:64-94). */
void cfo_assemble_ddt(const cfo_mesh* m, const double* oldField, double dt,
                      int sign, cfo_ldu* A);

/* Laplacian non-orthogonal corrector (laplacianScheme.cpp:106-154) for one
 * component: adds into rhs[n]. */
void cfo_nonorth_correct(const cfo_mesh* m, const double* gammaFace,
                         const double* x, const double* lin_c,
                         const double* lin_k, int ncomp, int comp,
                         int gradCorrected, int sign, double* rhs);

/* CSR in the reference layout (Matrix.cpp:56-117): rows sorted by column. */
typedef struct cfo_csr {
    int64_t n, nnz;
    int64_t* rowStart; /* [n+1] */
    int64_t* cols;     /* [nnz] */
    double*  vals;     /* [nnz] */
} cfo_csr;

/* Build the CSR pattern of the face graph + diagonal and fill values of one
 * component from an LDU.  Caller allocates rowStart[n+1], cols/vals[n+2F]. */
void cfo_csr_from_ldu(const cfo_mesh* m, const cfo_ldu* A, int comp, cfo_csr* out);

/* SparseMatrix::product (Matrix.cpp:119-133). */
void cfo_spmv(const cfo_csr* A, const double* x, double* y);

typedef struct cfo_report {
    int iterations;
    double initialResidual, finalResidual;
    int converged;
} cfo_report;

/* solveCG (Matrix.cpp:153-222), Jacobi-preconditioned; x in/out.
 * resHist (optional, [maxIter+1]) gets ||r||inf per iteration. */
cfo_report cfo_solve_cg(const cfo_csr* A, const double* b, double* x,
                        double tol, int maxIter, double* resHist);

/* solveBiCGStab (Matrix.cpp:224-319). */
cfo_report cfo_solve_bicgstab(const cfo_csr* A, const double* b, double* x,
                              double tol, int maxIter);

/* ------------------------------------------------------ flux operators */
/* updatePhiInternal (pressureCorrection.cpp:10-25); U [n][3]. */
void cfo_update_phi_internal(const cfo_mesh* m, const double* U, double* phi);
/* surfaceScalarField::correctBoundaryFlux (volFields.cpp:294-306). */
void cfo_correct_boundary_flux(const cfo_mesh* m, const double* U,
                               const double* lin_c, const double* lin_k,
                               double* phi);
/* correctFluxes (pressureCorrection.cpp:37-88); corrected = laplacian
 * scheme is "corrected". */
void cfo_correct_fluxes(const cfo_mesh* m, const double* p,
                        const double* gradP, const double* rAUf,
                        const double* lin_c, const double* lin_k,
                        int corrected, double* phi);
/* fvc::divPhi (interpolate.cpp:138-151). */
void cfo_div_phi(const cfo_mesh* m, const double* phi, double* out);
/* continuityError (pressureCorrection.cpp:90-125). */
double cfo_continuity_error(const cfo_mesh* m, const double* phi);

/* ---------------------------------------------------------- drivers */
typedef struct cfo_schemes {
    int divScheme;        /* 0 upwind, 1 linear, 2 blend */
    int laplacianCorrected;
    int gradCorrected;
    double blendCoeff;
} cfo_schemes;

typedef struct cfo_simple_controls {
    int maxIterations;
    double continuityTol, pcorrTol, momentumTol;
    double linearTol;
    int linearMaxIter, nonOrthCorrectors;
    double relaxU, relaxP, pRelaxStart;
    double bodyForce[3];
} cfo_simple_controls;

typedef struct cfo_simple_result {
    int iterations;
    double finalContinuityError, finalMomentumResidual;
    int converged;
    int diverged;
} cfo_simple_result;

/* Boundary conditions of a field, per boundary face. */
typedef struct cfo_bc {
    const int32_t* type;  /* [nB] */
    const double* value;  /* [nB][ncomp] */
} cfo_bc;

/* solveIncompressibleSimple (simpleSolver.cpp:33-394), laminar, no energy
 * hook, steady (no ddt).  U [n][3], p [n], phi [nFaces] in/out.
 * hist (optional) receives per iteration {contErr, errAfter, pDeltaMax, uRes}. */
cfo_simple_result cfo_simple(const cfo_mesh* m, double* U, double* p,
                             double* phi, double nu, const cfo_schemes* s,
                             const cfo_simple_controls* ctl,
                             cfo_bc Ubc, cfo_bc pbc, double* hist);

typedef struct cfo_piso_controls {
    double dt;
    int correctors;
    double continuityTol, linearTol;
    int linearMaxIter, nonOrthCorrectors;
    double bodyForce[3];
    /* not in the reference: with no fixedValue pressure face, project each
     * pass' pressure right-hand side to zero sum (the compatibility
     * condition of the singular Neumann system); 0 = the reference */
    int projectNeumann;
} cfo_piso_controls;

typedef struct cfo_piso_result {
    int correctorsUsed;
    double predictedContinuityError, continuityError, pressureCorrectionMax;
    int converged;
} cfo_piso_result;

/* solveTransientIncompressiblePiso (pisoSolver.cpp:13-235), one step. */
cfo_piso_result cfo_piso_step(const cfo_mesh* m, double* U, double* p,
                              double* phi, double nu, const cfo_schemes* s,
                              const cfo_piso_controls* ctl, cfo_bc Ubc,
                              cfo_bc pbc);

#ifdef __cplusplus
}
#endif
#endif
