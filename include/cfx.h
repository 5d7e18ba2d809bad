/*
 * cfx.h -- C ABI of the MI355X SIMPLE pressure-velocity hot path.
 *
 * Plain pointers and sizes only.  Conventions:
 *  - every function returns an int status: CFX_OK (0) or a CFX_E_* code;
 *    cfx_last_error(ctx) returns the message of the last failure on ctx;
 *  - pointers named d_* are DEVICE pointers (cfx_malloc or any HIP
 *    allocation on the context's device); everything else is host memory,
 *    borrowed for the duration of the call;
 *  - cell vectors on the device are SoA [3][nCells]; boundary
 *    linearisations are lin_c/lin_k [nBoundaryFaces][ncomp] with
 *    Tf = c * Tp + k (the reference's volScalarField/volVectorField::linearise);
 *  - one HIP stream per context; a context is not thread safe, separate
 *    contexts are independent.
 *
 * Each declaration names the reference interface it replaces
 * (reference file:line).
 */
#ifndef CFX_H
#define CFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFX_OK 0
#define CFX_E_INVALID 1     /* bad argument / shape */
#define CFX_E_HIP 2         /* HIP runtime error */
#define CFX_E_STATE 3       /* object in the wrong state */
#define CFX_E_NUMERIC 4     /* non-finite values where the reference throws */
#define CFX_E_INTERNAL 5

typedef struct cfx_ctx cfx_ctx;
typedef struct cfx_mesh cfx_mesh;
typedef struct cfx_matrix cfx_matrix;
typedef struct cfx_csr cfx_csr;

/* SolverReport (src/common/include/cf/common/Matrix.h:57-63) plus
 * extensions: true FP64 residuals, refinement sweeps, preconditioner
 * fallbacks, device time. */
typedef struct cfx_report {
    int iterations;
    double initialResidual;
    double finalResidual;
    int converged;
    int breakdown;          /* 1: pAp <= 0 (not SPD); 2: BiCGStab t.t < 1e-300 */
    double l2Residual;      /* ||b - A x||2 / ||b||2, FP64 (mixed precision) */
    double trueResidual;    /* ||b - A x||inf / ||b||inf, FP64 (mixed precision) */
    int outerIterations;    /* FP64 refinement sweeps */
    int precondFallbacks;   /* learned preconditioner rejected -> Jacobi */
    double ms;              /* device time of the solve */
} cfx_report;

/* CFX_PRECOND_LEARNED: multigrid flexible CG with the attached network's
 * finest-level weights (CFX_E_STATE without one); CFX_PRECOND_MG: the same
 * solver with the multigrid's own weights (cfx_mg_params constants or the
 * per-row arrays of *_set_mg_weights), any network ignored */
enum { CFX_PRECOND_JACOBI = 0, CFX_PRECOND_DIC = 1, CFX_PRECOND_LEARNED = 2, CFX_PRECOND_MG = 3 };
enum { CFX_FP64 = 0, CFX_FP32_IR = 1, CFX_FP32_MG = 2 /* drivers' pressure solve: multigrid FCG */ };
enum { CFX_SOLVER_CG = 0, CFX_SOLVER_BICGSTAB = 1 };

/* ------------------------------------------------------------ context */
/* No reference counterpart (the reference is single-process CPU code). */
int cfx_create(int device, cfx_ctx** out);
int cfx_destroy(cfx_ctx* ctx);
const char* cfx_last_error(const cfx_ctx* ctx);
int cfx_synchronize(cfx_ctx* ctx);
/* name: gcnArch string, cus: compute units */
int cfx_device_info(cfx_ctx* ctx, char* name, size_t nameLen, int* cus);
/* what a plain stream reaches on this device, the yardstick of the bandwidth
 * rooflines: the best over a few grid sizes of a read-only pass and of a copy
 * of `bytes` (>= 1 MiB), each timed over `reps` launches on the context's
 * stream; GB/s of bytes read (copy: read + written).  No reference
 * counterpart. */
int cfx_stream_bandwidth(cfx_ctx* ctx, size_t bytes, int reps, double* readGBps, double* copyGBps);
int cfx_malloc(cfx_ctx* ctx, size_t bytes, void** d_out);
int cfx_free(cfx_ctx* ctx, void* d_ptr);
int cfx_memcpy_htod(cfx_ctx* ctx, void* d_dst, const void* src, size_t bytes);
int cfx_memcpy_dtoh(cfx_ctx* ctx, void* dst, const void* d_src, size_t bytes);
int cfx_memcpy_dtod(cfx_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
/* zero-fill by kernel (safe inside captured graphs) */
int cfx_zero(cfx_ctx* ctx, void* d_ptr, size_t bytes);
/* weighted sums over index segments, on the device: out[2 s] = sum x[i] w[i], out[2 s + 1] = sum w[i]
 * for i = d_idx[d_segStart[s] .. d_segStart[s + 1]) (d_w null: weights 1); out is host [2 nSeg].
 * The section means and boundary fluxes of a detail region (coupledNetwork.cpp:1154-1200) without
 * copying its fields to the host; no counterpart in the reference. */
int cfx_segment_sums(cfx_ctx* ctx, int nSeg, const int64_t* d_segStart, const int64_t* d_idx, const double* d_x,
                     const double* d_w, double* out);

/* ------------------------------------------------------------- mesh */
/* Device mesh from a reference polyMesh's arrays
 * (src/3d/include/cf/fvm/PolyMesh.h:17-103): owner/neighbour (int64 label),
 * faceSf, faceCentres, cellCentres, cellVolumes, deltaCoeffVec, faceAreas
 * (Field<Vector> AoS [n][3]), patches as start/size.  Builds the face
 * streams, cell->face gathers, geometry-derived face coefficients and the
 * fixed CSR pattern (the reference rebuilds it by a sort on every solve,
 * src/common/src/Matrix.cpp:56-117). */
int cfx_mesh_create(cfx_ctx* ctx, int64_t nCells, int64_t nInternal, int64_t nFaces,
                    const int64_t* owner, const int64_t* neighbour, const double* Sf,
                    const double* Cf, const double* C, const double* V, const double* d,
                    const double* magSf, int64_t nPatches, const int64_t* patchStart,
                    const int64_t* patchSize, cfx_mesh** out);
int cfx_mesh_destroy(cfx_mesh* mesh);
/* sizes: {nCells, nInternal, nFaces, nnz, nRowBlocks} */
int cfx_mesh_sizes(cfx_mesh* mesh, int64_t sizes[5]);
/* host copy of the CSR pattern: rowPtr[nCells+1], cols[nnz] (int32) */
int cfx_mesh_pattern(cfx_mesh* mesh, int32_t* rowPtr, int32_t* cols);
/* makeInternalOwnerWeights (src/3d/src/discrete/interpolate.cpp:28-35) */
int cfx_owner_weights(cfx_mesh* mesh, double* d_w);

/* ------------------------------------------------------- field ops */
/* fvc::grad (src/3d/include/cf/fvm/FvMatrix.h:122-125, gradScheme.cpp:15-96):
 * corrected != 0 -> "corrected" (2 skew passes); else "linear", where
 * quirk != 0 reproduces volScalarField::faceValue on internal faces
 * (volFields.cpp:101-111) as fvc::grad(volScalarField) does. */
int cfx_grad(cfx_mesh* mesh, const double* d_x, const double* d_linc, const double* d_link,
             int ncomp, int comp, int corrected, int quirk, double* d_grad3);
/* harmonicFaceGamma with a uniform value (laplacianScheme.cpp:37-62) */
int cfx_harmonic_gamma_uniform(cfx_mesh* mesh, double gamma, double* d_gammaFace);
/* interpolateToFaces (src/3d/include/cf/fvm/PressureCorrection.h:28) */
int cfx_interpolate_to_faces(cfx_mesh* mesh, const double* d_cell, double* d_face);
/* updatePhiInternal (PressureCorrection.h:23) */
int cfx_update_phi_internal(cfx_mesh* mesh, const double* d_U3, double* d_phi);
/* surfaceScalarField::correctBoundaryFlux (src/3d/include/cf/fvm/Fields.h:144) */
int cfx_correct_boundary_flux(cfx_mesh* mesh, const double* d_U3, const double* d_linc3,
                              const double* d_link3, double* d_phi);
/* correctFluxes, the Rhie-Chow closure (PressureCorrection.h:15-20) */
int cfx_correct_fluxes(cfx_mesh* mesh, const double* d_p, const double* d_gradP3,
                       const double* d_rAUf, const double* d_linc, const double* d_link,
                       int corrected, double* d_phi);
/* fvc::divPhi (FvMatrix.h:131) */
int cfx_div_phi(cfx_mesh* mesh, const double* d_phi, double* d_out);
/* continuityError (PressureCorrection.h:31) */
int cfx_continuity_error(cfx_mesh* mesh, const double* d_phi, double* err);

/* ----------------------------------------------------------- fvMatrix */
/* fvMatrix (src/3d/include/cf/fvm/FvMatrix.h:42-101) on the fixed pattern. */
int cfx_matrix_create(cfx_mesh* mesh, int ncomp, cfx_matrix** out);
int cfx_matrix_destroy(cfx_matrix* A);
int cfx_matrix_reset(cfx_matrix* A);
/* fvm::laplacian / fvm::laplacianFace (FvMatrix.h:170-175, laplacianScheme.cpp:67-104) */
int cfx_assemble_laplacian(cfx_matrix* A, const double* d_gammaFace, const double* d_linc,
                           const double* d_link, int sign);
/* fvm::div (FvMatrix.h:166, divScheme.cpp:11-62): scheme 0 upwind, 1 linear, 2 blend */
int cfx_assemble_div(cfx_matrix* A, const double* d_phi, int scheme, double blendCoeff,
                     const double* d_linc, const double* d_link, int sign);
/* fvm::ddt (FvMatrix.h:162-163, divScheme.cpp:64-94); d_old is [ncomp][n] */
int cfx_assemble_ddt(cfx_matrix* A, const double* d_old, double dt, int sign);
/* fvMatrix::addRhs over all cells of one component: rhs[comp] += d_v */
int cfx_matrix_add_rhs(cfx_matrix* A, int comp, const double* d_v);
/* host copies of diag/rhs [ncomp][n] and upper/lower [nInternal] */
int cfx_matrix_get(cfx_matrix* A, double* diag, double* rhs, double* upper, double* lower);
/* device pointers of diag/rhs [ncomp][n] */
int cfx_matrix_device_ptrs(cfx_matrix* A, double** d_diag, double** d_rhs);
/* the laplacian corrector run by fvMatrix::solve before each pass
 * (laplacianScheme.cpp:111-154, fvMatrix.cpp:86-91): adds into d_rhs */
int cfx_nonorth_correct(cfx_mesh* mesh, const double* d_gammaFace, const double* d_x,
                        const double* d_linc, const double* d_link, int ncomp, int comp,
                        int gradCorrected, int sign, double* d_rhs);
/* fvMatrix::solve (FvMatrix.h:77, fvMatrix.cpp:73-129): builds the CSR
 * values, then (nonOrth+1 passes when d_gammaFace != NULL, else 1) restores
 * the rhs, applies the corrector from the current solution and solves every
 * component (CG if solver == CFX_SOLVER_CG else BiCGStab) from d_x
 * ([ncomp][n], initial guess and result).  rep gets the last solve. */
int cfx_matrix_solve(cfx_matrix* A, double* d_x, int solver, double tol, int maxIter, int nonOrth,
                     const double* d_gammaFace, const double* d_linc, const double* d_link,
                     int sign, int gradCorrected, int precond, int precision, cfx_report* rep);
/* Schemes::relaxation of this matrix (FvMatrix.h:30-38, default 1): after the
 * final pass of cfx_matrix_solve each component becomes
 * xOld + relaxation (x - xOld), xOld its value before that pass
 * (fvMatrix.cpp:110-118).  relaxation in (0, 1]. */
int cfx_matrix_set_relaxation(cfx_matrix* A, double relaxation);
/* CFX_FP32_IR solves of this matrix: the factor by which the FP32 recurrence
 * residual must fall below the last true FP64 residual before the next
 * reliable update (0, the default: each solver's own, as
 * cfx_csr_set_inner_tol).  No counterpart in the reference. */
int cfx_matrix_set_inner_tol(cfx_matrix* A, double innerTol);
/* explicitOperator (FvMatrix.h:86, fvMatrix.cpp:131-146): h = rhs - (A x - D x) */
int cfx_matrix_explicit_operator(cfx_matrix* A, const double* d_x, int comp, double* d_h);
/* y = A x with the matrix's current CSR values of one component */
int cfx_matrix_spmv(cfx_matrix* A, int comp, const double* d_x, double* d_y);

/* ------------------------------------------------- linear-solver seam */
/* SparseMatrix (Matrix.h:13-55) uploaded from reference CSR arrays
 * (int64 rowStart/cols, rows sorted by column). */
int cfx_csr_create(cfx_ctx* ctx, int64_t n, const int64_t* rowStart, const int64_t* cols,
                   const double* vals, cfx_csr** out);
int cfx_csr_destroy(cfx_csr* A);
/* replace the values (same pattern) */
int cfx_csr_set_values(cfx_csr* A, const double* vals);
/* the same from a device array (stream-ordered, no host synchronisation); no counterpart in the
 * reference, whose matrices live in host memory */
int cfx_csr_set_values_device(cfx_csr* A, const double* d_vals);
/* SparseMatrix::product (Matrix.h:47, Matrix.cpp:119-133) */
int cfx_csr_spmv(cfx_csr* A, const double* d_x, double* d_y);
/* solveCG / solveBiCGStab / solveWithBackend (Matrix.h:67-89): host b, x
 * (x = initial guess and result), tol/maxIter as the reference.  precision
 * CFX_FP32_IR (an extension): CG with FP32 Krylov vectors and FP64 reliable
 * updates, or BiCGStab in FP32 with FP64 iterative refinement; both decide
 * convergence on the true FP64 residual with the reference's stop test. */
int cfx_csr_solve(cfx_csr* A, const double* b, double* x, double tol, int maxIter, int solver,
                  int precond, int precision, cfx_report* rep);
/* the same with device-resident b and x */
int cfx_csr_solve_device(cfx_csr* A, const double* d_b, double* d_x, double tol, int maxIter,
                         int solver, int precond, int precision, cfx_report* rep);
/* per-iteration ||r||inf of the last FP64 CG solve: hist[0..iterations] */
int cfx_csr_residual_history(cfx_csr* A, double* hist, int maxLen);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, code run under stream capture, harnesses that
 * forbid hipMalloc): the FP64 Jacobi-PCG of cfx_csr_solve_device (CG,
 * CFX_PRECOND_JACOBI, CFX_FP64; solveCG's stop test and guards) on a device
 * CSR system the caller owns -- int32 d_rowPtr[n+1], d_col[nnz], FP64
 * d_val[nnz], d_b[n], d_x[n] (initial guess and result) -- with every
 * temporary in a library-owned static device array of fixed size: nothing is
 * allocated.  At most CFX_NOALLOC_MAX_ROWS rows (larger: CFX_E_INVALID).
 * Each GPU holds one such array (about 21 MB) until the process ends; all
 * calls on a GPU share it, so calls that overlap on one GPU race.  The
 * result equals cfx_csr_solve_device's to rounding (the row blocks differ,
 * so dot products are grouped differently); rep->ms is host wall time. */
#define CFX_NOALLOC_MAX_ROWS 524288
int cfx_pcg_solve_noalloc(cfx_ctx* ctx, int n, int nnz, const int* d_rowPtr, const int* d_col,
                          const double* d_val, const double* d_b, double* d_x, double tol, int maxIter,
                          cfx_report* rep);
/* CFX_FP32_IR: the factor by which the FP32 recurrence residual must fall
 * below the last true FP64 residual before the next reliable update; 0 (the
 * default) takes each solver's own: 1e-2 for Jacobi-PCG, 1e-4 for the
 * flexible CG of the multigrid, learned and DIC preconditioners (whose every
 * update costs one more preconditioner application) */
int cfx_csr_set_inner_tol(cfx_csr* A, double innerTol);

/* ------------------------------- multigrid / learned preconditioner */
/* precond = CFX_PRECOND_LEARNED (solver CFX_SOLVER_CG, precision
 * CFX_FP32_IR): flexible CG (alpha = p.r/p.q, beta = -z.q/p.q) preconditioned
 * by one aggregation-multigrid V-cycle, FP32 Krylov vectors with FP64
 * reliable updates and the reference's stop test on true FP64 residuals; on
 * a breakdown or a non-finite value it continues with Jacobi-PCG
 * (report.precondFallbacks = 1).  The finest level takes per-row weights
 * a (pre-smoother a/D), s (prolongation scale), c (post-smoother c/D) and
 * t (restriction scale, coarse rhs = P^T diag(t) r): constants from
 * cfx_mg_params, or per-row values from the message-passing network
 * (cfx_mpnn_*) or any device array.  a = c and s = t give a symmetric
 * cycle; a = 1, s = 0, c = 0 gives z = D^-1 r exactly (Jacobi).  The reference has no counterpart: it offers
 * Jacobi-PCG and Jacobi-BiCGStab only (Matrix.cpp:153-319). */
typedef struct cfx_mg_params {
    float omegaPre, scale, omegaPost;   /* finest-level a, s, c */
    float restrictScale;                /* finest-level t */
    float coarseOmega, coarseScale;     /* coarse levels: a = c, s */
    int coarseSweeps;                   /* damped Jacobi sweeps on the coarsest level */
    float coarsestOmega;
    int maxLevels, coarseMax;           /* hierarchy depth cap, coarsest size cap (<= 2048; <= 128: exact solve) */
    int kcycleDepth;                    /* coarse levels 1..kcycleDepth use the K-cycle (0: V-cycle) */
    int fusion;                         /* bit 0: the cycle of the small coarsest levels (below the K-cycle
                                           levels) in one workgroup launch; bit 1: prolongation and K-cycle
                                           combination fused into the post-smoothing of coarse levels up to
                                           65536 rows; bit 2: that small-level cycle run from LDS by every
                                           workgroup of the post-smoothing of the level above it (no z.q
                                           partials there); bit 3: the K-cycle's second right-hand side
                                           formed inside the pre-smoothing; bit 4: the K-cycle combination of
                                           a level of at most 256 workgroups formed inside the prolongation;
                                           bit 5: the restriction from a coarse level (whose rows are
                                           numbered aggregate by aggregate) formed by its pre-smoothing;
                                           bit 6: the finest level's smoothing sweeps and the flexible CG's
                                           product on copies of its values in 64-row slices, stored
                                           slot-major (where cfx_csr_mg_slices is not 0)
                                           (default 125 = bits 0, 2, 3, 4, 5, 6; bit 1 measured slower);
                                           results are bit for bit those of fusion 0 */
    int smoothFine, smoothCoarse;       /* smoothing sweeps before and after the coarse correction (the same
                                           number, 1..8) on the finest level and on level 1 of the K-cycle
                                           levels (default 4 and 2; the levels below the K-cycle levels take
                                           one) */
    int smoothDeep;                     /* the same on the K-cycle levels below level 1 (1..8) */
} cfx_mg_params;
int cfx_mg_default_params(cfx_mg_params* p);
/* flexible CG replays one captured iteration per step (hipGraph, 1) or launches
 * its kernels one by one (0, the default: measured faster); both give identical results */
int cfx_set_graphs(cfx_ctx* ctx, int enable);
/* the drivers' and fvMatrix solves of vector equations with FP64 BiCGStab
 * (fvMatrix.cpp:93-122 solves component after component): all components in
 * shared launches (default 1) or one after another (0); both give identical
 * results.  No counterpart in the reference. */
int cfx_set_component_batching(cfx_ctx* ctx, int enable);
int cfx_csr_set_mg(cfx_csr* A, const cfx_mg_params* p);
/* per-row finest-level a, s, c, t (device FP32 [n] each, or null for the
 * constant); the pointers are kept and read whenever the hierarchy's values
 * are set up (after cfx_csr_set_values or this call) */
int cfx_csr_set_mg_weights(cfx_csr* A, const float* d_pre, const float* d_scale, const float* d_post,
                           const float* d_restrict);
/* the drivers' pressure solves with pPrecision CFX_FP32_MG use the mesh's
 * multigrid constants and, when attached, its network; every `every`-th
 * pressure system (CSR values, rhs) is copied into d_val [max][nnz],
 * d_rhs [max][n] (training data; every <= 0 stops) */
int cfx_mesh_set_mg(cfx_mesh* m, const cfx_mg_params* p);
int cfx_mesh_set_pressure_dump(cfx_mesh* m, double* d_val, double* d_rhs, int every, int max);
int cfx_mesh_pressure_dump_count(cfx_mesh* m, int* count);
/* the same for an fvMatrix: cfx_matrix_solve with precond
 * CFX_PRECOND_LEARNED (CG, CFX_FP32_IR) rebuilds the Galerkin values from
 * the assembled values at each call */
int cfx_matrix_set_mg(cfx_matrix* A, const cfx_mg_params* p);
int cfx_matrix_set_mg_weights(cfx_matrix* A, const float* d_pre, const float* d_scale, const float* d_post,
                              const float* d_restrict);
/* hierarchy sizes: rows[l], nnz[l] for l < min(maxOut, levels) */
int cfx_csr_mg_info(cfx_csr* A, int* levels, int64_t* rows, int64_t* nnz, int maxOut);
/* first level whose cycle, with every coarser level, fits one workgroup (cfx_mg_params.fusion bit 0
 * runs it as one launch below the K-cycle levels); -1: none */
int cfx_csr_mg_tail_level(cfx_csr* A, int* level);
/* slots of the finest level's 64-row sliced layout (cfx_mg_params.fusion bit 6; nonzeros plus padding), 0
 * where its pattern does not allow it (an int16 column offset overflows, or padding above a quarter) */
int cfx_csr_mg_slices(cfx_csr* A, int64_t* slots);

/* precond = CFX_PRECOND_DIC (solver CG, precision CFX_FP32_IR): flexible CG
 * preconditioned by diagonal incomplete Cholesky, D_ii = a_ii - sum_{j<i}
 * a_ij^2 / D_jj (OpenFOAM's DIC; no counterpart in the reference), with
 * level-scheduled substitutions (one kernel per level of the ordering's
 * dependency graph).  cfx_csr_dic_info gives the level count and the FP64
 * factor D (host [n]); cfx_csr_dic_apply is z = M^-1 r on FP32 device vectors. */
int cfx_csr_dic_info(cfx_csr* A, int* levels, double* diag);
int cfx_csr_dic_apply(cfx_csr* A, const float* d_r, float* d_z);
/* level l of the hierarchy (host outputs, each may be null): CSR rowPtr
 * [rows+1], col [nnz], FP64 Galerkin values [nnz], and agg [rows] = the
 * level-(l+1) row of each row (not written on the coarsest level) */
int cfx_csr_mg_level(cfx_csr* A, int l, int* rowPtr, int* col, double* val, int* agg);
/* one V-cycle z = M r on device FP32 vectors */
int cfx_csr_mg_apply(cfx_csr* A, const float* d_r, float* d_z);

/* ------------------------------------------- message-passing network */
/* The learned part of the multigrid preconditioner (no counterpart in the
 * reference): a 3-layer message-passing network on the matrix graph (hidden
 * width 32, MFMA with FP16 or FP32 storage and FP32 accumulation) mapping
 * per-row inputs (normalised diagonal/off-diagonal coefficients, face
 * non-orthogonality, |r0|) to tied finest-level weights a (pre = post
 * smoother) and t (restriction = prolongation scale); an all-zero output
 * layer gives a = 2/3, t = sqrt(1.8).  Trained on the GPU with the
 * project's own forward and backward kernels against the A-norm error left
 * by one preconditioned step (optimal step length) on smooth error samples. */
typedef struct cfx_mpnn cfx_mpnn;
int cfx_mpnn_num_params(void);
int cfx_mpnn_create(cfx_ctx* ctx, uint64_t seed, int fp16, cfx_mpnn** out);
int cfx_mpnn_destroy(cfx_mpnn* net);
/* flat FP32 parameters (host): W_in[8][32], b_in[32], 3 x (W[96][32], b[32]),
 * W_out[32][2], b_out[2] */
int cfx_mpnn_get_params(cfx_mpnn* net, float* host);
int cfx_mpnn_set_params(cfx_mpnn* net, const float* host);
/* a, t (device FP32 [n]) for a CSR system; d_geo per slot and d_r0 per row
 * may be null; train = 1 keeps the activations for cfx_mpnn_backward */
int cfx_mpnn_forward_csr(cfx_mpnn* net, cfx_csr* A, const double* d_geo, const double* d_r0, float* d_a,
                         float* d_t, int train);
/* d_grad (device FP32 [num_params]) = dJ/dparams for dJ/da, dJ/dt of the last training forward */
int cfx_mpnn_backward(cfx_mpnn* net, const float* d_ga, const float* d_gt, float* d_grad);
int cfx_mpnn_adam(cfx_mpnn* net, const float* d_grad, float lr);
typedef struct cfx_train_options {
    int steps;        /* Adam steps, systems visited round robin */
    int samples;      /* error samples per step */
    int smooth;       /* cycles applied to a random vector to form a sample */
    float lr;
    uint64_t seed;
} cfx_train_options;
/* trains on count systems (the hierarchy constants of systems[0]); d_geo and
 * d_r0 are arrays of per-system device pointers (or null); lossHist [steps] */
int cfx_mpnn_train(cfx_mpnn* net, cfx_csr* const* systems, const double* const* d_geo,
                   const double* const* d_r0, int count, const cfx_train_options* opt, double* lossHist);
/* mean objective over `samples` error samples, with the network or (0) with the constants */
int cfx_mpnn_evaluate(cfx_mpnn* net, cfx_csr* A, const double* d_geo, const double* d_r0,
                      const cfx_train_options* opt, int samples, int withNetwork, double* loss);
/* the objective and its gradient for one error sample: with z = M A e,
 * L = 1 - (z.A e)^2 / ((e.A e)(z.A z)) = min_alpha |e - alpha z|_A^2 / |e|_A^2,
 * tied per-row weights a, t (linear V-cycle below level 1); ga, gt accumulate */
int cfx_csr_mg_loss_grad(cfx_csr* A, const float* d_e, const float* d_a, const float* d_t, float* d_ga,
                         float* d_gt, double* loss);
/* CFX_PRECOND_LEARNED solves take their finest-level weights from the
 * network (inputs from the matrix, d_geo and the right-hand side); null detaches */
int cfx_csr_set_mpnn(cfx_csr* A, cfx_mpnn* net, const double* d_geo);
int cfx_matrix_set_mpnn(cfx_matrix* A, cfx_mpnn* net);
int cfx_mesh_set_mpnn(cfx_mesh* m, cfx_mpnn* net);

/* ----------------------------------------------------- batched solves */
/* K independent systems (the pressure systems of many detail regions)
 * solved by Jacobi-PCG together, every system with its own scalars and stop
 * test (solveCG's, Matrix.cpp:153-222).  Three-launch form (default): each
 * of the three PCG kernels is one launch over all systems per iteration,
 * replayed as a captured hipGraph, the host polling between batches of
 * iterations.  Persistent form (persistent = 1): the whole batch is one
 * launch, captured as a hipGraph; its workgroups claim chunks of whichever
 * systems have work, and each system retires at its own stop test.  Both
 * forms, with the default schedule {4, 2048, 1, 16, 0, 2, 0} or any other
 * pollEvery, graph, persistent, claim and wgs, give each system the result
 * of cfx_csr_solve_device (FP64 Jacobi-PCG) bit for bit; vecElems and wgCap
 * change the summation grouping.  sched = null picks the schedule with the
 * autotuner, which times both forms.  The reference solves its detail blocks one after
 * another (coupledNetwork.cpp:1461-1475); it has no batched solver. */
typedef struct cfx_batch_schedule {
    int vecElems;   /* elements per thread of the vector kernels */
    int wgCap;      /* SpMV workgroups per system at most (<= 2048) */
    int graph;      /* 1: replay the captured iteration (persistent: the captured solve) */
    int pollEvery;  /* three-launch form: iterations between host polls */
    int persistent; /* 1: one launch for the whole batch; 0: three launches per iteration */
    int claim;      /* persistent: about this many claims per workgroup and step (>= 1); sets the chunks
                       each claim takes */
    int wgs;        /* persistent: workgroups of the launch (0: as many as fit the device at once) */
} cfx_batch_schedule;
typedef struct cfx_batch_stats {
    double ms;              /* device time of the whole batch */
    int maxIterations;
    cfx_batch_schedule used;
} cfx_batch_stats;
/* d_b[k], d_x[k]: device FP64 vectors of system k (x = initial guess and result);
 * reps[k] and finishMs[k] (device time at which system k's stop was seen) may be null */
int cfx_batch_solve(cfx_ctx* ctx, int count, cfx_csr* const* systems, const double* const* d_b,
                    double* const* d_x, double tol, int maxIter, const cfx_batch_schedule* sched,
                    cfx_report* reps, double* finishMs, cfx_batch_stats* stats);
/* times the candidate schedules for `iters` iterations each from the given
 * b and x (no stop test; the x are left as they were)
 * and returns the fastest; cand/candMs (may be null) receive up to maxCand timings */
int cfx_batch_autotune(cfx_ctx* ctx, int count, cfx_csr* const* systems, const double* const* d_b,
                       double* const* d_x, int iters, cfx_batch_schedule* best, cfx_batch_schedule* cand,
                       double* candMs, int maxCand);

/* ------------------------------------------------------------ drivers */
/* Schemes (src/3d/include/cf/fvm/FvMatrix.h:30-38): div 0 upwind, 1 linear,
 * 2 blend; laplacian "corrected" (1) or "none" (0); grad "corrected" (1)
 * or "linear" (0). */
typedef struct cfx_schemes {
    int div;
    int laplacianCorrected;
    int gradCorrected;
    double blendCoeff;
} cfx_schemes;

/* SimpleControls (src/3d/include/cf/fvm/SimpleSolver.h:13-64), steady,
 * laminar; pPrecision selects the pressure solve (CFX_FP64 = the
 * reference's Jacobi-PCG, CFX_FP32_IR = FP32 inner + FP64 refinement). */
typedef struct cfx_simple_controls {
    int maxIterations;
    double continuityTol, pcorrTol, momentumTol, linearTol;
    int linearMaxIter, nonOrthCorrectors;
    double relaxU, relaxP, pRelaxStart;
    double bodyForce[3];
    int pPrecision;
    /* momentum solve: CFX_FP64 (the reference's BiCGStab) or CFX_FP32_IR (FP32 BiCGStab with FP64
     * iterative refinement; stop test on the true FP64 residual); not in the reference */
    int uPrecision;
} cfx_simple_controls;

/* SimpleResult (SimpleSolver.h:66-72) + divergence flag and device time */
typedef struct cfx_simple_result {
    int iterations;
    double finalContinuityError, finalMomentumResidual;
    int converged, diverged;
    double ms;
} cfx_simple_result;

/* PisoControls / PisoStepResult (src/3d/include/cf/fvm/PisoSolver.h:12-49) */
typedef struct cfx_piso_controls {
    double dt;
    int correctors;
    double continuityTol, linearTol;
    int linearMaxIter, nonOrthCorrectors;
    double bodyForce[3];
    int pPrecision;
    /* not in the reference: with no fixedValue pressure face (a closed
     * cavity), project each pass' pressure right-hand side to zero sum, the
     * compatibility condition of the singular Neumann system; without it the
     * non-orthogonal correction of a jittered or tet mesh makes the system
     * inconsistent and the reference's PISO diverges */
    int projectNeumann;
    int uPrecision; /* momentum solve, as in cfx_simple_controls */
} cfx_piso_controls;

typedef struct cfx_piso_result {
    int correctorsUsed;
    double predictedContinuityError, continuityError, pressureCorrectionMax;
    int converged;
    double ms;
} cfx_piso_result;

/* volScalarField/volVectorField::linearise (src/3d/src/fields/volFields.cpp:
 * 117-147, 246-283) for every boundary face: type (host, [nB]: 0 fixedValue,
 * 1 fixedGradient, 2 zeroGradient/outlet, 3 symmetry; for vectors 3 also
 * covers "wall") and value (host, [nB][ncomp]) -> d_linc, d_link. */
int cfx_linearise(cfx_mesh* mesh, int ncomp, const int32_t* type, const double* value, double* d_linc,
                  double* d_link);
/* solveIncompressibleSimple (src/3d/include/cf/fvm/SimpleSolver.h:119-127):
 * d_U [3][n], d_p [n], d_phi [nFaces] in/out on the device; boundary
 * conditions per boundary face (host) as for cfx_linearise; hist (host,
 * optional, [maxIterations][4]) gets {continuity, corrected continuity,
 * max |dp|, dU} per iteration (the reference's log line). */
int cfx_simple_solve(cfx_mesh* mesh, double nu, const cfx_schemes* schemes, const cfx_simple_controls* ctl,
                     const int32_t* Utype, const double* Uvalue, const int32_t* ptype, const double* pvalue,
                     double* d_U, double* d_p, double* d_phi, cfx_simple_result* result, double* hist);
/* solveTransientIncompressiblePiso (src/3d/include/cf/fvm/PisoSolver.h:92-100), one step */
int cfx_piso_step(cfx_mesh* mesh, double nu, const cfx_schemes* schemes, const cfx_piso_controls* ctl,
                  const int32_t* Utype, const double* Uvalue, const int32_t* ptype, const double* pvalue,
                  double* d_U, double* d_p, double* d_phi, cfx_piso_result* result);

/* ----------------------------------------------------------- timing */
/* Average device time of one SpMV launch (y = A x) over `reps` launches,
 * measured with HIP events on the context's stream.  precision CFX_FP64
 * uses the FP64 values, CFX_FP32_IR the FP32 copy (x, y then float).
 * No reference counterpart (measurement only). */
int cfx_csr_time_spmv(cfx_csr* A, const void* d_x, void* d_y, int reps, int precision, double* msPerLaunch);
int cfx_matrix_time_spmv(cfx_matrix* A, int comp, const void* d_x, void* d_y, int reps, int precision,
                         double* msPerLaunch);
/* average device time (ms, HIP events on the context's stream) of the
 * flexible CG's fine-level kernel on this matrix -- q = A p with the FP64
 * values on FP32 vectors plus the p.q and p.r partials, the kernel the
 * multigrid-preconditioned solves (CFX_PRECOND_LEARNED) run once per
 * iteration -- over reps launches on the solver's own vectors.  A
 * measurement helper with no reference counterpart. */
int cfx_matrix_time_fcg_spmv(cfx_matrix* A, int reps, double* ms);
/* measurement mode (no reference counterpart): with enable != 0 the
 * multigrid-preconditioned solves launch their kernels one by one and time
 * every fine-level k_fcg_A launch with HIP events on the context's stream;
 * cfx_matrix_fcg_probe returns the average device time (ms) and the count of
 * the launches that ran in the last such solve on this matrix's size (those
 * queued after the stop decision return at once and are not counted). */
int cfx_set_fcg_probe(cfx_ctx* ctx, int enable);
int cfx_matrix_fcg_probe(cfx_matrix* A, double* avgMs, int* launches);
/* the same probe, on the multigrid's finest-level post-smoothing sweeps of the last flexible-CG solve (set
 * by cfx_set_fcg_probe); and that sweep timed over reps back-to-back launches (on the hierarchy of the last
 * multigrid solve, else on one built on the matrix pattern by this call).  Measurement helpers with no
 * counterpart in the reference. */
int cfx_matrix_smooth_probe(cfx_matrix* A, double* avgMs, int* launches);
int cfx_matrix_time_smooth(cfx_matrix* A, int reps, double* ms);

#ifdef __cplusplus
}
#endif
#endif
