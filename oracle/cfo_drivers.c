/*
 * cfo_drivers.c -- oracle: fvMatrix::solve pass loop, steady SIMPLE and one
 * transient PISO step.  TEST INFRASTRUCTURE ONLY (see cfo.h).
 */
#include "cfo.h"
#include "cfo_vec.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SMAX(a, b) ((a) < (b) ? (b) : (a))
#define SMIN(a, b) ((b) < (a) ? (b) : (a))

static double* dalloc(int64_t n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

void cfo_ldu_reset(cfo_ldu* A)
{
    for (int64_t i = 0; i < (int64_t)A->ncomp * A->n; ++i) { A->diag[i] = 0.0; A->rhs[i] = 0.0; }
    /* -0.0 is the additive identity: the first contribution is stored
     * unchanged, as the reference stores the first triplet. */
    for (int64_t f = 0; f < A->F; ++f) { A->upper[f] = -0.0; A->lower[f] = -0.0; }
}

/* The laplacian corrector registered by assembleLaplacian
 * (laplacianScheme.cpp:111-154). */
typedef struct corrector {
    int active;
    const double* gammaFace;
    const double* lin_c;
    const double* lin_k;
    int sign;
    int gradCorrected;
} corrector;

/* fvMatrix::solve (fvMatrix.cpp:73-129).  x [ncomp][n] in/out; rhsOut
 * (optional, [ncomp][n]) receives rhs_ as it stands after the solve (the
 * base rhs plus the final pass' corrections), which explicitOperator reads. */
static cfo_report fvmatrix_solve_p(const cfo_mesh* m, const cfo_ldu* A, double* x,
                                   const corrector* corr, double tol, int maxIter,
                                   int nonOrth, int preferCG, double* rhsOut,
                                   cfo_csr* csrOut, int projectMean);

static cfo_report fvmatrix_solve(const cfo_mesh* m, const cfo_ldu* A, double* x,
                                 const corrector* corr, double tol, int maxIter,
                                 int nonOrth, int preferCG, double* rhsOut,
                                 cfo_csr* csrOut /* [ncomp], may be NULL */)
{
    return fvmatrix_solve_p(m, A, x, corr, tol, maxIter, nonOrth, preferCG, rhsOut, csrOut, 0);
}

/* projectMean: subtract the mean of each pass' right-hand side (singular
 * Neumann systems) */
static cfo_report fvmatrix_solve_p(const cfo_mesh* m, const cfo_ldu* A, double* x,
                                   const corrector* corr, double tol, int maxIter,
                                   int nonOrth, int preferCG, double* rhsOut,
                                   cfo_csr* csrOut, int projectMean)
{
    const int64_t n = m->nCells;
    const int nc = A->ncomp;
    const int64_t nnzMax = n + 2 * m->nInternal;
    cfo_csr* mats = csrOut ? csrOut : (cfo_csr*)calloc((size_t)nc, sizeof(cfo_csr));
    for (int c = 0; c < nc; ++c) {
        if (!csrOut) {
            mats[c].rowStart = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
            mats[c].cols = (int64_t*)malloc(sizeof(int64_t) * (size_t)nnzMax);
            mats[c].vals = (double*)malloc(sizeof(double) * (size_t)nnzMax);
        }
        cfo_csr_from_ldu(m, A, c, &mats[c]);
    }
    double* rhs = dalloc((int64_t)nc * n);
    const int passes = (corr && corr->active) ? SMAX(1, nonOrth + 1) : 1;
    cfo_report last = {0, 0, 0, 0};
    for (int pass = 0; pass < passes; ++pass) {
        memcpy(rhs, A->rhs, sizeof(double) * (size_t)(nc * n));
        if (corr && corr->active)
            for (int c = 0; c < nc; ++c)
                cfo_nonorth_correct(m, corr->gammaFace, x + (int64_t)c * n, corr->lin_c,
                                    corr->lin_k, nc, c, corr->gradCorrected, corr->sign,
                                    rhs + (int64_t)c * n);
        if (projectMean)
            for (int c = 0; c < nc; ++c) {
                double sum = 0;
                for (int64_t i = 0; i < n; ++i) sum += rhs[(int64_t)c * n + i];
                const double mean = sum / (double)n;
                for (int64_t i = 0; i < n; ++i) rhs[(int64_t)c * n + i] -= mean;
            }
        for (int c = 0; c < nc; ++c) {
            if (preferCG) last = cfo_solve_cg(&mats[c], rhs + (int64_t)c * n, x + (int64_t)c * n, tol, maxIter, NULL);
            else last = cfo_solve_bicgstab(&mats[c], rhs + (int64_t)c * n, x + (int64_t)c * n, tol, maxIter);
        }
    }
    if (rhsOut) memcpy(rhsOut, rhs, sizeof(double) * (size_t)(nc * n));
    free(rhs);
    if (!csrOut) {
        for (int c = 0; c < nc; ++c) { free(mats[c].rowStart); free(mats[c].cols); free(mats[c].vals); }
        free(mats);
    }
    return last;
}

typedef struct lin_set { double* c; double* k; } lin_set;

static lin_set make_lin(const cfo_mesh* m, int ncomp, const int32_t* type, const double* value, int* rc)
{
    const int64_t nB = m->nFaces - m->nInternal;
    lin_set l = {dalloc(nB * ncomp), dalloc(nB * ncomp)};
    const int r = cfo_linearise(m, ncomp, type, value, l.c, l.k);
    if (r && rc) *rc = r;
    return l;
}

/* HbyA's boundary conditions (simpleSolver.cpp:175-187, pisoSolver.cpp:142-150):
 * fixedValue (with U's face values) where U is fixedValue, else zeroGradient. */
static lin_set hbya_lin(const cfo_mesh* m, cfo_bc Ubc)
{
    const int64_t nB = m->nFaces - m->nInternal;
    int32_t* t = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nB > 0 ? nB : 1));
    for (int64_t b = 0; b < nB; ++b)
        t[b] = Ubc.type[b] == CFO_BC_FIXED_VALUE ? CFO_BC_FIXED_VALUE : CFO_BC_ZERO_GRADIENT;
    lin_set l = make_lin(m, 3, t, Ubc.value, NULL);
    free(t);
    return l;
}

static int fields_finite(const double* U, const double* p, int64_t n)
{
    for (int64_t i = 0; i < 3 * n; ++i) if (!isfinite(U[i])) return 0;
    for (int64_t i = 0; i < n; ++i) if (!isfinite(p[i])) return 0;
    return 1;
}

static void ldu_alloc(cfo_ldu* A, int64_t n, int64_t F, int nc)
{
    A->n = n; A->F = F; A->ncomp = nc;
    A->diag = dalloc((int64_t)nc * n); A->rhs = dalloc((int64_t)nc * n);
    A->upper = dalloc(F); A->lower = dalloc(F);
    cfo_ldu_reset(A);
}
static void ldu_free(cfo_ldu* A) { free(A->diag); free(A->rhs); free(A->upper); free(A->lower); }

/* solveIncompressibleSimple, simpleSolver.cpp:33-394 (laminar, no energy,
 * steady). */
cfo_simple_result cfo_simple(const cfo_mesh* m, double* U, double* p,
                             double* phi, double nu, const cfo_schemes* s,
                             const cfo_simple_controls* ctl,
                             cfo_bc Ubc, cfo_bc pbc, double* hist)
{
    const int64_t n = m->nCells, nI = m->nInternal, nF = m->nFaces;
    cfo_simple_result result = {0, 0, 0, 0, 0};
    int rc = 0;
    lin_set Ul = make_lin(m, 3, Ubc.type, Ubc.value, &rc);
    lin_set pl = make_lin(m, 1, pbc.type, pbc.value, &rc);
    lin_set Hl = hbya_lin(m, Ubc);
    if (rc) { result.diverged = -1; goto out0; }

    double contLag1 = INFINITY, contLag2 = INFINITY;
    double* gradPn = dalloc(3 * n);
    double* gradPFull = dalloc(3 * n);
    double* x = dalloc(3 * n);
    double* xc = dalloc(3 * n);   /* [comp][n] solve vector */
    double* UOld = dalloc(3 * n);
    double* Hfree = dalloc(3 * n);
    double* rAUr = dalloc(n);
    double* rAUf = dalloc(nF);
    double* divPhi = dalloc(n);
    double* pNew = dalloc(n);
    double* gammaU = dalloc(nF);
    cfo_harmonic_gamma_uniform(m, nu, gammaU);
    int needReference = 1;
    for (int64_t b = 0; b < nF - nI; ++b) if (pbc.type[b] == CFO_BC_FIXED_VALUE) needReference = 0;

    for (int iter = 1; iter <= ctl->maxIterations; ++iter) {
        cfo_grad(m, p, pl.c, pl.k, 1, 0, s->gradCorrected, 1, gradPn);

        /* momentum */
        cfo_ldu UE;
        ldu_alloc(&UE, n, nI, 3);
        cfo_assemble_div(m, phi, s->divScheme, s->blendCoeff, Ul.c, Ul.k, 1, &UE);
        cfo_assemble_laplacian(m, gammaU, Ul.c, Ul.k, -1, &UE);
        for (int64_t c = 0; c < n; ++c) {
            const V3 force = v3(ctl->bodyForce[0], ctl->bodyForce[1], ctl->bodyForce[2]);
            const V3 rhs = vmul(vsub(vadd(force, vzero()), vload(gradPn + 3 * c)), m->V[c]);
            for (int k = 0; k < 3; ++k) UE.rhs[k * n + c] += vget(rhs, k);
        }
        corrector uc = {s->laplacianCorrected, gammaU, Ul.c, Ul.k, -1, s->gradCorrected};
        memset(xc, 0, sizeof(double) * 3 * (size_t)n);
        fvmatrix_solve(m, &UE, xc, &uc, ctl->linearTol, ctl->linearMaxIter,
                       ctl->nonOrthCorrectors, 0, NULL, NULL);
        memcpy(UOld, U, sizeof(double) * 3 * (size_t)n);

        for (int k = 1; k < 3; ++k)
            for (int64_t c = 0; c < n; ++c) {
                const double d0 = UE.diag[c], dk = UE.diag[k * n + c];
                if (fabs(d0 - dk) > 1e-10 * (fabs(d0) + fabs(dk) + 1.0)) { result.diverged = -2; ldu_free(&UE); goto out; }
            }
        for (int64_t c = 0; c < n; ++c) rAUr[c] = m->V[c] / SMAX(UE.diag[c], 1e-300);
        ldu_free(&UE);

        /* HbyA */
        for (int k = 0; k < 3; ++k)
            for (int64_t c = 0; c < n; ++c)
                Hfree[3 * c + k] = (1.0 - ctl->relaxU) * UOld[3 * c + k] + ctl->relaxU * xc[k * n + c];

        /* predictor flux */
        cfo_update_phi_internal(m, Hfree, phi);
        cfo_correct_boundary_flux(m, Hfree, Hl.c, Hl.k, phi);

        if (needReference) {
            /* adjustPhi, simpleSolver.cpp:211-247 */
            double netSum = 0, freeSum = 0;
            for (int64_t pi = 0; pi < m->nPatches; ++pi) {
                const int64_t st = m->patchStart[pi], sz = m->patchSize[pi];
                if (sz == 0) continue;
                const int fixedU = Ubc.type[st - nI] == CFO_BC_FIXED_VALUE;
                double sum = 0;
                for (int64_t f = st; f < st + sz; ++f) sum += phi[f];
                netSum += sum;
                if (!fixedU) freeSum += sum;
            }
            if (fabs(freeSum) > 1e-30) {
                const double scale = (freeSum - netSum) / freeSum;
                for (int64_t pi = 0; pi < m->nPatches; ++pi) {
                    const int64_t st = m->patchStart[pi], sz = m->patchSize[pi];
                    if (sz == 0 || Ubc.type[st - nI] == CFO_BC_FIXED_VALUE) continue;
                    for (int64_t f = st; f < st + sz; ++f) phi[f] *= scale;
                }
            }
        }

        const double contErr = cfo_continuity_error(m, phi);
        cfo_div_phi(m, phi, divPhi);

        /* pressure */
        cfo_interpolate_to_faces(m, rAUr, rAUf);
        cfo_ldu PE;
        ldu_alloc(&PE, n, nI, 1);
        cfo_assemble_laplacian(m, rAUf, pl.c, pl.k, -1, &PE);
        for (int64_t c = 0; c < n; ++c) PE.rhs[c] += -divPhi[c];
        corrector pc = {s->laplacianCorrected, rAUf, pl.c, pl.k, -1, s->gradCorrected};
        memset(pNew, 0, sizeof(double) * (size_t)n);
        fvmatrix_solve(m, &PE, pNew, &pc, ctl->linearTol, ctl->linearMaxIter,
                       ctl->nonOrthCorrectors, 1, NULL, NULL);
        ldu_free(&PE);

        /* correct */
        cfo_grad(m, pNew, pl.c, pl.k, 1, 0, s->gradCorrected, 1, gradPFull);
        cfo_correct_fluxes(m, pNew, gradPFull, rAUf, pl.c, pl.k, s->laplacianCorrected, phi);

        double pStep = 0, pScale = 0;
        for (int64_t c = 0; c < n; ++c) {
            pStep = SMAX(pStep, fabs(pNew[c] - p[c]));
            pScale = SMAX(pScale, fabs(p[c]));
        }
        const double aP = ctl->pRelaxStart > 0
            ? SMIN(ctl->relaxP, SMAX(ctl->pRelaxStart, 0.15 * pScale / SMAX(pStep, 1e-300)))
            : ctl->relaxP;
        double pDeltaMax = 0;
        for (int64_t c = 0; c < n; ++c) {
            const double dp = aP * (pNew[c] - p[c]);
            pDeltaMax = SMAX(pDeltaMax, fabs(dp));
            p[c] += dp;
            const V3 dg = vsub(vload(gradPFull + 3 * c), vload(gradPn + 3 * c));
            vstore(U + 3 * c, vsub(vload(Hfree + 3 * c), vmul(dg, aP * rAUr[c])));
        }
        const double errAfter = cfo_continuity_error(m, phi);
        double uMax = 0, dUMax = 0;
        for (int64_t c = 0; c < n; ++c) {
            uMax = SMAX(uMax, vmag(vload(U + 3 * c)));
            dUMax = SMAX(dUMax, vmag(vsub(vload(U + 3 * c), vload(UOld + 3 * c))));
        }
        const double uRes = dUMax / SMAX(uMax, 1e-300);
        if (hist) {
            hist[4 * (iter - 1) + 0] = contErr;
            hist[4 * (iter - 1) + 1] = errAfter;
            hist[4 * (iter - 1) + 2] = pDeltaMax;
            hist[4 * (iter - 1) + 3] = uRes;
        }
        result.iterations = iter;
        result.finalContinuityError = contErr;
        result.finalMomentumResidual = uRes;
        if (!isfinite(contErr) || !isfinite(pDeltaMax) || !fields_finite(U, p, n)) {
            result.diverged = 1;
            break;
        }
        const int settled = contErr <= contLag2 * 1.01 + 1e-12;
        const int momentumSettled = !(ctl->momentumTol > 0) || uRes <= ctl->momentumTol;
        if (momentumSettled &&
            (contErr < ctl->continuityTol ||
             (iter >= 5 && ctl->pcorrTol > 0 && pDeltaMax < ctl->pcorrTol && settled))) {
            result.converged = 1;
            break;
        }
        contLag2 = contLag1;
        contLag1 = contErr;
    }
out:
    free(gradPn); free(gradPFull); free(x); free(xc); free(UOld); free(Hfree);
    free(rAUr); free(rAUf); free(divPhi); free(pNew); free(gammaU);
out0:
    free(Ul.c); free(Ul.k); free(pl.c); free(pl.k); free(Hl.c); free(Hl.k);
    return result;
}

/* solveTransientIncompressiblePiso, pisoSolver.cpp:13-235 (laminar). */
cfo_piso_result cfo_piso_step(const cfo_mesh* m, double* U, double* p,
                              double* phi, double nu, const cfo_schemes* s,
                              const cfo_piso_controls* ctl, cfo_bc Ubc,
                              cfo_bc pbc)
{
    const int64_t n = m->nCells, nI = m->nInternal, nF = m->nFaces;
    cfo_piso_result res = {0, 0, 0, 0, 0};
    if (ctl->dt <= 0) { res.correctorsUsed = -1; return res; }
    int rc = 0;
    lin_set Ul = make_lin(m, 3, Ubc.type, Ubc.value, &rc);
    lin_set pl = make_lin(m, 1, pbc.type, pbc.value, &rc);
    lin_set Hl = hbya_lin(m, Ubc);
    int needRefP = 1;  /* no fixedValue pressure face: a singular Neumann system */
    for (int64_t b = 0; b < nF - nI; ++b)
        if (pbc.type[b] == 0) needRefP = 0;

    double* gradP = dalloc(3 * n);
    double* gradPn = dalloc(3 * n);
    double* UOld = dalloc(3 * n);
    double* xc = dalloc(3 * n);
    double* rhsFinal = dalloc(3 * n);
    double* rAU = dalloc(n);
    double* rAUf = dalloc(nF);
    double* HbyA = dalloc(3 * n);
    double* divPhi = dalloc(n);
    double* pBefore = dalloc(n);
    double* h = dalloc(n);
    double* ax = dalloc(n);
    double* gammaU = dalloc(nF);
    cfo_csr mats[3];
    const int64_t nnzMax = n + 2 * nI;
    for (int c = 0; c < 3; ++c) {
        mats[c].rowStart = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
        mats[c].cols = (int64_t*)malloc(sizeof(int64_t) * (size_t)nnzMax);
        mats[c].vals = (double*)malloc(sizeof(double) * (size_t)nnzMax);
    }
    cfo_harmonic_gamma_uniform(m, nu, gammaU);
    if (rc) { res.correctorsUsed = -2; goto out; }

    cfo_grad(m, p, pl.c, pl.k, 1, 0, s->gradCorrected, 1, gradP);
    memcpy(gradPn, gradP, sizeof(double) * 3 * (size_t)n);
    memcpy(UOld, U, sizeof(double) * 3 * (size_t)n);

    cfo_ldu UE;
    ldu_alloc(&UE, n, nI, 3);
    cfo_assemble_ddt(m, UOld, ctl->dt, 1, &UE);
    cfo_assemble_div(m, phi, s->divScheme, s->blendCoeff, Ul.c, Ul.k, 1, &UE);
    cfo_assemble_laplacian(m, gammaU, Ul.c, Ul.k, -1, &UE);
    for (int64_t c = 0; c < n; ++c) {
        const V3 force = vadd(vadd(v3(ctl->bodyForce[0], ctl->bodyForce[1], ctl->bodyForce[2]), vzero()), vzero());
        const V3 rhs = vmul(vsub(force, vload(gradP + 3 * c)), m->V[c]);
        for (int k = 0; k < 3; ++k) UE.rhs[k * n + c] += vget(rhs, k);
    }
    corrector uc = {s->laplacianCorrected, gammaU, Ul.c, Ul.k, -1, s->gradCorrected};
    fvmatrix_solve(m, &UE, xc, &uc, ctl->linearTol, ctl->linearMaxIter,
                   ctl->nonOrthCorrectors, 0, rhsFinal, mats);
    for (int k = 0; k < 3; ++k)
        for (int64_t c = 0; c < n; ++c) U[3 * c + k] = xc[k * n + c];

    for (int k = 1; k < 3; ++k)
        for (int64_t c = 0; c < n; ++c) {
            const double d0 = UE.diag[c], dk = UE.diag[k * n + c];
            if (fabs(d0 - dk) > 1e-10 * (fabs(d0) + fabs(dk) + 1.0)) { res.correctorsUsed = -3; goto out2; }
        }
    for (int64_t c = 0; c < n; ++c) rAU[c] = m->V[c] / SMAX(UE.diag[c], 1e-300);

    /* rebuildHbyA (pisoSolver.cpp:121-151) via explicitOperator (fvMatrix.cpp:131-146) */
#define REBUILD_HBYA() do { \
        for (int k = 0; k < 3; ++k) { \
            for (int64_t c = 0; c < n; ++c) xc[c] = U[3 * c + k]; \
            memcpy(h, rhsFinal + (int64_t)k * n, sizeof(double) * (size_t)n); \
            cfo_spmv(&mats[k], xc, ax); \
            for (int64_t c = 0; c < n; ++c) h[c] -= ax[c] - UE.diag[k * n + c] * xc[c]; \
            for (int64_t c = 0; c < n; ++c) \
                HbyA[3 * c + k] = rAU[c] * (h[c] + m->V[c] * gradPn[3 * c + k]) / m->V[c]; \
        } } while (0)
    REBUILD_HBYA();
    cfo_interpolate_to_faces(m, rAU, rAUf);

    for (int corr = 1; corr <= ctl->correctors; ++corr) {
        if (corr > 1) REBUILD_HBYA();
        cfo_update_phi_internal(m, HbyA, phi);
        cfo_correct_boundary_flux(m, HbyA, Hl.c, Hl.k, phi);
        res.predictedContinuityError = cfo_continuity_error(m, phi);
        cfo_div_phi(m, phi, divPhi);

        cfo_ldu PE;
        ldu_alloc(&PE, n, nI, 1);
        cfo_assemble_laplacian(m, rAUf, pl.c, pl.k, -1, &PE);
        for (int64_t c = 0; c < n; ++c) PE.rhs[c] += -divPhi[c];
        memcpy(pBefore, p, sizeof(double) * (size_t)n);
        corrector pc = {s->laplacianCorrected, rAUf, pl.c, pl.k, -1, s->gradCorrected};
        fvmatrix_solve_p(m, &PE, p, &pc, ctl->linearTol, ctl->linearMaxIter,
                         ctl->nonOrthCorrectors, 1, NULL, NULL, ctl->projectNeumann && needRefP);
        ldu_free(&PE);
        double pDeltaMax = 0;
        for (int64_t c = 0; c < n; ++c) pDeltaMax = SMAX(pDeltaMax, fabs(p[c] - pBefore[c]));
        res.pressureCorrectionMax = pDeltaMax;

        cfo_grad(m, p, pl.c, pl.k, 1, 0, s->gradCorrected, 1, gradP);
        for (int64_t c = 0; c < n; ++c)
            vstore(U + 3 * c, vsub(vload(HbyA + 3 * c), vmul(vload(gradP + 3 * c), rAU[c])));
        cfo_correct_fluxes(m, p, gradP, rAUf, pl.c, pl.k, s->laplacianCorrected, phi);
        res.correctorsUsed = corr;
        res.continuityError = cfo_continuity_error(m, phi);
        if (!isfinite(res.continuityError) || !isfinite(pDeltaMax)) break;
        if (res.continuityError < ctl->continuityTol) { res.converged = 1; break; }
    }
#undef REBUILD_HBYA
out2:
    ldu_free(&UE);
out:
    for (int c = 0; c < 3; ++c) { free(mats[c].rowStart); free(mats[c].cols); free(mats[c].vals); }
    free(gradP); free(gradPn); free(UOld); free(xc); free(rhsFinal); free(rAU); free(rAUf);
    free(HbyA); free(divPhi); free(pBefore); free(h); free(ax); free(gammaU);
    free(Ul.c); free(Ul.k); free(pl.c); free(pl.k); free(Hl.c); free(Hl.k);
    return res;
}
