/*
 * cfo_fv.c -- oracle: boundary linearisation, interpolation, gradients,
 * implicit assembly, CSR, Krylov solvers, flux operators.
 * TEST INFRASTRUCTURE ONLY (see cfo.h).
 */
#include "cfo.h"
#include "cfo_vec.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NB(m) ((m)->nFaces - (m)->nInternal)
/* std::max / std::min semantics (NaN in the second argument is dropped) */
#define SMAX(a, b) ((a) < (b) ? (b) : (a))
#define SMIN(a, b) ((b) < (a) ? (b) : (a))

/* ---------------------------------------------------- linearisation */
/* dominantAxis, volFields.cpp:23-38 */
static int dominant_axis(V3 n)
{
    const double ax = fabs(n.x), ay = fabs(n.y), az = fabs(n.z);
    if (ax >= ay && ax >= az) return ax < 1.0 - 1e-6 ? -1 : 0;
    if (ay >= az) return ay < 1.0 - 1e-6 ? -1 : 1;
    return az < 1.0 - 1e-6 ? -1 : 2;
}

/* volScalarField::linearise (volFields.cpp:117-147) and
 * volVectorField::linearise (:246-283). */
int cfo_linearise(const cfo_mesh* m, int ncomp, const int32_t* type,
                  const double* value, double* lin_c, double* lin_k)
{
    const int64_t nI = m->nInternal, nB = NB(m);
    for (int64_t b = 0; b < nB; ++b) {
        const int64_t f = nI + b;
        const V3 sf = vload(m->Sf + 3 * f);
        for (int c = 0; c < ncomp; ++c) {
            const double v = value[b * ncomp + c];
            double cc = 0, kk = 0;
            switch (type[b]) {
            case CFO_BC_FIXED_VALUE: cc = 0; kk = v; break;
            case CFO_BC_FIXED_GRADIENT: {
                /* normalDistance, volFields.cpp:14-21 */
                const double a = vmag(sf);
                const V3 nrm = vdiv(sf, a);
                cc = 1; kk = v * vdot(vload(m->dvec + 3 * f), nrm);
                break;
            }
            case CFO_BC_ZERO_GRADIENT: cc = 1; kk = 0; break;
            case CFO_BC_SYMMETRY:
                if (ncomp == 1) { cc = 1; kk = 0; }
                else {
                    const int k = dominant_axis(vdiv(sf, vmag(sf)));
                    if (k < 0) return -1;
                    cc = (c == k) ? 0.0 : 1.0; kk = 0;
                }
                break;
            default: return -2;
            }
            lin_c[b * ncomp + c] = cc;
            lin_k[b * ncomp + c] = kk;
        }
    }
    return 0;
}

/* ------------------------------------------------------ interpolation */
void cfo_owner_weights(const cfo_mesh* m, double* w)
{
    for (int64_t f = 0; f < m->nInternal; ++f) {
        const V3 mf = vload(m->Cf + 3 * f);
        const V3 cp = vload(m->C + 3 * m->owner[f]);
        const V3 d = vload(m->dvec + 3 * f);
        const double d2 = vdot(d, d);
        double nw = vdot(vsub(mf, cp), d) / d2;
        nw = nw < 0.0 ? 0.0 : (nw > 1.0 ? 1.0 : nw);
        w[f] = 1.0 - nw;
    }
}

/* makeFaceSkewVectors, interpolate.cpp:37-48 */
static void skew_vectors(const cfo_mesh* m, const double* ow, V3* skew)
{
    for (int64_t f = 0; f < m->nInternal; ++f) {
        const V3 co = vload(m->C + 3 * m->owner[f]);
        const V3 cn = vload(m->C + 3 * m->neighbour[f]);
        const double nw = 1.0 - ow[f];
        skew[f] = vsub(vload(m->Cf + 3 * f), vadd(co, vmul(vsub(cn, co), nw)));
    }
}

/* volScalarField::faceValue on an internal face, volFields.cpp:101-111: the
 * clamped NEIGHBOUR fraction multiplies the owner value there. */
static double quirk_internal_face_value(const cfo_mesh* m, int64_t f, const double* data)
{
    const int64_t o = m->owner[f];
    const V3 d = vload(m->dvec + 3 * f);
    const double d2 = vdot(d, d);
    double w = 0.5;
    if (d2 > 0) {
        w = vdot(vsub(vload(m->Cf + 3 * f), vload(m->C + 3 * o)), d) / d2;
        w = w > 0.0 ? w : 0.0;
        w = w < 1.0 ? w : 1.0;
    }
    return w * data[o] + (1.0 - w) * data[m->neighbour[f]];
}

/* gaussGrad, gradScheme.cpp:15-96 */
void cfo_grad(const cfo_mesh* m, const double* data, const double* lin_c,
              const double* lin_k, int ncomp, int comp, int corrected,
              int quirkLinear, double* gradOut)
{
    const int64_t nC = m->nCells, nF = m->nFaces, nI = m->nInternal;
    double* ow = (double*)malloc(sizeof(double) * (size_t)(nI > 0 ? nI : 1));
    cfo_owner_weights(m, ow);
    V3* sum = (V3*)calloc((size_t)nC, sizeof(V3));
    V3* skew = corrected ? (V3*)malloc(sizeof(V3) * (size_t)(nI > 0 ? nI : 1)) : NULL;
    if (corrected) skew_vectors(m, ow, skew);

#define RECON(f) ((f) < nI \
        ? ((corrected || !quirkLinear) \
              ? ow[f] * data[m->owner[f]] + (1.0 - ow[f]) * data[m->neighbour[f]] \
              : quirk_internal_face_value(m, (f), data)) \
        : lin_c[((f) - nI) * ncomp + comp] * data[m->owner[f]] + lin_k[((f) - nI) * ncomp + comp])

    for (int64_t f = 0; f < nF; ++f) {
        const V3 flux = vmul(vload(m->Sf + 3 * f), RECON(f));
        sum[m->owner[f]] = vadd(sum[m->owner[f]], flux);
        if (f < nI) sum[m->neighbour[f]] = vsub(sum[m->neighbour[f]], flux);
    }
    for (int64_t c = 0; c < nC; ++c) vstore(gradOut + 3 * c, vdiv(sum[c], m->V[c]));

    if (corrected) {
        V3* gradFace = (V3*)malloc(sizeof(V3) * (size_t)nF);
        for (int pass = 0; pass < 2; ++pass) {
            for (int64_t f = 0; f < nF; ++f) {
                if (f < nI) {
                    const V3 gp = vload(gradOut + 3 * m->owner[f]);
                    const V3 gn = vload(gradOut + 3 * m->neighbour[f]);
                    gradFace[f] = vadd(vmul(gp, ow[f]), vmul(gn, 1.0 - ow[f]));
                } else {
                    gradFace[f] = vload(gradOut + 3 * m->owner[f]);
                }
            }
            memset(sum, 0, sizeof(V3) * (size_t)nC);
            for (int64_t f = 0; f < nF; ++f) {
                double tf = RECON(f);
                if (f < nI) tf += vdot(gradFace[f], skew[f]);
                const V3 flux = vmul(vload(m->Sf + 3 * f), tf);
                sum[m->owner[f]] = vadd(sum[m->owner[f]], flux);
                if (f < nI) sum[m->neighbour[f]] = vsub(sum[m->neighbour[f]], flux);
            }
            for (int64_t c = 0; c < nC; ++c) vstore(gradOut + 3 * c, vdiv(sum[c], m->V[c]));
        }
        free(gradFace);
    }
#undef RECON
    free(ow); free(sum); free(skew);
}

/* harmonicFaceGamma with a one-entry (uniform) gammaCell,
 * laplacianScheme.cpp:37-62 */
void cfo_harmonic_gamma_uniform(const cfo_mesh* m, double gamma, double* gf)
{
    for (int64_t f = 0; f < m->nFaces; ++f) {
        if (f < m->nInternal) {
            const double gp = gamma, gn = gamma;
            const V3 cP = vload(m->C + 3 * m->owner[f]);
            const V3 cN = vload(m->C + 3 * m->neighbour[f]);
            const V3 cf = vload(m->Cf + 3 * f);
            const double dP = vmag(vsub(cf, cP)), dN = vmag(vsub(cN, cf));
            const double tot = dP + dN;
            gf[f] = (tot > 0 && gp > 0 && gn > 0) ? tot * gp * gn / (dN * gp + dP * gn)
                                                  : 0.5 * (gp + gn);
        } else {
            gf[f] = gamma;
        }
    }
}

void cfo_interpolate_to_faces(const cfo_mesh* m, const double* cell, double* face)
{
    const int64_t nI = m->nInternal;
    double* ow = (double*)malloc(sizeof(double) * (size_t)(nI > 0 ? nI : 1));
    cfo_owner_weights(m, ow);
    for (int64_t f = 0; f < m->nFaces; ++f) {
        if (f < nI)
            face[f] = ow[f] * cell[m->owner[f]] + (1.0 - ow[f]) * cell[m->neighbour[f]];
        else
            face[f] = cell[m->owner[f]];
    }
    free(ow);
}

/* ------------------------------------------------------------ assembly */
/* assembleLaplacian static part, laplacianScheme.cpp:67-104 */
void cfo_assemble_laplacian(const cfo_mesh* m, const double* gammaFace,
                            const double* lin_c, const double* lin_k,
                            int sign, cfo_ldu* A)
{
    const int64_t nI = m->nInternal, n = m->nCells;
    const int nc = A->ncomp;
    for (int64_t f = 0; f < m->nFaces; ++f) {
        const V3 d = vload(m->dvec + 3 * f);
        const double oc = vdot(vload(m->Sf + 3 * f), d) / vdot(d, d);
        const double coef = gammaFace[f] * oc;
        const int64_t o = m->owner[f];
        if (f < nI) {
            const int64_t nb = m->neighbour[f];
            A->upper[f] += (double)sign * coef;
            A->lower[f] += (double)sign * coef;
            for (int c = 0; c < nc; ++c) {
                A->diag[c * n + o] += (double)(-sign) * coef;
                A->diag[c * n + nb] += (double)(-sign) * coef;
            }
        } else {
            const int64_t b = f - nI;
            for (int c = 0; c < nc; ++c) {
                const double cc = lin_c[b * nc + c], kk = lin_k[b * nc + c];
                A->diag[c * n + o] += (double)sign * coef * (cc - 1.0);
                A->rhs[c * n + o] += (double)(-sign) * coef * kk;
            }
        }
    }
}

/* fvm::div, divScheme.cpp:11-62 */
void cfo_assemble_div(const cfo_mesh* m, const double* phi, int scheme,
                      double blendCoeff, const double* lin_c,
                      const double* lin_k, int sign, cfo_ldu* A)
{
    const int64_t nI = m->nInternal, n = m->nCells;
    const int nc = A->ncomp;
    double* ow = (double*)malloc(sizeof(double) * (size_t)(nI > 0 ? nI : 1));
    cfo_owner_weights(m, ow);
    const double alpha = scheme == 2 ? blendCoeff : (scheme == 1 ? 1.0 : 0.0);
    for (int64_t f = 0; f < m->nFaces; ++f) {
        const int64_t o = m->owner[f];
        const double p = phi[f];
        if (f < nI) {
            const int64_t nb = m->neighbour[f];
            double wPo = 0.0, wNo = 0.0;
            if (p >= 0) wPo = 1.0; else wNo = 1.0;
            if (alpha > 0) {
                const double c = ow[f];
                wPo = alpha * c + (1.0 - alpha) * wPo;
                wNo = alpha * (1.0 - c) + (1.0 - alpha) * wNo;
            }
            const double fp = (double)sign * p * wPo;
            const double fn = (double)sign * p * wNo;
            A->upper[f] += fn;
            A->lower[f] += -fp;
            for (int c = 0; c < nc; ++c) {
                A->diag[c * n + o] += fp;
                A->diag[c * n + nb] += -fn;
            }
        } else {
            const int64_t b = f - nI;
            for (int c = 0; c < nc; ++c) {
                const double cc = lin_c[b * nc + c], kk = lin_k[b * nc + c];
                A->diag[c * n + o] += (double)sign * p * cc;
                A->rhs[c * n + o] += (double)(-sign) * p * kk;
            }
        }
    }
    free(ow);
}

void cfo_assemble_ddt(const cfo_mesh* m, const double* oldField, double dt,
                      int sign, cfo_ldu* A)
{
    const int64_t n = m->nCells;
    const int nc = A->ncomp;
    for (int64_t c = 0; c < n; ++c) {
        const double a = (double)sign * m->V[c] / dt;
        for (int k = 0; k < nc; ++k) {
            A->diag[k * n + c] += a;
            A->rhs[k * n + c] += a * oldField[c * nc + k];
        }
    }
}

/* non-orthogonal corrector, laplacianScheme.cpp:106-154 */
void cfo_nonorth_correct(const cfo_mesh* m, const double* gammaFace,
                         const double* x, const double* lin_c,
                         const double* lin_k, int ncomp, int comp,
                         int gradCorrected, int sign, double* rhs)
{
    const int64_t nC = m->nCells, nI = m->nInternal;
    double* g = (double*)malloc(sizeof(double) * 3 * (size_t)nC);
    cfo_grad(m, x, lin_c, lin_k, ncomp, comp, gradCorrected, 0, g);
    double* ow = (double*)malloc(sizeof(double) * (size_t)(nI > 0 ? nI : 1));
    cfo_owner_weights(m, ow);
    for (int64_t f = 0; f < m->nFaces; ++f) {
        V3 gradFace;
        if (f < nI) {
            const V3 gp = vload(g + 3 * m->owner[f]);
            const V3 gn = vload(g + 3 * m->neighbour[f]);
            gradFace = vadd(vmul(gp, ow[f]), vmul(gn, 1.0 - ow[f]));
        } else {
            gradFace = vload(g + 3 * m->owner[f]);
        }
        const V3 d = vload(m->dvec + 3 * f);
        const double dm = vmag(d);
        if (dm <= 0) continue;
        const V3 dh = vdiv(d, dm);
        const V3 sfv = vload(m->Sf + 3 * f);
        const V3 cfVec = vsub(sfv, vmul(dh, vdot(sfv, dh)));
        const double corr = gammaFace[f] * vdot(cfVec, gradFace);
        rhs[m->owner[f]] += (double)(-sign) * corr;
        if (f < nI) rhs[m->neighbour[f]] += (double)sign * corr;
    }
    free(g); free(ow);
}

/* ----------------------------------------------------------------- CSR */
/* finishWithDiagonal, Matrix.cpp:56-117: per row the entries sorted by
 * column (lower entries from faces where the row is the neighbour, the
 * diagonal, upper entries from faces where it is the owner). */
void cfo_csr_from_ldu(const cfo_mesh* m, const cfo_ldu* A, int comp, cfo_csr* out)
{
    const int64_t n = m->nCells, nI = m->nInternal;
    int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t f = 0; f < nI; ++f) { ++cnt[m->owner[f]]; ++cnt[m->neighbour[f]]; }
    out->n = n;
    out->rowStart[0] = 0;
    for (int64_t r = 0; r < n; ++r) out->rowStart[r + 1] = out->rowStart[r] + cnt[r] + 1;
    out->nnz = out->rowStart[n];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t r = 0; r < n; ++r) {
        fill[r] = out->rowStart[r];
        out->cols[fill[r]] = r;
        out->vals[fill[r]] = A->diag[comp * n + r];
        ++fill[r];
    }
    for (int64_t f = 0; f < nI; ++f) {
        const int64_t o = m->owner[f], nb = m->neighbour[f];
        out->cols[fill[o]] = nb; out->vals[fill[o]] = A->upper[f]; ++fill[o];
        out->cols[fill[nb]] = o; out->vals[fill[nb]] = A->lower[f]; ++fill[nb];
    }
    /* sort each row by column (insertion sort, stable), merge duplicates */
    int64_t w = 0;
    int64_t rs = 0;
    for (int64_t r = 0; r < n; ++r) {
        const int64_t a = rs, e = out->rowStart[r + 1];
        for (int64_t i = a + 1; i < e; ++i) {
            const int64_t c = out->cols[i];
            const double v = out->vals[i];
            int64_t j = i - 1;
            while (j >= a && out->cols[j] > c) {
                out->cols[j + 1] = out->cols[j]; out->vals[j + 1] = out->vals[j]; --j;
            }
            out->cols[j + 1] = c; out->vals[j + 1] = v;
        }
        const int64_t start = w;
        for (int64_t i = a; i < e; ++i) {
            if (w > start && out->cols[w - 1] == out->cols[i]) out->vals[w - 1] += out->vals[i];
            else { out->cols[w] = out->cols[i]; out->vals[w] = out->vals[i]; ++w; }
        }
        rs = e;
        out->rowStart[r + 1] = w;
    }
    out->nnz = w;
    free(cnt); free(fill);
}

void cfo_spmv(const cfo_csr* A, const double* x, double* y)
{
    for (int64_t r = 0; r < A->n; ++r) {
        double s = 0;
        for (int64_t k = A->rowStart[r]; k < A->rowStart[r + 1]; ++k) s += A->vals[k] * x[A->cols[k]];
        y[r] = s;
    }
}

static double dotv(const double* a, const double* b, int64_t n)
{
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* infNorm, Matrix.cpp:33-42 */
static double inf_norm(const double* v, int64_t n)
{
    double m = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (isnan(v[i])) return NAN;
        const double a = fabs(v[i]);
        m = SMAX(m, a);
    }
    return m;
}

static void csr_diag(const cfo_csr* A, double* d)
{
    for (int64_t r = 0; r < A->n; ++r) {
        d[r] = 0.0;
        for (int64_t k = A->rowStart[r]; k < A->rowStart[r + 1]; ++k)
            if (A->cols[k] == r) { d[r] = A->vals[k]; break; }
    }
}

/* solveCG, Matrix.cpp:153-222 */
cfo_report cfo_solve_cg(const cfo_csr* A, const double* b, double* x,
                        double tol, int maxIter, double* resHist)
{
    const int64_t n = A->n;
    cfo_report rep = {0, 0.0, 0.0, 0};
    double* invD = (double*)malloc(sizeof(double) * (size_t)n);
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    double* z = (double*)malloc(sizeof(double) * (size_t)n);
    double* p = (double*)malloc(sizeof(double) * (size_t)n);
    double* Ap = (double*)malloc(sizeof(double) * (size_t)n);
    csr_diag(A, invD);
    for (int64_t i = 0; i < n; ++i) invD[i] = (fabs(invD[i]) > 1e-300) ? 1.0 / invD[i] : 1.0;
    cfo_spmv(A, x, Ap);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    rep.initialResidual = inf_norm(r, n);
    const double bnorm = SMAX(inf_norm(b, n), 1e-300);
    if (resHist) resHist[0] = rep.initialResidual;
    if (rep.initialResidual / bnorm < tol) {
        rep.converged = 1;
        rep.finalResidual = rep.initialResidual;
        goto done;
    }
    for (int64_t i = 0; i < n; ++i) z[i] = invD[i] * r[i];
    memcpy(p, z, sizeof(double) * (size_t)n);
    double rz = dotv(r, z, n);
    for (int it = 1; it <= maxIter; ++it) {
        cfo_spmv(A, p, Ap);
        const double pAp = dotv(p, Ap, n);
        if (pAp <= 0 || fabs(rz) < 1e-300) break;
        const double alpha = rz / pAp;
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        rep.iterations = it;
        rep.finalResidual = inf_norm(r, n);
        if (resHist) resHist[it] = rep.finalResidual;
        if (rep.finalResidual / bnorm < tol) { rep.converged = 1; break; }
        for (int64_t i = 0; i < n; ++i) z[i] = invD[i] * r[i];
        const double rzNew = dotv(r, z, n);
        const double beta = rzNew / rz;
        rz = rzNew;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
done:
    free(invD); free(r); free(z); free(p); free(Ap);
    return rep;
}

/* solveBiCGStab, Matrix.cpp:224-319 */
cfo_report cfo_solve_bicgstab(const cfo_csr* A, const double* b, double* x,
                              double tol, int maxIter)
{
    const int64_t n = A->n;
    cfo_report rep = {0, 0.0, 0.0, 0};
    const size_t sz = sizeof(double) * (size_t)n;
    double* invD = (double*)malloc(sz);
    double* r = (double*)malloc(sz);
    double* r0 = (double*)malloc(sz);
    double* p = (double*)calloc((size_t)n, sizeof(double));
    double* v = (double*)calloc((size_t)n, sizeof(double));
    double* s = (double*)calloc((size_t)n, sizeof(double));
    double* z = (double*)calloc((size_t)n, sizeof(double));
    double* sh = (double*)calloc((size_t)n, sizeof(double));
    double* t = (double*)calloc((size_t)n, sizeof(double));
    csr_diag(A, invD);
    for (int64_t i = 0; i < n; ++i) invD[i] = (invD[i] > 1e-300) ? 1.0 / invD[i] : 1.0;
    cfo_spmv(A, x, t);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - t[i];
    rep.initialResidual = inf_norm(r, n);
    const double bnorm = SMAX(inf_norm(b, n), 1e-300);
    if (rep.initialResidual / bnorm < tol) {
        rep.converged = 1;
        rep.finalResidual = rep.initialResidual;
        goto done;
    }
    memcpy(r0, r, sz);
    memset(t, 0, sz);
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    for (int it = 1; it <= maxIter; ++it) {
        const double rhoNew = dotv(r0, r, n);
        if (fabs(rhoNew) < 1e-300) break;
        const double beta = (rhoNew / rho) * (alpha / omega);
        rho = rhoNew;
        for (int64_t i = 0; i < n; ++i) {
            p[i] = r[i] + beta * (p[i] - omega * v[i]);
            z[i] = p[i] * invD[i];
        }
        cfo_spmv(A, z, v);
        const double denom = dotv(r0, v, n);
        if (fabs(denom) < 1e-300) break;
        alpha = rho / denom;
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * z[i];
            s[i] = r[i] - alpha * v[i];
        }
        for (int64_t i = 0; i < n; ++i) sh[i] = s[i] * invD[i];
        cfo_spmv(A, sh, t);
        const double tDenom = dotv(t, t, n);
        if (tDenom < 1e-300) {
            memcpy(r, s, sz);
            rep.iterations = it;
            rep.finalResidual = inf_norm(r, n);
            rep.converged = rep.finalResidual / bnorm < tol;
            break;
        }
        omega = dotv(t, s, n) / tDenom;
        for (int64_t i = 0; i < n; ++i) {
            x[i] += omega * sh[i];
            r[i] = s[i] - omega * t[i];
        }
        rep.iterations = it;
        rep.finalResidual = inf_norm(r, n);
        if (rep.finalResidual / bnorm < tol) { rep.converged = 1; break; }
    }
done:
    free(invD); free(r); free(r0); free(p); free(v); free(s); free(z); free(sh); free(t);
    return rep;
}

/* ------------------------------------------------------ flux operators */
void cfo_update_phi_internal(const cfo_mesh* m, const double* U, double* phi)
{
    const int64_t nI = m->nInternal;
    double* ow = (double*)malloc(sizeof(double) * (size_t)(nI > 0 ? nI : 1));
    cfo_owner_weights(m, ow);
    for (int64_t f = 0; f < nI; ++f) {
        const V3 cO = vload(U + 3 * m->owner[f]);
        const V3 cN = vload(U + 3 * m->neighbour[f]);
        const V3 uf = vadd(vmul(cO, ow[f]), vmul(cN, 1.0 - ow[f]));
        phi[f] = vdot(uf, vload(m->Sf + 3 * f));
    }
    free(ow);
}

void cfo_correct_boundary_flux(const cfo_mesh* m, const double* U,
                               const double* lin_c, const double* lin_k,
                               double* phi)
{
    const int64_t nI = m->nInternal;
    for (int64_t f = nI; f < m->nFaces; ++f) {
        const int64_t b = f - nI;
        const V3 tp = vload(U + 3 * m->owner[f]);
        const V3 uf = v3(lin_c[3 * b] * tp.x + lin_k[3 * b],
                         lin_c[3 * b + 1] * tp.y + lin_k[3 * b + 1],
                         lin_c[3 * b + 2] * tp.z + lin_k[3 * b + 2]);
        phi[f] = vdot(uf, vload(m->Sf + 3 * f));
    }
}

void cfo_correct_fluxes(const cfo_mesh* m, const double* p,
                        const double* gradP, const double* rAUf,
                        const double* lin_c, const double* lin_k,
                        int corrected, double* phi)
{
    const int64_t nI = m->nInternal;
    double* ow = (double*)malloc(sizeof(double) * (size_t)(nI > 0 ? nI : 1));
    cfo_owner_weights(m, ow);
    for (int64_t f = 0; f < m->nFaces; ++f) {
        const V3 dcv = vload(m->dvec + 3 * f);
        const V3 sf = vload(m->Sf + 3 * f);
        const double d2 = vdot(dcv, dcv);
        if (d2 <= 0) continue;
        const double coef = rAUf[f] * vdot(sf, dcv) / d2;
        double corr = 0;
        if (corrected) {
            const V3 dh = vdiv(dcv, sqrt(d2));
            const V3 cfVec = vsub(sf, vmul(dh, vdot(sf, dh)));
            V3 gradFace;
            if (f < nI)
                gradFace = vadd(vmul(vload(gradP + 3 * m->owner[f]), ow[f]),
                                vmul(vload(gradP + 3 * m->neighbour[f]), 1.0 - ow[f]));
            else
                gradFace = vload(gradP + 3 * m->owner[f]);
            corr = rAUf[f] * vdot(cfVec, gradFace);
        }
        if (f < nI) {
            const double dp = p[m->neighbour[f]] - p[m->owner[f]];
            phi[f] -= coef * dp + corr;
        } else {
            const int64_t b = f - nI;
            const double tp = p[m->owner[f]];
            phi[f] -= coef * ((lin_c[b] - 1.0) * tp + lin_k[b]) + corr;
        }
    }
    free(ow);
}

void cfo_div_phi(const cfo_mesh* m, const double* phi, double* out)
{
    for (int64_t c = 0; c < m->nCells; ++c) out[c] = 0.0;
    for (int64_t f = 0; f < m->nFaces; ++f) {
        out[m->owner[f]] += phi[f];
        if (f < m->nInternal) out[m->neighbour[f]] -= phi[f];
    }
}

double cfo_continuity_error(const cfo_mesh* m, const double* phi)
{
    const int64_t nC = m->nCells;
    double* scale = (double*)calloc((size_t)nC, sizeof(double));
    double* dv = (double*)malloc(sizeof(double) * (size_t)nC);
    for (int64_t f = 0; f < m->nFaces; ++f) {
        const double a = m->magSf[f];
        const int64_t o = m->owner[f];
        scale[o] = SMAX(scale[o], a);
        if (f < m->nInternal) {
            const int64_t nb = m->neighbour[f];
            scale[nb] = SMAX(scale[nb], a);
        }
    }
    cfo_div_phi(m, phi, dv);
    double err = 0;
    for (int64_t c = 0; c < nC; ++c) {
        const double e = fabs(dv[c]) / SMAX(scale[c], 1e-300);
        err = SMAX(err, e);
    }
    free(scale); free(dv);
    return err;
}
