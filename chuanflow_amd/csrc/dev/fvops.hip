// fvops.hip -- finite-volume operator kernels (face maps + per-cell gathers).
#include "fvops.h"

#include <algorithm>

namespace cfx
{
namespace
{

constexpr int B = 256;
inline int grid(long long n) { return std::max(1, divUp(n, B)); }
inline int gridCap(long long n) { return std::max(1, std::min(2048, divUp(n, B))); }

// Face value of cell data: bounded interpolation on internal faces (or the
// volScalarField::faceValue weights when `quirk`), BC linearisation on
// boundary faces.
__global__ void k_face_value(int nF, int nI, const int* __restrict__ owner,
                             const int* __restrict__ nbr, const double* __restrict__ ow,
                             const double* __restrict__ x, const double* __restrict__ linc,
                             const double* __restrict__ link, int ncomp, int comp, int quirk,
                             const double* __restrict__ Cf, const double* __restrict__ C,
                             const double* __restrict__ d, int n, double* __restrict__ out)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const int o = owner[f];
    if (f < nI)
    {
        const int nb = nbr[f];
        if (!quirk)
        {
            const double w = ow[f];
            out[f] = w * x[o] + (1.0 - w) * x[nb];
        }
        else
        {
            const double dx = d[f], dy = d[nF + f], dz = d[2 * nF + f];
            const double d2 = dx * dx + dy * dy + dz * dz;
            double w = 0.5;
            if (d2 > 0)
            {
                const double ex = Cf[f] - C[o], ey = Cf[nF + f] - C[n + o], ez = Cf[2 * nF + f] - C[2 * n + o];
                w = (ex * dx + ey * dy + ez * dz) / d2;
                w = w > 0.0 ? w : 0.0;
                w = w < 1.0 ? w : 1.0;
            }
            out[f] = w * x[o] + (1.0 - w) * x[nb];
        }
    }
    else
    {
        const int b = f - nI;
        out[f] = linc[b * ncomp + comp] * x[o] + link[b * ncomp + comp];
    }
}

// grad = (sum_f +-Sf tf) / V, faces in ascending order from zero.
__global__ void k_grad_gather(int n, int nF, const int* __restrict__ cfs, const int* __restrict__ cf,
                              const double* __restrict__ Sf, const double* __restrict__ tf,
                              const double* __restrict__ V, double* __restrict__ g)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double sx = 0, sy = 0, sz = 0;
    for (int k = cfs[c]; k < cfs[c + 1]; ++k)
    {
        const int e = cf[k], f = e >> 1;
        const double t = tf[f];
        const double fx = Sf[f] * t, fy = Sf[nF + f] * t, fz = Sf[2 * nF + f] * t;
        if (e & 1) { sx -= fx; sy -= fy; sz -= fz; }
        else { sx += fx; sy += fy; sz += fz; }
    }
    const double v = V[c];
    g[c] = sx / v;
    g[n + c] = sy / v;
    g[2 * n + c] = sz / v;
}

// tf = reconstructed value + (internal) gradFace . skew   (gradScheme.cpp:61-84)
__global__ void k_grad_skew_face(int nF, int nI, int n, const int* __restrict__ owner,
                                 const int* __restrict__ nbr, const double* __restrict__ ow,
                                 const double* __restrict__ skew, const double* __restrict__ g,
                                 const double* __restrict__ x, const double* __restrict__ linc,
                                 const double* __restrict__ link, int ncomp, int comp,
                                 double* __restrict__ tf)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const int o = owner[f];
    if (f < nI)
    {
        const int nb = nbr[f];
        const double w = ow[f], w1 = 1.0 - w;
        const double gx = g[o] * w + g[nb] * w1;
        const double gy = g[n + o] * w + g[n + nb] * w1;
        const double gz = g[2 * n + o] * w + g[2 * n + nb] * w1;
        double t = w * x[o] + (1.0 - w) * x[nb];
        t += gx * skew[f] + gy * skew[nI + f] + gz * skew[2 * nI + f];
        tf[f] = t;
    }
    else
    {
        const int b = f - nI;
        tf[f] = linc[b * ncomp + comp] * x[o] + link[b * ncomp + comp];
    }
}

__global__ void k_harmonic_gamma(int nF, int nI, int n, const int* __restrict__ owner,
                                 const int* __restrict__ nbr, const double* __restrict__ Cf,
                                 const double* __restrict__ C, double gamma, double* __restrict__ gf)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    if (f < nI)
    {
        const int o = owner[f], nb = nbr[f];
        const double gp = gamma, gn = gamma;
        const double ax = Cf[f] - C[o], ay = Cf[nF + f] - C[n + o], az = Cf[2 * nF + f] - C[2 * n + o];
        const double bx = C[nb] - Cf[f], by = C[n + nb] - Cf[nF + f], bz = C[2 * n + nb] - Cf[2 * nF + f];
        const double dP = sqrt(ax * ax + ay * ay + az * az), dN = sqrt(bx * bx + by * by + bz * bz);
        const double tot = dP + dN;
        gf[f] = (tot > 0 && gp > 0 && gn > 0) ? tot * gp * gn / (dN * gp + dP * gn) : 0.5 * (gp + gn);
    }
    else
    {
        gf[f] = gamma;
    }
}

__global__ void k_interp_faces(int nF, int nI, const int* __restrict__ owner, const int* __restrict__ nbr,
                               const double* __restrict__ ow, const double* __restrict__ cell,
                               double* __restrict__ face)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const int o = owner[f];
    face[f] = f < nI ? ow[f] * cell[o] + (1.0 - ow[f]) * cell[nbr[f]] : cell[o];
}

__global__ void k_ldu_reset(long long nd, int nI, double* __restrict__ diag, double* __restrict__ rhs,
                            double* __restrict__ up, double* __restrict__ lo)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < nd) { diag[i] = 0.0; rhs[i] = 0.0; }
    if (i < nI) { up[i] = -0.0; lo[i] = -0.0; }  // -0.0: additive identity
}

// laplacian: off-diagonals (laplacianScheme.cpp:82-91)
__global__ void k_lap_face(int nI, const double* __restrict__ gf, const double* __restrict__ sfd,
                           const double* __restrict__ d2, int sign, double* __restrict__ up,
                           double* __restrict__ lo)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nI) return;
    const double coef = gf[f] * (sfd[f] / d2[f]);
    up[f] += static_cast<double>(sign) * coef;
    lo[f] += static_cast<double>(sign) * coef;
}

// laplacian: diagonal and boundary rhs, gathered per cell (:87-102)
__global__ void k_lap_cell(int n, int nI, int ncomp, const int* __restrict__ cfs, const int* __restrict__ cf,
                           const double* __restrict__ gf, const double* __restrict__ sfd,
                           const double* __restrict__ d2, const double* __restrict__ linc,
                           const double* __restrict__ link, int sign, double* __restrict__ diag,
                           double* __restrict__ rhs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double ms = static_cast<double>(-sign), ps = static_cast<double>(sign);
    for (int comp = 0; comp < ncomp; ++comp)
    {
        double dg = diag[comp * n + c], rh = rhs[comp * n + c];
        for (int k = cfs[c]; k < cfs[c + 1]; ++k)
        {
            const int f = cf[k] >> 1;
            const double coef = gf[f] * (sfd[f] / d2[f]);
            if (f < nI)
            {
                dg += ms * coef;
            }
            else
            {
                const int b = f - nI;
                dg += ps * coef * (linc[b * ncomp + comp] - 1.0);
                rh += ms * coef * link[b * ncomp + comp];
            }
        }
        diag[comp * n + c] = dg;
        rhs[comp * n + c] = rh;
    }
}

// upwind/linear/blend weights (divScheme.cpp:26-42)
__device__ __forceinline__ void divWeights(double p, double alpha, double c, double& wPo, double& wNo)
{
    wPo = 0.0;
    wNo = 0.0;
    if (p >= 0) wPo = 1.0; else wNo = 1.0;
    if (alpha > 0)
    {
        wPo = alpha * c + (1.0 - alpha) * wPo;
        wNo = alpha * (1.0 - c) + (1.0 - alpha) * wNo;
    }
}

__global__ void k_div_face(int nI, const double* __restrict__ phi, const double* __restrict__ ow,
                           double alpha, int sign, double* __restrict__ up, double* __restrict__ lo)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nI) return;
    const double p = phi[f];
    double wPo, wNo;
    divWeights(p, alpha, ow[f], wPo, wNo);
    const double fp = static_cast<double>(sign) * p * wPo;
    const double fn = static_cast<double>(sign) * p * wNo;
    up[f] += fn;
    lo[f] += -fp;
}

__global__ void k_div_cell(int n, int nI, int ncomp, const int* __restrict__ cfs, const int* __restrict__ cf,
                           const double* __restrict__ phi, const double* __restrict__ ow, double alpha,
                           const double* __restrict__ linc, const double* __restrict__ link, int sign,
                           double* __restrict__ diag, double* __restrict__ rhs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int comp = 0; comp < ncomp; ++comp)
    {
        double dg = diag[comp * n + c], rh = rhs[comp * n + c];
        for (int k = cfs[c]; k < cfs[c + 1]; ++k)
        {
            const int e = cf[k], f = e >> 1;
            const double p = phi[f];
            if (f < nI)
            {
                double wPo, wNo;
                divWeights(p, alpha, ow[f], wPo, wNo);
                if (e & 1) dg += -(static_cast<double>(sign) * p * wNo);
                else dg += static_cast<double>(sign) * p * wPo;
            }
            else
            {
                const int b = f - nI;
                dg += static_cast<double>(sign) * p * linc[b * ncomp + comp];
                rh += static_cast<double>(-sign) * p * link[b * ncomp + comp];
            }
        }
        diag[comp * n + c] = dg;
        rhs[comp * n + c] = rh;
    }
}

__global__ void k_ddt(int n, int ncomp, const double* __restrict__ V, const double* __restrict__ old,
                      double dt, int sign, double* __restrict__ diag, double* __restrict__ rhs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double a = static_cast<double>(sign) * V[c] / dt;
    for (int k = 0; k < ncomp; ++k)
    {
        diag[k * n + c] += a;
        rhs[k * n + c] += a * old[k * n + c];
    }
}

__global__ void k_vol_source(int n, const double* __restrict__ V, double fx, double fy, double fz,
                             const double* __restrict__ g, double* __restrict__ rhs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double v = V[c];
    rhs[c] += ((fx + 0.0) - g[c]) * v;
    rhs[n + c] += ((fy + 0.0) - g[n + c]) * v;
    rhs[2 * n + c] += ((fz + 0.0) - g[2 * n + c]) * v;
}

__global__ void k_add_neg(int n, const double* __restrict__ v, double* __restrict__ rhs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) rhs[c] += -v[c];
}

__global__ void k_csr_offdiag(int nI, const int* __restrict__ slotU, const int* __restrict__ slotL,
                              const double* __restrict__ up, const double* __restrict__ lo,
                              double* __restrict__ val)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nI) return;
    val[slotU[f]] = up[f];
    val[slotL[f]] = lo[f];
}

__global__ void k_csr_diag(int n, const int* __restrict__ diagSlot, const double* __restrict__ diag,
                           double* __restrict__ val)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) val[diagSlot[c]] = diag[c];
}

// corr_f = gf * (kf . gradFace)   (laplacianScheme.cpp:124-146)
__global__ void k_nonorth_face(int nF, int nI, int n, const int* __restrict__ owner,
                               const int* __restrict__ nbr, const double* __restrict__ ow,
                               const double* __restrict__ kf, const double* __restrict__ gf,
                               const double* __restrict__ g, double* __restrict__ corr)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const int o = owner[f];
    double gx, gy, gz;
    if (f < nI)
    {
        const int nb = nbr[f];
        const double w = ow[f], w1 = 1.0 - w;
        gx = g[o] * w + g[nb] * w1;
        gy = g[n + o] * w + g[n + nb] * w1;
        gz = g[2 * n + o] * w + g[2 * n + nb] * w1;
    }
    else
    {
        gx = g[o];
        gy = g[n + o];
        gz = g[2 * n + o];
    }
    corr[f] = gf[f] * (kf[f] * gx + kf[nF + f] * gy + kf[2 * nF + f] * gz);
}

__global__ void k_nonorth_cell(int n, const int* __restrict__ cfs, const int* __restrict__ cf,
                               const double* __restrict__ corr, int sign, double* __restrict__ rhs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double r = rhs[c];
    const double ms = static_cast<double>(-sign), ps = static_cast<double>(sign);
    for (int k = cfs[c]; k < cfs[c + 1]; ++k)
    {
        const int e = cf[k];
        const double v = corr[e >> 1];
        r += (e & 1) ? ps * v : ms * v;
    }
    rhs[c] = r;
}

__global__ void k_update_phi(int nI, int nF, int n, const int* __restrict__ owner, const int* __restrict__ nbr,
                             const double* __restrict__ ow, const double* __restrict__ Sf,
                             const double* __restrict__ U, double* __restrict__ phi)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nI) return;
    const int o = owner[f], nb = nbr[f];
    const double w = ow[f], w1 = 1.0 - w;
    const double ux = U[o] * w + U[nb] * w1;
    const double uy = U[n + o] * w + U[n + nb] * w1;
    const double uz = U[2 * n + o] * w + U[2 * n + nb] * w1;
    phi[f] = ux * Sf[f] + uy * Sf[nF + f] + uz * Sf[2 * nF + f];
}

__global__ void k_boundary_flux(int nI, int nF, int n, const int* __restrict__ owner,
                                const double* __restrict__ Sf, const double* __restrict__ U,
                                const double* __restrict__ linc, const double* __restrict__ link,
                                double* __restrict__ phi)
{
    const int f = nI + blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const int o = owner[f], b = f - nI;
    const double ux = linc[3 * b] * U[o] + link[3 * b];
    const double uy = linc[3 * b + 1] * U[n + o] + link[3 * b + 1];
    const double uz = linc[3 * b + 2] * U[2 * n + o] + link[3 * b + 2];
    phi[f] = ux * Sf[f] + uy * Sf[nF + f] + uz * Sf[2 * nF + f];
}

// correctFluxes (pressureCorrection.cpp:50-87)
__global__ void k_correct_fluxes(int nF, int nI, int n, const int* __restrict__ owner,
                                 const int* __restrict__ nbr, const double* __restrict__ ow,
                                 const double* __restrict__ sfd, const double* __restrict__ d2,
                                 const double* __restrict__ kf, const double* __restrict__ p,
                                 const double* __restrict__ g, const double* __restrict__ rAUf,
                                 const double* __restrict__ linc, const double* __restrict__ link,
                                 int corrected, double* __restrict__ phi)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const double dd = d2[f];
    if (dd <= 0) return;
    const double coef = rAUf[f] * sfd[f] / dd;
    const int o = owner[f];
    double corr = 0;
    if (corrected)
    {
        double gx, gy, gz;
        if (f < nI)
        {
            const int nb = nbr[f];
            const double w = ow[f], w1 = 1.0 - w;
            gx = g[o] * w + g[nb] * w1;
            gy = g[n + o] * w + g[n + nb] * w1;
            gz = g[2 * n + o] * w + g[2 * n + nb] * w1;
        }
        else
        {
            gx = g[o];
            gy = g[n + o];
            gz = g[2 * n + o];
        }
        corr = rAUf[f] * (kf[f] * gx + kf[nF + f] * gy + kf[2 * nF + f] * gz);
    }
    if (f < nI)
    {
        const double dp = p[nbr[f]] - p[o];
        phi[f] -= coef * dp + corr;
    }
    else
    {
        const int b = f - nI;
        const double tp = p[o];
        phi[f] -= coef * ((linc[b] - 1.0) * tp + link[b]) + corr;
    }
}

__global__ void k_div_phi(int n, const int* __restrict__ cfs, const int* __restrict__ cf,
                          const double* __restrict__ phi, double* __restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double s = 0.0;
    for (int k = cfs[c]; k < cfs[c + 1]; ++k)
    {
        const int e = cf[k];
        if (e & 1) s -= phi[e >> 1];
        else s += phi[e >> 1];
    }
    out[c] = s;
}

// max_c |div_c| / max(localScale_c, 1e-300), NaN-dropping like std::max
__global__ __launch_bounds__(B) void k_cont_err(int n, const int* __restrict__ cfs, const int* __restrict__ cf,
                                                const double* __restrict__ phi,
                                                const double* __restrict__ scale, double* __restrict__ dv,
                                                double* result, unsigned int* ticket, double* part)
{
    __shared__ double red[B / kWave];
    __shared__ int flag;
    double m = 0.0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    {
        double s = 0.0;
        for (int k = cfs[c]; k < cfs[c + 1]; ++k)
        {
            const int e = cf[k];
            if (e & 1) s -= phi[e >> 1];
            else s += phi[e >> 1];
        }
        dv[c] = s;
        const double sc = scale[c];
        m = fmax(m, fabs(s) / (sc < 1e-300 ? 1e-300 : sc));
    }
    m = blockMax(m, red);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
    if (!lastBlockDone(ticket, &flag)) return;
    const double M = maxPartials(part, gridDim.x, red);
    if (threadIdx.x == 0) *result = M;
}

} // namespace

void gradGauss(DevMesh& m, const double* x, const double* linc, const double* link, int ncomp, int comp,
               bool corrected, bool quirk, double* g, double* tf)
{
    hipStream_t s = m.ctx->stream();
    k_face_value<<<grid(m.nF), B, 0, s>>>(m.nF, m.nI, m.owner.as<int>(), m.neighbour.as<int>(),
                                          m.ow.as<double>(), x, linc, link, ncomp, comp,
                                          (!corrected && quirk) ? 1 : 0, m.Cf.as<double>(),
                                          m.C.as<double>(), m.d.as<double>(), m.n, tf);
    k_grad_gather<<<grid(m.n), B, 0, s>>>(m.n, m.nF, m.cellFaceStart.as<int>(), m.cellFace.as<int>(),
                                          m.Sf.as<double>(), tf, m.V.as<double>(), g);
    if (corrected)
    {
        for (int pass = 0; pass < 2; ++pass)
        {
            k_grad_skew_face<<<grid(m.nF), B, 0, s>>>(m.nF, m.nI, m.n, m.owner.as<int>(),
                                                      m.neighbour.as<int>(), m.ow.as<double>(),
                                                      m.skew.as<double>(), g, x, linc, link, ncomp,
                                                      comp, tf);
            k_grad_gather<<<grid(m.n), B, 0, s>>>(m.n, m.nF, m.cellFaceStart.as<int>(),
                                                  m.cellFace.as<int>(), m.Sf.as<double>(), tf,
                                                  m.V.as<double>(), g);
        }
    }
    CFX_CHECK_LAUNCH();
}

void harmonicGammaUniform(DevMesh& m, double gamma, double* gf)
{
    k_harmonic_gamma<<<grid(m.nF), B, 0, m.ctx->stream()>>>(m.nF, m.nI, m.n, m.owner.as<int>(),
                                                            m.neighbour.as<int>(), m.Cf.as<double>(),
                                                            m.C.as<double>(), gamma, gf);
    CFX_CHECK_LAUNCH();
}

void interpolateToFaces(DevMesh& m, const double* cell, double* face)
{
    k_interp_faces<<<grid(m.nF), B, 0, m.ctx->stream()>>>(m.nF, m.nI, m.owner.as<int>(),
                                                          m.neighbour.as<int>(), m.ow.as<double>(), cell,
                                                          face);
    CFX_CHECK_LAUNCH();
}

void lduReset(DevMesh& m, const Ldu& A)
{
    const long long nd = static_cast<long long>(A.ncomp) * m.n;
    k_ldu_reset<<<grid(std::max<long long>(nd, m.nI)), B, 0, m.ctx->stream()>>>(nd, m.nI, A.diag, A.rhs,
                                                                              A.upper, A.lower);
    CFX_CHECK_LAUNCH();
}

void assembleLaplacian(DevMesh& m, const Ldu& A, const double* gf, const double* linc, const double* link,
                       int sign)
{
    hipStream_t s = m.ctx->stream();
    if (m.nI > 0)
        k_lap_face<<<grid(m.nI), B, 0, s>>>(m.nI, gf, m.sfd.as<double>(), m.d2.as<double>(), sign, A.upper,
                                            A.lower);
    k_lap_cell<<<grid(m.n), B, 0, s>>>(m.n, m.nI, A.ncomp, m.cellFaceStart.as<int>(), m.cellFace.as<int>(),
                                       gf, m.sfd.as<double>(), m.d2.as<double>(), linc, link, sign, A.diag,
                                       A.rhs);
    CFX_CHECK_LAUNCH();
}

void assembleDiv(DevMesh& m, const Ldu& A, const double* phi, int scheme, double blendCoeff,
                 const double* linc, const double* link, int sign)
{
    hipStream_t s = m.ctx->stream();
    const double alpha = scheme == 2 ? blendCoeff : (scheme == 1 ? 1.0 : 0.0);
    if (m.nI > 0)
        k_div_face<<<grid(m.nI), B, 0, s>>>(m.nI, phi, m.ow.as<double>(), alpha, sign, A.upper, A.lower);
    k_div_cell<<<grid(m.n), B, 0, s>>>(m.n, m.nI, A.ncomp, m.cellFaceStart.as<int>(), m.cellFace.as<int>(),
                                       phi, m.ow.as<double>(), alpha, linc, link, sign, A.diag, A.rhs);
    CFX_CHECK_LAUNCH();
}

void assembleDdt(DevMesh& m, const Ldu& A, const double* old3, double dt, int sign)
{
    k_ddt<<<grid(m.n), B, 0, m.ctx->stream()>>>(m.n, A.ncomp, m.V.as<double>(), old3, dt, sign, A.diag,
                                                A.rhs);
    CFX_CHECK_LAUNCH();
}

void addVolumeSource(DevMesh& m, const Ldu& A, const double f[3], const double* g)
{
    k_vol_source<<<grid(m.n), B, 0, m.ctx->stream()>>>(m.n, m.V.as<double>(), f[0], f[1], f[2], g, A.rhs);
    CFX_CHECK_LAUNCH();
}

void addNegRhs(DevMesh& m, const Ldu& A, const double* v)
{
    k_add_neg<<<grid(m.n), B, 0, m.ctx->stream()>>>(m.n, v, A.rhs);
    CFX_CHECK_LAUNCH();
}

void lduToCsr(DevMesh& m, const Ldu& A, int comp, double* val)
{
    hipStream_t s = m.ctx->stream();
    if (m.nI > 0)
        k_csr_offdiag<<<grid(m.nI), B, 0, s>>>(m.nI, m.slotU.as<int>(), m.slotL.as<int>(), A.upper, A.lower,
                                               val);
    k_csr_diag<<<grid(m.n), B, 0, s>>>(m.n, m.pattern.diagSlot.as<int>(), A.diag + static_cast<long long>(comp) * m.n,
                                       val);
    CFX_CHECK_LAUNCH();
}

void nonOrthCorrect(DevMesh& m, const double* gf, const double* x, const double* linc, const double* link,
                    int ncomp, int comp, bool gradCorrected, int sign, double* rhs, double* g, double* tf)
{
    gradGauss(m, x, linc, link, ncomp, comp, gradCorrected, false, g, tf);
    hipStream_t s = m.ctx->stream();
    k_nonorth_face<<<grid(m.nF), B, 0, s>>>(m.nF, m.nI, m.n, m.owner.as<int>(), m.neighbour.as<int>(),
                                            m.ow.as<double>(), m.kf.as<double>(), gf, g, tf);
    k_nonorth_cell<<<grid(m.n), B, 0, s>>>(m.n, m.cellFaceStart.as<int>(), m.cellFace.as<int>(), tf, sign,
                                           rhs);
    CFX_CHECK_LAUNCH();
}

void updatePhiInternal(DevMesh& m, const double* U, double* phi)
{
    if (m.nI == 0) return;
    k_update_phi<<<grid(m.nI), B, 0, m.ctx->stream()>>>(m.nI, m.nF, m.n, m.owner.as<int>(),
                                                        m.neighbour.as<int>(), m.ow.as<double>(),
                                                        m.Sf.as<double>(), U, phi);
    CFX_CHECK_LAUNCH();
}

void correctBoundaryFlux(DevMesh& m, const double* U, const double* linc, const double* link, double* phi)
{
    if (m.nB == 0) return;
    k_boundary_flux<<<grid(m.nB), B, 0, m.ctx->stream()>>>(m.nI, m.nF, m.n, m.owner.as<int>(),
                                                           m.Sf.as<double>(), U, linc, link, phi);
    CFX_CHECK_LAUNCH();
}

void correctFluxes(DevMesh& m, const double* p, const double* g, const double* rAUf, const double* linc,
                   const double* link, bool corrected, double* phi)
{
    k_correct_fluxes<<<grid(m.nF), B, 0, m.ctx->stream()>>>(
        m.nF, m.nI, m.n, m.owner.as<int>(), m.neighbour.as<int>(), m.ow.as<double>(), m.sfd.as<double>(),
        m.d2.as<double>(), m.kf.as<double>(), p, g, rAUf, linc, link, corrected ? 1 : 0, phi);
    CFX_CHECK_LAUNCH();
}

void divPhi(DevMesh& m, const double* phi, double* out)
{
    k_div_phi<<<grid(m.n), B, 0, m.ctx->stream()>>>(m.n, m.cellFaceStart.as<int>(), m.cellFace.as<int>(),
                                                    phi, out);
    CFX_CHECK_LAUNCH();
}

void continuityError(DevMesh& m, const double* phi, double* dv, double* dResult, double* part)
{
    // dResult[0] = result, the ticket lives right after it
    unsigned int* ticket = reinterpret_cast<unsigned int*>(dResult + 1);
    k_cont_err<<<gridCap(m.n), B, 0, m.ctx->stream()>>>(m.n, m.cellFaceStart.as<int>(), m.cellFace.as<int>(),
                                                        phi, m.localScale.as<double>(), dv, dResult, ticket,
                                                        part);
    CFX_CHECK_LAUNCH();
}

} // namespace cfx
