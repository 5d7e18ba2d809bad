/*
 * cfo_mesh.c -- oracle: RNG, BlockMesh restatement, jitter, polyMesh geometry.
 * TEST INFRASTRUCTURE ONLY (see cfo.h).
 */
#include "cfo.h"
#include "cfo_vec.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ RNG */
/* MT19937-64 as published by Matsumoto & Nishimura (mt19937-64.c, 2004). */
#define MT_NN 312
#define MT_MM 156
#define MT_MATRIX_A 0xB5026F5AA96619E9ULL
#define MT_UM 0xFFFFFFFF80000000ULL
#define MT_LM 0x7FFFFFFFULL

void cfo_mt64_seed(cfo_mt64* g, uint64_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < MT_NN; ++i)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->mti = MT_NN;
}

uint64_t cfo_mt64_next(cfo_mt64* g)
{
    static const uint64_t mag01[2] = {0ULL, MT_MATRIX_A};
    uint64_t x;
    if (g->mti >= MT_NN) {
        int i;
        for (i = 0; i < MT_NN - MT_MM; ++i) {
            x = (g->mt[i] & MT_UM) | (g->mt[i + 1] & MT_LM);
            g->mt[i] = g->mt[i + MT_MM] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        }
        for (; i < MT_NN - 1; ++i) {
            x = (g->mt[i] & MT_UM) | (g->mt[i + 1] & MT_LM);
            g->mt[i] = g->mt[i + (MT_MM - MT_NN)] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        }
        x = (g->mt[MT_NN - 1] & MT_UM) | (g->mt[0] & MT_LM);
        g->mt[MT_NN - 1] = g->mt[MT_MM - 1] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        g->mti = 0;
    }
    x = g->mt[g->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

uint64_t cfo_mt64_nth(uint64_t seed, int64_t n)
{
    cfo_mt64 g;
    cfo_mt64_seed(&g, seed);
    uint64_t v = 0;
    for (int64_t i = 0; i < n; ++i) v = cfo_mt64_next(&g);
    return v;
}

double cfo_mt64_uniform(cfo_mt64* g)
{
    return (double)(cfo_mt64_next(g) >> 11) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------ BlockMesh */
/* BlockMesh::create, blockMesh.cpp:36-189.  The reference deduplicates faces
 * through a std::map keyed on sorted vertex ids and keeps them in
 * first-encounter order; on the structured lattice that order is: for each
 * cell (k, j, i outermost to innermost) the faces zlo (only if k==0), zhi,
 * ylo (only if j==0), yhi, xlo (only if i==0), xhi, each with the vertex list
 * of the cell that registered it first (the lower cell id, which is also the
 * owner).  Internal faces first, then the boundary faces side by side
 * (xmin xmax ymin ymax zmin zmax, :124-185), each side in the same
 * first-encounter order. */
int cfo_blockmesh(int64_t nx, int64_t ny, int64_t nz,
                  const double mn[3], const double mx[3],
                  double* points, int64_t* faceVerts,
                  int64_t* owner, int64_t* neighbour,
                  int64_t* patchStart, int64_t* patchSize,
                  int64_t* nInternalOut)
{
    if (nx <= 0 || ny <= 0 || nz <= 0) return -1;
    const int64_t px = nx + 1, py = ny + 1;
#define PT(i, j, k) ((int64_t)(i) + px * ((int64_t)(j) + py * (int64_t)(k)))
#define CELL(i, j, k) ((int64_t)(i) + nx * ((int64_t)(j) + ny * (int64_t)(k)))
    for (int64_t k = 0; k <= nz; ++k)
        for (int64_t j = 0; j <= ny; ++j)
            for (int64_t i = 0; i <= nx; ++i) {
                double* q = points + 3 * PT(i, j, k);
                q[0] = mn[0] + (mx[0] - mn[0]) * (double)i / (double)nx;
                q[1] = mn[1] + (mx[1] - mn[1]) * (double)j / (double)ny;
                q[2] = mn[2] + (mx[2] - mn[2]) * (double)k / (double)nz;
            }

    const int64_t nCells = nx * ny * nz;
    const int64_t nFaces = 3 * nCells + nx * ny + ny * nz + nz * nx;
    /* first-encounter enumeration */
    int64_t* fv = (int64_t*)malloc(sizeof(int64_t) * 4 * (size_t)nFaces);
    int64_t* fo = (int64_t*)malloc(sizeof(int64_t) * (size_t)nFaces);
    int64_t* fn = (int64_t*)malloc(sizeof(int64_t) * (size_t)nFaces);
    int* side = (int*)malloc(sizeof(int) * (size_t)nFaces);
    int64_t cnt = 0;
#define ADD(a, b, c, d, own, nb, sd) do { \
        fv[4 * cnt] = (a); fv[4 * cnt + 1] = (b); fv[4 * cnt + 2] = (c); fv[4 * cnt + 3] = (d); \
        fo[cnt] = (own); fn[cnt] = (nb); side[cnt] = (sd); ++cnt; } while (0)
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                const int64_t c = CELL(i, j, k);
                const int64_t v0 = PT(i, j, k), v1 = PT(i + 1, j, k);
                const int64_t v2 = PT(i + 1, j + 1, k), v3 = PT(i, j + 1, k);
                const int64_t v4 = PT(i, j, k + 1), v5 = PT(i + 1, j, k + 1);
                const int64_t v6 = PT(i + 1, j + 1, k + 1), v7 = PT(i, j + 1, k + 1);
                if (k == 0) ADD(v0, v1, v2, v3, c, -1, 4);
                ADD(v4, v5, v6, v7, c, (k + 1 < nz) ? CELL(i, j, k + 1) : -1, 5);
                if (j == 0) ADD(v0, v1, v5, v4, c, -1, 2);
                ADD(v2, v3, v7, v6, c, (j + 1 < ny) ? CELL(i, j + 1, k) : -1, 3);
                if (i == 0) ADD(v0, v3, v7, v4, c, -1, 0);
                ADD(v1, v2, v6, v5, c, (i + 1 < nx) ? CELL(i + 1, j, k) : -1, 1);
            }
#undef ADD
    int64_t out = 0;
    for (int64_t f = 0; f < cnt; ++f) {
        if (fn[f] < 0) continue;
        memcpy(faceVerts + 4 * out, fv + 4 * f, 4 * sizeof(int64_t));
        owner[out] = fo[f];
        neighbour[out] = fn[f];
        ++out;
    }
    *nInternalOut = out;
    for (int s = 0; s < 6; ++s) {
        patchStart[s] = out;
        for (int64_t f = 0; f < cnt; ++f) {
            if (fn[f] >= 0 || side[f] != s) continue;
            memcpy(faceVerts + 4 * out, fv + 4 * f, 4 * sizeof(int64_t));
            owner[out] = fo[f];
            ++out;
        }
        patchSize[s] = out - patchStart[s];
    }
    free(fv); free(fo); free(fn); free(side);
#undef PT
#undef CELL
    return out == nFaces ? 0 : -2;
}

void cfo_jitter(int64_t nx, int64_t ny, int64_t nz,
                const double mn[3], const double mx[3],
                double amp, uint64_t seed, double* points)
{
    cfo_mt64 g;
    cfo_mt64_seed(&g, seed);
    const double h[3] = {(mx[0] - mn[0]) / (double)nx, (mx[1] - mn[1]) / (double)ny,
                         (mx[2] - mn[2]) / (double)nz};
    const int64_t px = nx + 1, py = ny + 1;
    for (int64_t k = 1; k < nz; ++k)
        for (int64_t j = 1; j < ny; ++j)
            for (int64_t i = 1; i < nx; ++i) {
                double* q = points + 3 * (i + px * (j + py * k));
                for (int c = 0; c < 3; ++c) {
                    const double u = cfo_mt64_uniform(&g);
                    q[c] += (2.0 * u - 1.0) * amp * h[c];
                }
            }
}

/* ------------------------------------------------------------ geometry */
/* polyMesh::calcGeometry, polyMesh.cpp:83-288, in the same operation order. */
int cfo_geometry(const cfo_mesh* m)
{
    const int64_t nF = m->nFaces, nI = m->nInternal, nC = m->nCells;
    const double* P = m->points;
#define PNT(v) vload(P + 3 * (v))
    for (int64_t f = 0; f < nF; ++f) {
        const int64_t* vs = m->faceVerts + m->faceStart[f];
        const int64_t nv = m->faceStart[f + 1] - m->faceStart[f];
        const V3 v0 = PNT(vs[0]);
        V3 n = vzero();
        for (int64_t i = 1; i + 1 < nv; ++i) {
            const V3 a = vsub(PNT(vs[i]), v0);
            const V3 b = vsub(PNT(vs[i + 1]), v0);
            n = vadd(n, vcross(a, b));
        }
        vstore(m->Sf + 3 * f, vmul(n, 0.5));

        int exactMean = (nv == 3);
        if (nv == 4) exactMean = veq(vadd(PNT(vs[0]), PNT(vs[2])), vadd(PNT(vs[1]), PNT(vs[3])));
        V3 c = vzero();
        if (exactMean) {
            for (int64_t i = 0; i < nv; ++i) c = vadd(c, PNT(vs[i]));
            c = vdiv(c, (double)nv);
        } else {
            const double wSum = vdot(n, n);
            if (wSum > 0) {
                for (int64_t i = 1; i + 1 < nv; ++i) {
                    const V3 vi = PNT(vs[i]), vj = PNT(vs[i + 1]);
                    const double w = vdot(vcross(vsub(vi, v0), vsub(vj, v0)), n);
                    c = vadd(c, vmul(vadd(vadd(v0, vi), vj), w));
                }
                c = vdiv(c, 3.0 * wSum);
            } else {
                for (int64_t i = 0; i < nv; ++i) c = vadd(c, PNT(vs[i]));
                c = vdiv(c, (double)nv);
            }
        }
        vstore(m->Cf + 3 * f, c);
    }

    /* provisional centres (:160-182) */
    V3* est = (V3*)calloc((size_t)nC, sizeof(V3));
    int64_t* cnt = (int64_t*)calloc((size_t)nC, sizeof(int64_t));
    for (int64_t f = 0; f < nF; ++f) {
        const int64_t o = m->owner[f];
        est[o] = vadd(est[o], vload(m->Cf + 3 * f));
        ++cnt[o];
        if (f < nI) {
            const int64_t nb = m->neighbour[f];
            est[nb] = vadd(est[nb], vload(m->Cf + 3 * f));
            ++cnt[nb];
        }
    }
    for (int64_t c = 0; c < nC; ++c)
        if (cnt[c] > 0) est[c] = vdiv(est[c], (double)cnt[c]);

    /* orientation (:190-210) */
    for (int64_t f = 0; f < nF; ++f) {
        const int64_t o = m->owner[f];
        V3 d;
        if (f < nI) d = vsub(est[m->neighbour[f]], est[o]);
        else d = vsub(vload(m->Cf + 3 * f), est[o]);
        const V3 sf = vload(m->Sf + 3 * f);
        const double side = vdot(sf, d);
        const double tol = 1e-12 * vmag(sf) * vmag(d);
        if (side < -tol) vstore(m->Sf + 3 * f, vneg(sf));
    }
    for (int64_t f = 0; f < nF; ++f) m->magSf[f] = vmag(vload(m->Sf + 3 * f));

    /* volumes and centroids (:219-265) */
    V3* moment = (V3*)calloc((size_t)nC, sizeof(V3));
    for (int64_t c = 0; c < nC; ++c) m->V[c] = 0.0;
    for (int64_t f = 0; f < nF; ++f) {
        const int64_t* vs = m->faceVerts + m->faceStart[f];
        const int64_t nv = m->faceStart[f + 1] - m->faceStart[f];
        const V3 v0 = PNT(vs[0]);
        double fv = 0;
        V3 fm = vzero();
        for (int64_t i = 1; i + 1 < nv; ++i) {
            const V3 a = vsub(PNT(vs[i]), v0);
            const V3 b = vsub(PNT(vs[i + 1]), v0);
            const double tV = vdot(v0, vcross(a, b)) / 6.0;
            fv += tV;
            fm = vadd(fm, vmul(vdiv(vadd(vadd(v0, PNT(vs[i])), PNT(vs[i + 1])), 4.0), tV));
        }
        V3 raw = vzero();
        for (int64_t i = 1; i + 1 < nv; ++i)
            raw = vadd(raw, vcross(vsub(PNT(vs[i]), v0), vsub(PNT(vs[i + 1]), v0)));
        const double s = vdot(raw, vload(m->Sf + 3 * f)) >= 0 ? 1.0 : -1.0;
        const int64_t o = m->owner[f];
        m->V[o] += s * fv;
        moment[o] = vadd(moment[o], vmul(fm, s));
        if (f < nI) {
            const int64_t nb = m->neighbour[f];
            m->V[nb] -= s * fv;
            moment[nb] = vsub(moment[nb], vmul(fm, s));
        }
    }
    int rc = 0;
    for (int64_t c = 0; c < nC; ++c) {
        vstore(m->C + 3 * c, vdiv(moment[c], m->V[c]));
        m->V[c] = fabs(m->V[c]);
        if (m->V[c] <= 0) rc = -1;
    }

    /* deltas (:267-287) */
    for (int64_t f = 0; f < nF; ++f) {
        const int64_t o = m->owner[f];
        V3 d;
        if (f < nI) d = vsub(vload(m->C + 3 * m->neighbour[f]), vload(m->C + 3 * o));
        else d = vsub(vload(m->Cf + 3 * f), vload(m->C + 3 * o));
        vstore(m->dvec + 3 * f, d);
        const double dm = vmag(d);
        if (dm <= 0) rc = -1;
        m->invDist[f] = 1.0 / dm;
    }
#undef PNT
    free(est); free(cnt); free(moment);
    return rc;
}
