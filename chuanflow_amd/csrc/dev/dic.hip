// dic.hip -- level-scheduled DIC factorisation and substitutions.
#include "dic.h"

#include <algorithm>
#include <atomic>

namespace cfx
{
namespace
{

constexpr int B = 256;

// D_ii for the rows of one level (lower neighbours are in earlier levels)
__global__ void k_dic_factor(int cnt, const int* __restrict__ rows, CsrView A, const double* __restrict__ val,
                             double* __restrict__ d, float* __restrict__ d32)
{
    const int t = blockIdx.x * B + threadIdx.x;
    if (t >= cnt) return;
    const int i = rows[t];
    double aii = 0.0, s = 0.0;
    for (int k = A.rowPtr[i]; k < A.rowPtr[i + 1]; ++k)
    {
        const int j = A.col[k];
        if (j < i)
            s += val[k] * val[k] / d[j];
        else if (j == i)
            aii = val[k];
    }
    double di = aii - s;
    if (!(di > 1e-14 * fabs(aii))) di = aii;  // breakdown guard
    d[i] = di;
    d32[i] = static_cast<float>(di);
}

// forward: v_i = (r_i - sum_{j<i} a_ij v_j) / D_ii
__global__ void k_dic_fwd(int cnt, const int* __restrict__ rows, CsrView A, const float* __restrict__ val,
                          const float* __restrict__ d, const float* __restrict__ r, float* __restrict__ v,
                          const int* done)
{
    if (done && *done) return;
    const int t = blockIdx.x * B + threadIdx.x;
    if (t >= cnt) return;
    const int i = rows[t];
    float s = r[i];
    for (int k = A.rowPtr[i]; k < A.rowPtr[i + 1]; ++k)
    {
        const int j = A.col[k];
        if (j >= i) break;  // columns sorted: the lower part comes first
        s -= val[k] * v[j];
    }
    v[i] = s / d[i];
}

// backward (in place on v): z_i = v_i - (sum_{j>i} a_ij z_j) / D_ii
__global__ void k_dic_bwd(int cnt, const int* __restrict__ rows, CsrView A, const float* __restrict__ val,
                          const float* __restrict__ d, float* __restrict__ z, const int* done)
{
    if (done && *done) return;
    const int t = blockIdx.x * B + threadIdx.x;
    if (t >= cnt) return;
    const int i = rows[t];
    float s = 0.0f;
    for (int k = A.rowPtr[i + 1] - 1; k >= A.rowPtr[i]; --k)
    {
        const int j = A.col[k];
        if (j <= i) break;
        s += val[k] * z[j];
    }
    z[i] -= s / d[i];
}

__global__ void k_dic_cast(CsrView A, const double* __restrict__ val, float* __restrict__ v32, float* __restrict__ invA)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= A.n) return;
    for (int k = A.rowPtr[i]; k < A.rowPtr[i + 1]; ++k) v32[k] = static_cast<float>(val[k]);
    const double a = A.diagSlot[i] >= 0 ? val[A.diagSlot[i]] : 0.0;
    invA[i] = static_cast<float>(fabs(a) > 1e-300 ? 1.0 / a : 1.0);
}

// Gs partials of z.q (strided chunks)
__global__ void k_dic_dot(int n, const float* __restrict__ z, const float* __restrict__ q, double* __restrict__ part,
                          const int* done)
{
    __shared__ double red[B / kWave];
    if (done && *done) return;
    double v = 0.0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B)
        v += static_cast<double>(z[i]) * static_cast<double>(q[i]);
    v = blockSum(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

} // namespace

Dic::Dic(Context& ctx, const DevPattern& pat) : ctx_(ctx), pat_(pat)
{
    static std::atomic<std::uintptr_t> counter{1u << 20};
    id_ = ++counter;
    const int n = pat.n;
    const std::vector<int>& rp = pat.hRowPtr;
    const std::vector<int>& cl = pat.hCol;
    std::vector<int> level(n, 0);
    int nl = 0;
    for (int i = 0; i < n; ++i)
    {
        int L = 0;
        for (int k = rp[i]; k < rp[i + 1] && cl[k] < i; ++k) L = std::max(L, level[cl[k]] + 1);
        level[i] = L;
        nl = std::max(nl, L + 1);
    }
    levelPtr_.assign(nl + 1, 0);
    for (int i = 0; i < n; ++i) ++levelPtr_[level[i] + 1];
    for (int L = 0; L < nl; ++L) levelPtr_[L + 1] += levelPtr_[L];
    std::vector<int> rows(n), fill(levelPtr_.begin(), levelPtr_.end() - 1);
    for (int i = 0; i < n; ++i) rows[fill[level[i]]++] = i;
    hipStream_t s = ctx.stream();
    levelRows_ = upload(rows, s);
    d64_.alloc(sizeof(double) * n);
    d32_.alloc(sizeof(float) * n);
    val32_.alloc(sizeof(float) * std::max(1, pat.nnz));
    invA32_.alloc(sizeof(float) * n);
    ctx.sync();
}

void Dic::setMatrix(const double* val64)
{
    hipStream_t s = ctx_.stream();
    const CsrView A = pat_.view();
    k_dic_cast<<<divUp(A.n, B), B, 0, s>>>(A, val64, val32_.as<float>(), invA32_.as<float>());
    const int* rows = levelRows_.as<int>();
    for (int L = 0; L < levels(); ++L)
    {
        const int cnt = levelPtr_[L + 1] - levelPtr_[L];
        k_dic_factor<<<divUp(cnt, B), B, 0, s>>>(cnt, rows + levelPtr_[L], A, val64, d64_.as<double>(),
                                                 d32_.as<float>());
    }
    CFX_CHECK_LAUNCH();
}

void Dic::apply(const float* r, float* z, const float* q, double* partZQ, int Gs, int /*rbPerWg*/, const int* done)
{
    hipStream_t s = ctx_.stream();
    const CsrView A = pat_.view();
    const int* rows = levelRows_.as<int>();
    for (int L = 0; L < levels(); ++L)
    {
        const int cnt = levelPtr_[L + 1] - levelPtr_[L];
        k_dic_fwd<<<divUp(cnt, B), B, 0, s>>>(cnt, rows + levelPtr_[L], A, val32_.as<float>(), d32_.as<float>(), r, z,
                                              done);
    }
    for (int L = levels() - 1; L >= 0; --L)
    {
        const int cnt = levelPtr_[L + 1] - levelPtr_[L];
        k_dic_bwd<<<divUp(cnt, B), B, 0, s>>>(cnt, rows + levelPtr_[L], A, val32_.as<float>(), d32_.as<float>(), z,
                                              done);
    }
    if (partZQ) k_dic_dot<<<Gs, B, 0, s>>>(A.n, z, q, partZQ, done);
    CFX_CHECK_LAUNCH();
}

} // namespace cfx
