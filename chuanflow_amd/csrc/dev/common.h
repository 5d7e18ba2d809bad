// common.h -- device helpers shared by the CDNA4 kernels.
//
// Wavefronts are 64 lanes (gfx950).  All reductions here are deterministic:
// a fixed butterfly inside the wavefront, a fixed order across the waves of a
// workgroup, and a fixed order across workgroups in the finalising
// workgroup, so every run of a solve produces the same bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace cfx
{

constexpr int kWave = 64;

struct HipError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

#define CFX_HIP(call)                                                                        \
    do {                                                                                     \
        const hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                                \
            throw ::cfx::HipError(std::string(#call) + ": " + hipGetErrorString(e_) + " (" + \
                                  __FILE__ + ":" + std::to_string(__LINE__) + ")");          \
    } while (0)

#define CFX_CHECK_LAUNCH() CFX_HIP(hipGetLastError())

// ------------------------------------------------------------ wave level
__device__ __forceinline__ double waveSum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ float waveSum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
// max that drops NaN operands (IEEE maxNum): matches std::max(acc, x) with a
// NaN x, which keeps acc.
__device__ __forceinline__ double waveMax(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ int waveOr(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, kWave);
    return v;
}

// ------------------------------------------------------- workgroup level
// Result valid in thread 0.  `lds` needs blockDim.x / 64 entries.
template <typename T>
__device__ __forceinline__ T blockSum(T v, T* lds)
{
    v = waveSum(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) lds[w] = v;
    __syncthreads();
    T s = T(0);
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / kWave); ++i) s += lds[i];
    __syncthreads();
    return s;
}

__device__ __forceinline__ double blockMax(double v, double* lds)
{
    v = waveMax(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / kWave); ++i) s = fmax(s, lds[i]);
    __syncthreads();
    return s;
}

__device__ __forceinline__ int blockOr(int v, int* lds)
{
    v = waveOr(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) lds[w] = v;
    __syncthreads();
    int s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / kWave); ++i) s |= lds[i];
    __syncthreads();
    return s;
}

// --------------------------------------------------- grid finalisation
// Each workgroup publishes its partials, then takes a ticket; the workgroup
// that takes the last ticket sees every partial (release/acquire at agent
// scope, which spans all XCDs' L2s) and finalises.  It resets the ticket so
// the same counter serves the next launch (also inside a replayed graph).
__device__ __forceinline__ bool lastBlockDone(unsigned int* ticket, int* ldsFlag)
{
    __syncthreads();
    if (threadIdx.x == 0)
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // partials visible device-wide
        const unsigned int t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        const bool last = (t == gridDim.x * gridDim.y - 1);
        if (last)
        {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        *ldsFlag = last ? 1 : 0;
    }
    __syncthreads();
    return *ldsFlag != 0;
}

// Agent-scope loads for partials written by other workgroups in this launch.
__device__ __forceinline__ double loadAgent(const double* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int loadAgent(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum partials[0..n) in a fixed order by one workgroup; result in thread 0.
__device__ __forceinline__ double sumPartials(const double* partials, int n, double* lds)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += loadAgent(partials + i);
    return blockSum(s, lds);
}
__device__ __forceinline__ double maxPartials(const double* partials, int n, double* lds)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s = fmax(s, loadAgent(partials + i));
    return blockMax(s, lds);
}
__device__ __forceinline__ int orPartials(const int* partials, int n, int* lds)
{
    int s = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s |= loadAgent(partials + i);
    return blockOr(s, lds);
}

inline int divUp(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

// Consecutive workgroup ids are dealt round-robin to the 8 XCDs of the
// device; workgroup h of G works on chunk xcdRemap(h, G), so each XCD walks
// one contiguous range of chunks and its own L2 holds the rows that range
// gathers (neighbours of RCM-ordered rows are close in index).
__device__ __forceinline__ int xcdRemap(int h, int G)
{
    constexpr int X = 8;
    const int q = G / X, r = G % X;
    const int x = h % X, idx = h / X;
    return (x < r) ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

// CSR-stream staging of one row block's products: lds[k] = val[k0 + k] *
// x[col[k0 + k]] for k < nnz (as TL), lane t taking k = t, t + B, ... Each lane
// has four nonzeros' value, column and x loads in flight per step; every
// product is the one of the plain loop.
template <typename TL, typename TV, typename TX>
__device__ __forceinline__ void stageProducts(TL* lds, const TV* __restrict__ val, const int* __restrict__ col,
                                              const TX* __restrict__ x, int k0, int nnz)
{
    const int bd = static_cast<int>(blockDim.x);
    int k = static_cast<int>(threadIdx.x);
    for (; k + 3 * bd < nnz; k += 4 * bd)
    {
        TV v[4];
        int c[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            v[q] = val[k0 + k + q * bd];
            c[q] = col[k0 + k + q * bd];
        }
        TX xv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) xv[q] = x[c[q]];
#pragma unroll
        for (int q = 0; q < 4; ++q) lds[k + q * bd] = static_cast<TL>(v[q]) * static_cast<TL>(xv[q]);
    }
    for (; k < nnz; k += bd) lds[k] = static_cast<TL>(val[k0 + k]) * static_cast<TL>(x[col[k0 + k]]);
}

} // namespace cfx
