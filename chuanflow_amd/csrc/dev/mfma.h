// mfma.h -- 32x32 matrix-core tiles for one wavefront (CDNA3/CDNA4).
//
// acc(32x32, FP32) += A(32xK) * B(Kx32) for one 64-lane wavefront, with the
// operands read from LDS through accessor functors:
//   a(i, k) -> element of A (row i, column k), b(k, j) -> element of B.
// FP16 storage uses v_mfma_f32_32x32x8_f16 (K = 8 per instruction: lane l
// supplies A[l % 32][4 (l / 32) + 0..3] and B[4 (l / 32) + 0..3][l % 32]);
// FP32 storage uses v_mfma_f32_32x32x2_f32 (K = 2: lane l supplies
// A[l % 32][l / 32] and B[l / 32][l % 32]).  The accumulator element r of
// lane l is C[8 (r / 4) + 4 (l / 32) + (r % 4)][l % 32].
#pragma once

#include <hip/hip_runtime.h>

namespace cfx
{

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int mfmaRow(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
__device__ __forceinline__ int mfmaCol(int lane) { return lane & 31; }

template <typename T> struct Mfma;

template <> struct Mfma<_Float16>
{
    static constexpr int kStep = 8;
    // one instruction over columns k0 .. k0+7 of A / rows of B
    template <typename FA, typename FB>
    __device__ static __forceinline__ f32x16 step(f32x16 c, int k0, FA a, FB b)
    {
        const int lane = threadIdx.x & 63;
        const int i = lane & 31, kk = k0 + 4 * (lane >> 5);
        f16x4 av, bv;
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            av[j] = a(i, kk + j);
            bv[j] = b(kk + j, i);
        }
        return __builtin_amdgcn_mfma_f32_32x32x8f16(av, bv, c, 0, 0, 0);
    }
};

template <> struct Mfma<float>
{
    static constexpr int kStep = 2;
    template <typename FA, typename FB>
    __device__ static __forceinline__ f32x16 step(f32x16 c, int k0, FA a, FB b)
    {
        const int lane = threadIdx.x & 63;
        const int i = lane & 31, kk = k0 + (lane >> 5);
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a(i, kk), b(kk, i), c, 0, 0, 0);
    }
};

// acc += A(32xK) B(Kx32), K a multiple of the instruction step
template <typename T, int K, typename FA, typename FB>
__device__ __forceinline__ f32x16 mfmaTile(f32x16 acc, FA a, FB b)
{
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += Mfma<T>::kStep) acc = Mfma<T>::step(acc, k0, a, b);
    return acc;
}

__device__ __forceinline__ f32x16 zero16()
{
    f32x16 z;
#pragma unroll
    for (int r = 0; r < 16; ++r) z[r] = 0.0f;
    return z;
}

} // namespace cfx
