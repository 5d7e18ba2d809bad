// noalloc.h -- FP64 Jacobi-PCG whose scratch is a file-scope device array
// (no device allocation per call); see noalloc.hip.
#pragma once

#include "krylov.h"

#include <cstddef>

namespace cfx
{

// largest system the scratch holds
constexpr int kNoAllocMaxRows = 1 << 19;
// bytes of the scratch array on each GPU: row blocks and diagonal slots (int), inverse diagonal, residual,
// direction and product (double) per row, the reduction partials and the state, each 256-byte aligned
constexpr std::size_t noallocAlign(std::size_t b) { return (b + 255) / 256 * 256; }
constexpr std::size_t kNoAllocScratchBytes =
    noallocAlign(sizeof(int) * (static_cast<std::size_t>(kNoAllocMaxRows) + 1)) +
    noallocAlign(sizeof(int) * kNoAllocMaxRows) + 4 * noallocAlign(sizeof(double) * kNoAllocMaxRows) +
    noallocAlign(sizeof(double) * kPcgPartials) + noallocAlign(kPcgStateBytes);

// Takes the array's address on `device` (which must be current); once per device.
void noallocSetup(int device);

// solveCG on a device CSR system (int32 rowPtr[n+1], col[nnz], FP64 val[nnz]; b, x device vectors, x the
// initial guess and the result) with every temporary in the scratch array: no allocation.
SolveReport pcgNoAlloc(Context& ctx, int n, int nnz, const int* rowPtr, const int* col, const double* val,
                       const double* b, double* x, double tol, int maxIter);

} // namespace cfx
