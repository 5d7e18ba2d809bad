// noalloc.hip -- FP64 Jacobi-PCG that allocates nothing.
//
// For callers that cannot allocate device memory (code under stream capture,
// tight serving loops, harnesses that forbid hipMalloc): the solve of
// Krylov::cg (same kernels, same host loop) whose scratch -- CSR-stream row
// blocks, diagonal slots, inverse diagonal, residual, direction and product
// vectors, reduction partials and the iteration state -- lives in one
// file-scope device array, `scratch`, sized for systems of at most
// kNoAllocMaxRows rows.  Each GPU holds its own copy until the process ends;
// its address is taken once per device, with that device current, when a
// context is created, and host code uses only that address.  Every call on
// a GPU shares the array: calls that overlap on one GPU race.  The row
// blocks are a uniform partition (at most 256 rows, about 2048 nonzeros each)
// formed on the device, so the dot products are summed in a different
// grouping than a pattern built by the library: results agree with
// cfx_csr_solve_device to rounding, not bit for bit.
#include "noalloc.h"

#include "krylov.h"

#include <chrono>
#include <mutex>
#include <stdexcept>
#include <string>

namespace cfx
{
namespace
{
constexpr std::size_t align256(std::size_t b) { return noallocAlign(b); }
constexpr std::size_t kMaxN = static_cast<std::size_t>(kNoAllocMaxRows);
// layout of the array (byte offsets)
constexpr std::size_t kRowBlocksOff = 0;
constexpr std::size_t kDiagSlotOff = kRowBlocksOff + align256(sizeof(int) * (kMaxN + 1));
constexpr std::size_t kInvDOff = kDiagSlotOff + align256(sizeof(int) * kMaxN);
constexpr std::size_t kROff = kInvDOff + align256(sizeof(double) * kMaxN);
constexpr std::size_t kPOff = kROff + align256(sizeof(double) * kMaxN);
constexpr std::size_t kQOff = kPOff + align256(sizeof(double) * kMaxN);
constexpr std::size_t kPartOff = kQOff + align256(sizeof(double) * kMaxN);
constexpr std::size_t kStateOff = kPartOff + align256(sizeof(double) * kPcgPartials);
constexpr std::size_t kScratchBytes = kStateOff + align256(kPcgStateBytes);
static_assert(kScratchBytes == kNoAllocScratchBytes, "the scratch size in noalloc.h is stale");
} // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace
{
constexpr int kMaxDevices = 64;
std::mutex gMutex;
unsigned char* gScratch[kMaxDevices] = {};  // the array's address on each device

// rows [b R, (b + 1) R) form row block b
__global__ void k_uniform_blocks(int n, int R, int nb, int* __restrict__ rb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= nb) rb[i] = i * R < n ? i * R : n;
}

// slot of each row's diagonal entry, -1 if the row has none
__global__ void k_diag_slots(int n, const int* __restrict__ rowPtr, const int* __restrict__ col, int* __restrict__ ds)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int slot = -1;
    for (int k = rowPtr[i]; k < rowPtr[i + 1]; ++k)
        if (col[k] == i)
        {
            slot = k;
            break;
        }
    ds[i] = slot;
}
} // namespace

void noallocSetup(int device)
{
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("no-allocation solve: device id out of range");
    std::lock_guard<std::mutex> lock(gMutex);
    if (gScratch[device]) return;
    void* p = nullptr;
    CFX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    gScratch[device] = static_cast<unsigned char*>(p);
}

SolveReport pcgNoAlloc(Context& ctx, int n, int nnz, const int* rowPtr, const int* col, const double* val,
                       const double* b, double* x, double tol, int maxIter)
{
    if (n < 1 || n > kNoAllocMaxRows)
        throw std::invalid_argument("no-allocation solve: " + std::to_string(n) + " rows; the scratch holds 1 to " +
                                    std::to_string(kNoAllocMaxRows));
    if (nnz < 1) throw std::invalid_argument("no-allocation solve: the matrix has no nonzero");
    if (!rowPtr || !col || !val || !b || !x) throw std::invalid_argument("no-allocation solve: null array");
    if (maxIter < 0) throw std::invalid_argument("no-allocation solve: maxIter < 0");
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(gMutex);
        base = gScratch[ctx.device()];
    }
    if (!base) throw std::runtime_error("no-allocation solve: scratch not set up on this device");
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = ctx.stream();
    // rows per block: at most 256, about 2048 nonzeros (the LDS staging size) on average
    int R = 256;
    while (R > 1 && static_cast<long long>(R) * nnz > 2048LL * n) R /= 2;
    const int nb = divUp(n, R);
    int* rb = reinterpret_cast<int*>(base + kRowBlocksOff);
    int* ds = reinterpret_cast<int*>(base + kDiagSlotOff);
    double* invD = reinterpret_cast<double*>(base + kInvDOff);
    double* r = reinterpret_cast<double*>(base + kROff);
    k_uniform_blocks<<<divUp(nb + 1, 256), 256, 0, s>>>(n, R, nb, rb);
    k_diag_slots<<<divUp(n, 256), 256, 0, s>>>(n, rowPtr, col, ds);
    CFX_CHECK_LAUNCH();
    CsrView A;
    A.n = n;
    A.nnz = nnz;
    A.nRowBlocks = nb;
    A.rowPtr = rowPtr;
    A.col = col;
    A.diagSlot = ds;
    A.rowBlocks = rb;
    computeInvD(ctx, A, val, invD, false);
    const PcgWork w{reinterpret_cast<double*>(base + kPartOff), base + kStateOff, base + kPOff, base + kQOff, nullptr};
    SolveReport rep = pcgSolveWith(ctx, A, val, invD, b, x, tol, maxIter, w, r);
    rep.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rep;
}

} // namespace cfx
