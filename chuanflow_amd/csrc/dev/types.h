// types.h -- device-resident objects of the MI355X hot path.
//
// Layout in HBM (one GPU, everything resident for the whole solve):
//  * cell arrays SoA, [n] per component;
//  * face arrays SoA in the mesh's face order (RCM-renumbered meshes keep
//    internal faces sorted by (owner, neighbour), so face streams and the
//    cell gathers they feed are coalesced);
//  * the matrix pattern is CSR with int32 columns, rows sorted by column
//    (the reference SparseMatrix layout), plus face->slot maps for assembly;
//  * CSR-stream row blocks: each workgroup owns a contiguous row range whose
//    nonzeros fit in LDS, so the val/col streams are read fully coalesced
//    and each row is then summed from LDS in column order.
#pragma once

#include "common.h"

#include <memory>
#include <string>
#include <unordered_map>
#include <cstdint>
#include <vector>

namespace cfx
{

constexpr int kSpmvBlock = 256;      // threads = max rows per CSR-stream block
constexpr int kSpmvLds = 2048;       // max nonzeros staged per block
constexpr int kVecBlock = 256;
constexpr int kMaxPartials = 4096;   // grid size cap for reduction kernels

// Device allocation that frees itself.
struct DevBuf
{
    void* p = nullptr;
    std::size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(std::size_t b) { alloc(b); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(std::size_t b)
    {
        release();
        if (b == 0) b = 256;
        CFX_HIP(hipMalloc(&p, b));
        bytes = b;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

template <typename T>
DevBuf upload(const std::vector<T>& h, hipStream_t s)
{
    DevBuf b(h.size() * sizeof(T));
    if (!h.empty()) CFX_HIP(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return b;
}

// Scalars of one Krylov solve, device-resident so the iteration never waits
// on the host; the host polls `done` between batches of iterations.
struct SolveState
{
    double rz, pAp, alpha, beta, rnorm, bnorm, initialResidual, tol;
    double rho, omega, denom, tt, ts;  // BiCGStab
    double sum0;                       // scratch reduction result
    int iter, maxIter, done, converged, breakdown, nan;
    unsigned int ticket[8];
};

// Non-owning view of a CSR pattern on the device.
struct CsrView
{
    int n = 0, nnz = 0, nRowBlocks = 0;
    const int* rowPtr = nullptr;
    const int* col = nullptr;
    const int* diagSlot = nullptr;
    const int* rowBlocks = nullptr;
    // each column less the first row of its row block, as int16 (null unless every offset fits): the
    // CSR-stream staging reads 2 bytes per nonzero instead of 4
    const short* col16 = nullptr;
};

// A pattern's CSR-stream row blocks cut into slices of 64 consecutive rows
// (one wavefront), each slice stored slot-major: slot j of every row of the
// slice together, padded to the slice's longest row.  A wavefront's loads of
// one slot are contiguous, and its gathers of x fall on the neighbours of 64
// consecutive rows.  Columns are int16 offsets from the row block's first row;
// padding slots carry kSlicePad.
constexpr short kSlicePad = -32768;
struct SliceView
{
    const int* slicePtr = nullptr;    // [slices + 1] first slot of each slice (multiples of 64)
    const int* blockSlice = nullptr;  // [nRowBlocks + 1] first slice of each row block
    const short* col = nullptr;       // [slots]
};
// readable slots past the last slice (padding) in every slot array, for the
// unconditional loads of sliceRowSum
constexpr int kSliceTail = 8 * 64;

// Row sum of y = M v for this thread's row of row block blk, over the block's
// slices: the wave of threads 64 w .. 64 w + 63 sums the rows of the block's
// slice w, slot j of a row being its j-th nonzero, so each sum runs in column
// order from zero with the products TL(val) * TL(v), as the CSR-stream sum over
// staged products does (bit for bit); padding slots are skipped by a select.
// Eight slots in flight per lane, every load issued unconditionally and only
// the use masked.
template <typename TL, typename TV, typename TX>
__device__ __forceinline__ bool sliceRowSum(const CsrView& A, const SliceView& S, const TV* __restrict__ valM,
                                            const TX* __restrict__ v, int blk, int& row, TL& sum)
{
    const int r0 = A.rowBlocks[blk], r1 = A.rowBlocks[blk + 1];
    row = r0 + static_cast<int>(threadIdx.x);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    TL s = TL(0);
    if (r0 + wave * kWave < r1)
    {
        const int sl = S.blockSlice[blk] + wave;
        const int p0 = S.slicePtr[sl], w = (S.slicePtr[sl + 1] - p0) / kWave;  // the slice's width
        for (int j = 0; j < w; j += 8)
        {
            const int k = p0 + j * kWave + static_cast<int>(threadIdx.x & (kWave - 1));
            TV a[8];
            TX xv[8];
            int c[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
            {
                a[u] = valM[k + u * kWave];
                const int cu = S.col[k + u * kWave];
                c[u] = j + u < w ? cu : static_cast<int>(kSlicePad);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) xv[u] = v[r0 + (c[u] == kSlicePad ? 0 : c[u])];
#pragma unroll
            for (int u = 0; u < 8; ++u)
            {
                const TL t = s + static_cast<TL>(a[u]) * static_cast<TL>(xv[u]);
                s = c[u] != kSlicePad ? t : s;
            }
        }
    }
    sum = s;
    return row < r1;
}

// The same for a row block starting at row r0 of a pattern, through its int16
// column offsets where it has them (the same columns, so the same products).
template <typename TL, typename TV, typename TX>
__device__ __forceinline__ void stageProducts(TL* lds, const TV* __restrict__ val, const CsrView& A,
                                              const TX* __restrict__ x, int k0, int r0, int nnz)
{
    if (!A.col16)
    {
        stageProducts(lds, val, A.col, x, k0, nnz);
        return;
    }
    const short* __restrict__ c16 = A.col16;
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
            c[q] = r0 + c16[k0 + k + q * bd];
        }
        TX xv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) xv[q] = x[c[q]];
#pragma unroll
        for (int q = 0; q < 4; ++q) lds[k + q * bd] = static_cast<TL>(v[q]) * static_cast<TL>(xv[q]);
    }
    for (; k < nnz; k += bd) lds[k] = static_cast<TL>(val[k0 + k]) * static_cast<TL>(x[r0 + c16[k0 + k]]);
}


// Owning CSR pattern (int32 columns, rows sorted by column).
struct DevPattern
{
    int n = 0, nnz = 0, nRowBlocks = 0;
    std::uint64_t uid = 0;  // unique per build (caches keyed on a pattern compare this, not the address)
    DevBuf rowPtr, col, diagSlot, rowBlocks, col16;
    std::vector<int> hRowPtr, hCol;
    CsrView view() const
    {
        return {n,           nnz, nRowBlocks, rowPtr.as<int>(), col.as<int>(), diagSlot.as<int>(), rowBlocks.as<int>(),
                col16.p ? col16.as<short>() : nullptr};
    }
    // blocks: CSR-stream row blocks to use (empty: buildRowBlocks)
    void build(std::vector<int> rowPtrH, std::vector<int> colH, hipStream_t s, std::vector<int> blocks = {});
};

// Builds rowBlocks for CSR-stream from a host row pointer: each block holds
// at most kSpmvBlock rows and kSpmvLds nonzeros (a longer row gets a block
// of its own and is summed directly from global memory).
std::vector<int> buildRowBlocks(const std::vector<int>& rowPtr);
// The same for rows numbered group by group (group[i] non-decreasing): a
// block ends only where the group changes, so each block holds whole groups
// (at most kSpmvBlock rows; a group whose rows exceed kSpmvLds nonzeros gets
// a block of its own).  Empty when a group has more than kSpmvBlock rows.
std::vector<int> buildRowBlocksGrouped(const std::vector<int>& rowPtr, const std::vector<int>& group);

class Context;

// Mesh on the device: face streams, cell->face gathers, geometry-derived
// face coefficients and the matrix pattern with face->slot maps.
struct DevMesh
{
    Context* ctx = nullptr;
    int n = 0, nI = 0, nF = 0, nB = 0;
    int nPatches = 0;
    std::vector<long long> patchStart, patchSize;
    DevBuf owner, neighbour;               // int[nF], int[nI]
    DevBuf cellFaceStart, cellFace;        // int[n+1], int[2nI+nB]: face<<1 | isNeighbour
    DevBuf Sf, Cf, d, C;                   // SoA double[3][nF] / [3][n]
    DevBuf V, magSf, localScale;           // double[n], [nF], [n]
    DevBuf ow;                             // double[nI] owner weight
    DevBuf sfd, d2;                        // double[nF]: Sf.d and d.d
    DevBuf skew;                           // double[3][nI]
    DevBuf kf;                             // double[3][nF]: Sf - dh (Sf.dh)
    DevBuf faceGeom;                       // reserved
    // matrix pattern (face graph + diagonal)
    DevPattern pattern;
    DevBuf slotU, slotL;                   // int[nI]
    std::vector<double> hV;                // host copy of volumes
    // the SIMPLE/PISO drivers' work buffers, kept between calls (allocated by the first)
    std::shared_ptr<void> driverWork;
};

// Context: one HIP stream, one device, scratch for reductions.
class Context
{
public:
    explicit Context(int device);
    ~Context();
    int device() const { return device_; }
    hipStream_t stream() const { return stream_; }
    void sync() { CFX_HIP(hipStreamSynchronize(stream_)); }
    std::string lastError;

    // reduction scratch: partials + tickets (zeroed once)
    double* partials(int slot) { return partials_.as<double>() + slot * kMaxPartials; }
    int* ipartials() { return ipartials_.as<int>(); }
    unsigned int* tickets() { return tickets_.as<unsigned int>(); }
    SolveState* hostState() { return hostState_; }

    int numCUs() const { return numCUs_; }
    std::string deviceName;

private:
    int device_ = 0;
    int numCUs_ = 0;
    hipStream_t stream_ = nullptr;
    DevBuf partials_, ipartials_, tickets_;
    SolveState* hostState_ = nullptr;  // pinned
};

// A preconditioner of flexible CG in FP32: z = M r on device vectors and,
// when partZQ is non-null, Gs per-workgroup partials of z.q (Gs workgroups
// of rbPerWg row blocks, the grid of the CG kernels).  `done` (device, may
// be null) skips the work once the iteration has stopped.
class Preconditioner
{
public:
    virtual ~Preconditioner() = default;
    virtual int rows() const = 0;
    virtual std::uintptr_t id() const = 0;        // unique per set of device buffers
    virtual const float* val32() const = 0;       // FP32 copy of the matrix values
    virtual const float* invD32() const = 0;      // FP32 inverse diagonal
    virtual void apply(const float* r, float* z, const float* q, double* partZQ, int Gs, int rbPerWg,
                       const int* done) = 0;
    // measurement: while on, the finest-level post-smoothing launches carry HIP start and stop events (no
    // graph capture then); the result is the average device time of those that ran and their count
    virtual void setProbe(bool) {}
    virtual bool probeResult(double& avgMs, int& count) { return false; }
    // the matrix in 64-row slices with its FP64 values in slot order, where the preconditioner holds them
    // for exactly the values `val` (the pointer its setup read); false: none
    virtual bool fineSlices(const double* val, SliceView& S, const double*& sliceVal64) const { return false; }
};

} // namespace cfx
