// mg.h -- aggregation multigrid V-cycle used as the (learned) preconditioner
// of flexible CG.
//
// Hierarchy (built once per matrix pattern, on the host): greedy aggregation
// of the face graph (a root and its whole neighbourhood, then attachment of
// leftovers to an adjacent aggregate, then leftover groups), repeated until the
// coarsest level is small; coarse operators are the Galerkin products
// P^T A P with the piecewise-constant prolongation P, evaluated on the device
// as fixed-order sums of fine nonzeros (no atomics), so a new matrix on the
// same pattern only costs a few streaming kernels.
//
// V-cycle in FP32: on level l, x = Wpre b, r = b - A x (one SpMV on the
// column-scaled values A Wpre), restrict, recurse, x += S P e, then
// z = x + Wpost (b - A x).  The coarsest level (at most 128 rows) is solved
// exactly with a dense inverse formed in LDS at setup (a singular, pure
// Neumann coarsest operator A is inverted as A + gamma 1 1^T, which equals
// the pseudo-inverse on zero-mean right-hand sides); a larger coarsest level
// runs damped Jacobi sweeps in one workgroup.  Wpre = diag(a) D^-1, Wpost = diag(c) D^-1, the
// prolongation scale S = diag(s) and the restriction scale R = diag(t)
// (restrict P^T R r) are per-row: on the finest level they may come from the
// learned network, on coarse levels they are constants (a = c, t = 1).  With
// a = c and s = t the cycle is a symmetric operator; with a = 1, s = 0, c = 0
// it is exactly z = D^-1 b.
//
// K-cycle (Notay's Krylov-accelerated cycle): on the coarse levels
// 1..kcycleDepth the coarse correction is two steps of flexible CG on that
// level, preconditioned by the cycle of that level, instead of one cycle:
// c1 = M b, v1 = A c1, r2 = b - (c1.b / c1.v1) v1, c2 = M r2, v2 = A c2 and
// e = k1 c1 + k2 c2 with the FCG coefficients.  It keeps the convergence of
// the aggregation hierarchy nearly independent of its depth, and makes the
// preconditioner nonlinear, which flexible CG allows.
#pragma once

#include "types.h"

#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

namespace cfx
{

struct MgParams
{
    // defaults tuned on a 128^3 jittered pressure system with the 4/2 sweeps (a grid over the weights, the
    // transfer scales and the sweeps: 8 flexible-CG iterations against 13 with a = 2/3, s = 1.8, t = 1 and
    // 3/2 sweeps)
    float omegaPre = 0.8f;            // finest-level constants (used where no per-row values are given)
    float scale = 1.8f;
    float omegaPost = 0.8f;
    float restrictScale = 1.8f;
    float coarseOmega = 0.9f;         // coarse levels: pre = post smoother weight
    float coarseScale = 3.0f;
    int coarseSweeps = 30;            // damped Jacobi sweeps on a coarsest level above 128 rows
    float coarsestOmega = 2.0f / 3.0f;
    int maxLevels = 10, coarseMax = 128;
    int kcycleDepth = 2;              // coarse levels 1..kcycleDepth use the K-cycle (0: V-cycle)
    int smoothFine = 4;               // smoothing sweeps before and after the coarse correction: finest level,
    int smoothCoarse = 2;             // level 1 of the K-cycle levels,
    int smoothDeep = 2;               // the K-cycle levels below level 1 (the levels below them take one)
    int fusion = 125;                 // bit 0: the small coarse levels' cycle in one workgroup;
                                      // bit 1: prolongation (and K-cycle combination) fused into the
                                      // post-smoothing of coarse levels of at most 65536 rows;
                                      // bit 2: that cycle (the tail) run by every workgroup of the
                                      // post-smoothing of the level above it, from LDS;
                                      // bit 3: the K-cycle's r2 = b - coef v1 formed in the second
                                      // visit's pre-smoothing;
                                      // bit 4: the K-cycle combination of a level of at most 256
                                      // workgroups formed in the prolongation to the level above;
                                      // bit 5: the restriction from a coarse level formed by its
                                      // pre-smoothing (rows numbered aggregate by aggregate);
                                      // bit 6: the finest level's smoothing sweeps, and the flexible
                                      // CG's product, on copies of its values in 64-row slices,
                                      // slot-major (where its pattern allows them)
};

// One level of the single-workgroup coarse tail (device pointers).
struct TailLevel
{
    int n = 0;
    const int *rowPtr = nullptr, *col = nullptr;
    const float *valS = nullptr, *val = nullptr, *wPre = nullptr, *wPost = nullptr, *sc = nullptr, *invD = nullptr;
    const int* agg = nullptr;                       // coarse row of each row (to the next level)
    const int *rRowPtr = nullptr, *rCol = nullptr;  // P^T pattern of that transfer
    const float* rval = nullptr;
    const float* dense = nullptr;                   // coarsest: FP32 inverse, or null (Jacobi sweeps)
};
constexpr int kTailLevels = 4;
constexpr int kTailLds = 12288;  // LDS floats of the tail: 3 per row of every tail level
struct TailArgs
{
    int count = 0;    // levels in the tail, the coarsest included
    int sweeps = 0;   // coarsest Jacobi sweeps and weight (no dense inverse)
    float omega = 0.0f;
    TailLevel L[kTailLevels];
};
// The arrays of a tail level, as indexed in the tail's pointer table.
enum TailArray
{
    kTRowPtr, kTCol, kTValS, kTVal, kTWPre, kTWPost, kTSc, kTInvD, kTAgg, kTRRowPtr, kTRCol, kTRval, kTDense,
    kTailArrays
};
// Which tail arrays a launch copies into LDS first (one coalesced pass over
// all of them, many loads in flight), and where: word offset per
// (level, array), -1 = read from global memory.
constexpr int kTailUpArena = 36864;  // LDS words of the fused tail: its vectors and the staged arrays
constexpr int kTailChunks = kTailLevels * kTailArrays + 1;  // the arrays and the tail's right-hand side
struct TailStage
{
    int chunks = 0;
    int total4 = 0;                    // float4 slots over all chunks
    int start4[kTailChunks + 1];       // first float4 slot of each chunk
    int words[kTailChunks];
    int dst[kTailChunks];              // LDS word offset (a multiple of 4)
    const void* src[kTailChunks];
    int slot[kTailLevels][kTailArrays];
};

struct MgLevel
{
    int n = 0, nnz = 0;
    DevPattern pat;
    DevBuf val64, val32, valS32;   // FP64 Galerkin values, FP32 copy, A Wpre (FP32)
    DevBuf invD32, wPre32, wPost32, s32, t32;  // D^-1, a D^-1, c D^-1, prolongation / restriction scale
    DevBuf b, x, r, e;             // FP32: rhs, iterate, residual, coarse-correction result
    DevBuf t2;                     // FP32: second residual / smoothing buffer (sweeps > 1)
    DevBuf c1, c2, v1, r2, part;   // K-cycle vectors and 5 x kPartMax dot partials
    int G = 1, rbPerWg = 1;        // SpMV grid of this level (G <= kPartMax)
    // transfer to level l+1
    DevBuf agg;                    // int[n]: coarse row of each row
    DevBuf aggStart, aggMem;       // members of each coarse row (ascending)
    DevPattern rpat;               // P^T as a CSR pattern (rows = coarse rows, columns = members)
    DevBuf rval;                   // its values: the restriction scale of each member (1 off the finest level)
    DevBuf gStart, gPerm;          // coarse nonzero c = sum of fine nonzeros gPerm[gStart[c]..gStart[c+1])
    DevBuf dense;                  // coarsest level: FP32 inverse [n][n] (symmetric)
    std::vector<int> hAgg;
    // a coarse level above the coarsest: rows numbered aggregate by aggregate and row blocks of whole
    // aggregates, so its pre-smoothing can form the coarse right-hand side itself
    bool groupedRows = false;
};

class Multigrid : public Preconditioner
{
public:
    // Builds the hierarchy from a fine pattern (host copies of rowPtr/col).
    Multigrid(Context& ctx, const DevPattern& fine, int maxLevels = 10, int coarseMax = 128);
    bool denseCoarsest() const { return levels() > 1 && lv_.back().n <= kDenseMax; }
    static constexpr int kDenseMax = 128;
    const MgParams& params() const { return p_; }
    // unique per constructed hierarchy (device buffers never move after construction) and per
    // launch structure (K-cycle depth, fusion, coarsest sweeps): a captured cycle is valid while it holds
    std::uintptr_t id() const override { return (id_ << 16) | (structVersion_ & 0xffffu); }
    // first level whose cycle (with all coarser levels) runs in one workgroup; -1: none
    int tailLevel() const { return tailFrom_; }
    // slots of the finest level's sliced layout (0: none)
    long long fineSlots() const { return sl_.ok ? sl_.slots : 0; }
    bool fineSlices(const double* val, SliceView& S, const double*& sliceVal64) const override
    {
        if (!sliced(0) || val != fineVal64_) return false;
        S = sl_.view();
        sliceVal64 = sl_.val64.as<double>();
        return true;
    }
    int rows() const override { return lv_[0].n; }
    const float* val32() const override { return lv_[0].val32.as<float>(); }
    const float* invD32() const override { return lv_[0].invD32.as<float>(); }
    int levels() const { return static_cast<int>(lv_.size()); }
    int rows(int l) const { return lv_[l].n; }
    int nnz(int l) const { return lv_[l].nnz; }
    const std::vector<int>& aggregation(int l) const { return lv_[l].hAgg; }
    // device vectors of a level (training and tests)
    MgLevel& level(int l) { return lv_[l]; }
    const float* fineVal32() const { return lv_[0].val32.as<float>(); }
    const float* fineInvD32() const { return lv_[0].invD32.as<float>(); }

    // Galerkin values of every level from the fine FP64 values, FP32 copies,
    // inverse diagonals and the scaled values A Wpre.  pre/scale/post/restr
    // (device, FP32, per fine row; each may be null) are the finest-level
    // a, s, c, t; null takes the constant from `p`.
    void setMatrix(const double* fineVal64, const float* pre, const float* scale, const float* post,
                   const float* restr, const MgParams& p);
    // The two halves of setMatrix: what depends on the matrix values (Galerkin values, coarse-level setup
    // with the constants, the coarsest inverse, the finest FP32 copy), and what depends on the finest-level
    // weights (its inverse diagonal and weights, column-scaled values and restriction values), which reads
    // the finest values given to the last setValues.  New weights on unchanged values need only the second.
    void setValues(const double* fineVal64, const MgParams& p);
    void setWeights(const float* pre, const float* scale, const float* post, const float* restr, const MgParams& p);
    // z = V-cycle(r) on the finest level.  When partZQ is non-null also writes
    // per-workgroup partials of z.q (grid Gs with rbPerWg row blocks each).
    void apply(const float* r, float* z, const float* q, double* partZQ, int Gs, int rbPerWg,
               const int* done) override;

    // Training objective of the finest-level weights.  Requires setMatrix
    // with pre = post = a, scale = restr = t.  The coarse correction Q is the
    // cycle's own below level 1 (the K-cycle on levels 1..kcycleDepth); the
    // adjoint takes it as the linear symmetric operator it approximates
    // (exact for kcycleDepth = 0).  For an error sample e
    // and z = M A e, L = 1 - (z.A e)^2 / ((e.A e)(z.A z)) = min over alpha of
    // |e - alpha z|_A^2 / |e|_A^2: the error left by one preconditioned step
    // with the best step length (invariant to the scale of M, as the Krylov
    // iteration is).  Returns L and adds dL/da, dL/dt (per fine row) into
    // ga, gt.
    double lossGrad(const float* e, const float* a, const float* t, float* ga, float* gt);
    // e <- (I - M A) e, `sweeps` times, then e / max|e| (smooth error samples)
    void smoothSample(float* e, int sweeps);
    void setProbe(bool on) override;
    bool probeResult(double& avgMs, int& count) override;
    // average device time of one finest-level post-smoothing sweep over `reps` back-to-back launches (ms)
    double timeFineSmooth(int reps);

    static constexpr int kPartMax = 2048;

private:
    Context& ctx_;
    const DevPattern& fine_;
    MgParams p_;
    std::uintptr_t id_ = 0;
    std::uintptr_t structVersion_ = 0;
    const double* fineVal64_ = nullptr;  // the finest values of the last setValues
    int tailFrom_ = -1;
    std::vector<MgLevel> lv_;  // lv_[0] uses the fine pattern (pat unused)
    // out = cycle of level l applied to in (in, out distinct from the level's work vectors)
    // kcB/kcV non-null: the K-cycle's second visit, whose right-hand side in = kcB - coef kcV is formed
    // by the pre-smoothing kernel (fusion bit 3) and stored to `in`
    void visit(int l, const float* in, float* out, const float* q, double* partZQ, int Gs, int rbPerWg,
               const int* done, const float* kcB = nullptr, const float* kcV = nullptr);
    // coarse correction of level l >= 1: L.e from L.b.  With deferCombine a
    // K-cycle level leaves e = k1 c1 + k2 c2 unformed (its partials hold the
    // coefficients) and returns true
    bool solveLevel(int l, const int* done, bool deferCombine = false);
    // the same for l >= tailLevel(), below the K-cycle levels: one single-workgroup launch
    bool useTail(int l) const;
    bool tailAt(int l) const;  // levels l..last make a tail (whatever the fusion bits)
    TailArgs tailArgs(int l) const;
    void launchTail(int l, const int* done);
    // fusion bit 2: the tail below level l fused into l's post-smoothing; false where it does not apply
    // res: l's residual after its pre-smoothing; restricted: l's pre-smoothing formed the tail's right-hand
    // side (fusion bit 5); out: l's first post-smoothing sweep
    bool launchTailUp(int l, const float* in, float* out, const float* res, const double* partZQ, const int* done,
                      bool restricted);
    // launches the post-smoothing z = x + W (b - A x) of level l (with the probe's events on the finest level)
    void launchUp(int l, const CsrView& A, int Gs, int rbPerWg, const float* x, const float* b, float* z,
                  const float* q, double* partZQ, const int* done);
    // the finest level in 64-row slices (fusion bit 6): its FP32 and column-scaled values in slot order, the
    // CSR slot of every slot (-1: padding) to form them; ok where the pattern's offsets fit int16 and the
    // padding stays under a quarter of the nonzeros
    struct FineSlices
    {
        bool ok = false;
        long long slots = 0;
        DevBuf slicePtr, blockSlice, col, idx, val, valS, val64;
        SliceView view() const { return {slicePtr.as<int>(), blockSlice.as<int>(), col.as<short>()}; }
    } sl_;
    void buildSlices();
    bool sliced(int l) const { return l == 0 && sl_.ok && (p_.fusion & 64); }
    bool probeOn_ = false;
    std::vector<hipEvent_t> probeEv_;  // start/stop pairs of the probed launches
    int probeUsed_ = 0;
    DevBuf tv_[14];          // finest-level FP32 scratch of lossGrad / smoothSample
    std::vector<DevBuf> tk_; // lossGrad: the sweeps' D^-1 (b - A x) (2 (smoothFine - 1) vectors)
    DevBuf tpart_;
};

// The multigrid preconditioner of one matrix pattern: the hierarchy (rebuilt
// when maxLevels / coarseMax change), its parameters, the caller's per-row
// finest-level weights (device pointers, read at each setup) and whether the
// Galerkin values match the current matrix values.
struct MgCache
{
    std::unique_ptr<Multigrid> mg;
    MgParams p;
    const float *pre = nullptr, *scale = nullptr, *post = nullptr, *restr = nullptr;
    // the caller's own weights (null: the constants); a learned solve points pre..restr at the network's
    // outputs, a plain multigrid solve back at these
    const float *userPre = nullptr, *userScale = nullptr, *userPost = nullptr, *userRestr = nullptr;
    bool valid = false;         // the hierarchy holds the current matrix values (and weights)
    bool weightsValid = false;  // it holds the current finest-level weights

    // point the finest-level weights at the arrays (pre a, scale s, post c, restrict t); a change of
    // arrays invalidates the weights (new contents of the same arrays: the caller clears weightsValid)
    void useWeights(const float* a, const float* s, const float* c, const float* t)
    {
        if (a != pre || s != scale || c != post || t != restr) weightsValid = false;
        pre = a;
        scale = s;
        post = c;
        restr = t;
    }
    std::pair<int, int> built{-1, -1};

    Multigrid& ready(Context& ctx, const DevPattern& pat, const double* val)
    {
        if (!mg || built != std::make_pair(p.maxLevels, p.coarseMax))
        {
            mg.reset();
            mg = std::make_unique<Multigrid>(ctx, pat, p.maxLevels, p.coarseMax);
            built = {p.maxLevels, p.coarseMax};
            valid = false;
        }
        if (!valid)
        {
            mg->setMatrix(val, pre, scale, post, restr, p);
            valid = weightsValid = true;
        }
        else if (!weightsValid)
        {
            mg->setWeights(pre, scale, post, restr, p);
            weightsValid = true;
        }
        return *mg;
    }
};

} // namespace cfx
