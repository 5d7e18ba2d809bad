// train.hip -- training loop of the message-passing network on the GPU.
#include "train.h"

#include <algorithm>
#include <cmath>

namespace cfx
{
namespace
{

constexpr int B = 256;

__device__ __forceinline__ std::uint64_t mix64(std::uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// standard normal entries from a counter-based hash (Box-Muller)
__global__ void k_randn(int n, std::uint64_t seed, float* __restrict__ e)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const std::uint64_t h = mix64(seed ^ mix64(static_cast<std::uint64_t>(i)));
    const double u1 = (static_cast<double>(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = static_cast<double>(mix64(h) >> 11) * (1.0 / 9007199254740992.0);
    e[i] = static_cast<float>(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

__global__ void k_fill(int n, float* __restrict__ x, float v)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void k_scale(int n, float* __restrict__ x, float s)
{
    const int i = blockIdx.x * B + threadIdx.x;
    if (i < n) x[i] *= s;
}

__global__ void k_slot_geometry(int nI, const double* __restrict__ sfd, const double* __restrict__ magSf,
                                const double* __restrict__ d2, const int* __restrict__ slotU,
                                const int* __restrict__ slotL, double* __restrict__ geo)
{
    const int f = blockIdx.x * B + threadIdx.x;
    if (f >= nI) return;
    const double c = sfd[f] / (magSf[f] * sqrt(d2[f]));
    geo[slotU[f]] = 1.0 - c;
    geo[slotL[f]] = 1.0 - c;
}

} // namespace

void meshSlotGeometry(const DevMesh& m, double* geo)
{
    hipStream_t s = m.ctx->stream();
    CFX_HIP(hipMemsetAsync(geo, 0, sizeof(double) * m.pattern.nnz, s));
    if (m.nI > 0)
        k_slot_geometry<<<divUp(m.nI, B), B, 0, s>>>(m.nI, m.sfd.as<double>(), m.magSf.as<double>(), m.d2.as<double>(),
                                                     m.slotU.as<int>(), m.slotL.as<int>(), geo);
    CFX_CHECK_LAUNCH();
}

void learnedWeights(Mpnn& net, const DevPattern& pat, const double* val, const double* geo, const double* rhs,
                    float* a, float* t, bool sameGraph)
{
    net.setGraph(pat.view(), val, geo, rhs, sameGraph);
    net.forward(a, t, false);
}

void MpnnTrainer::reserve(int n)
{
    if (static_cast<std::size_t>(n) <= cap_) return;
    for (DevBuf* b : {&a_, &t_, &ga_, &gt_, &e_}) b->alloc(sizeof(float) * n);
    if (!grad_.p) grad_.alloc(sizeof(float) * MpnnShape::nParams);
    cap_ = n;
}

Multigrid& MpnnTrainer::hierarchy(std::size_t i, const TrainSystem& s, const MgParams& p)
{
    if (mg_.size() <= i)
    {
        mg_.resize(i + 1);
        mgPat_.resize(i + 1, 0);
    }
    if (!mg_[i] || mgPat_[i] != s.pat->uid || mg_[i]->params().maxLevels != p.maxLevels)
    {
        mg_[i].reset();
        mg_[i] = std::make_unique<Multigrid>(ctx_, *s.pat, p.maxLevels, p.coarseMax);
        mgPat_[i] = s.pat->uid;
    }
    return *mg_[i];
}

double MpnnTrainer::sampleLoss(Multigrid& mg, int n, std::uint64_t seed, int smooth, bool accumulate)
{
    hipStream_t s = ctx_.stream();
    k_randn<<<divUp(n, B), B, 0, s>>>(n, seed, e_.as<float>());
    CFX_CHECK_LAUNCH();
    mg.smoothSample(e_.as<float>(), smooth);
    (void)accumulate;
    return mg.lossGrad(e_.as<float>(), a_.as<float>(), t_.as<float>(), ga_.as<float>(), gt_.as<float>());
}

std::vector<double> MpnnTrainer::train(const std::vector<TrainSystem>& systems, const TrainOptions& opt)
{
    if (systems.empty()) throw std::invalid_argument("train: no systems");
    hipStream_t s = ctx_.stream();
    MgParams p = opt.mg;
    std::vector<double> hist;
    for (int step = 0; step < opt.steps; ++step)
    {
        const std::size_t i = static_cast<std::size_t>(step) % systems.size();
        const TrainSystem& sys = systems[i];
        const int n = sys.pat->n;
        reserve(n);
        Multigrid& mg = hierarchy(i, sys, p);
        net_.setGraph(sys.pat->view(), sys.val, sys.geo, sys.r0);
        net_.forward(a_.as<float>(), t_.as<float>(), true);
        mg.setMatrix(sys.val, a_.as<float>(), t_.as<float>(), a_.as<float>(), t_.as<float>(), p);
        k_fill<<<divUp(n, B), B, 0, s>>>(n, ga_.as<float>(), 0.0f);
        k_fill<<<divUp(n, B), B, 0, s>>>(n, gt_.as<float>(), 0.0f);
        double loss = 0.0;
        for (int q = 0; q < opt.samples; ++q)
            loss += sampleLoss(mg, n, opt.seed * 1000003ull + static_cast<std::uint64_t>(step) * 131ull + q,
                               opt.smooth, true);
        if (!std::isfinite(loss))
        {
            // a degenerate sample (or system): no update from this step
            hist.push_back(loss);
            continue;
        }
        const float inv = 1.0f / static_cast<float>(opt.samples);
        k_scale<<<divUp(n, B), B, 0, s>>>(n, ga_.as<float>(), inv);
        k_scale<<<divUp(n, B), B, 0, s>>>(n, gt_.as<float>(), inv);
        net_.backward(ga_.as<float>(), gt_.as<float>(), grad_.as<float>());
        net_.adam(grad_.as<float>(), opt.lr);
        hist.push_back(loss / opt.samples);
    }
    ctx_.sync();
    return hist;
}

double MpnnTrainer::evaluate(const TrainSystem& sys, const TrainOptions& opt, int samples, bool withNetwork)
{
    hipStream_t s = ctx_.stream();
    MgParams p = opt.mg;
    const int n = sys.pat->n;
    reserve(n);
    Multigrid mg(ctx_, *sys.pat, p.maxLevels, p.coarseMax);
    if (withNetwork)
    {
        net_.setGraph(sys.pat->view(), sys.val, sys.geo, sys.r0);
        net_.forward(a_.as<float>(), t_.as<float>(), false);
    }
    else
    {
        // the constants as per-row arrays (tied weights: a, t = sqrt(scale))
        k_fill<<<divUp(n, B), B, 0, s>>>(n, a_.as<float>(), p.omegaPre);
        k_fill<<<divUp(n, B), B, 0, s>>>(n, t_.as<float>(), std::sqrt(p.scale));
    }
    mg.setMatrix(sys.val, a_.as<float>(), t_.as<float>(), a_.as<float>(), t_.as<float>(), p);
    double loss = 0.0;
    for (int q = 0; q < samples; ++q)
    {
        k_fill<<<divUp(n, B), B, 0, s>>>(n, ga_.as<float>(), 0.0f);
        k_fill<<<divUp(n, B), B, 0, s>>>(n, gt_.as<float>(), 0.0f);
        loss += sampleLoss(mg, n, opt.seed * 7919ull + q, opt.smooth, false);
    }
    return loss / samples;
}

} // namespace cfx
