// train.h -- on-device training of the message-passing network against the
// multigrid objective, and the glue that turns a network into the finest
// level weights of a solve.
#pragma once

#include "mesh.h"
#include "mg.h"
#include "mpnn.h"

#include <cstdint>
#include <memory>
#include <vector>

namespace cfx
{

// one training system: a pattern with FP64 values, optional per-slot face
// geometry and optional right-hand side (the residual input of the network)
struct TrainSystem
{
    const DevPattern* pat = nullptr;
    const double* val = nullptr;
    const double* geo = nullptr;
    const double* r0 = nullptr;
};

struct TrainOptions
{
    int steps = 200;      // optimiser steps (systems visited round robin)
    int samples = 2;      // error samples per step
    int smooth = 4;       // cycles applied to a random vector to make a sample
    float lr = 1e-3f;     // Adam learning rate
    std::uint64_t seed = 1;
    MgParams mg;          // hierarchy / coarse constants (kcycleDepth forced to 0)
};

class MpnnTrainer
{
public:
    MpnnTrainer(Context& ctx, Mpnn& net) : ctx_(ctx), net_(net) {}
    // Adam steps over the systems; returns the mean objective of every step
    std::vector<double> train(const std::vector<TrainSystem>& systems, const TrainOptions& opt);
    // mean objective over `samples` fixed-seed error samples (no update);
    // withNetwork = false evaluates the constants of opt.mg instead
    double evaluate(const TrainSystem& sys, const TrainOptions& opt, int samples, bool withNetwork);

private:
    Context& ctx_;
    Mpnn& net_;
    std::vector<std::unique_ptr<Multigrid>> mg_;
    std::vector<std::uint64_t> mgPat_;  // DevPattern::uid each hierarchy was built for
    DevBuf a_, t_, ga_, gt_, e_, grad_;
    std::size_t cap_ = 0;
    Multigrid& hierarchy(std::size_t i, const TrainSystem& s, const MgParams& p);
    void reserve(int n);
    double sampleLoss(Multigrid& mg, int n, std::uint64_t seed, int smooth, bool accumulate);
};

// per-slot face geometry feature of a mesh's pattern: 1 - cos(Sf, d) on the
// two slots of every internal face, 0 on the diagonal
void meshSlotGeometry(const DevMesh& m, double* geo);

// weights of a learned solve: features of (pattern, values, geometry, rhs),
// network forward, tied weights a (pre = post) and t (restrict = prolong);
// sameGraph: the values are those of the network's previous graph (only the
// residual feature is formed again)
void learnedWeights(Mpnn& net, const DevPattern& pat, const double* val, const double* geo, const double* rhs,
                    float* a, float* t, bool sameGraph = false);

} // namespace cfx
