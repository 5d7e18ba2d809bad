// dic.h -- diagonal incomplete Cholesky (DIC) preconditioner, level scheduled.
//
// D_ii = a_ii - sum_{j<i} a_ij^2 / D_jj over the stored lower neighbours (the
// DIC factor of OpenFOAM's definition, rD = 1/D), M = (D + L) D^-1 (D + U);
// z = M^-1 r by a forward and a backward substitution.  Rows are grouped into
// levels of the dependency graph of the given ordering (a row's level is one
// more than the highest level of its lower neighbours), so each level is one
// parallel kernel; the number of levels, about the RCM bandwidth of the
// mesh graph, bounds the parallelism.  Factor in FP64, apply in FP32.
#pragma once

#include "types.h"

#include <vector>

namespace cfx
{

class Dic : public Preconditioner
{
public:
    Dic(Context& ctx, const DevPattern& pat);
    // factor from the FP64 values (a non-positive pivot falls back to a_ii)
    void setMatrix(const double* val64);
    int levels() const { return static_cast<int>(levelPtr_.size()) - 1; }
    const double* diag64() const { return d64_.as<double>(); }

    int rows() const override { return pat_.n; }
    std::uintptr_t id() const override { return id_; }
    const float* val32() const override { return val32_.as<float>(); }
    const float* invD32() const override { return invA32_.as<float>(); }
    void apply(const float* r, float* z, const float* q, double* partZQ, int Gs, int rbPerWg,
               const int* done) override;

private:
    Context& ctx_;
    const DevPattern& pat_;
    std::uintptr_t id_ = 0;
    std::vector<int> levelPtr_;          // rows of level L: levelRows[levelPtr[L] .. levelPtr[L+1])
    DevBuf levelRows_, d64_, d32_, val32_, invA32_;
};

} // namespace cfx
