// batch.h -- many independent pressure systems solved together.
//
// Jacobi-PCG on K systems at once: one launch of each of the three PCG
// kernels covers every system (workgroups are dealt to systems in
// proportion to their size, each workgroup runs the single-system kernel
// body on its system with that system's own scalars, partials and stop
// test), so a converged system drops out while the others continue.  The
// three-kernel iteration is captured once as a hipGraph and replayed.  With
// the default schedule every system's iterates are bit for bit those of the
// single-system solve; the schedule (elements per thread of the vector
// kernels, cap of SpMV workgroups per system, graph replay) can be picked by
// the autotuner, which times each candidate on the batch itself.
#pragma once

#include "krylov.h"

#include <map>
#include <vector>

namespace cfx
{

struct BatchSchedule
{
    int vecElems = 4;     // elements per thread of the vector kernels (Gv = n / (vecElems * 256))
    int wgCap = 2048;     // SpMV workgroups per system at most
    int graph = 1;        // replay the captured iteration
    int pollEvery = 16;   // iterations between host polls
};

struct BatchSystem
{
    CsrView A;
    const double* val = nullptr;
    const double* b = nullptr;
    double* x = nullptr;
};

struct BatchResult
{
    std::vector<SolveReport> reports;
    std::vector<double> finishMs;  // device time at which each system's stop was observed
    double ms = 0;                 // whole batch, device time
    int maxIterations = 0;
};

class BatchedPcg
{
public:
    explicit BatchedPcg(Context& ctx);
    ~BatchedPcg();
    BatchResult solve(const std::vector<BatchSystem>& sys, double tol, int maxIter, const BatchSchedule& s);
    // times every candidate schedule for `iters` iterations (x restored) and
    // returns the fastest; results are cached per batch shape
    BatchSchedule autotune(const std::vector<BatchSystem>& sys, int iters, std::vector<std::pair<BatchSchedule, double>>* log);

private:
    Context& ctx_;
    std::size_t capN_ = 0, capK_ = 0;
    DevBuf r_, p_, q_, invD_, part_, state_, sys_, tabS_, tabV_, xs_;
    hipEvent_t e0_ = nullptr, e1_ = nullptr;
    std::map<std::vector<int>, BatchSchedule> tuned_;
    void reserve(std::size_t n, std::size_t K);
};

} // namespace cfx
