// batch.h -- many independent pressure systems solved together.
//
// Jacobi-PCG on K systems at once.  Three-launch form (the default): one
// launch of each of the three PCG kernels per iteration covers every system
// (workgroups dealt to systems in proportion to their size), the iteration
// replayed as a captured hipGraph and the host polling between batches of
// iterations.  Persistent form: one launch, captured as a hipGraph, runs the
// whole batch -- its workgroups claim chunks of whichever systems have work,
// each system advancing through its steps on its own and retiring at its own
// stop test.  On a device whose XCDs keep separate L2s every claim costs an
// agent-scope acquire and release (L2 invalidate and write-back), which the
// three-launch form pays once per kernel boundary, so the autotuner normally
// keeps the three-launch form.  Either way a converged system drops out
// while the others continue.  Either way a chunk runs the single-system kernel body with
// that system's own scalars, partials and stop test, so with the default
// schedule every system's iterates are bit for bit those of the
// single-system solve; the schedule (form, elements per thread of the vector
// kernels, cap of SpMV workgroups per system, chunks per claim, graph
// replay) can be picked by the autotuner, which times each candidate on the
// batch itself.
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
    int graph = 1;        // replay the captured iteration (persistent: the captured solve)
    int pollEvery = 16;   // iterations between host polls (three-launch form)
    int persistent = 0;   // one launch for the whole batch (1) or three launches per iteration (0)
    int claim = 2;        // persistent: claims per workgroup and step (sets the chunks taken per claim)
    int wgs = 0;          // persistent: workgroups of the launch (0: as many as fit the device at once)
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
    DevBuf ctl_, chunks_, home_;  // persistent form: control words, chunk counts, first system per workgroup
    hipEvent_t e0_ = nullptr, e1_ = nullptr;
    std::map<std::vector<int>, BatchSchedule> tuned_;
    void reserve(std::size_t n, std::size_t K);
    void runPersistent(const std::vector<BatchSystem>& sys, const std::vector<int2>& grids, const BatchSchedule& sc,
                       double tol, int maxIter, std::size_t N, BatchResult& res, std::vector<double>& finishMs);
};

} // namespace cfx
