// cfx_api.hip -- the C ABI (include/cfx.h) over the device objects.
#include "../../include/cfx.h"

#include "dev/batch.h"
#include "dev/dic.h"
#include "dev/fvops.h"
#include "dev/krylov.h"
#include "dev/mesh.h"
#include "dev/mg.h"
#include "dev/noalloc.h"
#include "dev/simple.h"
#include "dev/train.h"

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>

using namespace cfx;

struct cfx_ctx
{
    std::unique_ptr<Context> ctx;
    std::string err;
    std::map<int, std::unique_ptr<Krylov>> krylov;  // workspace per system size
    bool graphs = false;  // flexible CG launches eagerly (measured faster than graph replay)
    bool batchComponents = true;
    bool probeFcgA = false;
    std::unique_ptr<BatchedPcg> batch;

    Krylov& solver(int n, int rowBlocks)
    {
        auto& k = krylov[n];
        // a workspace is sized for a row-block count; grow it for a matrix
        // of the same size with more CSR-stream blocks
        if (!k || k->maxBlocks() < rowBlocks) k = std::make_unique<Krylov>(*ctx, n, std::max(rowBlocks, 1));
        k->useGraphs = graphs;
        k->batchComponents = batchComponents;
        k->probeFcgA = probeFcgA;
        return *k;
    }
};

struct cfx_mpnn
{
    cfx_ctx* owner = nullptr;
    std::unique_ptr<Mpnn> net;
    std::unique_ptr<MpnnTrainer> trainer;
};

struct cfx_mpnn;

struct cfx_mesh
{
    cfx_ctx* owner = nullptr;
    std::unique_ptr<DevMesh> m;
    DevBuf faceTmp, gradTmp, cellTmp, result;
    // pressure solves of the drivers with pPrecision CFX_FP32_MG
    MgCache pmg;
    cfx_mpnn* net = nullptr;
    DevBuf geo, wa, wt;
    PressureDump dump;
    PressureSolver* pressureSolver();
    PressureSolver ps;
};

struct cfx_matrix
{
    cfx_mesh* mesh = nullptr;
    int ncomp = 1;
    DevBuf diag, rhs, rhsBase, upper, lower, val, val32, tmp;
    // values version: bumped by every call that changes the matrix values; a solve whose version equals
    // the last solve's reuses what depends on the values alone (multigrid hierarchy, network features).
    // Raw device access to the values (cfx_matrix_device_ptrs) turns the reuse off for good.
    std::uint64_t version = 1, solvedVersion = 0;
    bool rawAccess = false;
    double relaxation = 1.0;     // Schemes::relaxation of this matrix
    double innerTol = 0.0;       // reliable-update factor of its FP32-refined solves (0: each solver's default)
    DevBuf relaxOld;             // the solution before the final pass ([ncomp][n])
    MgCache mg;
    cfx_mpnn* net = nullptr;     // learned finest-level weights (precond LEARNED)
    DevBuf geo, wa, wt;          // per-slot face geometry, network outputs
    Ldu ldu() const
    {
        Ldu A;
        A.ncomp = ncomp;
        A.diag = diag.as<double>();
        A.rhs = rhs.as<double>();
        A.upper = upper.as<double>();
        A.lower = lower.as<double>();
        return A;
    }
};

struct cfx_csr
{
    cfx_ctx* owner = nullptr;
    DevPattern pattern;
    DevBuf val, val32, b, x, resHist;
    double innerTol = 0.0;  // reliable-update factor of the mixed-precision solves (0: each solver's default)
    int lastIters = -1;
    bool val32Valid = false;
    MgCache mg;
    cfx_mpnn* net = nullptr;     // learned finest-level weights (precond LEARNED)
    const double* geo = nullptr; // caller's per-slot geometry (optional)
    // values version (bumped by set_values) and the version the network's features were formed for
    std::uint64_t version = 1, netVersion = 0;
    DevBuf wa, wt;
    std::unique_ptr<Dic> dic;    // precond DIC
    bool dicValid = false;
    Dic& dicReady()
    {
        if (!dic) dic = std::make_unique<Dic>(*owner->ctx, pattern);
        if (!dicValid)
        {
            dic->setMatrix(val.as<double>());
            dicValid = true;
        }
        return *dic;
    }
};

namespace
{

int fail(cfx_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

// an object in the wrong state for the call (CFX_E_STATE)
struct StateError : std::logic_error
{
    using std::logic_error::logic_error;
};

template <typename F>
int guard(cfx_ctx* c, F&& f)
{
    if (!c) return CFX_E_INVALID;
    try
    {
        CFX_HIP(hipSetDevice(c->ctx->device()));
        f();
        return CFX_OK;
    }
    catch (const HipError& e)
    {
        return fail(c, CFX_E_HIP, e.what());
    }
    catch (const std::invalid_argument& e)
    {
        return fail(c, CFX_E_INVALID, e.what());
    }
    catch (const StateError& e)
    {
        return fail(c, CFX_E_STATE, e.what());
    }
    catch (const std::exception& e)
    {
        return fail(c, CFX_E_INTERNAL, e.what());
    }
}

void toReport(const SolveReport& s, cfx_report* r)
{
    if (!r) return;
    r->iterations = s.iterations;
    r->initialResidual = s.initialResidual;
    r->finalResidual = s.finalResidual;
    r->converged = s.converged;
    r->breakdown = s.breakdown;
    r->l2Residual = s.l2Residual;
    r->trueResidual = s.trueResidual;
    r->outerIterations = s.outerIterations;
    r->precondFallbacks = s.precondFallbacks;
    r->ms = s.ms;
}

void need(bool ok, const char* what)
{
    if (!ok) throw std::invalid_argument(what);
}

// segment sums out[s] = (sum x[i] w[i], sum w[i]) over i = idx[start[s] .. start[s+1]) (w null: 1),
// one workgroup per segment, fixed-order block reduction
__global__ __launch_bounds__(256) void k_segment_sums(const long long* __restrict__ start,
                                                      const long long* __restrict__ idx,
                                                      const double* __restrict__ x, const double* __restrict__ w,
                                                      double* __restrict__ out)
{
    __shared__ double red[256 / kWave];
    const int sgm = blockIdx.x;
    double num = 0.0, den = 0.0;
    for (long long k = start[sgm] + threadIdx.x; k < start[sgm + 1]; k += 256)
    {
        const long long i = idx[k];
        const double wi = w ? w[i] : 1.0;
        num += x[i] * wi;
        den += wi;
    }
    num = blockSum(num, red);
    den = blockSum(den, red);
    if (threadIdx.x == 0)
    {
        out[2 * sgm] = num;
        out[2 * sgm + 1] = den;
    }
}

__global__ void k_zero_bytes(unsigned char* p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_explicit_op(int n, const double* __restrict__ rhs, const double* __restrict__ ax,
                              const double* __restrict__ diag, const double* __restrict__ x,
                              double* __restrict__ h)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) h[i] = rhs[i] - (ax[i] - diag[i] * x[i]);
}

// streaming yardsticks for the rooflines: a read-only pass (16-byte loads, grid-stride, one partial per
// workgroup) and a copy, over n4 16-byte elements
__global__ __launch_bounds__(256) void k_stream_read(const float4* __restrict__ a, long long n4,
                                                     float* __restrict__ out)
{
    __shared__ float red[256];
    float s = 0.f;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n4; i += gridDim.x * 256LL)
    {
        const float4 v = a[i];
        s += (v.x + v.y) + (v.z + v.w);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b, long long n4)
{
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n4; i += gridDim.x * 256LL) b[i] = a[i];
}

__global__ void k_add_vec(int n, const double* __restrict__ v, double* __restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += v[i];
}

SolveReport solveOn(cfx_ctx* c, const CsrView& A, const double* val, DevBuf& val32, bool& val32Valid,
                    const double* b, double* x, double tol, int maxIter, int solver, int precond,
                    int precision, double innerTol, double* resHist)
{
    need(solver == CFX_SOLVER_CG || solver == CFX_SOLVER_BICGSTAB, "unknown solver");
    need(precision == CFX_FP64 || precision == CFX_FP32_IR, "unknown precision");
    need(precond == CFX_PRECOND_JACOBI, "preconditioner not available in this build");
    Krylov& K = c->solver(A.n, A.nRowBlocks);
    if (solver == CFX_SOLVER_BICGSTAB)
        return precision == CFX_FP64 ? K.bicgstab(A, val, b, x, tol, maxIter)
                                     : K.bicgstabMixed(A, val, b, x, tol, maxIter);
    if (precision == CFX_FP64) return K.cg(A, val, b, x, tol, maxIter, resHist);
    if (!val32Valid)
    {
        if (!val32.p || val32.bytes < sizeof(float) * A.nnz) val32.alloc(sizeof(float) * A.nnz);
        castToFloat(*c->ctx, val, val32.as<float>(), A.nnz);
        val32Valid = true;
    }
    return K.cgMixed(A, val, val32.as<float>(), b, x, tol, maxIter, innerTol, false);
}

} // namespace

PressureSolver* cfx_mesh::pressureSolver()
{
    DevMesh& dm = *m;
    ps.mg = &pmg;
    ps.net = nullptr;
    if (net)
    {
        if (!geo.p)
        {
            geo.alloc(sizeof(double) * dm.pattern.nnz);
            meshSlotGeometry(dm, geo.as<double>());
            wa.alloc(sizeof(float) * dm.n);
            wt.alloc(sizeof(float) * dm.n);
        }
        ps.net = net->net.get();
        ps.geo = geo.as<double>();
        ps.wa = wa.as<float>();
        ps.wt = wt.as<float>();
    }
    return &ps;
}

extern "C" {

int cfx_create(int device, cfx_ctx** out)
{
    if (!out) return CFX_E_INVALID;
    *out = nullptr;
    auto* c = new cfx_ctx;
    try
    {
        c->ctx = std::make_unique<Context>(device);
    }
    catch (const std::exception& e)
    {
        delete c;
        return CFX_E_HIP;
    }
    *out = c;
    return CFX_OK;
}

int cfx_destroy(cfx_ctx* c)
{
    if (!c) return CFX_E_INVALID;
    try
    {
        c->krylov.clear();
        c->ctx.reset();
    }
    catch (...)
    {
    }
    delete c;
    return CFX_OK;
}

const char* cfx_last_error(const cfx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int cfx_synchronize(cfx_ctx* c)
{
    return guard(c, [&] { c->ctx->sync(); });
}

int cfx_device_info(cfx_ctx* c, char* name, size_t len, int* cus)
{
    return guard(c, [&] {
        if (name && len) { std::strncpy(name, c->ctx->deviceName.c_str(), len - 1); name[len - 1] = 0; }
        if (cus) *cus = c->ctx->numCUs();
    });
}

int cfx_stream_bandwidth(cfx_ctx* c, size_t bytes, int reps, double* readGBps, double* copyGBps)
{
    return guard(c, [&] {
        need(bytes >= (1u << 20) && reps > 0, "stream bandwidth: at least 1 MiB and one repetition");
        hipStream_t s = c->ctx->stream();
        const long long n4 = static_cast<long long>(bytes / 16);
        DevBuf a(16 * static_cast<std::size_t>(n4)), b(16 * static_cast<std::size_t>(n4));
        constexpr int kGrids[] = {1024, 2048, 4096};
        DevBuf out(sizeof(float) * 4096);
        CFX_HIP(hipMemsetAsync(a.p, 0, a.bytes, s));
        hipEvent_t e0, e1;
        CFX_HIP(hipEventCreate(&e0));
        CFX_HIP(hipEventCreate(&e1));
        double bestR = 0, bestC = 0;
        // the best of a few grid sizes, each timed over reps launches after one warm-up
        for (int g : kGrids)
        {
            float ms = 0;
            k_stream_read<<<g, 256, 0, s>>>(a.as<float4>(), n4, out.as<float>());
            CFX_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) k_stream_read<<<g, 256, 0, s>>>(a.as<float4>(), n4, out.as<float>());
            CFX_HIP(hipEventRecord(e1, s));
            CFX_HIP(hipEventSynchronize(e1));
            CFX_HIP(hipEventElapsedTime(&ms, e0, e1));
            bestR = std::max(bestR, 16.0 * n4 * reps / (ms * 1e-3) / 1e9);
            k_stream_copy<<<g, 256, 0, s>>>(a.as<float4>(), b.as<float4>(), n4);
            CFX_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) k_stream_copy<<<g, 256, 0, s>>>(a.as<float4>(), b.as<float4>(), n4);
            CFX_HIP(hipEventRecord(e1, s));
            CFX_HIP(hipEventSynchronize(e1));
            CFX_HIP(hipEventElapsedTime(&ms, e0, e1));
            bestC = std::max(bestC, 32.0 * n4 * reps / (ms * 1e-3) / 1e9);
        }
        CFX_CHECK_LAUNCH();
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (readGBps) *readGBps = bestR;
        if (copyGBps) *copyGBps = bestC;
    });
}

int cfx_malloc(cfx_ctx* c, size_t bytes, void** out)
{
    return guard(c, [&] {
        need(out != nullptr, "null out");
        CFX_HIP(hipMalloc(out, bytes ? bytes : 1));
    });
}

int cfx_free(cfx_ctx* c, void* p)
{
    return guard(c, [&] {
        c->ctx->sync();
        CFX_HIP(hipFree(p));
    });
}

int cfx_memcpy_htod(cfx_ctx* c, void* dst, const void* src, size_t bytes)
{
    return guard(c, [&] {
        CFX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->ctx->stream()));
        c->ctx->sync();
    });
}

int cfx_memcpy_dtoh(cfx_ctx* c, void* dst, const void* src, size_t bytes)
{
    return guard(c, [&] {
        CFX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->ctx->stream()));
        c->ctx->sync();
    });
}

int cfx_memcpy_dtod(cfx_ctx* c, void* dst, const void* src, size_t bytes)
{
    return guard(c, [&] { CFX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->ctx->stream())); });
}

int cfx_zero(cfx_ctx* c, void* p, size_t bytes)
{
    return guard(c, [&] {
        if (!bytes) return;
        k_zero_bytes<<<std::max(1, std::min(4096, divUp((long long)bytes, 256))), 256, 0, c->ctx->stream()>>>(
            static_cast<unsigned char*>(p), bytes);
        CFX_CHECK_LAUNCH();
    });
}

int cfx_segment_sums(cfx_ctx* c, int nSeg, const int64_t* d_segStart, const int64_t* d_idx, const double* d_x,
                     const double* d_w, double* out)
{
    return guard(c, [&] {
        need(nSeg >= 0 && (nSeg == 0 || (d_segStart && d_idx && d_x && out)), "bad segment arguments");
        if (nSeg == 0) return;
        DevBuf res(sizeof(double) * 2 * nSeg);
        k_segment_sums<<<nSeg, 256, 0, c->ctx->stream()>>>(reinterpret_cast<const long long*>(d_segStart),
                                                           reinterpret_cast<const long long*>(d_idx), d_x, d_w,
                                                           res.as<double>());
        CFX_CHECK_LAUNCH();
        CFX_HIP(hipMemcpyAsync(out, res.p, sizeof(double) * 2 * nSeg, hipMemcpyDeviceToHost, c->ctx->stream()));
        c->ctx->sync();
    });
}

// ----------------------------------------------------------------- mesh
int cfx_mesh_create(cfx_ctx* c, int64_t nCells, int64_t nInternal, int64_t nFaces, const int64_t* owner,
                    const int64_t* neighbour, const double* Sf, const double* Cf, const double* C,
                    const double* V, const double* d, const double* magSf, int64_t nPatches,
                    const int64_t* patchStart, const int64_t* patchSize, cfx_mesh** out)
{
    return guard(c, [&] {
        need(out && owner && Sf && Cf && C && V && d && magSf, "null mesh array");
        need(nInternal == 0 || neighbour, "null neighbour");
        MeshDesc desc{nCells, nInternal, nFaces, reinterpret_cast<const long long*>(owner),
                      reinterpret_cast<const long long*>(neighbour), Sf, Cf, C, V, d, magSf, nPatches,
                      reinterpret_cast<const long long*>(patchStart), reinterpret_cast<const long long*>(patchSize)};
        auto* m = new cfx_mesh;
        m->owner = c;
        try
        {
            m->m = createDevMesh(*c->ctx, desc);
            m->faceTmp.alloc(sizeof(double) * nFaces);
            m->gradTmp.alloc(sizeof(double) * 3 * nCells);
            m->cellTmp.alloc(sizeof(double) * nCells);
            m->result.alloc(256);
            CFX_HIP(hipMemsetAsync(m->result.p, 0, 256, c->ctx->stream()));
            c->ctx->sync();
        }
        catch (...)
        {
            delete m;
            throw;
        }
        *out = m;
    });
}

int cfx_mesh_destroy(cfx_mesh* m)
{
    if (!m) return CFX_E_INVALID;
    cfx_ctx* c = m->owner;
    return guard(c, [&] {
        c->ctx->sync();
        delete m;
    });
}

int cfx_mesh_sizes(cfx_mesh* m, int64_t s[5])
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        s[0] = m->m->n;
        s[1] = m->m->nI;
        s[2] = m->m->nF;
        s[3] = m->m->pattern.nnz;
        s[4] = m->m->pattern.nRowBlocks;
    });
}

int cfx_mesh_pattern(cfx_mesh* m, int32_t* rowPtr, int32_t* cols)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        const auto& p = m->m->pattern;
        if (rowPtr) std::memcpy(rowPtr, p.hRowPtr.data(), sizeof(int) * p.hRowPtr.size());
        if (cols) std::memcpy(cols, p.hCol.data(), sizeof(int) * p.hCol.size());
    });
}

int cfx_owner_weights(cfx_mesh* m, double* d_w)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        if (m->m->nI)
            CFX_HIP(hipMemcpyAsync(d_w, m->m->ow.p, sizeof(double) * m->m->nI, hipMemcpyDeviceToDevice,
                                   m->owner->ctx->stream()));
    });
}

// ------------------------------------------------------------ field ops
int cfx_grad(cfx_mesh* m, const double* x, const double* linc, const double* link, int ncomp, int comp,
             int corrected, int quirk, double* g)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        need(ncomp >= 1 && comp >= 0 && comp < ncomp, "bad component");
        gradGauss(*m->m, x, linc, link, ncomp, comp, corrected != 0, quirk != 0, g, m->faceTmp.as<double>());
    });
}

int cfx_harmonic_gamma_uniform(cfx_mesh* m, double gamma, double* gf)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] { harmonicGammaUniform(*m->m, gamma, gf); });
}

int cfx_interpolate_to_faces(cfx_mesh* m, const double* cell, double* face)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] { interpolateToFaces(*m->m, cell, face); });
}

int cfx_update_phi_internal(cfx_mesh* m, const double* U3, double* phi)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] { updatePhiInternal(*m->m, U3, phi); });
}

int cfx_correct_boundary_flux(cfx_mesh* m, const double* U3, const double* linc, const double* link, double* phi)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] { correctBoundaryFlux(*m->m, U3, linc, link, phi); });
}

int cfx_correct_fluxes(cfx_mesh* m, const double* p, const double* g, const double* rAUf, const double* linc,
                       const double* link, int corrected, double* phi)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] { correctFluxes(*m->m, p, g, rAUf, linc, link, corrected != 0, phi); });
}

int cfx_div_phi(cfx_mesh* m, const double* phi, double* out)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] { divPhi(*m->m, phi, out); });
}

int cfx_continuity_error(cfx_mesh* m, const double* phi, double* err)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        double* r = m->result.as<double>();
        continuityError(*m->m, phi, m->cellTmp.as<double>(), r, m->owner->ctx->partials(0));
        CFX_HIP(hipMemcpyAsync(err, r, sizeof(double), hipMemcpyDeviceToHost, m->owner->ctx->stream()));
        m->owner->ctx->sync();
    });
}

// ------------------------------------------------------------- fvMatrix
int cfx_matrix_create(cfx_mesh* m, int ncomp, cfx_matrix** out)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        need(out && ncomp >= 1 && ncomp <= 3, "bad matrix arguments");
        auto* A = new cfx_matrix;
        A->mesh = m;
        A->ncomp = ncomp;
        const std::size_t n = m->m->n, nI = std::max(1, m->m->nI), nnz = m->m->pattern.nnz;
        A->diag.alloc(sizeof(double) * ncomp * n);
        A->rhs.alloc(sizeof(double) * ncomp * n);
        A->rhsBase.alloc(sizeof(double) * ncomp * n);
        A->tmp.alloc(sizeof(double) * n);
        A->upper.alloc(sizeof(double) * nI);
        A->lower.alloc(sizeof(double) * nI);
        A->val.alloc(sizeof(double) * ncomp * nnz);
        lduReset(*m->m, A->ldu());
        *out = A;
    });
}

int cfx_matrix_destroy(cfx_matrix* A)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        A->mesh->owner->ctx->sync();
        delete A;
    });
}

int cfx_matrix_reset(cfx_matrix* A)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        ++A->version;
        lduReset(*A->mesh->m, A->ldu());
    });
}

int cfx_assemble_laplacian(cfx_matrix* A, const double* gf, const double* linc, const double* link, int sign)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        ++A->version;
        assembleLaplacian(*A->mesh->m, A->ldu(), gf, linc, link, sign);
    });
}

int cfx_assemble_div(cfx_matrix* A, const double* phi, int scheme, double blend, const double* linc,
                     const double* link, int sign)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        need(scheme >= 0 && scheme <= 2, "div scheme must be 0 upwind, 1 linear, 2 blend");
        ++A->version;
        assembleDiv(*A->mesh->m, A->ldu(), phi, scheme, blend, linc, link, sign);
    });
}

int cfx_assemble_ddt(cfx_matrix* A, const double* old, double dt, int sign)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        need(dt > 0, "fvm::ddt: non-positive time step");
        ++A->version;
        assembleDdt(*A->mesh->m, A->ldu(), old, dt, sign);
    });
}

int cfx_matrix_add_rhs(cfx_matrix* A, int comp, const double* v)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        need(comp >= 0 && comp < A->ncomp, "bad component");
        const int n = A->mesh->m->n;
        k_add_vec<<<divUp(n, 256), 256, 0, A->mesh->owner->ctx->stream()>>>(n, v, A->rhs.as<double>() + (long long)comp * n);
        CFX_CHECK_LAUNCH();
    });
}

int cfx_matrix_get(cfx_matrix* A, double* diag, double* rhs, double* upper, double* lower)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        hipStream_t s = A->mesh->owner->ctx->stream();
        const std::size_t n = A->mesh->m->n, nI = A->mesh->m->nI;
        if (diag) CFX_HIP(hipMemcpyAsync(diag, A->diag.p, sizeof(double) * A->ncomp * n, hipMemcpyDeviceToHost, s));
        if (rhs) CFX_HIP(hipMemcpyAsync(rhs, A->rhs.p, sizeof(double) * A->ncomp * n, hipMemcpyDeviceToHost, s));
        if (upper && nI) CFX_HIP(hipMemcpyAsync(upper, A->upper.p, sizeof(double) * nI, hipMemcpyDeviceToHost, s));
        if (lower && nI) CFX_HIP(hipMemcpyAsync(lower, A->lower.p, sizeof(double) * nI, hipMemcpyDeviceToHost, s));
        A->mesh->owner->ctx->sync();
    });
}

int cfx_matrix_device_ptrs(cfx_matrix* A, double** d_diag, double** d_rhs)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        A->rawAccess = true;  // the caller may change the diagonal at any time
        if (d_diag) *d_diag = A->diag.as<double>();
        if (d_rhs) *d_rhs = A->rhs.as<double>();
    });
}

int cfx_nonorth_correct(cfx_mesh* m, const double* gf, const double* x, const double* linc, const double* link,
                        int ncomp, int comp, int gradCorrected, int sign, double* rhs)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        nonOrthCorrect(*m->m, gf, x, linc, link, ncomp, comp, gradCorrected != 0, sign, rhs,
                       m->gradTmp.as<double>(), m->faceTmp.as<double>());
    });
}

int cfx_matrix_solve(cfx_matrix* A, double* x, int solver, double tol, int maxIter, int nonOrth,
                     const double* gf, const double* linc, const double* link, int sign, int gradCorrected,
                     int precond, int precision, cfx_report* rep)
{
    if (!A) return CFX_E_INVALID;
    cfx_ctx* c = A->mesh->owner;
    return guard(c, [&] {
        DevMesh& m = *A->mesh->m;
        need(solver == CFX_SOLVER_CG || solver == CFX_SOLVER_BICGSTAB, "unknown solver");
        need(precision == CFX_FP64 || precision == CFX_FP32_IR, "unknown precision");
        need(precond == CFX_PRECOND_JACOBI || precond == CFX_PRECOND_LEARNED || precond == CFX_PRECOND_MG,
             "preconditioner not available in this build");
        const bool multigrid = precond == CFX_PRECOND_LEARNED || precond == CFX_PRECOND_MG;
        need(!multigrid || (solver == CFX_SOLVER_CG && precision == CFX_FP32_IR),
             "the multigrid preconditioner runs inside flexible CG in FP32 with FP64 refinement");
        if (precond == CFX_PRECOND_LEARNED && !A->net)
            throw StateError("CFX_PRECOND_LEARNED needs a network attached (cfx_matrix_set_mpnn); "
                             "CFX_PRECOND_MG solves with the multigrid weights alone");
        FvSolveArgs a;
        if (multigrid)
        {
            a.mg = &A->mg;
            if (precond == CFX_PRECOND_MG)
                A->mg.useWeights(A->mg.userPre, A->mg.userScale, A->mg.userPost, A->mg.userRestr);
            else
            {
                if (!A->geo.p)
                {
                    A->geo.alloc(sizeof(double) * m.pattern.nnz);
                    meshSlotGeometry(m, A->geo.as<double>());
                    A->wa.alloc(sizeof(float) * m.n);
                    A->wt.alloc(sizeof(float) * m.n);
                }
                a.net = A->net->net.get();
                a.geo = A->geo.as<double>();
                a.wa = A->wa.as<float>();
                a.wt = A->wt.as<float>();
            }
        }
        a.solver = solver;
        a.tol = tol;
        a.maxIter = maxIter;
        a.nonOrth = nonOrth;
        a.gammaFace = gf;
        a.linc = linc;
        a.link = link;
        a.sign = sign;
        a.gradCorrected = gradCorrected != 0;
        a.precision = precision;
        a.relaxation = A->relaxation;
        a.delta = A->innerTol;
        need(!multigrid || A->ncomp == 1, "the multigrid preconditioner solves scalar matrices");
        a.sameMatrix = !A->rawAccess && A->version == A->solvedVersion;
        A->solvedVersion = A->version;
        if (A->relaxation < 1.0)
        {
            const std::size_t need = sizeof(double) * A->ncomp * m.n;
            if (!A->relaxOld.p || A->relaxOld.bytes < need) A->relaxOld.alloc(need);
            a.relaxOld = A->relaxOld.as<double>();
        }
        const SolveReport last = fvSolve(m, c->solver(m.n, m.pattern.nRowBlocks), A->ldu(), A->val.as<double>(),
                                         A->rhsBase.as<double>(), A->val32, x, a, A->mesh->gradTmp.as<double>(),
                                         A->mesh->faceTmp.as<double>());
        toReport(last, rep);
    });
}

int cfx_matrix_set_relaxation(cfx_matrix* A, double relaxation)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        need(relaxation > 0.0 && relaxation <= 1.0 && std::isfinite(relaxation), "relaxation must be in (0, 1]");
        A->relaxation = relaxation;
    });
}

int cfx_matrix_set_inner_tol(cfx_matrix* A, double t)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        need(t == 0 || (t > 0 && t < 1), "inner tolerance must be 0 (the solver's default) or in (0,1)");
        A->innerTol = t;
    });
}

int cfx_matrix_explicit_operator(cfx_matrix* A, const double* x, int comp, double* h)
{
    if (!A) return CFX_E_INVALID;
    cfx_ctx* c = A->mesh->owner;
    return guard(c, [&] {
        DevMesh& m = *A->mesh->m;
        need(comp >= 0 && comp < A->ncomp, "bad component");
        Krylov& K = c->solver(m.n, m.pattern.nRowBlocks);
        K.spmv(m.pattern.view(), A->val.as<double>() + (long long)comp * m.pattern.nnz, x, A->tmp.as<double>());
        k_explicit_op<<<divUp(m.n, 256), 256, 0, c->ctx->stream()>>>(
            m.n, A->rhs.as<double>() + (long long)comp * m.n, A->tmp.as<double>(),
            A->diag.as<double>() + (long long)comp * m.n, x, h);
        CFX_CHECK_LAUNCH();
    });
}

int cfx_matrix_spmv(cfx_matrix* A, int comp, const double* x, double* y)
{
    if (!A) return CFX_E_INVALID;
    cfx_ctx* c = A->mesh->owner;
    return guard(c, [&] {
        DevMesh& m = *A->mesh->m;
        need(comp >= 0 && comp < A->ncomp, "bad component");
        c->solver(m.n, m.pattern.nRowBlocks).spmv(m.pattern.view(), A->val.as<double>() + (long long)comp * m.pattern.nnz, x, y);
    });
}

// ------------------------------------------------------ linear-solver seam
int cfx_csr_create(cfx_ctx* c, int64_t n, const int64_t* rowStart, const int64_t* cols, const double* vals,
                   cfx_csr** out)
{
    return guard(c, [&] {
        need(out && rowStart && cols && vals && n > 0, "bad CSR arguments");
        need(n < (1LL << 30) && rowStart[n] < (1LL << 31) - 1, "CSR too large for int32 indices");
        std::vector<int> rp(n + 1), cl(rowStart[n]);
        for (int64_t i = 0; i <= n; ++i) rp[i] = static_cast<int>(rowStart[i]);
        for (int64_t r = 0; r < n; ++r)
        {
            need(rowStart[r + 1] >= rowStart[r], "rowStart not monotone");
            for (int64_t k = rowStart[r]; k < rowStart[r + 1]; ++k)
            {
                need(cols[k] >= 0 && cols[k] < n, "column out of range");
                need(k == rowStart[r] || cols[k] > cols[k - 1], "columns not sorted within a row");
                cl[k] = static_cast<int>(cols[k]);
            }
        }
        auto* A = new cfx_csr;
        A->owner = c;
        try
        {
            A->pattern.build(std::move(rp), std::move(cl), c->ctx->stream());
            A->val.alloc(sizeof(double) * std::max(1, A->pattern.nnz));
            A->b.alloc(sizeof(double) * n);
            A->x.alloc(sizeof(double) * n);
            CFX_HIP(hipMemcpyAsync(A->val.p, vals, sizeof(double) * A->pattern.nnz, hipMemcpyHostToDevice,
                                   c->ctx->stream()));
            c->ctx->sync();
        }
        catch (...)
        {
            delete A;
            throw;
        }
        *out = A;
    });
}

int cfx_csr_destroy(cfx_csr* A)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        A->owner->ctx->sync();
        delete A;
    });
}

int cfx_csr_set_values(cfx_csr* A, const double* vals)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        CFX_HIP(hipMemcpyAsync(A->val.p, vals, sizeof(double) * A->pattern.nnz, hipMemcpyHostToDevice,
                               A->owner->ctx->stream()));
        ++A->version;
        A->val32Valid = false;
        A->mg.valid = false;
        A->dicValid = false;
        A->owner->ctx->sync();
    });
}

int cfx_csr_set_values_device(cfx_csr* A, const double* d_vals)
{
    if (!A || !d_vals) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        CFX_HIP(hipMemcpyAsync(A->val.p, d_vals, sizeof(double) * A->pattern.nnz, hipMemcpyDeviceToDevice,
                               A->owner->ctx->stream()));
        ++A->version;
        A->val32Valid = false;
        A->mg.valid = false;
        A->dicValid = false;
    });
}

int cfx_csr_spmv(cfx_csr* A, const double* x, double* y)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        A->owner->solver(A->pattern.n, A->pattern.nRowBlocks).spmv(A->pattern.view(), A->val.as<double>(), x, y);
    });
}

int cfx_csr_solve_device(cfx_csr* A, const double* b, double* x, double tol, int maxIter, int solver, int precond,
                         int precision, cfx_report* rep)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        need(maxIter >= 0, "maxIter < 0");
        double* hist = nullptr;
        if (solver == CFX_SOLVER_CG && precision == CFX_FP64)
        {
            if (!A->resHist.p || A->resHist.bytes < sizeof(double) * (maxIter + 1))
                A->resHist.alloc(sizeof(double) * (maxIter + 1));
            hist = A->resHist.as<double>();
        }
        SolveReport s;
        if (precond == CFX_PRECOND_DIC)
        {
            need(solver == CFX_SOLVER_CG, "DIC runs inside (flexible) CG");
            need(precision == CFX_FP32_IR, "DIC runs in FP32 with FP64 refinement");
            s = A->owner->solver(A->pattern.n, A->pattern.nRowBlocks)
                    .fcgMixed(A->pattern.view(), A->val.as<double>(), A->dicReady(), b, x, tol, maxIter, A->innerTol,
                              nullptr);
        }
        else if (precond == CFX_PRECOND_LEARNED || precond == CFX_PRECOND_MG)
        {
            need(solver == CFX_SOLVER_CG, "the multigrid preconditioner runs inside flexible CG");
            need(precision == CFX_FP32_IR, "the multigrid preconditioner runs in FP32 with FP64 refinement");
            if (precond == CFX_PRECOND_LEARNED && !A->net)
                throw StateError("CFX_PRECOND_LEARNED needs a network attached (cfx_csr_set_mpnn); "
                                 "CFX_PRECOND_MG solves with the multigrid weights alone");
            Context& ctx = *A->owner->ctx;
            if (precond == CFX_PRECOND_MG)
                A->mg.useWeights(A->mg.userPre, A->mg.userScale, A->mg.userPost, A->mg.userRestr);
            else
            {
                if (!A->wa.p)
                {
                    A->wa.alloc(sizeof(float) * A->pattern.n);
                    A->wt.alloc(sizeof(float) * A->pattern.n);
                }
                // features from the matrix alone are reused while its values are those of the last learned solve
                learnedWeights(*A->net->net, A->pattern, A->val.as<double>(), A->geo, b, A->wa.as<float>(),
                               A->wt.as<float>(), A->netVersion == A->version);
                A->netVersion = A->version;
                A->mg.useWeights(A->wa.as<float>(), A->wt.as<float>(), A->wa.as<float>(), A->wt.as<float>());
                A->mg.weightsValid = false;  // new contents
            }
            Multigrid& mg = A->mg.ready(ctx, A->pattern, A->val.as<double>());
            s = A->owner->solver(A->pattern.n, A->pattern.nRowBlocks)
                    .fcgMixed(A->pattern.view(), A->val.as<double>(), mg, b, x, tol, maxIter, A->innerTol, nullptr);
        }
        else
            s = solveOn(A->owner, A->pattern.view(), A->val.as<double>(), A->val32, A->val32Valid, b, x, tol, maxIter,
                        solver, precond, precision, A->innerTol, hist);
        A->lastIters = hist ? s.iterations : -1;
        toReport(s, rep);
    });
}

int cfx_pcg_solve_noalloc(cfx_ctx* c, int n, int nnz, const int* d_rowPtr, const int* d_col, const double* d_val,
                          const double* d_b, double* d_x, double tol, int maxIter, cfx_report* rep)
{
    static_assert(CFX_NOALLOC_MAX_ROWS == kNoAllocMaxRows, "the header's row limit differs from the library's");
    return guard(c, [&] {
        const SolveReport s = pcgNoAlloc(*c->ctx, n, nnz, d_rowPtr, d_col, d_val, d_b, d_x, tol, maxIter);
        toReport(s, rep);
    });
}

int cfx_csr_solve(cfx_csr* A, const double* b, double* x, double tol, int maxIter, int solver, int precond,
                  int precision, cfx_report* rep)
{
    if (!A) return CFX_E_INVALID;
    const std::size_t bytes = sizeof(double) * A->pattern.n;
    int rc = guard(A->owner, [&] {
        CFX_HIP(hipMemcpyAsync(A->b.p, b, bytes, hipMemcpyHostToDevice, A->owner->ctx->stream()));
        CFX_HIP(hipMemcpyAsync(A->x.p, x, bytes, hipMemcpyHostToDevice, A->owner->ctx->stream()));
    });
    if (rc) return rc;
    rc = cfx_csr_solve_device(A, A->b.as<double>(), A->x.as<double>(), tol, maxIter, solver, precond, precision, rep);
    if (rc) return rc;
    return guard(A->owner, [&] {
        CFX_HIP(hipMemcpyAsync(x, A->x.p, bytes, hipMemcpyDeviceToHost, A->owner->ctx->stream()));
        A->owner->ctx->sync();
    });
}

int cfx_csr_residual_history(cfx_csr* A, double* hist, int maxLen)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        need(A->lastIters >= 0, "no FP64 CG history recorded");
        const int len = std::min(maxLen, A->lastIters + 1);
        // entry 0 is the initial residual: filled from the report path
        CFX_HIP(hipMemcpyAsync(hist, A->resHist.p, sizeof(double) * len, hipMemcpyDeviceToHost, A->owner->ctx->stream()));
        A->owner->ctx->sync();
    });
}

int cfx_set_graphs(cfx_ctx* c, int enable)
{
    if (!c) return CFX_E_INVALID;
    c->graphs = enable != 0;
    return CFX_OK;
}

int cfx_set_component_batching(cfx_ctx* c, int enable)
{
    if (!c) return CFX_E_INVALID;
    c->batchComponents = enable != 0;
    return CFX_OK;
}

int cfx_mg_default_params(cfx_mg_params* p)
{
    if (!p) return CFX_E_INVALID;
    const MgParams d;
    *p = cfx_mg_params{d.omegaPre, d.scale, d.omegaPost, d.restrictScale, d.coarseOmega, d.coarseScale,
                       d.coarseSweeps, d.coarsestOmega, d.maxLevels, d.coarseMax, d.kcycleDepth, d.fusion,
                       d.smoothFine, d.smoothCoarse, d.smoothDeep};
    return CFX_OK;
}

extern "C++" {  // helpers of the entry points: C++ linkage
namespace
{
void setMgParams(MgCache& mg, const cfx_mg_params* p)
{
    need(p->maxLevels >= 1 && p->maxLevels <= 16, "maxLevels must be in [1,16]");
    need(p->coarseMax >= 1 && p->coarseMax <= 2048, "coarseMax must be in [1,2048]");
    need(p->coarseSweeps >= 0, "coarseSweeps < 0");
    MgParams& m = mg.p;
    m.omegaPre = p->omegaPre;
    m.scale = p->scale;
    m.omegaPost = p->omegaPost;
    m.restrictScale = p->restrictScale;
    m.coarseOmega = p->coarseOmega;
    m.coarseScale = p->coarseScale;
    m.coarseSweeps = p->coarseSweeps;
    m.coarsestOmega = p->coarsestOmega;
    m.maxLevels = p->maxLevels;
    m.coarseMax = p->coarseMax;
    need(p->kcycleDepth >= 0, "kcycleDepth < 0");
    m.kcycleDepth = p->kcycleDepth;
    m.fusion = p->fusion;
    need(p->smoothFine >= 1 && p->smoothFine <= 8 && p->smoothCoarse >= 1 && p->smoothCoarse <= 8 &&
             p->smoothDeep >= 1 && p->smoothDeep <= 8,
         "smoothFine, smoothCoarse and smoothDeep must be in [1,8]");
    m.smoothFine = p->smoothFine;
    m.smoothCoarse = p->smoothCoarse;
    m.smoothDeep = p->smoothDeep;
    mg.valid = false;
}

void setMgWeights(MgCache& mg, const float* pre, const float* scale, const float* post, const float* restr)
{
    mg.userPre = pre;
    mg.userScale = scale;
    mg.userPost = post;
    mg.userRestr = restr;
    mg.pre = pre;
    mg.scale = scale;
    mg.post = post;
    mg.restr = restr;
    mg.weightsValid = false;
}
} // namespace
} // extern "C++"

int cfx_csr_set_mg(cfx_csr* A, const cfx_mg_params* p)
{
    if (!A || !p) return CFX_E_INVALID;
    return guard(A->owner, [&] { setMgParams(A->mg, p); });
}

int cfx_csr_set_mg_weights(cfx_csr* A, const float* pre, const float* scale, const float* post, const float* restr)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->owner, [&] { setMgWeights(A->mg, pre, scale, post, restr); });
}

int cfx_matrix_set_mg(cfx_matrix* A, const cfx_mg_params* p)
{
    if (!A || !p) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] { setMgParams(A->mg, p); });
}

int cfx_matrix_set_mg_weights(cfx_matrix* A, const float* pre, const float* scale, const float* post,
                              const float* restr)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] { setMgWeights(A->mg, pre, scale, post, restr); });
}

int cfx_csr_mg_info(cfx_csr* A, int* levels, int64_t* rows, int64_t* nnz, int maxOut)
{
    if (!A || !levels) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        Multigrid& mg = A->mg.ready(*A->owner->ctx, A->pattern, A->val.as<double>());
        *levels = mg.levels();
        for (int l = 0; l < std::min(maxOut, mg.levels()); ++l)
        {
            if (rows) rows[l] = mg.rows(l);
            if (nnz) nnz[l] = mg.nnz(l);
        }
    });
}

int cfx_csr_mg_tail_level(cfx_csr* A, int* level)
{
    if (!A || !level) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        Multigrid& mg = A->mg.ready(*A->owner->ctx, A->pattern, A->val.as<double>());
        *level = mg.tailLevel();
    });
}

int cfx_csr_mg_slices(cfx_csr* A, int64_t* slots)
{
    if (!A || !slots) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        Multigrid& mg = A->mg.ready(*A->owner->ctx, A->pattern, A->val.as<double>());
        *slots = mg.fineSlots();
    });
}

int cfx_csr_mg_level(cfx_csr* A, int l, int* rowPtr, int* col, double* val, int* agg)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        Context& ctx = *A->owner->ctx;
        Multigrid& mg = A->mg.ready(ctx, A->pattern, A->val.as<double>());
        need(l >= 0 && l < mg.levels(), "level out of range");
        MgLevel& L = mg.level(l);
        const DevPattern& P = l == 0 ? A->pattern : L.pat;
        hipStream_t s = ctx.stream();
        if (rowPtr) std::memcpy(rowPtr, P.hRowPtr.data(), sizeof(int) * (L.n + 1));
        if (col) std::memcpy(col, P.hCol.data(), sizeof(int) * L.nnz);
        if (val)
            CFX_HIP(hipMemcpyAsync(val, l == 0 ? A->val.p : L.val64.p, sizeof(double) * L.nnz, hipMemcpyDeviceToHost, s));
        if (agg && l + 1 < mg.levels()) std::memcpy(agg, L.hAgg.data(), sizeof(int) * L.n);
        ctx.sync();
    });
}

int cfx_csr_mg_apply(cfx_csr* A, const float* r, float* z)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        Context& ctx = *A->owner->ctx;
        Multigrid& mg = A->mg.ready(ctx, A->pattern, A->val.as<double>());
        mg.apply(r, z, nullptr, nullptr, A->pattern.nRowBlocks, 1, nullptr);
        ctx.sync();
    });
}

int cfx_csr_dic_info(cfx_csr* A, int* levels, double* diag)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        Dic& d = A->dicReady();
        if (levels) *levels = d.levels();
        if (diag)
        {
            CFX_HIP(hipMemcpyAsync(diag, d.diag64(), sizeof(double) * A->pattern.n, hipMemcpyDeviceToHost,
                                   A->owner->ctx->stream()));
            A->owner->ctx->sync();
        }
    });
}

int cfx_csr_dic_apply(cfx_csr* A, const float* r, float* z)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        A->dicReady().apply(r, z, nullptr, nullptr, 1, 1, nullptr);
        A->owner->ctx->sync();
    });
}

int cfx_csr_set_inner_tol(cfx_csr* A, double t)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        need(t == 0 || (t > 0 && t < 1), "inner tolerance must be 0 (the solver's default) or in (0,1)");
        A->innerTol = t;
    });
}

extern "C++" {  // helpers of the entry points: C++ linkage
namespace
{
double timeSpmv(cfx_ctx* c, const CsrView& A, const double* val, DevBuf& val32, bool& v32ok, const void* x, void* y,
                int reps, int precision)
{
    need(reps > 0, "reps must be positive");
    Krylov& K = c->solver(A.n, A.nRowBlocks);
    hipStream_t s = c->ctx->stream();
    if (precision == CFX_FP32_IR && !v32ok)
    {
        if (!val32.p || val32.bytes < sizeof(float) * A.nnz) val32.alloc(sizeof(float) * A.nnz);
        castToFloat(*c->ctx, val, val32.as<float>(), A.nnz);
        v32ok = true;
    }
    hipEvent_t e0, e1;
    CFX_HIP(hipEventCreate(&e0));
    CFX_HIP(hipEventCreate(&e1));
    CFX_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i)
    {
        if (precision == CFX_FP32_IR)
            K.spmv32(A, val32.as<float>(), static_cast<const float*>(x), static_cast<float*>(y));
        else
            K.spmv(A, val, static_cast<const double*>(x), static_cast<double*>(y));
    }
    CFX_HIP(hipEventRecord(e1, s));
    CFX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    CFX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}
} // namespace
} // extern "C++"

int cfx_csr_time_spmv(cfx_csr* A, const void* x, void* y, int reps, int precision, double* ms)
{
    if (!A || !ms) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        *ms = timeSpmv(A->owner, A->pattern.view(), A->val.as<double>(), A->val32, A->val32Valid, x, y, reps,
                       precision);
    });
}

int cfx_matrix_time_spmv(cfx_matrix* A, int comp, const void* x, void* y, int reps, int precision, double* ms)
{
    if (!A || !ms) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        need(comp >= 0 && comp < A->ncomp, "bad component");
        DevMesh& m = *A->mesh->m;
        bool ok = false;
        *ms = timeSpmv(A->mesh->owner, m.pattern.view(), A->val.as<double>() + (long long)comp * m.pattern.nnz,
                       A->val32, ok, x, y, reps, precision);
    });
}

int cfx_matrix_time_fcg_spmv(cfx_matrix* A, int reps, double* ms)
{
    if (!A || !ms) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        need(A->ncomp == 1, "the flexible CG solves scalar matrices");
        DevMesh& m = *A->mesh->m;
        *ms = A->mesh->owner->solver(m.pattern.n, m.pattern.nRowBlocks)
                  .timeFcgSpmv(m.pattern.view(), A->val.as<double>(), reps, A->mg.mg.get());
    });
}

int cfx_set_fcg_probe(cfx_ctx* c, int enable)
{
    if (!c) return CFX_E_INVALID;
    c->probeFcgA = enable != 0;
    for (auto& k : c->krylov)
        if (k.second) k.second->probeFcgA = c->probeFcgA;
    return CFX_OK;
}

int cfx_matrix_fcg_probe(cfx_matrix* A, double* avgMs, int* launches)
{
    if (!A || !avgMs || !launches) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        DevMesh& m = *A->mesh->m;
        Krylov& k = A->mesh->owner->solver(m.pattern.n, m.pattern.nRowBlocks);
        *avgMs = k.probeMs;
        *launches = k.probeCount;
    });
}

int cfx_matrix_smooth_probe(cfx_matrix* A, double* avgMs, int* launches)
{
    if (!A || !avgMs || !launches) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        DevMesh& m = *A->mesh->m;
        Krylov& k = A->mesh->owner->solver(m.pattern.n, m.pattern.nRowBlocks);
        *avgMs = k.probeSmoothMs;
        *launches = k.probeSmoothCount;
    });
}

int cfx_matrix_time_smooth(cfx_matrix* A, int reps, double* ms)
{
    if (!A || !ms) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        need(A->ncomp == 1, "the multigrid preconditioner solves scalar matrices");
        DevMesh& m = *A->mesh->m;
        // the hierarchy of the last multigrid solve, else one built on the pattern now (the time depends on
        // the pattern alone)
        Multigrid& mg = A->mg.mg ? *A->mg.mg : A->mg.ready(*m.ctx, m.pattern, A->val.as<double>());
        *ms = mg.timeFineSmooth(reps);
    });
}

int cfx_linearise(cfx_mesh* m, int ncomp, const int32_t* type, const double* value, double* linc, double* link)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        need(ncomp == 1 || ncomp == 3, "ncomp must be 1 or 3");
        const int nB = m->m->nB;
        DevBuf t(sizeof(int) * std::max(1, nB)), v(sizeof(double) * ncomp * std::max(1, nB));
        hipStream_t s = m->owner->ctx->stream();
        CFX_HIP(hipMemcpyAsync(t.p, type, sizeof(int) * nB, hipMemcpyHostToDevice, s));
        CFX_HIP(hipMemcpyAsync(v.p, value, sizeof(double) * ncomp * nB, hipMemcpyHostToDevice, s));
        need(lineariseBc(*m->m, ncomp, t.as<int>(), v.as<double>(), linc, link),
             "fields: symmetry normal not axis aligned, or unknown bc type");
    });
}

extern "C++" {  // helpers of the entry points: C++ linkage
namespace
{
Schemes toSchemes(const cfx_schemes* s)
{
    Schemes o;
    if (s)
    {
        o.div = s->div;
        o.laplacianCorrected = s->laplacianCorrected;
        o.gradCorrected = s->gradCorrected;
        o.blendCoeff = s->blendCoeff;
    }
    return o;
}
} // namespace
} // extern "C++"

int cfx_simple_solve(cfx_mesh* m, double nu, const cfx_schemes* schemes, const cfx_simple_controls* ctl,
                     const int32_t* Utype, const double* Uvalue, const int32_t* ptype, const double* pvalue,
                     double* U, double* p, double* phi, cfx_simple_result* result, double* hist)
{
    if (!m || !ctl) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        need(m->m->nB == 0 || (Utype && Uvalue && ptype && pvalue), "null boundary conditions");
        need(ctl->pPrecision == CFX_FP64 || ctl->pPrecision == CFX_FP32_IR || ctl->pPrecision == CFX_FP32_MG,
             "unknown pressure precision");
        SimpleControls c;
        c.maxIterations = ctl->maxIterations;
        c.continuityTol = ctl->continuityTol;
        c.pcorrTol = ctl->pcorrTol;
        c.momentumTol = ctl->momentumTol;
        c.linearTol = ctl->linearTol;
        c.linearMaxIter = ctl->linearMaxIter;
        c.nonOrthCorrectors = ctl->nonOrthCorrectors;
        c.relaxU = ctl->relaxU;
        c.relaxP = ctl->relaxP;
        c.pRelaxStart = ctl->pRelaxStart;
        for (int k = 0; k < 3; ++k) c.bodyForce[k] = ctl->bodyForce[k];
        c.pPrecision = ctl->pPrecision;
        need(ctl->uPrecision == CFX_FP64 || ctl->uPrecision == CFX_FP32_IR, "uPrecision must be FP64 or FP32_IR");
        c.uPrecision = ctl->uPrecision;
        c.pSolver = ctl->pPrecision == CFX_FP32_MG ? m->pressureSolver() : nullptr;
        c.dump = &m->dump;
        const SimpleResult r = simpleSolve(*m->m, m->owner->solver(m->m->n, m->m->pattern.nRowBlocks), nu,
                                           toSchemes(schemes), c, Utype, Uvalue, ptype, pvalue, U, p, phi, hist);
        if (result)
        {
            result->iterations = r.iterations;
            result->finalContinuityError = r.finalContinuityError;
            result->finalMomentumResidual = r.finalMomentumResidual;
            result->converged = r.converged;
            result->diverged = r.diverged;
            result->ms = r.ms;
        }
    });
}

int cfx_piso_step(cfx_mesh* m, double nu, const cfx_schemes* schemes, const cfx_piso_controls* ctl,
                  const int32_t* Utype, const double* Uvalue, const int32_t* ptype, const double* pvalue, double* U,
                  double* p, double* phi, cfx_piso_result* result)
{
    if (!m || !ctl) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        need(m->m->nB == 0 || (Utype && Uvalue && ptype && pvalue), "null boundary conditions");
        need(ctl->pPrecision == CFX_FP64 || ctl->pPrecision == CFX_FP32_IR || ctl->pPrecision == CFX_FP32_MG,
             "unknown pressure precision");
        PisoControls c;
        c.dt = ctl->dt;
        c.correctors = ctl->correctors;
        c.continuityTol = ctl->continuityTol;
        c.linearTol = ctl->linearTol;
        c.linearMaxIter = ctl->linearMaxIter;
        c.nonOrthCorrectors = ctl->nonOrthCorrectors;
        for (int k = 0; k < 3; ++k) c.bodyForce[k] = ctl->bodyForce[k];
        c.projectNeumann = ctl->projectNeumann != 0;
        c.pPrecision = ctl->pPrecision;
        need(ctl->uPrecision == CFX_FP64 || ctl->uPrecision == CFX_FP32_IR, "uPrecision must be FP64 or FP32_IR");
        c.uPrecision = ctl->uPrecision;
        c.pSolver = ctl->pPrecision == CFX_FP32_MG ? m->pressureSolver() : nullptr;
        c.dump = &m->dump;
        const PisoResult r = pisoStep(*m->m, m->owner->solver(m->m->n, m->m->pattern.nRowBlocks), nu,
                                      toSchemes(schemes), c, Utype, Uvalue, ptype, pvalue, U, p, phi);
        if (result)
        {
            result->correctorsUsed = r.correctorsUsed;
            result->predictedContinuityError = r.predictedContinuityError;
            result->continuityError = r.continuityError;
            result->pressureCorrectionMax = r.pressureCorrectionMax;
            result->converged = r.converged;
            result->ms = r.ms;
        }
    });
}

int cfx_mesh_set_mg(cfx_mesh* m, const cfx_mg_params* p)
{
    if (!m || !p) return CFX_E_INVALID;
    return guard(m->owner, [&] { setMgParams(m->pmg, p); });
}

int cfx_mesh_set_pressure_dump(cfx_mesh* m, double* val, double* rhs, int every, int max)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        need(every <= 0 || (val && rhs && max >= 0), "dump buffers missing");
        m->dump = PressureDump{val, rhs, every, max, 0, 0};
    });
}

int cfx_mesh_pressure_dump_count(cfx_mesh* m, int* count)
{
    if (!m || !count) return CFX_E_INVALID;
    *count = m->dump.count;
    return CFX_OK;
}

// ------------------------------------------------- message-passing network
int cfx_mpnn_num_params(void) { return MpnnShape::nParams; }

int cfx_mpnn_create(cfx_ctx* c, uint64_t seed, int fp16, cfx_mpnn** out)
{
    if (!c || !out) return CFX_E_INVALID;
    *out = nullptr;
    return guard(c, [&] {
        auto* m = new cfx_mpnn;
        m->owner = c;
        try
        {
            m->net = std::make_unique<Mpnn>(*c->ctx, seed, fp16 != 0);
            m->trainer = std::make_unique<MpnnTrainer>(*c->ctx, *m->net);
        }
        catch (...)
        {
            delete m;
            throw;
        }
        *out = m;
    });
}

int cfx_mpnn_destroy(cfx_mpnn* m)
{
    if (!m) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        m->owner->ctx->sync();
        delete m;
    });
}

int cfx_mpnn_get_params(cfx_mpnn* m, float* host)
{
    if (!m || !host) return CFX_E_INVALID;
    return guard(m->owner, [&] { m->net->getParams(host); });
}

int cfx_mpnn_set_params(cfx_mpnn* m, const float* host)
{
    if (!m || !host) return CFX_E_INVALID;
    return guard(m->owner, [&] { m->net->setParams(host); });
}

int cfx_mpnn_forward_csr(cfx_mpnn* m, cfx_csr* A, const double* geo, const double* r0, float* a, float* t,
                         int train)
{
    if (!m || !A || A->owner != m->owner) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        m->net->setGraph(A->pattern.view(), A->val.as<double>(), geo, r0);
        m->net->forward(a, t, train != 0);
    });
}

int cfx_mpnn_backward(cfx_mpnn* m, const float* ga, const float* gt, float* grad)
{
    if (!m || !ga || !gt || !grad) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        need(m->net->n() > 0, "no forward pass recorded");
        m->net->backward(ga, gt, grad);
    });
}

int cfx_mpnn_adam(cfx_mpnn* m, const float* grad, float lr)
{
    if (!m || !grad) return CFX_E_INVALID;
    return guard(m->owner, [&] { m->net->adam(grad, lr); });
}

extern "C++" {  // helpers of the entry points: C++ linkage
namespace
{
TrainOptions toTrain(const cfx_train_options* o, const MgParams& mg)
{
    TrainOptions t;
    t.mg = mg;
    if (o)
    {
        t.steps = o->steps;
        t.samples = o->samples;
        t.smooth = o->smooth;
        t.lr = o->lr;
        t.seed = o->seed;
    }
    need(t.steps >= 0 && t.samples >= 1 && t.smooth >= 0 && t.lr > 0, "bad training options");
    return t;
}
} // namespace
} // extern "C++"

int cfx_mpnn_train(cfx_mpnn* m, cfx_csr* const* systems, const double* const* geo, const double* const* r0, int count,
                   const cfx_train_options* opt, double* lossHist)
{
    if (!m || !systems || count <= 0) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        std::vector<TrainSystem> sys(count);
        for (int i = 0; i < count; ++i)
        {
            need(systems[i] && systems[i]->owner == m->owner, "bad training system");
            sys[i].pat = &systems[i]->pattern;
            sys[i].val = systems[i]->val.as<double>();
            sys[i].geo = geo ? geo[i] : nullptr;
            sys[i].r0 = r0 ? r0[i] : nullptr;
        }
        const std::vector<double> h = m->trainer->train(sys, toTrain(opt, systems[0]->mg.p));
        if (lossHist) std::copy(h.begin(), h.end(), lossHist);
    });
}

int cfx_mpnn_evaluate(cfx_mpnn* m, cfx_csr* A, const double* geo, const double* r0, const cfx_train_options* opt,
                      int samples, int withNetwork, double* loss)
{
    if (!m || !A || !loss || samples <= 0) return CFX_E_INVALID;
    return guard(m->owner, [&] {
        TrainSystem s;
        s.pat = &A->pattern;
        s.val = A->val.as<double>();
        s.geo = geo;
        s.r0 = r0;
        *loss = m->trainer->evaluate(s, toTrain(opt, A->mg.p), samples, withNetwork != 0);
    });
}

int cfx_csr_mg_loss_grad(cfx_csr* A, const float* e, const float* a, const float* t, float* ga, float* gt,
                         double* loss)
{
    if (!A || !e || !a || !t || !ga || !gt || !loss) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        const MgParams& p = A->mg.p;
            Multigrid mg(*A->owner->ctx, A->pattern, p.maxLevels, p.coarseMax);
        mg.setMatrix(A->val.as<double>(), a, t, a, t, p);
        *loss = mg.lossGrad(e, a, t, ga, gt);
    });
}

int cfx_csr_set_mpnn(cfx_csr* A, cfx_mpnn* m, const double* geo)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->owner, [&] {
        need(!m || m->owner == A->owner, "network of another context");
        A->net = m;
        A->geo = geo;
        A->netVersion = 0;  // new network or geometry: its features are formed again
        if (!m) A->mg.useWeights(A->mg.userPre, A->mg.userScale, A->mg.userPost, A->mg.userRestr);
    });
}

int cfx_mesh_set_mpnn(cfx_mesh* M, cfx_mpnn* net)
{
    if (!M) return CFX_E_INVALID;
    return guard(M->owner, [&] {
        need(!net || net->owner == M->owner, "network of another context");
        M->net = net;
    });
}

int cfx_matrix_set_mpnn(cfx_matrix* A, cfx_mpnn* m)
{
    if (!A) return CFX_E_INVALID;
    return guard(A->mesh->owner, [&] {
        need(!m || m->owner == A->mesh->owner, "network of another context");
        A->net = m;
        A->solvedVersion = 0;  // the next solve forms the network's features again
        if (!m) A->mg.useWeights(A->mg.userPre, A->mg.userScale, A->mg.userPost, A->mg.userRestr);
    });
}

// ------------------------------------------------------- batched solves
extern "C++" {  // helpers of the entry points: C++ linkage
namespace
{
std::vector<BatchSystem> batchSystems(cfx_ctx* c, int count, cfx_csr* const* systems, const double* const* b,
                                      double* const* x)
{
    need(count > 0 && systems && x, "bad batch");
    std::vector<BatchSystem> v(count);
    for (int k = 0; k < count; ++k)
    {
        need(systems[k] && systems[k]->owner == c, "batch system of another context");
        v[k].A = systems[k]->pattern.view();
        v[k].val = systems[k]->val.as<double>();
        v[k].b = b ? b[k] : nullptr;
        v[k].x = x[k];
    }
    return v;
}
} // namespace
} // extern "C++"

static cfx_batch_schedule toSchedule(const BatchSchedule& b)
{
    return cfx_batch_schedule{b.vecElems, b.wgCap, b.graph, b.pollEvery, b.persistent, b.claim, b.wgs};
}

int cfx_batch_solve(cfx_ctx* c, int count, cfx_csr* const* systems, const double* const* b, double* const* x,
                    double tol, int maxIter, const cfx_batch_schedule* sched, cfx_report* reps, double* finishMs,
                    cfx_batch_stats* stats)
{
    if (!c) return CFX_E_INVALID;
    return guard(c, [&] {
        need(b != nullptr, "null right-hand sides");
        need(maxIter >= 0, "maxIter < 0");
        if (!c->batch) c->batch = std::make_unique<BatchedPcg>(*c->ctx);
        const std::vector<BatchSystem> sys = batchSystems(c, count, systems, b, x);
        BatchSchedule sc;
        if (sched)
        {
            sc.vecElems = sched->vecElems;
            sc.wgCap = sched->wgCap;
            sc.graph = sched->graph;
            sc.pollEvery = sched->pollEvery;
            sc.persistent = sched->persistent;
            sc.claim = sched->claim;
            sc.wgs = sched->wgs;
        }
        else
            sc = c->batch->autotune(sys, 20, nullptr);
        const BatchResult r = c->batch->solve(sys, tol, maxIter, sc);
        for (int k = 0; k < count; ++k)
        {
            if (reps) toReport(r.reports[k], reps + k);
            if (finishMs) finishMs[k] = r.finishMs[k];
        }
        if (stats)
        {
            stats->ms = r.ms;
            stats->maxIterations = r.maxIterations;
            stats->used = toSchedule(sc);
        }
    });
}

int cfx_batch_autotune(cfx_ctx* c, int count, cfx_csr* const* systems, const double* const* b, double* const* x,
                       int iters, cfx_batch_schedule* best, cfx_batch_schedule* cand, double* candMs, int maxCand)
{
    if (!c || !best || !b) return CFX_E_INVALID;
    return guard(c, [&] {
        need(iters > 0, "iters must be positive");
        if (!c->batch) c->batch = std::make_unique<BatchedPcg>(*c->ctx);
        std::vector<BatchSystem> sys = batchSystems(c, count, systems, b, x);
        std::vector<std::pair<BatchSchedule, double>> log;
        const BatchSchedule b = c->batch->autotune(sys, iters, &log);
        *best = toSchedule(b);
        for (int k = 0; k < std::min<int>(maxCand, static_cast<int>(log.size())); ++k)
        {
            if (cand) cand[k] = toSchedule(log[k].first);
            if (candMs) candMs[k] = log[k].second;
        }
    });
}

} // extern "C"

