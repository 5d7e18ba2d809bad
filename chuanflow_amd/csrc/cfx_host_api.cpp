// cfx_host_api.cpp -- C ABI of the host mesh layer (include/cfx_host.h).
#include "../../include/cfx_host.h"

#include "../../include/cfx.h"
#include "cf/DetailBlock.h"
#include "cf/MeshIO.h"
#include "cf/PolyMesh.h"

#include <memory>

#include <cstring>
#include <vector>
#include <thread>
#include <set>
#include <atomic>
#include <string>

struct cfx_hmesh
{
    cf::polyMesh mesh;
};

struct cfx_detail
{
    std::unique_ptr<cf::DetailBlock> blk;
};

namespace
{
thread_local std::string g_err;

template <typename F>
int guardHost(F&& f)
{
    try
    {
        f();
        return 0;
    }
    catch (const std::invalid_argument& e)
    {
        g_err = e.what();
        return 1;
    }
    catch (const cf::Error& e)
    {
        g_err = e.what();
        return 1;
    }
    catch (const std::exception& e)
    {
        g_err = e.what();
        return 5;
    }
}
} // namespace

extern "C" {

const char* cfx_hmesh_last_error(void) { return g_err.c_str(); }

int cfx_hmesh_box(int64_t nx, int64_t ny, int64_t nz, const double mn[3], const double mx[3], double jitter,
                  uint64_t seed, int tets, cfx_hmesh** out)
{
    return guardHost([&] {
        if (!out || !mn || !mx) throw std::invalid_argument("null argument");
        const cf::Vec3 a{mn[0], mn[1], mn[2]}, b{mx[0], mx[1], mx[2]};
        auto* h = new cfx_hmesh;
        try
        {
            h->mesh = tets ? cf::BlockMesh::createTets(a, b, nx, ny, nz, {}, jitter, seed)
                           : cf::BlockMesh::create(a, b, nx, ny, nz, {}, jitter, seed);
        }
        catch (...)
        {
            delete h;
            throw;
        }
        *out = h;
    });
}

int cfx_hmesh_channel(int64_t nx, int64_t ny, int64_t nz, const double mn[3], const double mx[3], double jitter,
                      uint64_t seed, cfx_hmesh** out)
{
    return guardHost([&] {
        if (!out || !mn || !mx) throw std::invalid_argument("null argument");
        const cf::Vec3 a{mn[0], mn[1], mn[2]}, b{mx[0], mx[1], mx[2]};
        cf::BlockSides sides;
        sides.xmin = {"inlet", "inlet"};
        sides.xmax = {"outlet", "outlet"};
        sides.ymin = {"ymin", "wall"};
        sides.ymax = {"ymax", "wall"};
        sides.zmin = {"zmin", "wall"};
        sides.zmax = {"zmax", "wall"};
        auto* h = new cfx_hmesh;
        try
        {
            h->mesh = cf::BlockMesh::create(a, b, nx, ny, nz, sides, jitter, seed);
        }
        catch (...)
        {
            delete h;
            throw;
        }
        *out = h;
    });
}

int cfx_detail_create(cfx_ctx* ctx, const cfx_hmesh* mesh, double nu, double rho, const char* shape,
                      cfx_detail** out)
{
    return guardHost([&] {
        if (!ctx || !mesh || !shape || !out) throw std::invalid_argument("null argument");
        auto* d = new cfx_detail;
        try
        {
            d->blk = std::make_unique<cf::DetailBlock>(ctx, mesh->mesh, nu, rho, shape);
        }
        catch (...)
        {
            delete d;
            throw;
        }
        *out = d;
    });
}

int cfx_detail_destroy(cfx_detail* d)
{
    delete d;
    return 0;
}

int cfx_detail_loss_at(cfx_detail* d, double q, const cfx_simple_controls* ctl, cfx_detail_measure* m)
{
    return guardHost([&] {
        if (!d || !ctl) throw std::invalid_argument("null argument");
        d->blk->lossAt(q, *ctl);
        if (m)
        {
            const cf::BlockMeasure& b = d->blk->lastMeasure();
            *m = cfx_detail_measure{b.cells, b.sweeps, b.converged, b.imposedQ, b.inletArea, b.plugSpeed,
                                    b.dpPa, b.fluxIn, b.fluxOut};
        }
    });
}

int cfx_detail_loss_at_many(int count, cfx_detail* const* blocks, const double* q, const cfx_simple_controls* ctl,
                            int threads, cfx_detail_measure* m)
{
    return guardHost([&] {
        if (count < 0 || (count > 0 && (!blocks || !q || !ctl))) throw std::invalid_argument("null argument");
        if (threads < 1) throw std::invalid_argument("threads must be >= 1");
        std::set<cfx_ctx*> ctxs;
        for (int i = 0; i < count; ++i)
        {
            if (!blocks[i]) throw std::invalid_argument("null block");
            ctxs.insert(blocks[i]->blk->context());
        }
        if (threads > 1 && static_cast<int>(ctxs.size()) != count)
            throw std::invalid_argument("concurrent lossAt needs one context (stream) per block");
        std::atomic<int> next{0};
        std::vector<std::string> errs(static_cast<std::size_t>(count));
        auto work = [&] {
            for (int i; (i = next++) < count;)
            {
                try
                {
                    blocks[i]->blk->lossAt(q[i], *ctl);
                    if (m)
                    {
                        const cf::BlockMeasure& b = blocks[i]->blk->lastMeasure();
                        m[i] = cfx_detail_measure{b.cells, b.sweeps, b.converged, b.imposedQ, b.inletArea,
                                                  b.plugSpeed, b.dpPa, b.fluxIn, b.fluxOut};
                    }
                }
                catch (const std::exception& e)
                {
                    errs[static_cast<std::size_t>(i)] = e.what();
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < std::min(threads, count); ++t) pool.emplace_back(work);
        work();
        for (auto& t : pool) t.join();
        for (int i = 0; i < count; ++i)
            if (!errs[static_cast<std::size_t>(i)].empty())
                throw cf::Error("block " + std::to_string(i) + ": " + errs[static_cast<std::size_t>(i)]);
    });
}

int cfx_detail_info(const cfx_detail* d, int* axis, int* flowSign, double* inletArea, int64_t* nSections,
                    int64_t* nInletFaces)
{
    return guardHost([&] {
        if (!d) throw std::invalid_argument("null argument");
        if (axis) *axis = d->blk->axis();
        if (flowSign) *flowSign = d->blk->flowSign();
        if (inletArea) *inletArea = d->blk->inletArea();
        if (nSections) *nSections = static_cast<int64_t>(d->blk->nSections());
        if (nInletFaces) *nInletFaces = static_cast<int64_t>(d->blk->inletFaces().size());
    });
}

int cfx_detail_inlet(const cfx_detail* d, int64_t* faces, double* weights)
{
    return guardHost([&] {
        if (!d) throw std::invalid_argument("null argument");
        const auto& f = d->blk->inletFaces();
        const auto& w = d->blk->inletWeights();
        for (std::size_t j = 0; j < f.size(); ++j)
        {
            if (faces) faces[j] = f[j];
            if (weights) weights[j] = w[j];
        }
    });
}

int cfx_detail_fields(const cfx_detail* d, double* U, double* p, double* phi)
{
    return guardHost([&] {
        if (!d) throw std::invalid_argument("null argument");
        d->blk->fields(U, p, phi);
    });
}

cfx_mesh* cfx_detail_mesh(const cfx_detail* d) { return d ? d->blk->deviceMesh() : nullptr; }

int cfx_hmesh_obstacle_channel(int64_t nx, int64_t ny, int64_t nz, const double mn[3], const double mx[3],
                               double jitter, uint64_t seed, double cx, double cy, double radius, cfx_hmesh** out)
{
    return guardHost([&] {
        if (!out || !mn || !mx) throw std::invalid_argument("null argument");
        if (!(radius > 0)) throw std::invalid_argument("radius must be positive");
        const cf::Vec3 a{mn[0], mn[1], mn[2]}, b{mx[0], mx[1], mx[2]};
        auto* h = new cfx_hmesh;
        try
        {
            h->mesh = cf::obstacleChannel(a, b, nx, ny, nz, jitter, seed, cx, cy, radius);
        }
        catch (...)
        {
            delete h;
            throw;
        }
        *out = h;
    });
}

int cfx_hmesh_from_arrays(int64_t nPoints, const double* points, int64_t nFaces, const int64_t* faceStart,
                          const int64_t* faceVerts, int64_t nInternal, const int64_t* owner,
                          const int64_t* neighbour, int64_t nPatches, const int64_t* patchStart,
                          const int64_t* patchSize, cfx_hmesh** out)
{
    return guardHost([&] {
        if (!out || !points || !faceStart || !faceVerts || !owner) throw std::invalid_argument("null argument");
        std::vector<double> p(points, points + 3 * nPoints);
        std::vector<cf::label> fs(faceStart, faceStart + nFaces + 1), fv(faceVerts, faceVerts + faceStart[nFaces]),
            ow(owner, owner + nFaces), nb(neighbour, neighbour + nInternal);
        for (cf::label v : fv)
            if (v < 0 || v >= nPoints) throw std::invalid_argument("face vertex out of range");
        std::vector<cf::polyMesh::Patch> patches;
        for (int64_t i = 0; i < nPatches; ++i)
            patches.push_back({"patch" + std::to_string(i), "patch", patchStart[i], patchSize[i]});
        auto* h = new cfx_hmesh;
        try
        {
            h->mesh = cf::polyMesh(std::move(p), std::move(fs), std::move(fv), std::move(ow), std::move(nb),
                                   std::move(patches));
        }
        catch (...)
        {
            delete h;
            throw;
        }
        *out = h;
    });
}

int cfx_hmesh_free(cfx_hmesh* h)
{
    delete h;
    return 0;
}

int cfx_hmesh_sizes(const cfx_hmesh* h, int64_t s[6])
{
    return guardHost([&] {
        if (!h) throw std::invalid_argument("null mesh");
        const cf::polyMesh& m = h->mesh;
        s[0] = m.nPoints();
        s[1] = m.nFaces();
        s[2] = m.nInternalFaces();
        s[3] = m.nCells();
        s[4] = m.nPatches();
        s[5] = static_cast<int64_t>(m.faceVerts().size());
    });
}

int cfx_hmesh_get(const cfx_hmesh* h, double* points, int64_t* faceStart, int64_t* faceVerts, int64_t* owner,
                  int64_t* neighbour, int64_t* patchStart, int64_t* patchSize)
{
    return guardHost([&] {
        if (!h) throw std::invalid_argument("null mesh");
        const cf::polyMesh& m = h->mesh;
        auto cp = [](auto* dst, const auto& v) {
            if (dst && !v.empty()) std::memcpy(dst, v.data(), sizeof(v[0]) * v.size());
        };
        cp(points, m.points());
        cp(faceStart, m.faceStart());
        cp(faceVerts, m.faceVerts());
        cp(owner, m.owner());
        cp(neighbour, m.neighbour());
        for (cf::label i = 0; i < m.nPatches(); ++i)
        {
            if (patchStart) patchStart[i] = m.patches()[i].start;
            if (patchSize) patchSize[i] = m.patches()[i].size;
        }
    });
}

int cfx_hmesh_geometry(const cfx_hmesh* h, double* Sf, double* Cf, double* magSf, double* C, double* V, double* d,
                       double* invDist)
{
    return guardHost([&] {
        if (!h) throw std::invalid_argument("null mesh");
        const cf::polyMesh& m = h->mesh;
        auto cp = [](double* dst, const std::vector<double>& v) {
            if (dst && !v.empty()) std::memcpy(dst, v.data(), sizeof(double) * v.size());
        };
        cp(Sf, m.faceSf());
        cp(Cf, m.faceCentres());
        cp(magSf, m.faceAreas());
        cp(C, m.cellCentres());
        cp(V, m.cellVolumes());
        cp(d, m.deltaCoeffVec());
        cp(invDist, m.deltaCoeff());
    });
}

int cfx_hmesh_shuffle(cfx_hmesh* h, uint64_t seed)
{
    return guardHost([&] {
        if (!h) throw std::invalid_argument("null mesh");
        h->mesh.shuffle(seed);
    });
}

int cfx_hmesh_renumber_rcm(cfx_hmesh* h, int64_t* newId)
{
    return guardHost([&] {
        if (!h) throw std::invalid_argument("null mesh");
        const std::vector<cf::label> id = h->mesh.rcmOrdering();
        h->mesh.renumberCells(id);
        if (newId) std::memcpy(newId, id.data(), sizeof(int64_t) * id.size());
    });
}

int cfx_hmesh_bandwidth(const cfx_hmesh* h, int64_t* bw)
{
    return guardHost([&] {
        if (!h || !bw) throw std::invalid_argument("null argument");
        *bw = h->mesh.bandwidth();
    });
}


// ------------------------------------------------------------ mesh files
int cfx_hmesh_read_openfoam(const char* dir, const char* patchTypes, cfx_hmesh** out)
{
    return guardHost([&] {
        if (!dir || !out) throw std::invalid_argument("null argument");
        *out = new cfx_hmesh{cf::readOpenFoam(dir, cf::parsePatchTypes(patchTypes ? patchTypes : ""))};
    });
}

int cfx_hmesh_read_gmsh(const char* path, const char* patchTypes, cfx_hmesh** out)
{
    return guardHost([&] {
        if (!path || !out) throw std::invalid_argument("null argument");
        *out = new cfx_hmesh{cf::readGmsh(path, cf::parsePatchTypes(patchTypes ? patchTypes : ""))};
    });
}

int cfx_hmesh_patch(const cfx_hmesh* h, int64_t i, char* name, int64_t nameCap, char* type, int64_t typeCap)
{
    return guardHost([&] {
        if (!h) throw std::invalid_argument("null mesh");
        if (i < 0 || i >= h->mesh.nPatches()) throw std::invalid_argument("patch index out of range");
        const auto& p = h->mesh.patches()[static_cast<std::size_t>(i)];
        auto copy = [](const std::string& s, char* dst, int64_t cap) {
            if (!dst) return;
            if (cap <= static_cast<int64_t>(s.size())) throw std::invalid_argument("patch string buffer too small");
            std::memcpy(dst, s.c_str(), s.size() + 1);
        };
        copy(p.name, name, nameCap);
        copy(p.type, type, typeCap);
    });
}

int cfx_vtu_write(const cfx_hmesh* h, const char* path, int nFields, const char* const* names,
                  const int* components, const double* const* data)
{
    return guardHost([&] {
        if (!h || !path) throw std::invalid_argument("null argument");
        if (nFields < 0 || (nFields > 0 && (!names || !components || !data)))
            throw std::invalid_argument("bad field arrays");
        std::vector<cf::VtuField> f;
        for (int k = 0; k < nFields; ++k) f.push_back({names[k] ? names[k] : "", components[k], data[k]});
        cf::writeVtu(path, h->mesh, f);
    });
}

} // extern "C"
