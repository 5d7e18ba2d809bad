// mesh.hip -- device mesh construction and the execution context.
#include "mesh.h"
#include "noalloc.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace cfx
{

Context::Context(int device) : device_(device)
{
    int count = 0;
    CFX_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) throw HipError("cfx: device " + std::to_string(device) + " not present");
    CFX_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    CFX_HIP(hipGetDeviceProperties(&prop, device));
    numCUs_ = prop.multiProcessorCount;
    deviceName = prop.gcnArchName;
    // default priority, non-blocking w.r.t. the null stream
    CFX_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    partials_.alloc(sizeof(double) * 4 * kMaxPartials);
    ipartials_.alloc(sizeof(int) * kMaxPartials);
    tickets_.alloc(sizeof(double) * 64);
    CFX_HIP(hipMemsetAsync(tickets_.p, 0, sizeof(double) * 64, stream_));
    CFX_HIP(hipHostMalloc(reinterpret_cast<void**>(&hostState_), 4 * sizeof(SolveState), hipHostMallocDefault));
    CFX_HIP(hipStreamSynchronize(stream_));
    noallocSetup(device);  // the no-allocation solve's scratch address on this device (device current)
}

Context::~Context()
{
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (hostState_) (void)hipHostFree(hostState_);
    partials_.release();
    ipartials_.release();
    tickets_.release();
    if (stream_) (void)hipStreamDestroy(stream_);
}

namespace
{
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

template <typename T>
std::vector<double> soa3(const T* aos, long long n)
{
    std::vector<double> out(static_cast<std::size_t>(3 * n));
    for (long long i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) out[c * n + i] = aos[3 * i + c];
    return out;
}
} // namespace

std::unique_ptr<DevMesh> createDevMesh(Context& ctx, const MeshDesc& d)
{
    const long long n = d.nCells, nI = d.nInternal, nF = d.nFaces;
    if (n <= 0 || nF <= 0 || nI < 0 || nI > nF) throw std::invalid_argument("mesh: bad sizes");
    if (n >= (1LL << 30) || nF >= (1LL << 30)) throw std::invalid_argument("mesh: too large for int32 face streams");
    auto m = std::make_unique<DevMesh>();
    m->ctx = &ctx;
    m->n = static_cast<int>(n);
    m->nI = static_cast<int>(nI);
    m->nF = static_cast<int>(nF);
    m->nB = static_cast<int>(nF - nI);
    m->nPatches = static_cast<int>(d.nPatches);
    m->patchStart.assign(d.patchStart, d.patchStart + d.nPatches);
    m->patchSize.assign(d.patchSize, d.patchSize + d.nPatches);

    std::vector<int> own(nF), nbr(nI);
    for (long long f = 0; f < nF; ++f)
    {
        if (d.owner[f] < 0 || d.owner[f] >= n) throw std::invalid_argument("mesh: owner out of range");
        own[f] = static_cast<int>(d.owner[f]);
    }
    for (long long f = 0; f < nI; ++f)
    {
        if (d.neighbour[f] < 0 || d.neighbour[f] >= n || d.neighbour[f] == d.owner[f])
            throw std::invalid_argument("mesh: neighbour out of range");
        nbr[f] = static_cast<int>(d.neighbour[f]);
    }

    // geometry-derived face coefficients, in the reference's operation order
    std::vector<double> ow(std::max<long long>(nI, 1), 0.5), sfd(nF), d2(nF), skew(3 * std::max<long long>(nI, 1)),
        kf(3 * nF), scale(n, 0.0);
    for (long long f = 0; f < nF; ++f)
    {
        const double* dv = d.d + 3 * f;
        const double* sf = d.Sf + 3 * f;
        sfd[f] = dot3(sf, dv);
        d2[f] = dot3(dv, dv);
        // cfVec = Sf - dh (Sf . dh), dh = d / |d|   (laplacianScheme.cpp:141-145)
        const double dm = std::sqrt(d2[f]);
        const double dh[3] = {dv[0] / dm, dv[1] / dm, dv[2] / dm};
        const double s = dot3(sf, dh);
        for (int c = 0; c < 3; ++c) kf[c * nF + f] = sf[c] - dh[c] * s;
        // local face scale (pressureCorrection.cpp:103-116)
        const double a = d.magSf[f];
        const long long o = d.owner[f];
        scale[o] = (scale[o] < a) ? a : scale[o];
        if (f < nI)
        {
            const long long nb = d.neighbour[f];
            scale[nb] = (scale[nb] < a) ? a : scale[nb];
        }
    }
    for (long long f = 0; f < nI; ++f)
    {
        // ownerWeight (interpolate.cpp:15-24)
        const double* cf = d.Cf + 3 * f;
        const double* co = d.C + 3 * d.owner[f];
        const double* cn = d.C + 3 * d.neighbour[f];
        const double e[3] = {cf[0] - co[0], cf[1] - co[1], cf[2] - co[2]};
        double nw = dot3(e, d.d + 3 * f) / d2[f];
        nw = nw < 0.0 ? 0.0 : (nw > 1.0 ? 1.0 : nw);
        ow[f] = 1.0 - nw;
        // makeFaceSkewVectors (interpolate.cpp:37-47)
        const double w = 1.0 - ow[f];
        for (int c = 0; c < 3; ++c) skew[c * nI + f] = cf[c] - (co[c] + (cn[c] - co[c]) * w);
    }

    // cell -> faces, ascending face order, neighbour side flagged in bit 0
    std::vector<int> cfs(n + 1, 0), cfl(2 * nI + (nF - nI));
    for (long long f = 0; f < nF; ++f)
    {
        ++cfs[own[f] + 1];
        if (f < nI) ++cfs[nbr[f] + 1];
    }
    for (long long c = 0; c < n; ++c) cfs[c + 1] += cfs[c];
    {
        std::vector<int> fill(cfs.begin(), cfs.end() - 1);
        for (long long f = 0; f < nF; ++f)
        {
            cfl[fill[own[f]]++] = static_cast<int>(f << 1);
            if (f < nI) cfl[fill[nbr[f]]++] = static_cast<int>((f << 1) | 1);
        }
    }

    // CSR pattern: row r holds its diagonal and one entry per internal face,
    // sorted by column; slotU/slotL map each face to its two entries.
    std::vector<int> rowPtr(n + 1, 0);
    for (long long f = 0; f < nI; ++f) { ++rowPtr[own[f] + 1]; ++rowPtr[nbr[f] + 1]; }
    for (long long r = 0; r < n; ++r) rowPtr[r + 1] += rowPtr[r] + 1;
    const long long nnz = rowPtr[n];
    std::vector<int> col(nnz);
    std::vector<long long> ref(nnz);  // -1 diag, 2f upper, 2f+1 lower
    {
        std::vector<int> fill(rowPtr.begin(), rowPtr.end() - 1);
        for (long long r = 0; r < n; ++r) { col[fill[r]] = static_cast<int>(r); ref[fill[r]] = -1; ++fill[r]; }
        for (long long f = 0; f < nI; ++f)
        {
            col[fill[own[f]]] = nbr[f]; ref[fill[own[f]]] = 2 * f; ++fill[own[f]];
            col[fill[nbr[f]]] = own[f]; ref[fill[nbr[f]]] = 2 * f + 1; ++fill[nbr[f]];
        }
    }
    std::vector<int> slotU(std::max<long long>(nI, 1)), slotL(std::max<long long>(nI, 1));
    std::vector<std::pair<int, long long>> row;
    for (long long r = 0; r < n; ++r)
    {
        row.clear();
        for (int k = rowPtr[r]; k < rowPtr[r + 1]; ++k) row.emplace_back(col[k], ref[k]);
        std::sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (int k = rowPtr[r], i = 0; k < rowPtr[r + 1]; ++k, ++i)
        {
            if (i > 0 && row[i].first == row[i - 1].first)
                throw std::invalid_argument("mesh: two faces join the same pair of cells");
            col[k] = row[i].first;
            const long long q = row[i].second;
            if (q >= 0) ((q & 1) ? slotL : slotU)[q >> 1] = k;
        }
    }

    hipStream_t s = ctx.stream();
    m->owner = upload(own, s);
    m->neighbour = upload(nbr, s);
    m->cellFaceStart = upload(cfs, s);
    m->cellFace = upload(cfl, s);
    m->Sf = upload(soa3(d.Sf, nF), s);
    m->Cf = upload(soa3(d.Cf, nF), s);
    m->d = upload(soa3(d.d, nF), s);
    m->C = upload(soa3(d.C, n), s);
    m->hV.assign(d.V, d.V + n);
    m->V = upload(m->hV, s);
    m->magSf = upload(std::vector<double>(d.magSf, d.magSf + nF), s);
    m->localScale = upload(scale, s);
    m->ow = upload(ow, s);
    m->sfd = upload(sfd, s);
    m->d2 = upload(d2, s);
    m->skew = upload(skew, s);
    m->kf = upload(kf, s);
    m->slotU = upload(slotU, s);
    m->slotL = upload(slotL, s);
    m->pattern.build(std::move(rowPtr), std::move(col), s);
    ctx.sync();
    return m;
}

} // namespace cfx
