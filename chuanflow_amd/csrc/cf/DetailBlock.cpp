// DetailBlock.cpp -- the detail region's inlet profile, boundary
// conditions, GPU SIMPLE and section measurement.
#include "DetailBlock.h"

#include <initializer_list>

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace cf
{
namespace
{
double mag(const Vec3& v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
double comp(const Vec3& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
Vec3 faceVec(const std::vector<double>& a, label f) { return {a[3 * f], a[3 * f + 1], a[3 * f + 2]}; }
} // namespace

void DetailBlock::check(int rc) const
{
    if (rc != CFX_OK) throw Error(std::string("detailBlock: ") + cfx_last_error(ctx_));
}

Vec3 DetailBlock::patchArea(label patch) const
{
    const auto& p = mesh_.patches()[static_cast<std::size_t>(patch)];
    Vec3 s;
    for (label f = p.start; f < p.start + p.size; ++f)
    {
        const Vec3 v = faceVec(mesh_.faceSf(), f);
        s.x += v.x;
        s.y += v.y;
        s.z += v.z;
    }
    return s;
}

DetailBlock::DetailBlock(cfx_ctx* ctx, const polyMesh& mesh, double nu, double rho, const std::string& shape)
    : ctx_(ctx), mesh_(mesh), nu_(nu), rho_(rho), shape_(shape)
{
    for (label p = 0; p < mesh_.nPatches(); ++p)
    {
        const std::string& t = mesh_.patches()[static_cast<std::size_t>(p)].type;
        if (t == "inlet" && inletPatch_ < 0) inletPatch_ = p;
        else if (t == "outlet" && outletPatch_ < 0) outletPatch_ = p;
    }
    if (inletPatch_ < 0 || outletPatch_ < 0)
        throw Error("detailBlock: the detailed region needs one patch of type inlet and one of type outlet");
    const Vec3 in = patchArea(inletPatch_), out = patchArea(outletPatch_);
    inletArea_ = mag(in);
    if (!(inletArea_ > 0)) throw Error("detailBlock: the inlet patch has no area");
    if (mag(Vec3{in.x + out.x, in.y + out.y, in.z + out.z}) > 1e-6 * inletArea_)
        throw Error("detailBlock: the inlet and outlet patches are not opposed");
    // the inlet area vector points out of the region: its dominant component is the axis
    const double nx = in.x / inletArea_, ny = in.y / inletArea_, nz = in.z / inletArea_;
    axis_ = 0;
    if (std::abs(ny) > std::abs(nx)) axis_ = 1;
    if (std::abs(nz) > std::abs(axis_ == 0 ? nx : ny)) axis_ = 2;
    flowSign_ = comp(Vec3{nx, ny, nz}, axis_) < 0 ? 1 : -1;
    buildSections();
    buildInletProfile();

    // device mesh and fields (zero velocity, zero pressure)
    const label n = mesh_.nCells(), nF = mesh_.nFaces(), nI = mesh_.nInternalFaces(), nB = nF - nI;
    std::vector<int64_t> own(mesh_.owner().begin(), mesh_.owner().end());
    std::vector<int64_t> nei(mesh_.neighbour().begin(), mesh_.neighbour().end());
    std::vector<int64_t> ps, psz;
    for (const auto& p : mesh_.patches())
    {
        ps.push_back(p.start);
        psz.push_back(p.size);
    }
    check(cfx_mesh_create(ctx_, n, nI, nF, own.data(), nei.data(), mesh_.faceSf().data(), mesh_.faceCentres().data(),
                          mesh_.cellCentres().data(), mesh_.cellVolumes().data(), mesh_.deltaCoeffVec().data(),
                          mesh_.faceAreas().data(), static_cast<int64_t>(ps.size()), ps.data(), psz.data(), &dm_));
    void* p = nullptr;
    check(cfx_malloc(ctx_, sizeof(double) * 3 * n, &p));
    dU_ = static_cast<double*>(p);
    check(cfx_malloc(ctx_, sizeof(double) * n, &p));
    dp_ = static_cast<double*>(p);
    check(cfx_malloc(ctx_, sizeof(double) * std::max<label>(nF, 1), &p));
    dphi_ = static_cast<double*>(p);
    check(cfx_malloc(ctx_, sizeof(double) * 3 * std::max<label>(nB, 1), &p));
    dlc_ = static_cast<double*>(p);
    check(cfx_malloc(ctx_, sizeof(double) * 3 * std::max<label>(nB, 1), &p));
    dlk_ = static_cast<double*>(p);
    check(cfx_zero(ctx_, dU_, sizeof(double) * 3 * n));
    check(cfx_zero(ctx_, dp_, sizeof(double) * n));
    check(cfx_zero(ctx_, dphi_, sizeof(double) * std::max<label>(nF, 1)));

    // what measure() reads, resident on the device
    auto upload = [&](const void* src, std::size_t bytes) {
        void* d = nullptr;
        check(cfx_malloc(ctx_, std::max<std::size_t>(bytes, 8), &d));
        if (bytes) check(cfx_memcpy_htod(ctx_, d, src, bytes));
        return d;
    };
    dV_ = static_cast<double*>(upload(mesh_.cellVolumes().data(), sizeof(double) * n));
    const auto& s0 = sections_.front();
    const auto& s1 = sections_.back();
    std::vector<int64_t> cellIdx(s0.begin(), s0.end());
    cellIdx.insert(cellIdx.end(), s1.begin(), s1.end());
    const int64_t cellSeg[3] = {0, static_cast<int64_t>(s0.size()), static_cast<int64_t>(cellIdx.size())};
    dCellIdx_ = static_cast<int64_t*>(upload(cellIdx.data(), sizeof(int64_t) * cellIdx.size()));
    dCellSeg_ = static_cast<int64_t*>(upload(cellSeg, sizeof(cellSeg)));
    const auto& pin = mesh_.patches()[static_cast<std::size_t>(inletPatch_)];
    const auto& pout = mesh_.patches()[static_cast<std::size_t>(outletPatch_)];
    std::vector<int64_t> faceIdx;
    for (label f = pin.start; f < pin.start + pin.size; ++f) faceIdx.push_back(f);
    for (label f = pout.start; f < pout.start + pout.size; ++f) faceIdx.push_back(f);
    const int64_t faceSeg[3] = {0, static_cast<int64_t>(pin.size), static_cast<int64_t>(faceIdx.size())};
    dFaceIdx_ = static_cast<int64_t*>(upload(faceIdx.data(), sizeof(int64_t) * faceIdx.size()));
    dFaceSeg_ = static_cast<int64_t*>(upload(faceSeg, sizeof(faceSeg)));
}

DetailBlock::~DetailBlock()
{
    for (void* q : std::initializer_list<void*>{dU_, dp_, dphi_, dlc_, dlk_, dV_, dCellSeg_, dCellIdx_, dFaceSeg_,
                                                 dFaceIdx_})
        if (q) (void)cfx_free(ctx_, q);
    if (dm_) (void)cfx_mesh_destroy(dm_);
}

// cross-sections of cells in the order the flow meets them: cells whose
// centre coordinate along the axis lies within 1e-9 of the extent of the
// section's first (coordinate, id)-ordered cell (coupledNetwork.cpp:644-671)
void DetailBlock::buildSections()
{
    const auto& P = mesh_.points();
    double lo = std::numeric_limits<double>::max(), hi = -lo;
    for (std::size_t i = axis_; i < P.size(); i += 3)
    {
        lo = std::min(lo, P[i]);
        hi = std::max(hi, P[i]);
    }
    const double band = 1e-9 * std::max(std::abs(hi - lo), 1e-30);
    const auto& C = mesh_.cellCentres();
    std::vector<label> order(static_cast<std::size_t>(mesh_.nCells()));
    std::iota(order.begin(), order.end(), 0);
    auto coord = [&](label c) { return C[3 * c + axis_]; };
    std::sort(order.begin(), order.end(), [&](label a, label b) {
        const double ka = coord(a), kb = coord(b);
        if (ka != kb) return ka < kb;
        return a < b;
    });
    if (flowSign_ < 0) std::reverse(order.begin(), order.end());
    std::size_t g = 0;
    while (g < order.size())
    {
        const double anchor = coord(order[g]);
        std::size_t h = g + 1;
        while (h < order.size() && std::abs(coord(order[h]) - anchor) <= band) ++h;
        sections_.emplace_back(order.begin() + static_cast<std::ptrdiff_t>(g),
                               order.begin() + static_cast<std::ptrdiff_t>(h));
        g = h;
    }
    if (sections_.size() < 2) throw Error("detailBlock: fewer than two cross-sections to measure a drop between");
}

// inlet face weights over the wall-bounded transverse axes, normalised to
// an area integral equal to the section area (coupledNetwork.cpp:768-876)
void DetailBlock::buildInletProfile()
{
    if (shape_ == "flat") return;
    if (shape_ != "parabolic" && shape_ != "power1over7")
        throw Error("detailBlock: the inlet profile '" + shape_ + "' is not one this block can write");
    int bounds[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    for (label p = 0; p < mesh_.nPatches(); ++p)
    {
        if (mesh_.patches()[static_cast<std::size_t>(p)].type != "wall") continue;
        const Vec3 sf = patchArea(p);
        if (!(mag(sf) > 0)) continue;
        int k = 0;
        for (int c = 1; c < 3; ++c)
            if (std::abs(comp(sf, c)) > std::abs(comp(sf, k))) k = c;
        if (k == axis_) continue;
        bounds[k][comp(sf, k) < 0 ? 0 : 1] = 1;
    }
    std::vector<int> axes;
    for (int k = 0; k < 3; ++k)
        if (k != axis_ && bounds[k][0] && bounds[k][1]) axes.push_back(k);
    if (axes.empty()) throw Error("detailBlock: no wall-confined transverse direction for the inlet profile");
    const auto& P = mesh_.points();
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k)
    {
        lo[k] = std::numeric_limits<double>::max();
        hi[k] = -lo[k];
    }
    for (std::size_t i = 0; i < P.size(); ++i)
    {
        lo[i % 3] = std::min(lo[i % 3], P[i]);
        hi[i % 3] = std::max(hi[i % 3], P[i]);
    }
    for (int k : axes)
        if (!(hi[k] > lo[k])) throw Error("detailBlock: the inlet section has no extent to spread a profile over");
    const auto& pin = mesh_.patches()[static_cast<std::size_t>(inletPatch_)];
    const auto& fc = mesh_.faceCentres();
    std::vector<double> w(static_cast<std::size_t>(pin.size), 0.0);
    double wMin = std::numeric_limits<double>::max();
    for (label j = 0; j < pin.size; ++j)
    {
        const label u = pin.start + j;
        double value = 1.0;
        for (int k : axes)
        {
            const double mid = 0.5 * (hi[k] + lo[k]), half = 0.5 * (hi[k] - lo[k]);
            const double eta = (fc[3 * u + k] - mid) / half;
            value *= shape_ == "parabolic" ? (1.0 - eta * eta) : std::pow(1.0 - std::abs(eta), 1.0 / 7.0);
        }
        w[static_cast<std::size_t>(j)] = value;
        wMin = std::min(wMin, value);
    }
    if (!(wMin > 0)) throw Error("detailBlock: the inlet profile writes a zero or reversed weight");
    double sumWA = 0, sumA = 0;
    for (label j = 0; j < pin.size; ++j)
    {
        const double a = mesh_.faceAreas()[static_cast<std::size_t>(pin.start + j)];
        sumWA += w[static_cast<std::size_t>(j)] * a;
        sumA += a;
    }
    if (!(sumA > 0) || !(sumWA > 0)) throw Error("detailBlock: the inlet section integrates to no area or shape");
    if (std::abs(sumA - inletArea_) > 1e-9 * inletArea_)
        throw Error("detailBlock: the inlet face areas do not add up to the patch's area vector");
    const double norm = sumWA / sumA;
    for (label j = 0; j < pin.size; ++j)
    {
        inletFaces_.push_back(pin.start + j);
        inletWeights_.push_back(w[static_cast<std::size_t>(j)] / norm);
    }
}

// per boundary face conditions (coupledNetwork.cpp:878-926): inlet U fixed
// (into the region), outlet p fixed 0, walls no-slip, anything else
// zeroGradient for both fields
void DetailBlock::applyBoundary(double q, std::vector<int32_t>& Ut, std::vector<double>& Uv,
                                std::vector<int32_t>& pt, std::vector<double>& pv)
{
    const label nI = mesh_.nInternalFaces(), nB = mesh_.nFaces() - nI;
    Ut.assign(static_cast<std::size_t>(nB), 2);
    pt.assign(static_cast<std::size_t>(nB), 2);
    Uv.assign(static_cast<std::size_t>(3 * nB), 0.0);
    pv.assign(static_cast<std::size_t>(nB), 0.0);
    const double speed = flowSign_ * q / inletArea_;
    for (label p = 0; p < mesh_.nPatches(); ++p)
    {
        const auto& pd = mesh_.patches()[static_cast<std::size_t>(p)];
        for (label f = pd.start; f < pd.start + pd.size; ++f)
        {
            const std::size_t b = static_cast<std::size_t>(f - nI);
            if (p == inletPatch_)
            {
                Ut[b] = 0;
                Uv[3 * b + axis_] = speed;
            }
            else if (p == outletPatch_)
                pt[b] = 0;
            else if (pd.type == "wall")
                Ut[b] = 0;
        }
    }
    for (std::size_t j = 0; j < inletFaces_.size(); ++j)
    {
        const std::size_t b = static_cast<std::size_t>(inletFaces_[j] - nI);
        Uv[3 * b + axis_] = speed * inletWeights_[j];
    }
}

double DetailBlock::lossAt(double q, const cfx_simple_controls& ctl)
{
    if (!(q > 0)) throw Error("detailBlock: asked for the drop at a non-positive rate");
    std::vector<int32_t> Ut, pt;
    std::vector<double> Uv, pv;
    applyBoundary(q, Ut, Uv, pt, pv);
    // phi on the boundary from the imposed velocity (phi_.correctBoundaryFlux(U_))
    check(cfx_linearise(dm_, 3, Ut.data(), Uv.data(), dlc_, dlk_));
    check(cfx_correct_boundary_flux(dm_, dU_, dlc_, dlk_, dphi_));
    cfx_schemes sch{0, 1, 1, 0.9};
    cfx_simple_result r{};
    check(cfx_simple_solve(dm_, nu_, &sch, &ctl, Ut.data(), Uv.data(), pt.data(), pv.data(), dU_, dp_, dphi_, &r,
                           nullptr));
    measure(r, q);
    if (!std::isfinite(measure_.dpPa)) throw Error("detailBlock: non-finite pressure drop");
    return measure_.dpPa;
}

// section-mean pressures and boundary fluxes (coupledNetwork.cpp:1154-1200),
// reduced on the device: the fields never leave it
void DetailBlock::measure(const cfx_simple_result& r, double q)
{
    measure_ = BlockMeasure{};
    measure_.cells = mesh_.nCells();
    measure_.sweeps = r.iterations;
    measure_.converged = r.converged;
    measure_.imposedQ = q;
    measure_.inletArea = inletArea_;
    measure_.plugSpeed = q / inletArea_;
    double sec[4], flux[4];
    check(cfx_segment_sums(ctx_, 2, dCellSeg_, dCellIdx_, dp_, dV_, sec));
    check(cfx_segment_sums(ctx_, 2, dFaceSeg_, dFaceIdx_, dphi_, nullptr, flux));
    auto mean = [&](int s) { return rho_ * (sec[2 * s + 1] > 0 ? sec[2 * s] / sec[2 * s + 1] : 0.0); };
    measure_.dpPa = mean(0) - mean(1);
    measure_.fluxIn = -flux[0];
    measure_.fluxOut = flux[2];
}

void DetailBlock::fields(double* U, double* p, double* phi) const
{
    const label n = mesh_.nCells();
    if (U) check(cfx_memcpy_dtoh(ctx_, U, dU_, sizeof(double) * 3 * n));
    if (p) check(cfx_memcpy_dtoh(ctx_, p, dp_, sizeof(double) * n));
    if (phi) check(cfx_memcpy_dtoh(ctx_, phi, dphi_, sizeof(double) * mesh_.nFaces()));
}

} // namespace cf
