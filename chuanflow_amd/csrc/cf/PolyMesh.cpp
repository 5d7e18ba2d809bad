// PolyMesh.cpp -- host mesh topology, geometry and generators.
#include "PolyMesh.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>

namespace cf
{
namespace
{

// Component-wise vector arithmetic, evaluated left to right exactly as the
// reference Vector3 does (dot = (xx + yy) + zz, s*v == v*s), so geometry
// matches the reference bit for bit.
struct V
{
    double x, y, z;
};
inline V ld(const double* p) { return {p[0], p[1], p[2]}; }
inline void st(double* p, V a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator*(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V operator/(V a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double mag(V a) { return std::sqrt(dot(a, a)); }
inline bool same(V a, V b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

} // namespace

polyMesh::polyMesh(std::vector<double> points, std::vector<label> faceStart,
                   std::vector<label> faceVerts, std::vector<label> owner,
                   std::vector<label> neighbour, std::vector<Patch> patches)
    : points_(std::move(points)), faceStart_(std::move(faceStart)), faceVerts_(std::move(faceVerts)),
      owner_(std::move(owner)), neighbour_(std::move(neighbour)), patches_(std::move(patches))
{
    if (faceStart_.size() != owner_.size() + 1) CF_FATAL("polyMesh: faceStart size != nFaces + 1");
    if (neighbour_.size() > owner_.size()) CF_FATAL("polyMesh: more neighbours than faces");
    for (label o : owner_) nCells_ = std::max(nCells_, o + 1);
    for (label n : neighbour_) nCells_ = std::max(nCells_, n + 1);
    for (label f = 0; f < nInternalFaces(); ++f)
        if (owner_[f] == neighbour_[f]) CF_FATAL("polyMesh: face " + std::to_string(f) + " owns itself");
    facePatchMap_.assign(static_cast<std::size_t>(nFaces()), -1);
    for (label i = 0; i < nPatches(); ++i)
    {
        const Patch& p = patches_[i];
        for (label f = p.start; f < p.start + p.size; ++f)
        {
            if (f < nInternalFaces() || f >= nFaces())
                CF_FATAL("polyMesh: patch '" + p.name + "' covers face " + std::to_string(f));
            if (facePatchMap_[f] >= 0)
                CF_FATAL("polyMesh: boundary face " + std::to_string(f) + " in two patches");
            facePatchMap_[f] = i;
        }
    }
    for (label f = nInternalFaces(); f < nFaces(); ++f)
        if (facePatchMap_[f] < 0) CF_FATAL("polyMesh: boundary face " + std::to_string(f) + " is not in any patch");
    calcGeometry();
}

label polyMesh::patchIdByName(const std::string& name) const
{
    for (label i = 0; i < nPatches(); ++i)
        if (patches_[i].name == name) return i;
    CF_FATAL("polyMesh: no patch named '" + name + "'");
}

// Geometry in the operation order of polyMesh::calcGeometry
// (src/3d/src/mesh/polyMesh.cpp:83-288).
void polyMesh::calcGeometry()
{
    const label nF = nFaces(), nI = nInternalFaces(), nC = nCells_;
    const double* P = points_.data();
    Sf_.assign(3 * nF, 0.0);
    Cf_.assign(3 * nF, 0.0);
    for (label f = 0; f < nF; ++f)
    {
        const label* vs = faceVerts_.data() + faceStart_[f];
        const label nv = faceStart_[f + 1] - faceStart_[f];
        if (nv < 3) CF_FATAL("polyMesh: face " + std::to_string(f) + " has fewer than 3 vertices");
        const V v0 = ld(P + 3 * vs[0]);
        V n{0, 0, 0};
        for (label i = 1; i + 1 < nv; ++i) n = n + cross(ld(P + 3 * vs[i]) - v0, ld(P + 3 * vs[i + 1]) - v0);
        st(&Sf_[3 * f], n * 0.5);
        const bool exactMean =
            nv == 3 || (nv == 4 && same(ld(P + 3 * vs[0]) + ld(P + 3 * vs[2]), ld(P + 3 * vs[1]) + ld(P + 3 * vs[3])));
        V c{0, 0, 0};
        const double wSum = dot(n, n);
        if (exactMean || !(wSum > 0))
        {
            for (label i = 0; i < nv; ++i) c = c + ld(P + 3 * vs[i]);
            c = c / static_cast<double>(nv);
        }
        else
        {
            for (label i = 1; i + 1 < nv; ++i)
            {
                const V vi = ld(P + 3 * vs[i]), vj = ld(P + 3 * vs[i + 1]);
                const double w = dot(cross(vi - v0, vj - v0), n);
                c = c + ((v0 + vi) + vj) * w;
            }
            c = c / (3.0 * wSum);
        }
        st(&Cf_[3 * f], c);
    }

    std::vector<V> est(static_cast<std::size_t>(nC), V{0, 0, 0});
    std::vector<label> cnt(static_cast<std::size_t>(nC), 0);
    for (label f = 0; f < nF; ++f)
    {
        const label o = owner_[f];
        est[o] = est[o] + ld(&Cf_[3 * f]);
        ++cnt[o];
        if (f < nI)
        {
            const label nb = neighbour_[f];
            est[nb] = est[nb] + ld(&Cf_[3 * f]);
            ++cnt[nb];
        }
    }
    for (label c = 0; c < nC; ++c)
        if (cnt[c] > 0) est[c] = est[c] / static_cast<double>(cnt[c]);

    for (label f = 0; f < nF; ++f)
    {
        const label o = owner_[f];
        const V d = f < nI ? est[neighbour_[f]] - est[o] : ld(&Cf_[3 * f]) - est[o];
        const V sf = ld(&Sf_[3 * f]);
        const double side = dot(sf, d);
        const double tol = 1e-12 * mag(sf) * mag(d);
        if (side < -tol) st(&Sf_[3 * f], V{-sf.x, -sf.y, -sf.z});
    }
    magSf_.resize(nF);
    for (label f = 0; f < nF; ++f) magSf_[f] = mag(ld(&Sf_[3 * f]));

    V_.assign(nC, 0.0);
    std::vector<V> moment(static_cast<std::size_t>(nC), V{0, 0, 0});
    for (label f = 0; f < nF; ++f)
    {
        const label* vs = faceVerts_.data() + faceStart_[f];
        const label nv = faceStart_[f + 1] - faceStart_[f];
        const V v0 = ld(P + 3 * vs[0]);
        double fv = 0;
        V fm{0, 0, 0};
        V raw{0, 0, 0};
        for (label i = 1; i + 1 < nv; ++i)
        {
            const V pi = ld(P + 3 * vs[i]), pj = ld(P + 3 * vs[i + 1]);
            const double tV = dot(v0, cross(pi - v0, pj - v0)) / 6.0;
            fv += tV;
            fm = fm + (((v0 + pi) + pj) / 4.0) * tV;
        }
        for (label i = 1; i + 1 < nv; ++i)
            raw = raw + cross(ld(P + 3 * vs[i]) - v0, ld(P + 3 * vs[i + 1]) - v0);
        const double s = dot(raw, ld(&Sf_[3 * f])) >= 0 ? 1.0 : -1.0;
        const label o = owner_[f];
        V_[o] += s * fv;
        moment[o] = moment[o] + fm * s;
        if (f < nI)
        {
            const label nb = neighbour_[f];
            V_[nb] -= s * fv;
            moment[nb] = moment[nb] - fm * s;
        }
    }
    C_.assign(3 * nC, 0.0);
    for (label c = 0; c < nC; ++c)
    {
        st(&C_[3 * c], moment[c] / V_[c]);
        V_[c] = std::abs(V_[c]);
        if (!(V_[c] > 0)) CF_FATAL("polyMesh: degenerate cell volume for cell " + std::to_string(c));
    }

    d_.assign(3 * nF, 0.0);
    invDist_.assign(nF, 0.0);
    for (label f = 0; f < nF; ++f)
    {
        const label o = owner_[f];
        const V d = f < nI ? ld(&C_[3 * neighbour_[f]]) - ld(&C_[3 * o]) : ld(&Cf_[3 * f]) - ld(&C_[3 * o]);
        st(&d_[3 * f], d);
        const double dm = mag(d);
        if (!(dm > 0)) CF_FATAL("polyMesh: zero owner-neighbour distance at face " + std::to_string(f));
        invDist_[f] = 1.0 / dm;
    }
}

label polyMesh::bandwidth() const
{
    label bw = 0;
    for (label f = 0; f < nInternalFaces(); ++f) bw = std::max(bw, std::abs(neighbour_[f] - owner_[f]));
    return bw;
}

void polyMesh::renumberCells(const std::vector<label>& newId)
{
    if (static_cast<label>(newId.size()) != nCells_) CF_FATAL("renumberCells: permutation size");
    {
        std::vector<char> seen(static_cast<std::size_t>(nCells_), 0);
        for (label v : newId)
        {
            if (v < 0 || v >= nCells_ || seen[v]) CF_FATAL("renumberCells: not a permutation");
            seen[v] = 1;
        }
    }
    const label nI = nInternalFaces(), nF = nFaces();
    // internal faces: (owner, neighbour) after relabelling, sorted
    std::vector<label> order(static_cast<std::size_t>(nI));
    std::vector<label> o2(nF), n2(nI);
    for (label f = 0; f < nI; ++f)
    {
        const label a = newId[owner_[f]], b = newId[neighbour_[f]];
        o2[f] = std::min(a, b);
        n2[f] = std::max(a, b);
    }
    for (label f = nI; f < nF; ++f) o2[f] = newId[owner_[f]];
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](label a, label b) {
        return o2[a] != o2[b] ? o2[a] < o2[b] : n2[a] < n2[b];
    });
    std::vector<label> faceOrder(order);
    for (const Patch& p : patches_)
    {
        std::vector<label> po(static_cast<std::size_t>(p.size));
        std::iota(po.begin(), po.end(), p.start);
        std::stable_sort(po.begin(), po.end(), [&](label a, label b) { return o2[a] < o2[b]; });
        faceOrder.insert(faceOrder.end(), po.begin(), po.end());
    }
    std::vector<label> fs(1, 0), fv, ow, nb;
    fs.reserve(nF + 1);
    fv.reserve(faceVerts_.size());
    ow.reserve(nF);
    nb.reserve(nI);
    for (label k = 0; k < nF; ++k)
    {
        const label f = faceOrder[k];
        fv.insert(fv.end(), faceVerts_.begin() + faceStart_[f], faceVerts_.begin() + faceStart_[f + 1]);
        fs.push_back(static_cast<label>(fv.size()));
        ow.push_back(o2[f]);
        if (k < nI) nb.push_back(n2[f]);
    }
    faceStart_.swap(fs);
    faceVerts_.swap(fv);
    owner_.swap(ow);
    neighbour_.swap(nb);
    calcGeometry();
}

std::vector<label> polyMesh::rcmOrdering() const
{
    const label n = nCells_, nI = nInternalFaces();
    std::vector<label> deg(static_cast<std::size_t>(n) + 1, 0);
    for (label f = 0; f < nI; ++f) { ++deg[owner_[f] + 1]; ++deg[neighbour_[f] + 1]; }
    std::vector<label> start(deg);
    for (label i = 0; i < n; ++i) start[i + 1] += start[i];
    std::vector<label> adj(static_cast<std::size_t>(start[n]));
    {
        std::vector<label> fill(start.begin(), start.end() - 1);
        for (label f = 0; f < nI; ++f)
        {
            adj[fill[owner_[f]]++] = neighbour_[f];
            adj[fill[neighbour_[f]]++] = owner_[f];
        }
    }
    auto degree = [&](label c) { return start[c + 1] - start[c]; };
    // neighbour lists sorted by (degree, id) once
    for (label c = 0; c < n; ++c)
        std::sort(adj.begin() + start[c], adj.begin() + start[c + 1], [&](label a, label b) {
            return degree(a) != degree(b) ? degree(a) < degree(b) : a < b;
        });

    std::vector<label> order;
    order.reserve(n);
    std::vector<char> visited(static_cast<std::size_t>(n), 0);
    std::vector<label> level(static_cast<std::size_t>(n), -1);

    // BFS returning (last level width, eccentricity, a min-degree node of the last level)
    auto bfsLevels = [&](label root, label& ecc, label& far) {
        std::vector<label> q{root};
        std::fill(level.begin(), level.end(), -1);
        level[root] = 0;
        for (std::size_t h = 0; h < q.size(); ++h)
        {
            const label c = q[h];
            for (label k = start[c]; k < start[c + 1]; ++k)
                if (level[adj[k]] < 0) { level[adj[k]] = level[c] + 1; q.push_back(adj[k]); }
        }
        ecc = level[q.back()];
        far = q.back();
        for (label c : q)
            if (level[c] == ecc && (degree(c) < degree(far) || (degree(c) == degree(far) && c < far))) far = c;
    };

    for (label seed = 0; seed < n; ++seed)
    {
        if (visited[seed]) continue;
        // component: pick min-degree node, then George-Liu pseudo-peripheral iteration
        std::vector<label> comp{seed};
        visited[seed] = 1;
        for (std::size_t h = 0; h < comp.size(); ++h)
            for (label k = start[comp[h]]; k < start[comp[h] + 1]; ++k)
                if (!visited[adj[k]]) { visited[adj[k]] = 1; comp.push_back(adj[k]); }
        label root = comp[0];
        for (label c : comp)
            if (degree(c) < degree(root) || (degree(c) == degree(root) && c < root)) root = c;
        label ecc = 0, far = root;
        bfsLevels(root, ecc, far);
        for (int it = 0; it < 8; ++it)
        {
            label ecc2 = 0, far2 = far;
            bfsLevels(far, ecc2, far2);
            if (ecc2 <= ecc) break;
            root = far;
            ecc = ecc2;
            far = far2;
        }
        // Cuthill-McKee from root
        const std::size_t base = order.size();
        std::vector<char> inq(0);
        order.push_back(root);
        level[root] = -2;  // reuse as "queued" mark
        for (label c : comp) if (c != root) level[c] = -1;
        for (std::size_t h = base; h < order.size(); ++h)
        {
            const label c = order[h];
            for (label k = start[c]; k < start[c + 1]; ++k)
                if (level[adj[k]] != -2) { level[adj[k]] = -2; order.push_back(adj[k]); }
        }
        std::reverse(order.begin() + static_cast<std::ptrdiff_t>(base), order.end());
    }
    std::vector<label> newId(static_cast<std::size_t>(n));
    // components were reversed in place; number in final order
    for (label i = 0; i < n; ++i) newId[order[i]] = i;
    return newId;
}

void polyMesh::shuffle(std::uint64_t seed)
{
    std::mt19937_64 g(seed);
    auto below = [&g](label k) { return static_cast<label>(g() % static_cast<std::uint64_t>(k)); };
    const label nI = nInternalFaces(), nF = nFaces();
    std::vector<label> perm(static_cast<std::size_t>(nF));
    std::iota(perm.begin(), perm.end(), 0);
    for (label i = nI - 1; i > 0; --i) std::swap(perm[i], perm[below(i + 1)]);
    for (const Patch& p : patches_)
        for (label i = p.size - 1; i > 0; --i) std::swap(perm[p.start + i], perm[p.start + below(i + 1)]);
    std::vector<label> cellPerm(static_cast<std::size_t>(nCells_));
    std::iota(cellPerm.begin(), cellPerm.end(), 0);
    for (label i = nCells_ - 1; i > 0; --i) std::swap(cellPerm[i], cellPerm[below(i + 1)]);

    std::vector<label> fs(1, 0), fv, ow(nF), nb(nI);
    fv.reserve(faceVerts_.size());
    for (label k = 0; k < nF; ++k)
    {
        const label f = perm[k];
        fv.insert(fv.end(), faceVerts_.begin() + faceStart_[f], faceVerts_.begin() + faceStart_[f + 1]);
        fs.push_back(static_cast<label>(fv.size()));
        if (k < nI)
        {
            const label a = cellPerm[owner_[f]], b = cellPerm[neighbour_[f]];
            ow[k] = std::min(a, b);
            nb[k] = std::max(a, b);
        }
        else
        {
            ow[k] = cellPerm[owner_[f]];
        }
    }
    faceStart_.swap(fs);
    faceVerts_.swap(fv);
    owner_.swap(ow);
    neighbour_.swap(nb);
    calcGeometry();
}

void jitterLattice(std::vector<double>& points, label nx, label ny, label nz, const Vec3& mn,
                   const Vec3& mx, double amp, std::uint64_t seed)
{
    std::mt19937_64 g(seed);
    const double h[3] = {(mx.x - mn.x) / static_cast<double>(nx), (mx.y - mn.y) / static_cast<double>(ny),
                         (mx.z - mn.z) / static_cast<double>(nz)};
    const label px = nx + 1, py = ny + 1;
    for (label k = 1; k < nz; ++k)
        for (label j = 1; j < ny; ++j)
            for (label i = 1; i < nx; ++i)
            {
                double* q = &points[3 * (i + px * (j + py * k))];
                for (int c = 0; c < 3; ++c)
                {
                    const double u = static_cast<double>(g() >> 11) * (1.0 / 9007199254740992.0);
                    q[c] += (2.0 * u - 1.0) * amp * h[c];
                }
            }
}

namespace
{
std::vector<double> latticePoints(const Vec3& mn, const Vec3& mx, label nx, label ny, label nz)
{
    const label px = nx + 1, py = ny + 1;
    std::vector<double> pts(static_cast<std::size_t>(3 * px * py * (nz + 1)));
    for (label k = 0; k <= nz; ++k)
        for (label j = 0; j <= ny; ++j)
            for (label i = 0; i <= nx; ++i)
            {
                double* q = &pts[3 * (i + px * (j + py * k))];
                q[0] = mn.x + (mx.x - mn.x) * static_cast<double>(i) / static_cast<double>(nx);
                q[1] = mn.y + (mx.y - mn.y) * static_cast<double>(j) / static_cast<double>(ny);
                q[2] = mn.z + (mx.z - mn.z) * static_cast<double>(k) / static_cast<double>(nz);
            }
    return pts;
}
} // namespace

// Structured enumeration of BlockMesh::create's first-encounter face order:
// per cell (k, j, i outermost first) zlo (k==0 only), zhi, ylo (j==0 only),
// yhi, xlo (i==0 only), xhi; a face belongs to the first cell visiting it,
// which is the lower id and hence the owner.
polyMesh BlockMesh::create(const Vec3& mn, const Vec3& mx, label nx, label ny, label nz,
                           const Sides& sides, double jitter, std::uint64_t seed)
{
    if (nx <= 0 || ny <= 0 || nz <= 0) CF_FATAL("blockMesh: cell counts must be positive");
    std::vector<double> pts = latticePoints(mn, mx, nx, ny, nz);
    if (jitter != 0.0) jitterLattice(pts, nx, ny, nz, mn, mx, jitter, seed);
    const label px = nx + 1, py = ny + 1;
    auto P = [&](label i, label j, label k) { return i + px * (j + py * k); };
    auto Cid = [&](label i, label j, label k) { return i + nx * (j + ny * k); };
    const label nC = nx * ny * nz;
    const label nF = 3 * nC + nx * ny + ny * nz + nz * nx;
    const label nI = 3 * nC - nx * ny - ny * nz - nz * nx;
    std::vector<label> fv(4 * nF), ow(nF), nb(nI);
    std::vector<label> sideCount(6, 0);
    // boundary faces go to per-side lists in first-encounter order
    std::vector<std::vector<label>> sideFaces(6);
    std::vector<std::vector<label>> sideOwner(6);
    label fi = 0;
    auto emit = [&](label a, label b, label c, label d, label own, label nbr, int side) {
        if (nbr >= 0)
        {
            fv[4 * fi] = a; fv[4 * fi + 1] = b; fv[4 * fi + 2] = c; fv[4 * fi + 3] = d;
            ow[fi] = own;
            nb[fi] = nbr;
            ++fi;
        }
        else
        {
            sideFaces[side].insert(sideFaces[side].end(), {a, b, c, d});
            sideOwner[side].push_back(own);
        }
    };
    for (label k = 0; k < nz; ++k)
        for (label j = 0; j < ny; ++j)
            for (label i = 0; i < nx; ++i)
            {
                const label c = Cid(i, j, k);
                const label v0 = P(i, j, k), v1 = P(i + 1, j, k), v2 = P(i + 1, j + 1, k), v3 = P(i, j + 1, k);
                const label v4 = P(i, j, k + 1), v5 = P(i + 1, j, k + 1), v6 = P(i + 1, j + 1, k + 1),
                            v7 = P(i, j + 1, k + 1);
                if (k == 0) emit(v0, v1, v2, v3, c, -1, 4);
                emit(v4, v5, v6, v7, c, k + 1 < nz ? Cid(i, j, k + 1) : -1, 5);
                if (j == 0) emit(v0, v1, v5, v4, c, -1, 2);
                emit(v2, v3, v7, v6, c, j + 1 < ny ? Cid(i, j + 1, k) : -1, 3);
                if (i == 0) emit(v0, v3, v7, v4, c, -1, 0);
                emit(v1, v2, v6, v5, c, i + 1 < nx ? Cid(i + 1, j, k) : -1, 1);
            }
    if (fi != nI) CF_FATAL("blockMesh: internal face count mismatch");
    const std::pair<std::string, std::string>* names[6] = {&sides.xmin, &sides.xmax, &sides.ymin,
                                                           &sides.ymax, &sides.zmin, &sides.zmax};
    std::vector<polyMesh::Patch> patches;
    for (int s = 0; s < 6; ++s)
    {
        const label sz = static_cast<label>(sideOwner[s].size());
        if (sz == 0) continue;
        patches.push_back({names[s]->first, names[s]->second, fi, sz});
        std::copy(sideFaces[s].begin(), sideFaces[s].end(), fv.begin() + 4 * fi);
        std::copy(sideOwner[s].begin(), sideOwner[s].end(), ow.begin() + fi);
        fi += sz;
    }
    std::vector<label> fs(static_cast<std::size_t>(nF) + 1);
    for (label f = 0; f <= nF; ++f) fs[f] = 4 * f;
    return polyMesh(std::move(pts), std::move(fs), std::move(fv), std::move(ow), std::move(nb),
                    std::move(patches));
}

// Kuhn (Freudenthal) split: the tet for axis permutation (a,b,c) is
// {corner 000, 000+e_a, 000+e_a+e_b, 111}.  Every hex uses the same
// diagonal, so shared faces match across hexes.  Faces are deduplicated by
// sorting (sorted-vertex key, tet, local face); internal faces keep the
// order of their first registration (tet order), owner = lower tet id.
polyMesh BlockMesh::createTets(const Vec3& mn, const Vec3& mx, label nx, label ny, label nz,
                               const Sides& sides, double jitter, std::uint64_t seed)
{
    if (nx <= 0 || ny <= 0 || nz <= 0) CF_FATAL("blockMesh: cell counts must be positive");
    std::vector<double> pts = latticePoints(mn, mx, nx, ny, nz);
    if (jitter != 0.0) jitterLattice(pts, nx, ny, nz, mn, mx, jitter, seed);
    const label px = nx + 1, py = ny + 1;
    auto P = [&](label i, label j, label k) { return i + px * (j + py * k); };
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    const label nHex = nx * ny * nz, nT = 6 * nHex;
    std::vector<label> tv(4 * nT);
    for (label k = 0; k < nz; ++k)
        for (label j = 0; j < ny; ++j)
            for (label i = 0; i < nx; ++i)
            {
                const label h = i + nx * (j + ny * k);
                for (int t = 0; t < 6; ++t)
                {
                    label c[3] = {i, j, k};
                    label* out = &tv[4 * (6 * h + t)];
                    out[0] = P(c[0], c[1], c[2]);
                    c[perms[t][0]] += 1;
                    out[1] = P(c[0], c[1], c[2]);
                    c[perms[t][1]] += 1;
                    out[2] = P(c[0], c[1], c[2]);
                    out[3] = P(i + 1, j + 1, k + 1);
                }
            }
    // 4 faces per tet: drop vertex q
    struct Rec
    {
        label a, b, c;  // sorted key
        label tet;
        int q;
    };
    std::vector<Rec> recs(static_cast<std::size_t>(4 * nT));
    for (label t = 0; t < nT; ++t)
        for (int q = 0; q < 4; ++q)
        {
            label v[3], m = 0;
            for (int r = 0; r < 4; ++r)
                if (r != q) v[m++] = tv[4 * t + r];
            std::sort(v, v + 3);
            recs[4 * t + q] = {v[0], v[1], v[2], t, q};
        }
    std::vector<label> idx(recs.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](label x, label y) {
        const Rec &A = recs[x], &B = recs[y];
        if (A.a != B.a) return A.a < B.a;
        if (A.b != B.b) return A.b < B.b;
        if (A.c != B.c) return A.c < B.c;
        return x < y;
    });
    // pair up
    std::vector<label> firstOf(recs.size(), -1);  // registration index -> partner
    std::vector<char> isBoundary(recs.size(), 0);
    for (std::size_t s = 0; s < idx.size();)
    {
        std::size_t e = s + 1;
        while (e < idx.size() && recs[idx[e]].a == recs[idx[s]].a && recs[idx[e]].b == recs[idx[s]].b &&
               recs[idx[e]].c == recs[idx[s]].c)
            ++e;
        if (e - s == 1) isBoundary[idx[s]] = 1;
        else if (e - s == 2) { firstOf[idx[s]] = idx[s + 1]; firstOf[idx[s + 1]] = -2; }
        else CF_FATAL("tet mesh: non-manifold face");
        s = e;
    }
    auto faceVertsOf = [&](label r, label* out) {
        const label t = r / 4;
        const int q = static_cast<int>(r % 4);
        int m = 0;
        for (int k = 0; k < 4; ++k)
            if (k != q) out[m++] = tv[4 * t + k];
    };
    std::vector<label> fv, ow, nb;
    const label nI = (4 * nT - static_cast<label>(std::count(isBoundary.begin(), isBoundary.end(), 1))) / 2;
    fv.reserve(3 * (4 * nT));
    for (label r = 0; r < 4 * nT; ++r)
    {
        if (firstOf[r] < 0) continue;
        label v[3];
        faceVertsOf(r, v);
        fv.insert(fv.end(), v, v + 3);
        ow.push_back(r / 4);
        nb.push_back(firstOf[r] / 4);
    }
    if (static_cast<label>(nb.size()) != nI) CF_FATAL("tet mesh: internal face count mismatch");
    // boundary faces by side (centroid on a box plane)
    const std::pair<std::string, std::string>* names[6] = {&sides.xmin, &sides.xmax, &sides.ymin,
                                                           &sides.ymax, &sides.zmin, &sides.zmax};
    const double lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
    std::vector<std::vector<label>> side(6);
    for (label r = 0; r < 4 * nT; ++r)
    {
        if (!isBoundary[r]) continue;
        label v[3];
        faceVertsOf(r, v);
        int s = -1;
        for (int ax = 0; ax < 3 && s < 0; ++ax)
        {
            const double c = (pts[3 * v[0] + ax] + pts[3 * v[1] + ax] + pts[3 * v[2] + ax]) / 3.0;
            const double eps = 1e-12 * (1.0 + std::abs(hi[ax] - lo[ax]));
            if (std::abs(c - lo[ax]) < eps) s = 2 * ax;
            else if (std::abs(c - hi[ax]) < eps) s = 2 * ax + 1;
        }
        if (s < 0) CF_FATAL("tet mesh: boundary face not on any side");
        side[s].push_back(r);
    }
    std::vector<polyMesh::Patch> patches;
    for (int s = 0; s < 6; ++s)
    {
        if (side[s].empty()) continue;
        patches.push_back({names[s]->first, names[s]->second, static_cast<label>(ow.size()),
                           static_cast<label>(side[s].size())});
        for (label r : side[s])
        {
            label v[3];
            faceVertsOf(r, v);
            fv.insert(fv.end(), v, v + 3);
            ow.push_back(r / 4);
        }
    }
    std::vector<label> fs(ow.size() + 1);
    for (std::size_t f = 0; f <= ow.size(); ++f) fs[f] = static_cast<label>(3 * f);
    return polyMesh(std::move(pts), std::move(fs), std::move(fv), std::move(ow), std::move(nb),
                    std::move(patches));
}

polyMesh removeCells(const polyMesh& m, const std::vector<char>& keep, const std::string& name,
                     const std::string& type)
{
    const label nC = m.nCells(), nI = m.nInternalFaces(), nF = m.nFaces();
    if (static_cast<label>(keep.size()) != nC) throw Error("removeCells: keep mask size mismatch");
    std::vector<label> newId(nC, -1);
    label k = 0;
    for (label c = 0; c < nC; ++c)
        if (keep[c]) newId[c] = k++;
    if (k == 0) throw Error("removeCells: no cell left");
    const auto& fs = m.faceStart();
    const auto& fv = m.faceVerts();
    std::vector<label> ofs{0}, ofv, ow, nb;
    auto addFace = [&](label f, bool reverse, label owner) {
        const label a = fs[f], b = fs[f + 1];
        if (!reverse)
            for (label q = a; q < b; ++q) ofv.push_back(fv[q]);
        else
        {
            ofv.push_back(fv[a]);  // same first vertex, opposite winding
            for (label q = b - 1; q > a; --q) ofv.push_back(fv[q]);
        }
        ofs.push_back(static_cast<label>(ofv.size()));
        ow.push_back(owner);
    };
    std::vector<label> cut;  // internal faces with one kept cell
    for (label f = 0; f < nI; ++f)
    {
        const label o = m.owner()[f], n = m.neighbour()[f];
        if (keep[o] && keep[n])
        {
            addFace(f, false, newId[o]);
            nb.push_back(newId[n]);
        }
        else if (keep[o] || keep[n])
            cut.push_back(f);
    }
    std::vector<polyMesh::Patch> patches;
    label start = static_cast<label>(nb.size());
    for (const auto& p : m.patches())
    {
        polyMesh::Patch q{p.name, p.type, start, 0};
        for (label f = p.start; f < p.start + p.size; ++f)
            if (keep[m.owner()[f]])
            {
                addFace(f, false, newId[m.owner()[f]]);
                ++q.size;
            }
        start += q.size;
        patches.push_back(q);
    }
    polyMesh::Patch q{name, type, start, static_cast<label>(cut.size())};
    for (label f : cut)
    {
        const label o = m.owner()[f], n = m.neighbour()[f];
        if (keep[o])
            addFace(f, false, newId[o]);
        else
            addFace(f, true, newId[n]);
    }
    patches.push_back(q);
    (void)nF;
    return polyMesh(m.points(), std::move(ofs), std::move(ofv), std::move(ow), std::move(nb), std::move(patches));
}

polyMesh obstacleChannel(const Vec3& mn, const Vec3& mx, label nx, label ny, label nz, double jitter,
                         std::uint64_t seed, double cx, double cy, double radius)
{
    const polyMesh box = BlockMesh::create(mn, mx, nx, ny, nz, {}, jitter, seed);
    const auto& C = box.cellCentres();
    std::vector<char> keep(box.nCells(), 1);
    for (label c = 0; c < box.nCells(); ++c)
    {
        const double dx = C[3 * c] - cx, dy = C[3 * c + 1] - cy;
        if (dx * dx + dy * dy < radius * radius) keep[c] = 0;
    }
    return removeCells(box, keep, "obstacle", "wall");
}

} // namespace cf
