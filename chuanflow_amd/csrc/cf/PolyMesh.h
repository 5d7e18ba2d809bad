// PolyMesh.h -- host-side polyhedral mesh of the MI355X hot path.
//
// Mirrors the reference mesh API (src/3d/include/cf/fvm/PolyMesh.h:17-103,
// BlockMesh.h:31-53) so a caller of the reference keeps its mesh code:
// owner/neighbour face addressing with internal faces first and owner <
// neighbour, patches as contiguous boundary-face ranges, and the same
// geometry (area vectors oriented owner->neighbour, face centroids, exact
// cell volumes/centroids by the divergence theorem, owner->neighbour delta
// vectors), computed in the reference's operation order so every derived
// quantity is bit-identical.
//
// Differences by design:
//  * nCells is explicit (max over owner AND neighbour + 1): after cell
//    renumbering the last cell need not own a face;
//  * faces are stored CSR (faceStart/faceVerts), not vector<vector>;
//  * geometry is computed eagerly once, no lazy `mutable` cache;
//  * mesh generators for the benchmark workloads (seeded vertex jitter,
//    Kuhn 6-tet split, seeded face/cell shuffle, RCM renumbering).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace cf
{

using label = std::int64_t;
using scalar = double;

/// Mirrors cf::Error / cfFatal (src/common/include/cf/common/Error.h:9-17).
class Error : public std::runtime_error
{
public:
    explicit Error(const std::string& what) : std::runtime_error(what) {}
};

#define CF_FATAL(msg) throw ::cf::Error(std::string(__FILE__) + ":" + std::to_string(__LINE__) + " " + (msg))

struct Vec3
{
    double x = 0, y = 0, z = 0;
};

class polyMesh
{
public:
    struct Patch
    {
        std::string name;
        std::string type;
        label start = 0;
        label size = 0;
    };

    polyMesh() = default;
    /// faces as CSR: vertices of face f are faceVerts[faceStart[f] .. faceStart[f+1]).
    polyMesh(std::vector<double> points, std::vector<label> faceStart, std::vector<label> faceVerts,
             std::vector<label> owner, std::vector<label> neighbour, std::vector<Patch> patches);

    label nPoints() const { return static_cast<label>(points_.size() / 3); }
    label nCells() const { return nCells_; }
    label nFaces() const { return static_cast<label>(owner_.size()); }
    label nInternalFaces() const { return static_cast<label>(neighbour_.size()); }
    label nBoundaryFaces() const { return nFaces() - nInternalFaces(); }
    label nPatches() const { return static_cast<label>(patches_.size()); }

    const std::vector<double>& points() const { return points_; }
    const std::vector<label>& faceStart() const { return faceStart_; }
    const std::vector<label>& faceVerts() const { return faceVerts_; }
    const std::vector<label>& owner() const { return owner_; }
    const std::vector<label>& neighbour() const { return neighbour_; }
    const std::vector<Patch>& patches() const { return patches_; }
    label patchIdByName(const std::string& name) const;

    // geometry, flat [n][3] AoS like the reference Field<Vector>
    const std::vector<double>& cellVolumes() const { return V_; }
    const std::vector<double>& cellCentres() const { return C_; }
    const std::vector<double>& faceCentres() const { return Cf_; }
    const std::vector<double>& faceSf() const { return Sf_; }
    const std::vector<double>& faceAreas() const { return magSf_; }
    const std::vector<double>& deltaCoeffVec() const { return d_; }
    const std::vector<double>& deltaCoeff() const { return invDist_; }

    /// Face -> patch id for boundary faces, -1 for internal faces.
    const std::vector<label>& facePatchMap() const { return facePatchMap_; }

    /// Renumber cells by `newId[old]` and re-sort faces: internal faces by
    /// (owner, neighbour) with owner < neighbour, boundary faces by owner
    /// within each patch.  Recomputes geometry.
    void renumberCells(const std::vector<label>& newId);

    /// Reverse Cuthill-McKee order of the cell graph: returns newId[old].
    /// Start at a pseudo-peripheral node (George-Liu) of minimum degree;
    /// neighbours visited by increasing degree, ties by lower cell id.
    std::vector<label> rcmOrdering() const;

    /// max |owner - neighbour| over internal faces (matrix bandwidth).
    label bandwidth() const;

    /// Seeded shuffle of internal faces, of boundary faces within each
    /// patch, and of cell ids (MT19937-64 Fisher-Yates).  Recomputes geometry.
    void shuffle(std::uint64_t seed);

    void calcGeometry();

private:
    std::vector<double> points_;
    std::vector<label> faceStart_, faceVerts_, owner_, neighbour_;
    std::vector<Patch> patches_;
    label nCells_ = 0;
    std::vector<label> facePatchMap_;

    std::vector<double> Sf_, Cf_, magSf_, C_, V_, d_, invDist_;
};

/// Side patch descriptor, as BlockSides (BlockMesh.h:13-21).
struct BlockSides
{
    std::pair<std::string, std::string> xmin{"xmin", "wall"};
    std::pair<std::string, std::string> xmax{"xmax", "wall"};
    std::pair<std::string, std::string> ymin{"ymin", "wall"};
    std::pair<std::string, std::string> ymax{"ymax", "wall"};
    std::pair<std::string, std::string> zmin{"zmin", "symmetry"};
    std::pair<std::string, std::string> zmax{"zmax", "symmetry"};
};

struct BlockMesh
{
    using Sides = BlockSides;
    /// BlockMesh::create (blockMesh.cpp:36-189): same points, faces, face
    /// order, owner/neighbour and patches, without the std::map.
    static polyMesh create(const Vec3& min, const Vec3& max, label nx, label ny, label nz,
                           const Sides& sides = {}, double jitter = 0.0, std::uint64_t seed = 0);
    /// Kuhn split of the same lattice into 6 tets per hex around the
    /// (0,0,0)->(1,1,1) diagonal (conforming).  Triangular faces.
    static polyMesh createTets(const Vec3& min, const Vec3& max, label nx, label ny, label nz,
                               const Sides& sides = {}, double jitter = 0.0, std::uint64_t seed = 0);
};

/// The sub-mesh of the cells with keep[c] != 0: internal faces between two
/// kept cells stay internal (same order), boundary faces of kept cells stay
/// in their patches, and faces between a kept and a removed cell become a new
/// patch (name, type) appended last, oriented out of the kept cell.  Cells
/// are renumbered compactly in their old order; points are kept.
polyMesh removeCells(const polyMesh& m, const std::vector<char>& keep, const std::string& name,
                     const std::string& type);

/// Channel block with a cylindrical obstacle: BlockMesh::create (seeded
/// jitter) minus the cells whose centre lies within `radius` of the axis
/// parallel to z through (cx, cy); the cut faces form patch "obstacle" (wall).
polyMesh obstacleChannel(const Vec3& min, const Vec3& max, label nx, label ny, label nz, double jitter,
                         std::uint64_t seed, double cx, double cy, double radius);

/// Seeded interior-vertex jitter of a structured (nx+1)(ny+1)(nz+1) lattice:
/// x += (2u-1)*amp*h per component, u the top 53 bits of successive
/// MT19937-64 outputs, interior points in lattice order (i fastest).
void jitterLattice(std::vector<double>& points, label nx, label ny, label nz, const Vec3& min,
                   const Vec3& max, double amp, std::uint64_t seed);

} // namespace cf
