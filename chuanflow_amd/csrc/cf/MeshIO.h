// MeshIO.h -- on-disk meshes into the host polyMesh, and a VTU writer.
//
// Mirrors the reference's readers and writer (OpenFoamMesh.h, GmshMesh.h,
// VtkWriter.h under src/3d/include/cf/fvm/):
//  * readOpenFoam: the ASCII polyMesh directory (points, faces, owner,
//    neighbour, boundary), taken as written -- the face order, owner and
//    neighbour are the file's; patch types wall/patch/"67" -> wall,
//    empty/symmetryPlane/symmetry -> symmetry, anything else verbatim;
//  * readGmsh: ASCII .msh 2.2 and 4.1 with first-order tets, hexes and
//    prisms; faces recovered from the volume connectivity and deduplicated
//    by vertex set, internal faces first in order of first appearance,
//    boundary faces grouped by patch (the physical group of the matching
//    surface element, or the bounding-box plane when the file names no
//    surface) in order of first appearance;
//  * writeVtu: ASCII VTK unstructured grid, every cell a polyhedron (type 42)
//    with its face stream, plus cell scalars and vectors.
// Malformed input throws cf::Error with the file and the reason; nothing is
// truncated or guessed.  The mesh can then be renumbered (RCM) and uploaded
// like any generated mesh.
#pragma once

#include "PolyMesh.h"

#include <map>
#include <string>
#include <utility>
#include <vector>

namespace cf
{

using PatchTypes = std::map<std::string, std::string>;

/// `dir` is the polyMesh directory or a case root holding constant/polyMesh.
polyMesh readOpenFoam(const std::string& dir, const PatchTypes& patchTypes = {});

/// Patch types default to "wall" for names absent from `patchTypes`.
polyMesh readGmsh(const std::string& path, const PatchTypes& patchTypes = {});

struct VtuField
{
    std::string name;
    int components = 1;           // 1 (scalar) or 3 (vector, AoS [n][3])
    const double* data = nullptr; // nCells * components values
};

void writeVtu(const std::string& path, const polyMesh& mesh, const std::vector<VtuField>& fields = {});

/// "name=type;name=type" -> map (empty string: no overrides)
PatchTypes parsePatchTypes(const std::string& spec);

} // namespace cf
