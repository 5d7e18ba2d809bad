// mesh.h -- building the device mesh from reference-layout host arrays.
#pragma once

#include "types.h"

#include <memory>

namespace cfx
{

// Host arrays in the reference layout (label = int64, Field<Vector> AoS).
struct MeshDesc
{
    long long nCells, nInternal, nFaces;
    const long long* owner;      // [nFaces]
    const long long* neighbour;  // [nInternal]
    const double* Sf;            // [nFaces][3]
    const double* Cf;            // [nFaces][3]
    const double* C;             // [nCells][3]
    const double* V;             // [nCells]
    const double* d;             // [nFaces][3] deltaCoeffVec
    const double* magSf;         // [nFaces]
    long long nPatches;
    const long long* patchStart;
    const long long* patchSize;
};

std::unique_ptr<DevMesh> createDevMesh(Context& ctx, const MeshDesc& d);

} // namespace cfx
