/*
 * cfx_host.h -- host-side mesh layer of the MI355X hot path (no GPU needed).
 *
 * A cfx_hmesh is the library's C++ polyMesh: the reference's mesh model
 * (src/3d/include/cf/fvm/PolyMesh.h:17-103) with the same geometry, plus the
 * generators of the benchmark workloads.  Same status codes as cfx.h.
 */
#ifndef CFX_HOST_H
#define CFX_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "cfx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cfx_hmesh cfx_hmesh;

/* BlockMesh::create (src/3d/include/cf/fvm/BlockMesh.h:31-53,
 * blockMesh.cpp:36-189): same points, faces, face order and patches
 * (xmin xmax ymin ymax zmin zmax).  jitter > 0 moves strictly interior
 * vertices by (2u-1)*jitter*h per component, u from MT19937-64(seed).
 * tets != 0 splits every hex into 6 Kuhn tetrahedra (no reference
 * counterpart). */
int cfx_hmesh_box(int64_t nx, int64_t ny, int64_t nz, const double mn[3], const double mx[3],
                  double jitter, uint64_t seed, int tets, cfx_hmesh** out);
/* polyMesh(points, faces, owner, neighbour, patches) (PolyMesh.h:29-33) with
 * faces as CSR (faceStart[nFaces+1], faceVerts). */
/* channel block (BlockMesh patches xmin xmax ymin ymax zmin zmax, seeded
 * jitter) minus the cells whose centre is within `radius` of the z-parallel
 * axis through (cx, cy); the cut faces form a 7th patch "obstacle" (wall).
 * No counterpart in the reference (its blockMesh makes boxes only). */
int cfx_hmesh_obstacle_channel(int64_t nx, int64_t ny, int64_t nz, const double mn[3], const double mx[3],
                               double jitter, uint64_t seed, double cx, double cy, double radius, cfx_hmesh** out);
/* channel block for detail regions: patches "inlet" (x min, type inlet),
 * "outlet" (x max, type outlet) and four walls */
int cfx_hmesh_channel(int64_t nx, int64_t ny, int64_t nz, const double mn[3], const double mx[3],
                      double jitter, uint64_t seed, cfx_hmesh** out);

/* ---------------------------------------------------- detail regions */
/* DetailBlock (src/coupling/include/cf/coupling/CoupledNetwork.h,
 * coupledNetwork.cpp:644-926, 1080-1200), laminar incompressible: inlet and
 * outlet are the patches typed inlet / outlet; the inlet profile is "flat",
 * "parabolic" (prod(1 - eta^2) over the wall-bounded transverse axes) or
 * "power1over7", normalised to the section area; cfx_detail_loss_at imposes
 * rate q, runs the GPU SIMPLE (warm-started from the block's fields) and
 * returns the drop of the volume-weighted section-mean pressure (Pa) between
 * the first and the last cross-section. */
typedef struct cfx_detail cfx_detail;
typedef struct cfx_detail_measure {
    int64_t cells;
    int sweeps, converged;
    double imposedQ, inletArea, plugSpeed;
    double dpPa, fluxIn, fluxOut;
} cfx_detail_measure;
int cfx_detail_create(cfx_ctx* ctx, const cfx_hmesh* mesh, double nu, double rho, const char* shape,
                      cfx_detail** out);
int cfx_detail_destroy(cfx_detail* d);
int cfx_detail_loss_at(cfx_detail* d, double q, const cfx_simple_controls* ctl, cfx_detail_measure* m);
/* lossAt of `count` blocks (rates q[i], measures into m[i]; m may be null), the coupling loop's
 * serial pass over its blocks (coupledNetwork.cpp:1461-1475) run by `threads` host threads; with
 * threads > 1 every block needs its own context, so the blocks' kernels run on concurrent streams.
 * An error in any block is reported after all have finished.  No counterpart in the reference. */
int cfx_detail_loss_at_many(int count, cfx_detail* const* blocks, const double* q, const cfx_simple_controls* ctl,
                            int threads, cfx_detail_measure* m);
int cfx_detail_info(const cfx_detail* d, int* axis, int* flowSign, double* inletArea, int64_t* nSections,
                    int64_t* nInletFaces);
/* inlet faces and their profile weights (flat: none) */
int cfx_detail_inlet(const cfx_detail* d, int64_t* faces, double* weights);
/* host copies: U [3][n] SoA, p [n], phi [nFaces] (each may be null) */
int cfx_detail_fields(const cfx_detail* d, double* U, double* p, double* phi);
/* the block's device mesh (pressure dumps, multigrid settings) */
cfx_mesh* cfx_detail_mesh(const cfx_detail* d);

/* OpenFoamMesh::fromDirectory (OpenFoamMesh.h:40-43): the ASCII polyMesh files of `dir` (or of
 * dir/constant/polyMesh) as written; patch types wall/patch/"67" -> wall, empty/symmetryPlane ->
 * symmetry, others verbatim.  patchTypes: "name=type;name=type" overrides (null or "" for none). */
int cfx_hmesh_read_openfoam(const char* dir, const char* patchTypes, cfx_hmesh** out);
/* GmshMesh::fromFile (GmshMesh.h:30-31): ASCII .msh 2.2 / 4.1, first-order tets, hexes, prisms;
 * faces from the cell connectivity (internal faces first, boundary faces grouped by physical
 * surface or bounding-box plane); patch types from patchTypes, default "wall". */
int cfx_hmesh_read_gmsh(const char* path, const char* patchTypes, cfx_hmesh** out);
/* name and type of patch i (nul-terminated; buffers of the given capacity, each may be null) */
int cfx_hmesh_patch(const cfx_hmesh* h, int64_t i, char* name, int64_t nameCap, char* type, int64_t typeCap);
/* VtkWriter (VtkWriter.h:11-24): ASCII .vtu, every cell a polyhedron (type 42, with its face stream
 * and face offsets), plus nFields cell fields: names[k], components[k] (1 or 3; vectors AoS [n][3]),
 * data[k] (host) */
int cfx_vtu_write(const cfx_hmesh* h, const char* path, int nFields, const char* const* names,
                  const int* components, const double* const* data);

int cfx_hmesh_from_arrays(int64_t nPoints, const double* points, int64_t nFaces,
                          const int64_t* faceStart, const int64_t* faceVerts, int64_t nInternal,
                          const int64_t* owner, const int64_t* neighbour, int64_t nPatches,
                          const int64_t* patchStart, const int64_t* patchSize, cfx_hmesh** out);
int cfx_hmesh_free(cfx_hmesh* h);
/* {nPoints, nFaces, nInternal, nCells, nPatches, nFaceVerts} */
int cfx_hmesh_sizes(const cfx_hmesh* h, int64_t sizes[6]);
int cfx_hmesh_get(const cfx_hmesh* h, double* points, int64_t* faceStart, int64_t* faceVerts,
                  int64_t* owner, int64_t* neighbour, int64_t* patchStart, int64_t* patchSize);
/* polyMesh geometry (polyMesh.cpp:83-288), AoS [n][3] */
int cfx_hmesh_geometry(const cfx_hmesh* h, double* Sf, double* Cf, double* magSf, double* C,
                       double* V, double* d, double* invDist);
/* seeded shuffle of internal faces, boundary faces within patches, and
 * cell ids (no reference counterpart) */
int cfx_hmesh_shuffle(cfx_hmesh* h, uint64_t seed);
/* reverse Cuthill-McKee renumbering; newId[old] written if non-NULL; faces
 * re-sorted by (owner, neighbour) (no reference counterpart) */
int cfx_hmesh_renumber_rcm(cfx_hmesh* h, int64_t* newId);
int cfx_hmesh_bandwidth(const cfx_hmesh* h, int64_t* bw);
const char* cfx_hmesh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
