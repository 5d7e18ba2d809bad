// ref_driver.cpp -- C-ABI shim over the REFERENCE's own public C++ API,
// linked with the reference hot-path sources compiled in place (see the
// Makefile's `ref` target).  TEST INFRASTRUCTURE ONLY: it exists so the
// oracle restatement can be compared with the reference build itself on
// small inputs.  It never travels into the product library.
//
// Meshes are passed as flat arrays (points [n][3], faces as CSR, owner,
// neighbour, patches as start/size), boundary conditions per patch as
// (type code, value[3]) with the codes: 0 fixedValue, 1 fixedGradient,
// 2 zeroGradient, 3 symmetry, 4 wall (vector only).
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "cf/common/Matrix.h"
#include "cf/fvm/BlockMesh.h"
#include "cf/fvm/Fields.h"
#include "cf/fvm/FvMatrix.h"
#include "cf/fvm/GmshMesh.h"
#include "cf/fvm/OpenFoamMesh.h"
#include "cf/fvm/PisoSolver.h"
#include "cf/fvm/PolyMesh.h"
#include "cf/fvm/PressureCorrection.h"
#include "cf/fvm/SimpleSolver.h"
#include "cf/fvm/VtkWriter.h"

using namespace cf;

namespace
{
const char* bcName(int code)
{
    switch (code)
    {
    case 0: return "fixedValue";
    case 1: return "fixedGradient";
    case 2: return "zeroGradient";
    case 3: return "symmetry";
    case 4: return "wall";
    default: return "unknown";
    }
}

polyMesh* buildMesh(int64_t nPoints, const double* pts, int64_t nFaces,
                    const int64_t* faceStart, const int64_t* faceVerts,
                    int64_t nInternal, const int64_t* owner, const int64_t* nbr,
                    int64_t nPatches, const int64_t* pStart, const int64_t* pSize)
{
    Field<Vector> points(static_cast<std::size_t>(nPoints));
    for (int64_t i = 0; i < nPoints; ++i) points[i] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    std::vector<std::vector<label>> faces(static_cast<std::size_t>(nFaces));
    for (int64_t f = 0; f < nFaces; ++f)
        faces[f].assign(faceVerts + faceStart[f], faceVerts + faceStart[f + 1]);
    Field<label> own(owner, owner + nFaces), nb(nbr, nbr + nInternal);
    std::vector<polyMesh::Patch> patches;
    for (int64_t p = 0; p < nPatches; ++p)
        patches.push_back({"patch" + std::to_string(p), "patch", pStart[p], pSize[p]});
    return new polyMesh(std::move(points), std::move(faces), std::move(own), std::move(nb),
                        std::move(patches));
}

void setScalarBC(volScalarField& f, const int32_t* type, const double* value)
{
    for (label p = 0; p < f.mesh().nPatches(); ++p) f.setPatchValue(p, bcName(type[p]), value[p]);
}

void setVectorBC(volVectorField& f, const int32_t* type, const double* value)
{
    for (label p = 0; p < f.mesh().nPatches(); ++p)
        f.setPatchValue(p, bcName(type[p]), {value[3 * p], value[3 * p + 1], value[3 * p + 2]});
}

Schemes schemesOf(int div, int lapCorrected, int gradScheme)
{
    Schemes s;
    s.div = div == 0 ? "upwind" : (div == 1 ? "linear" : "blend");
    s.laplacian = lapCorrected ? "corrected" : "none";
    s.grad = gradScheme == 1 ? "corrected" : (gradScheme == 2 ? "leastSquares" : "linear");
    return s;
}
} // namespace

extern "C" {

void* cfr_mesh_new(int64_t nPoints, const double* pts, int64_t nFaces,
                   const int64_t* faceStart, const int64_t* faceVerts,
                   int64_t nInternal, const int64_t* owner, const int64_t* nbr,
                   int64_t nPatches, const int64_t* pStart, const int64_t* pSize)
{
    try
    {
        return buildMesh(nPoints, pts, nFaces, faceStart, faceVerts, nInternal, owner, nbr,
                         nPatches, pStart, pSize);
    }
    catch (const std::exception&)
    {
        return nullptr;
    }
}

void cfr_mesh_free(void* m) { delete static_cast<polyMesh*>(m); }

// The reference's mesh readers: kind 0 OpenFoamMesh::fromDirectory, 1 GmshMesh::fromFile, with
// patch types "name=type;..." (may be empty).  Returns the mesh handle (null on error, message in
// cfr_read_error).
static std::string g_readError;
const char* cfr_read_error() { return g_readError.c_str(); }

void* cfr_read_mesh(int kind, const char* path, const char* patchTypes)
{
    try
    {
        std::map<std::string, std::string> types;
        std::string spec(patchTypes ? patchTypes : "");
        std::size_t i = 0;
        while (i < spec.size())
        {
            std::size_t j = spec.find(';', i);
            if (j == std::string::npos) j = spec.size();
            const std::string item = spec.substr(i, j - i);
            const std::size_t eq = item.find('=');
            if (eq != std::string::npos) types[item.substr(0, eq)] = item.substr(eq + 1);
            i = j + 1;
        }
        polyMesh m = kind == 0 ? OpenFoamMesh::fromDirectory(path, types) : GmshMesh::fromFile(path, types);
        return new polyMesh(std::move(m));
    }
    catch (const std::exception& e)
    {
        g_readError = e.what();
        return nullptr;
    }
}

// sizes: nPoints, nFaces, nInternal, nCells, nPatches, total face vertices
void cfr_mesh_sizes(void* mh, int64_t* s)
{
    const polyMesh& m = *static_cast<polyMesh*>(mh);
    int64_t fv = 0;
    for (const auto& f : m.faces()) fv += static_cast<int64_t>(f.size());
    s[0] = m.nPoints();
    s[1] = m.nFaces();
    s[2] = m.nInternalFaces();
    s[3] = m.nCells();
    s[4] = m.nPatches();
    s[5] = fv;
}

// arrays (caller-sized from cfr_mesh_sizes); patch names and types as "name=type;..." into buf
int cfr_mesh_arrays(void* mh, double* points, int64_t* faceStart, int64_t* faceVerts, int64_t* owner,
                    int64_t* nbr, int64_t* pStart, int64_t* pSize, char* buf, int64_t cap)
{
    const polyMesh& m = *static_cast<polyMesh*>(mh);
    for (label i = 0; i < m.nPoints(); ++i)
    {
        points[3 * i] = m.points()[i].x;
        points[3 * i + 1] = m.points()[i].y;
        points[3 * i + 2] = m.points()[i].z;
    }
    int64_t k = 0;
    faceStart[0] = 0;
    for (label f = 0; f < m.nFaces(); ++f)
    {
        for (const label v : m.faces()[f]) faceVerts[k++] = v;
        faceStart[f + 1] = k;
        owner[f] = m.owner()[f];
        if (f < m.nInternalFaces()) nbr[f] = m.neighbour()[f];
    }
    std::string names;
    for (label p = 0; p < m.nPatches(); ++p)
    {
        pStart[p] = m.patches()[p].start;
        pSize[p] = m.patches()[p].size;
        names += m.patches()[p].name + "=" + m.patches()[p].type + ";";
    }
    if (static_cast<int64_t>(names.size()) + 1 > cap) return -1;
    std::memcpy(buf, names.c_str(), names.size() + 1);
    return 0;
}

// VtkWriter with one cell scalar and one cell vector (AoS [n][3]); null skips a field
int cfr_write_vtu(void* mh, const char* path, const char* sname, const double* s, const char* vname,
                  const double* v)
{
    try
    {
        const polyMesh& m = *static_cast<polyMesh*>(mh);
        Field<scalar> sf(static_cast<std::size_t>(m.nCells()));
        Field<Vector> vf(static_cast<std::size_t>(m.nCells()));
        VtkWriter w(m);
        if (s)
        {
            for (label c = 0; c < m.nCells(); ++c) sf[c] = s[c];
            w.addCellScalar(sname, sf);
        }
        if (v)
        {
            for (label c = 0; c < m.nCells(); ++c) vf[c] = {v[3 * c], v[3 * c + 1], v[3 * c + 2]};
            w.addCellVector(vname, vf);
        }
        w.write(path);
        return 0;
    }
    catch (const std::exception& e)
    {
        g_readError = e.what();
        return -1;
    }
}

int64_t cfr_mesh_ncells(void* m) { return static_cast<polyMesh*>(m)->nCells(); }

// BlockMesh::create output sizes for (nx,ny,nz) are fixed; arrays caller-allocated.
int cfr_blockmesh(int64_t nx, int64_t ny, int64_t nz, const double* mn, const double* mx,
                  double* points, int64_t* faceVerts, int64_t* owner, int64_t* nbr,
                  int64_t* pStart, int64_t* pSize, int64_t* nInternal)
{
    polyMesh m = BlockMesh::create({mn[0], mn[1], mn[2]}, {mx[0], mx[1], mx[2]}, nx, ny, nz);
    for (label i = 0; i < m.nPoints(); ++i)
    {
        points[3 * i] = m.points()[i].x;
        points[3 * i + 1] = m.points()[i].y;
        points[3 * i + 2] = m.points()[i].z;
    }
    for (label f = 0; f < m.nFaces(); ++f)
    {
        if (m.faces()[f].size() != 4) return -1;
        for (int k = 0; k < 4; ++k) faceVerts[4 * f + k] = m.faces()[f][k];
        owner[f] = m.owner()[f];
        if (f < m.nInternalFaces()) nbr[f] = m.neighbour()[f];
    }
    *nInternal = m.nInternalFaces();
    for (label p = 0; p < m.nPatches(); ++p)
    {
        pStart[p] = m.patches()[p].start;
        pSize[p] = m.patches()[p].size;
    }
    return static_cast<int>(m.nPatches());
}

void cfr_geometry(void* mh, double* Sf, double* Cf, double* magSf, double* C, double* V,
                  double* d, double* invDist)
{
    const polyMesh& m = *static_cast<polyMesh*>(mh);
    for (label f = 0; f < m.nFaces(); ++f)
    {
        for (int k = 0; k < 3; ++k)
        {
            Sf[3 * f + k] = m.faceSf()[f][k];
            Cf[3 * f + k] = m.faceCentres()[f][k];
            d[3 * f + k] = m.deltaCoeffVec()[f][k];
        }
        magSf[f] = m.faceAreas()[f];
        invDist[f] = m.deltaCoeff()[f];
    }
    for (label c = 0; c < m.nCells(); ++c)
    {
        for (int k = 0; k < 3; ++k) C[3 * c + k] = m.cellCentres()[c][k];
        V[c] = m.cellVolumes()[c];
    }
}

void cfr_owner_weights(void* mh, double* w)
{
    const Field<scalar> ow = makeInternalOwnerWeights(*static_cast<polyMesh*>(mh));
    std::memcpy(w, ow.data(), sizeof(double) * ow.size());
}

// fvc::grad(volScalarField, scheme); gradScheme 0 linear, 1 corrected.
void cfr_grad(void* mh, const double* data, const int32_t* bcType, const double* bcValue,
              int gradScheme, double* out)
{
    const polyMesh& m = *static_cast<polyMesh*>(mh);
    volScalarField T(m, "T");
    setScalarBC(T, bcType, bcValue);
    std::memcpy(T.internalField().data(), data, sizeof(double) * m.nCells());
    const Field<Vector> g = fvc::grad(T, gradScheme == 1 ? "corrected" : "linear");
    for (label c = 0; c < m.nCells(); ++c)
        for (int k = 0; k < 3; ++k) out[3 * c + k] = g[c][k];
}

// fvm::laplacianFace(eqn, gammaFace, scheme, sign) on a scalar unknown, then
// solve(Controls{tol, maxIter, nonOrth}) from the initial guess x (in/out).
// Outputs aDiag, rhs_ after solve (= h(0)), and offU = -(Off * u) for a
// probe vector u (= h(u) - h(0)).
int cfr_laplacian_face_relaxed(void* mh, const double* gammaFace, const int32_t* bcType,
                               const double* bcValue, int lapCorrected, int gradScheme, int sign,
                               const double* extraRhs, double tol, int maxIter, int nonOrth, double* x,
                               double* diagOut, double* rhsOut, const double* u, double* offU,
                               int* iters, double relaxation)
{
    const polyMesh& m = *static_cast<polyMesh*>(mh);
    volScalarField T(m, "T");
    setScalarBC(T, bcType, bcValue);
    fvMatrix eq(m, viewOf(T), "TEqn");
    eq.schemes() = schemesOf(0, lapCorrected, gradScheme);
    eq.schemes().relaxation = relaxation;
    std::memcpy(eq.solution()[0].data(), x, sizeof(double) * m.nCells());
    Field<scalar> gf(gammaFace, gammaFace + m.nFaces());
    fvm::laplacianFace(eq, gf, lapCorrected ? "corrected" : "none", sign);
    if (extraRhs)
        for (label c = 0; c < m.nCells(); ++c) eq.addRhs(c, 0, extraRhs[c]);
    const SolverReport rep = eq.solve(fvMatrix::Controls{tol, maxIter, nonOrth});
    *iters = rep.iterations;
    std::memcpy(x, eq.solution()[0].data(), sizeof(double) * m.nCells());
    std::memcpy(diagOut, eq.aDiag(0).data(), sizeof(double) * m.nCells());
    Field<scalar> h0, hu;
    eq.explicitOperator(Field<scalar>(static_cast<std::size_t>(m.nCells()), 0.0), 0, h0);
    eq.explicitOperator(Field<scalar>(u, u + m.nCells()), 0, hu);
    for (label c = 0; c < m.nCells(); ++c)
    {
        rhsOut[c] = h0[c];
        offU[c] = hu[c] - h0[c];
    }
    return rep.converged ? 1 : 0;
}

int cfr_laplacian_face(void* mh, const double* gammaFace, const int32_t* bcType,
                       const double* bcValue, int lapCorrected, int gradScheme, int sign,
                       const double* extraRhs, double tol, int maxIter, int nonOrth, double* x,
                       double* diagOut, double* rhsOut, const double* u, double* offU,
                       int* iters)
{
    return cfr_laplacian_face_relaxed(mh, gammaFace, bcType, bcValue, lapCorrected, gradScheme, sign, extraRhs, tol,
                                      maxIter, nonOrth, x, diagOut, rhsOut, u, offU, iters, 1.0);
}

int cfr_solve_cg(int64_t n, const int64_t* rowStart, const int64_t* cols, const double* vals,
                 const double* b, double* x, double tol, int maxIter, int bicg, double* resOut)
{
    SparseMatrix::Builder bld(n);
    std::vector<double> diag(static_cast<std::size_t>(n), 0.0);
    for (int64_t r = 0; r < n; ++r)
        for (int64_t k = rowStart[r]; k < rowStart[r + 1]; ++k)
        {
            if (cols[k] == r) diag[r] += vals[k];
            else bld.add(r, cols[k], vals[k]);
        }
    const SparseMatrix A = bld.finishWithDiagonal(diag);
    Field<double> bb(b, b + n), xx(x, x + n);
    const SolverReport rep = bicg ? solveBiCGStab(A, bb, xx, tol, maxIter)
                                  : solveCG(A, bb, xx, tol, maxIter);
    std::memcpy(x, xx.data(), sizeof(double) * n);
    resOut[0] = rep.initialResidual;
    resOut[1] = rep.finalResidual;
    resOut[2] = rep.converged ? 1.0 : 0.0;
    return rep.iterations;
}

void cfr_flux_ops(void* mh, const double* U, const double* p, const double* gradP,
                  const double* rAUcell, const int32_t* UbcType, const double* UbcValue,
                  const int32_t* pbcType, const double* pbcValue, int lapCorrected,
                  double* phiOut, double* rAUfOut, double* divOut, double* contErr)
{
    const polyMesh& m = *static_cast<polyMesh*>(mh);
    volVectorField Uf(m, "U");
    setVectorBC(Uf, UbcType, UbcValue);
    for (label c = 0; c < m.nCells(); ++c) Uf.internalField()[c] = {U[3 * c], U[3 * c + 1], U[3 * c + 2]};
    volScalarField pf(m, "p");
    setScalarBC(pf, pbcType, pbcValue);
    std::memcpy(pf.internalField().data(), p, sizeof(double) * m.nCells());
    surfaceScalarField phi(m, "phi", 0.0);
    updatePhiInternal(m, Uf, phi);
    phi.correctBoundaryFlux(Uf);
    const Field<scalar> rAUf = interpolateToFaces(m, Field<scalar>(rAUcell, rAUcell + m.nCells()));
    Field<Vector> g(static_cast<std::size_t>(m.nCells()));
    for (label c = 0; c < m.nCells(); ++c) g[c] = {gradP[3 * c], gradP[3 * c + 1], gradP[3 * c + 2]};
    Schemes s;
    s.laplacian = lapCorrected ? "corrected" : "none";
    contErr[0] = continuityError(m, phi);
    correctFluxes(m, pf, g, rAUf, s, phi);
    contErr[1] = continuityError(m, phi);
    const Field<scalar> dv = fvc::divPhi(m, phi.values());
    std::memcpy(phiOut, phi.values().data(), sizeof(double) * m.nFaces());
    std::memcpy(rAUfOut, rAUf.data(), sizeof(double) * m.nFaces());
    std::memcpy(divOut, dv.data(), sizeof(double) * m.nCells());
}

// solveIncompressibleSimple on the given mesh; U [n][3], p [n], phi [nFaces]
// in/out (phi's boundary is reset from U first, as simpleFoamApp does).
// ctl: maxIter, continuityTol, pcorrTol, momentumTol, linearTol,
// linearMaxIter, nonOrth, relaxU, relaxP, pRelaxStart, fx, fy, fz.
int cfr_simple(void* mh, const int32_t* UbcType, const double* UbcValue, const int32_t* pbcType,
               const double* pbcValue, double nu, int div, int lapCorrected, int gradScheme,
               const double* ctl, double* U, double* p, double* phiIO, double* resultOut)
{
    polyMesh& m = *static_cast<polyMesh*>(mh);
    volVectorField Uf(m, "U");
    setVectorBC(Uf, UbcType, UbcValue);
    for (label c = 0; c < m.nCells(); ++c) Uf.internalField()[c] = {U[3 * c], U[3 * c + 1], U[3 * c + 2]};
    volScalarField pf(m, "p");
    setScalarBC(pf, pbcType, pbcValue);
    std::memcpy(pf.internalField().data(), p, sizeof(double) * m.nCells());
    surfaceScalarField phi(m, "phi", 0.0);
    std::memcpy(phi.values().data(), phiIO, sizeof(double) * m.nFaces());
    SimpleControls sc;
    sc.maxIterations = static_cast<int>(ctl[0]);
    sc.continuityTol = ctl[1];
    sc.pcorrTol = ctl[2];
    sc.momentumTol = ctl[3];
    sc.linearTol = ctl[4];
    sc.linearMaxIter = static_cast<int>(ctl[5]);
    sc.nonOrthCorrectors = static_cast<int>(ctl[6]);
    sc.relaxU = ctl[7];
    sc.relaxP = ctl[8];
    sc.pRelaxStart = ctl[9];
    sc.bodyForce = {ctl[10], ctl[11], ctl[12]};
    const SimpleResult r =
        solveIncompressibleSimple(m, Uf, pf, phi, nu, schemesOf(div, lapCorrected, gradScheme), sc);
    for (label c = 0; c < m.nCells(); ++c)
        for (int k = 0; k < 3; ++k) U[3 * c + k] = Uf.internalField()[c][k];
    std::memcpy(p, pf.internalField().data(), sizeof(double) * m.nCells());
    std::memcpy(phiIO, phi.values().data(), sizeof(double) * m.nFaces());
    resultOut[0] = r.iterations;
    resultOut[1] = r.finalContinuityError;
    resultOut[2] = r.finalMomentumResidual;
    resultOut[3] = r.converged ? 1.0 : 0.0;
    return 0;
}

// One PISO step.  ctl: dt, correctors, continuityTol, linearTol,
// linearMaxIter, nonOrth, fx, fy, fz.
int cfr_piso_step(void* mh, const int32_t* UbcType, const double* UbcValue,
                  const int32_t* pbcType, const double* pbcValue, double nu, int div,
                  int lapCorrected, int gradScheme, const double* ctl, double* U, double* p,
                  double* phiIO, double* resultOut)
{
    polyMesh& m = *static_cast<polyMesh*>(mh);
    volVectorField Uf(m, "U");
    setVectorBC(Uf, UbcType, UbcValue);
    for (label c = 0; c < m.nCells(); ++c) Uf.internalField()[c] = {U[3 * c], U[3 * c + 1], U[3 * c + 2]};
    volScalarField pf(m, "p");
    setScalarBC(pf, pbcType, pbcValue);
    std::memcpy(pf.internalField().data(), p, sizeof(double) * m.nCells());
    surfaceScalarField phi(m, "phi", 0.0);
    std::memcpy(phi.values().data(), phiIO, sizeof(double) * m.nFaces());
    PisoControls pc;
    pc.dt = ctl[0];
    pc.correctors = static_cast<int>(ctl[1]);
    pc.continuityTol = ctl[2];
    pc.linearTol = ctl[3];
    pc.linearMaxIter = static_cast<int>(ctl[4]);
    pc.nonOrthCorrectors = static_cast<int>(ctl[5]);
    pc.bodyForce = {ctl[6], ctl[7], ctl[8]};
    const PisoStepResult r = solveTransientIncompressiblePiso(
        m, Uf, pf, phi, nu, schemesOf(div, lapCorrected, gradScheme), pc);
    for (label c = 0; c < m.nCells(); ++c)
        for (int k = 0; k < 3; ++k) U[3 * c + k] = Uf.internalField()[c][k];
    std::memcpy(p, pf.internalField().data(), sizeof(double) * m.nCells());
    std::memcpy(phiIO, phi.values().data(), sizeof(double) * m.nFaces());
    resultOut[0] = r.correctorsUsed;
    resultOut[1] = r.predictedContinuityError;
    resultOut[2] = r.continuityError;
    resultOut[3] = r.pressureCorrectionMax;
    resultOut[4] = r.converged ? 1.0 : 0.0;
    return 0;
}

} // extern "C"
