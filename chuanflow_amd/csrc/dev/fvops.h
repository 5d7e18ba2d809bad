// fvops.h -- finite-volume operators on the device mesh.
//
// Every scatter of the reference face loops ("owner += x; neighbour -= x"
// in ascending face order) is done here as a per-cell gather over the
// cell's faces in ascending face order, starting from the same value, so
// the floating-point sums are the reference's, bit for bit, without atomics.
//
// Conventions: cell vectors are SoA [3][n]; boundary linearisations
// (Tf = c*Tp + k) are lin_c/lin_k [nB][ncomp] with `comp` selecting the
// column; every pointer is a device pointer.
#pragma once

#include "types.h"

namespace cfx
{

struct DevMatrix;  // fwd

// fvc::grad (gradScheme.cpp:15-96).  corrected: bounded interpolation + 2
// skew passes; quirk: volScalarField::faceValue on internal faces for the
// uncorrected scheme (volFields.cpp:101-111).  tmpFace: double[nF].
void gradGauss(DevMesh& m, const double* x, const double* linc, const double* link, int ncomp,
               int comp, bool corrected, bool quirk, double* grad3, double* tmpFace);

// harmonicFaceGamma for a uniform cell value (laplacianScheme.cpp:37-62).
void harmonicGammaUniform(DevMesh& m, double gamma, double* gf);

// interpolateToFaces (pressureCorrection.cpp:27-35).
void interpolateToFaces(DevMesh& m, const double* cell, double* face);

// Implicit terms added into an LDU (diag/rhs [ncomp][n], upper/lower [nI]).
struct Ldu
{
    int ncomp = 1;
    double* diag = nullptr;
    double* rhs = nullptr;
    double* upper = nullptr;
    double* lower = nullptr;
};
void lduReset(DevMesh& m, const Ldu& A);
void assembleLaplacian(DevMesh& m, const Ldu& A, const double* gammaFace, const double* linc,
                       const double* link, int sign);
void assembleDiv(DevMesh& m, const Ldu& A, const double* phi, int scheme, double blendCoeff,
                 const double* linc, const double* link, int sign);
void assembleDdt(DevMesh& m, const Ldu& A, const double* old3, double dt, int sign);
// rhs[c][k] += (V * ((f + 0) - g))[k]  (simpleSolver.cpp:104-111, pisoSolver.cpp:53-64)
void addVolumeSource(DevMesh& m, const Ldu& A, const double force[3], const double* grad3);
// rhs[c] += -v[c]
void addNegRhs(DevMesh& m, const Ldu& A, const double* v);
// CSR values of component `comp` from the LDU (finishWithDiagonal layout).
void lduToCsr(DevMesh& m, const Ldu& A, int comp, double* val);
// Laplacian non-orthogonal corrector for one component (laplacianScheme.cpp:111-154):
// rhs += corrections computed from x.
void nonOrthCorrect(DevMesh& m, const double* gammaFace, const double* x, const double* linc,
                    const double* link, int ncomp, int comp, bool gradCorrected, int sign,
                    double* rhs, double* grad3Tmp, double* faceTmp);

// Flux operators
void updatePhiInternal(DevMesh& m, const double* U3, double* phi);                     // :10-25
void correctBoundaryFlux(DevMesh& m, const double* U3, const double* linc3,
                         const double* link3, double* phi);                           // volFields.cpp:294-306
void correctFluxes(DevMesh& m, const double* p, const double* gradP3, const double* rAUf,
                   const double* linc, const double* link, bool corrected, double* phi);  // :37-88
void divPhi(DevMesh& m, const double* phi, double* out);                                  // interpolate.cpp:138-151
// continuityError (pressureCorrection.cpp:90-125); writes the result to *dResult (device)
void continuityError(DevMesh& m, const double* phi, double* divTmp, double* dResult, double* part);

} // namespace cfx
