"""ctypes access to the CPU oracle (``_build/libcfo.so``) and, where it was
built, to the reference build itself (``_ref/libcfref.so``).

TEST INFRASTRUCTURE ONLY.  Importable from tests/, from
``__graft_entry__.smoke()`` and from bench.py's cpu_baseline leg, as the
checker; never from the product package.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libcfo.so")
REF_PATH = os.path.join(HERE, "_ref", "libcfref.so")

BC_FIXED_VALUE, BC_FIXED_GRADIENT, BC_ZERO_GRADIENT, BC_SYMMETRY = 0, 1, 2, 3

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def _ptr(a, t=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(t))


class _Mesh(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("nPoints", "nFaces", "nInternal", "nCells", "nPatches")] + [
        ("points", C.POINTER(C.c_double)),
        ("faceStart", C.POINTER(C.c_int64)),
        ("faceVerts", C.POINTER(C.c_int64)),
        ("owner", C.POINTER(C.c_int64)),
        ("neighbour", C.POINTER(C.c_int64)),
        ("patchStart", C.POINTER(C.c_int64)),
        ("patchSize", C.POINTER(C.c_int64)),
    ] + [(n, C.POINTER(C.c_double)) for n in ("Sf", "Cf", "magSf", "C", "V", "dvec", "invDist")]


class _Ldu(C.Structure):
    _fields_ = [("n", C.c_int64), ("F", C.c_int64), ("ncomp", C.c_int)] + [
        (n, C.POINTER(C.c_double)) for n in ("diag", "rhs", "upper", "lower")]


class _Csr(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("rowStart", C.POINTER(C.c_int64)),
                ("cols", C.POINTER(C.c_int64)), ("vals", C.POINTER(C.c_double))]


class Report(C.Structure):
    _fields_ = [("iterations", C.c_int), ("initialResidual", C.c_double),
                ("finalResidual", C.c_double), ("converged", C.c_int)]


class Schemes(C.Structure):
    _fields_ = [("divScheme", C.c_int), ("laplacianCorrected", C.c_int),
                ("gradCorrected", C.c_int), ("blendCoeff", C.c_double)]


class SimpleControls(C.Structure):
    _fields_ = [("maxIterations", C.c_int), ("continuityTol", C.c_double),
                ("pcorrTol", C.c_double), ("momentumTol", C.c_double),
                ("linearTol", C.c_double), ("linearMaxIter", C.c_int),
                ("nonOrthCorrectors", C.c_int), ("relaxU", C.c_double),
                ("relaxP", C.c_double), ("pRelaxStart", C.c_double),
                ("bodyForce", C.c_double * 3)]


class SimpleResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("finalContinuityError", C.c_double),
                ("finalMomentumResidual", C.c_double), ("converged", C.c_int),
                ("diverged", C.c_int)]


class PisoControls(C.Structure):
    _fields_ = [("dt", C.c_double), ("correctors", C.c_int), ("continuityTol", C.c_double),
                ("linearTol", C.c_double), ("linearMaxIter", C.c_int),
                ("nonOrthCorrectors", C.c_int), ("bodyForce", C.c_double * 3),
                ("projectNeumann", C.c_int)]


class PisoResult(C.Structure):
    _fields_ = [("correctorsUsed", C.c_int), ("predictedContinuityError", C.c_double),
                ("continuityError", C.c_double), ("pressureCorrectionMax", C.c_double),
                ("converged", C.c_int)]


class _Bc(C.Structure):
    _fields_ = [("type", C.POINTER(C.c_int32)), ("value", C.POINTER(C.c_double))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        L.cfo_mt64_nth.restype = C.c_uint64
        L.cfo_mt64_nth.argtypes = [C.c_uint64, C.c_int64]
        L.cfo_solve_cg.restype = Report
        L.cfo_solve_bicgstab.restype = Report
        L.cfo_continuity_error.restype = C.c_double
        L.cfo_simple.restype = SimpleResult
        L.cfo_piso_step.restype = PisoResult
        L.cfo_simple.argtypes = [C.POINTER(_Mesh), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.c_double, C.POINTER(Schemes),
                                 C.POINTER(SimpleControls), _Bc, _Bc, C.POINTER(C.c_double)]
        L.cfo_piso_step.argtypes = [C.POINTER(_Mesh), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.c_double, C.POINTER(Schemes),
                                    C.POINTER(PisoControls), _Bc, _Bc]
        _lib = L
    return _lib


def have_ref() -> bool:
    return os.path.exists(REF_PATH)


_ref = None


def ref():
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError("reference build not present")
        R = C.CDLL(REF_PATH)
        R.cfr_mesh_new.restype = C.c_void_p
        R.cfr_mesh_new.argtypes = [C.c_int64, _f64p, C.c_int64, _i64p, _i64p, C.c_int64, _i64p,
                                   _i64p, C.c_int64, _i64p, _i64p]
        R.cfr_mesh_free.argtypes = [C.c_void_p]
        R.cfr_mesh_ncells.restype = C.c_int64
        R.cfr_mesh_ncells.argtypes = [C.c_void_p]
        for name in ("cfr_geometry", "cfr_owner_weights", "cfr_grad", "cfr_laplacian_face", "cfr_laplacian_face_relaxed",
                     "cfr_flux_ops", "cfr_simple", "cfr_piso_step"):
            getattr(R, name).argtypes = None
        _ref = R
    return _ref


# ------------------------------------------------------------------ meshes
@dataclass
class Mesh:
    """A polyMesh as flat arrays plus the oracle's geometry."""
    points: np.ndarray
    faceStart: np.ndarray
    faceVerts: np.ndarray
    owner: np.ndarray
    neighbour: np.ndarray
    patchStart: np.ndarray
    patchSize: np.ndarray
    nCells: int
    geom: dict = field(default_factory=dict)

    @property
    def nFaces(self):
        return len(self.owner)

    @property
    def nInternal(self):
        return len(self.neighbour)

    @property
    def nB(self):
        return self.nFaces - self.nInternal

    def struct(self) -> _Mesh:
        g = self.geom
        s = _Mesh(len(self.points), self.nFaces, self.nInternal, self.nCells, len(self.patchStart),
                  _ptr(self.points), _ptr(self.faceStart, C.c_int64), _ptr(self.faceVerts, C.c_int64),
                  _ptr(self.owner, C.c_int64), _ptr(self.neighbour, C.c_int64),
                  _ptr(self.patchStart, C.c_int64), _ptr(self.patchSize, C.c_int64),
                  *[_ptr(g[k]) for k in ("Sf", "Cf", "magSf", "C", "V", "dvec", "invDist")])
        self._keep = s
        return s

    def compute_geometry(self):
        nF, nC = self.nFaces, self.nCells
        self.geom = dict(Sf=np.zeros((nF, 3)), Cf=np.zeros((nF, 3)), magSf=np.zeros(nF),
                         C=np.zeros((nC, 3)), V=np.zeros(nC), dvec=np.zeros((nF, 3)),
                         invDist=np.zeros(nF))
        rc = lib().cfo_geometry(C.byref(self.struct()))
        if rc != 0:
            raise ValueError("degenerate mesh geometry")
        return self

    def patch_of_boundary_face(self):
        """patch id per boundary face."""
        out = np.empty(self.nB, np.int64)
        for p, (s, n) in enumerate(zip(self.patchStart, self.patchSize)):
            out[s - self.nInternal:s - self.nInternal + n] = p
        return out


def blockmesh(nx, ny, nz, mn=(0, 0, 0), mx=(1, 1, 1), jitter=0.0, seed=0) -> Mesh:
    nP = (nx + 1) * (ny + 1) * (nz + 1)
    nC = nx * ny * nz
    nF = 3 * nC + nx * ny + ny * nz + nz * nx
    pts = np.zeros((nP, 3))
    fv = np.zeros(4 * nF, np.int64)
    own = np.zeros(nF, np.int64)
    nb = np.zeros(nF, np.int64)
    ps = np.zeros(6, np.int64)
    pz = np.zeros(6, np.int64)
    nI = C.c_int64(0)
    mn_a = np.asarray(mn, np.float64)
    mx_a = np.asarray(mx, np.float64)
    rc = lib().cfo_blockmesh(C.c_int64(nx), C.c_int64(ny), C.c_int64(nz), _ptr(mn_a), _ptr(mx_a),
                             _ptr(pts), _ptr(fv, C.c_int64), _ptr(own, C.c_int64),
                             _ptr(nb, C.c_int64), _ptr(ps, C.c_int64), _ptr(pz, C.c_int64),
                             C.byref(nI))
    if rc != 0:
        raise ValueError("blockmesh failed")
    if jitter:
        lib().cfo_jitter(C.c_int64(nx), C.c_int64(ny), C.c_int64(nz), _ptr(mn_a), _ptr(mx_a),
                         C.c_double(jitter), C.c_uint64(seed), _ptr(pts))
    keep = pz > 0
    return Mesh(pts, np.arange(0, 4 * nF + 1, 4, dtype=np.int64), fv, own, nb[:nI.value].copy(),
                ps[keep].copy(), pz[keep].copy(), nC)


def mt64_nth(seed: int, n: int) -> int:
    return int(lib().cfo_mt64_nth(C.c_uint64(seed), C.c_int64(n)))


# ------------------------------------------------------------- operators
def linearise(m: Mesh, ncomp, types, values):
    types = np.ascontiguousarray(types, np.int32)
    values = np.ascontiguousarray(values, np.float64).reshape(m.nB, ncomp)
    lc = np.zeros((m.nB, ncomp))
    lk = np.zeros((m.nB, ncomp))
    rc = lib().cfo_linearise(C.byref(m.struct()), ncomp, _ptr(types, C.c_int32), _ptr(values),
                             _ptr(lc), _ptr(lk))
    if rc:
        raise ValueError(f"linearise failed ({rc})")
    return lc, lk


def owner_weights(m: Mesh):
    w = np.zeros(max(1, m.nInternal))
    lib().cfo_owner_weights(C.byref(m.struct()), _ptr(w))
    return w[:m.nInternal]


def grad(m: Mesh, data, lc, lk, ncomp=1, comp=0, corrected=True, quirk=True):
    out = np.zeros((m.nCells, 3))
    data = np.ascontiguousarray(data, np.float64)
    lib().cfo_grad(C.byref(m.struct()), _ptr(data), _ptr(np.ascontiguousarray(lc)),
                   _ptr(np.ascontiguousarray(lk)), ncomp, comp, int(corrected), int(quirk), _ptr(out))
    return out


def harmonic_gamma_uniform(m: Mesh, gamma):
    gf = np.zeros(m.nFaces)
    lib().cfo_harmonic_gamma_uniform(C.byref(m.struct()), C.c_double(gamma), _ptr(gf))
    return gf


def interpolate_to_faces(m: Mesh, cell):
    out = np.zeros(m.nFaces)
    lib().cfo_interpolate_to_faces(C.byref(m.struct()), _ptr(np.ascontiguousarray(cell, np.float64)),
                                   _ptr(out))
    return out


class Ldu:
    def __init__(self, m: Mesh, ncomp=1):
        self.n, self.F, self.ncomp = m.nCells, m.nInternal, ncomp
        self.diag = np.zeros((ncomp, m.nCells))
        self.rhs = np.zeros((ncomp, m.nCells))
        self.upper = np.full(max(1, m.nInternal), -0.0)
        self.lower = np.full(max(1, m.nInternal), -0.0)

    def struct(self):
        s = _Ldu(self.n, self.F, self.ncomp, _ptr(self.diag), _ptr(self.rhs), _ptr(self.upper),
                 _ptr(self.lower))
        self._keep = s
        return s


def assemble_laplacian(m: Mesh, A: Ldu, gammaFace, lc, lk, sign=-1):
    lib().cfo_assemble_laplacian(C.byref(m.struct()), _ptr(np.ascontiguousarray(gammaFace)),
                                 _ptr(np.ascontiguousarray(lc)), _ptr(np.ascontiguousarray(lk)),
                                 sign, C.byref(A.struct()))


def assemble_div(m: Mesh, A: Ldu, phi, lc, lk, scheme=0, blend=0.9, sign=1):
    lib().cfo_assemble_div(C.byref(m.struct()), _ptr(np.ascontiguousarray(phi)), scheme,
                           C.c_double(blend), _ptr(np.ascontiguousarray(lc)),
                           _ptr(np.ascontiguousarray(lk)), sign, C.byref(A.struct()))


def assemble_ddt(m: Mesh, A: Ldu, old, dt, sign=1):
    lib().cfo_assemble_ddt(C.byref(m.struct()), _ptr(np.ascontiguousarray(old, np.float64)),
                           C.c_double(dt), sign, C.byref(A.struct()))


def nonorth_correct(m: Mesh, gammaFace, x, lc, lk, ncomp=1, comp=0, gradCorrected=True, sign=-1,
                    rhs=None):
    rhs = np.zeros(m.nCells) if rhs is None else rhs
    lib().cfo_nonorth_correct(C.byref(m.struct()), _ptr(np.ascontiguousarray(gammaFace)),
                              _ptr(np.ascontiguousarray(x, np.float64)),
                              _ptr(np.ascontiguousarray(lc)), _ptr(np.ascontiguousarray(lk)),
                              ncomp, comp, int(gradCorrected), sign, _ptr(rhs))
    return rhs


class Csr:
    def __init__(self, n, nnz):
        self.rowStart = np.zeros(n + 1, np.int64)
        self.cols = np.zeros(nnz, np.int64)
        self.vals = np.zeros(nnz)
        self.n = n

    def struct(self):
        s = _Csr(self.n, len(self.cols), _ptr(self.rowStart, C.c_int64), _ptr(self.cols, C.c_int64),
                 _ptr(self.vals))
        self._keep = s
        return s

    def trim(self):
        nnz = int(self.rowStart[-1])
        self.cols = self.cols[:nnz].copy()
        self.vals = self.vals[:nnz].copy()
        return self

    def todense(self):
        A = np.zeros((self.n, self.n))
        for r in range(self.n):
            for k in range(self.rowStart[r], self.rowStart[r + 1]):
                A[r, self.cols[k]] += self.vals[k]
        return A


def csr_from_ldu(m: Mesh, A: Ldu, comp=0) -> Csr:
    out = Csr(m.nCells, m.nCells + 2 * m.nInternal)
    s = out.struct()
    lib().cfo_csr_from_ldu(C.byref(m.struct()), C.byref(A.struct()), comp, C.byref(s))
    return out.trim()


def spmv(A: Csr, x):
    y = np.zeros(A.n)
    lib().cfo_spmv(C.byref(A.struct()), _ptr(np.ascontiguousarray(x, np.float64)), _ptr(y))
    return y


def solve_cg(A: Csr, b, x0=None, tol=1e-8, maxIter=2000, hist=False):
    x = np.zeros(A.n) if x0 is None else np.array(x0, np.float64)
    h = np.zeros(maxIter + 1) if hist else None
    rep = lib().cfo_solve_cg(C.byref(A.struct()), _ptr(np.ascontiguousarray(b, np.float64)), _ptr(x),
                             C.c_double(tol), C.c_int(maxIter), _ptr(h))
    if hist:
        return x, rep, h[:rep.iterations + 1]
    return x, rep


def solve_bicgstab(A: Csr, b, x0=None, tol=1e-8, maxIter=2000):
    x = np.zeros(A.n) if x0 is None else np.array(x0, np.float64)
    rep = lib().cfo_solve_bicgstab(C.byref(A.struct()), _ptr(np.ascontiguousarray(b, np.float64)),
                                   _ptr(x), C.c_double(tol), C.c_int(maxIter))
    return x, rep


def update_phi_internal(m: Mesh, U, phi):
    lib().cfo_update_phi_internal(C.byref(m.struct()), _ptr(np.ascontiguousarray(U, np.float64)),
                                  _ptr(phi))


def correct_boundary_flux(m: Mesh, U, lc, lk, phi):
    lib().cfo_correct_boundary_flux(C.byref(m.struct()), _ptr(np.ascontiguousarray(U, np.float64)),
                                    _ptr(np.ascontiguousarray(lc)), _ptr(np.ascontiguousarray(lk)),
                                    _ptr(phi))


def correct_fluxes(m: Mesh, p, gradP, rAUf, lc, lk, corrected, phi):
    lib().cfo_correct_fluxes(C.byref(m.struct()), _ptr(np.ascontiguousarray(p, np.float64)),
                             _ptr(np.ascontiguousarray(gradP, np.float64)),
                             _ptr(np.ascontiguousarray(rAUf)), _ptr(np.ascontiguousarray(lc)),
                             _ptr(np.ascontiguousarray(lk)), int(corrected), _ptr(phi))


def div_phi(m: Mesh, phi):
    out = np.zeros(m.nCells)
    lib().cfo_div_phi(C.byref(m.struct()), _ptr(np.ascontiguousarray(phi)), _ptr(out))
    return out


def continuity_error(m: Mesh, phi):
    return lib().cfo_continuity_error(C.byref(m.struct()), _ptr(np.ascontiguousarray(phi)))


def _bc(types, values):
    t = np.ascontiguousarray(types, np.int32)
    v = np.ascontiguousarray(values, np.float64)
    return _Bc(_ptr(t, C.c_int32), _ptr(v)), (t, v)


def simple(m: Mesh, U, p, phi, nu, schemes: Schemes, ctl: SimpleControls, Ubc, pbc, hist=False):
    """Ubc/pbc: (types[nB], values[nB][ncomp]) per boundary face."""
    ub, k1 = _bc(*Ubc)
    pb, k2 = _bc(*pbc)
    h = np.zeros(4 * ctl.maxIterations) if hist else None
    r = lib().cfo_simple(C.byref(m.struct()), _ptr(U), _ptr(p), _ptr(phi), C.c_double(nu),
                         C.byref(schemes), C.byref(ctl), ub, pb, _ptr(h))
    if hist:
        return r, h.reshape(-1, 4)[:r.iterations]
    return r


def piso_step(m: Mesh, U, p, phi, nu, schemes: Schemes, ctl: PisoControls, Ubc, pbc):
    ub, k1 = _bc(*Ubc)
    pb, k2 = _bc(*pbc)
    return lib().cfo_piso_step(C.byref(m.struct()), _ptr(U), _ptr(p), _ptr(phi), C.c_double(nu),
                               C.byref(schemes), C.byref(ctl), ub, pb)


# ---------------------------------------------------- reference build shim
class RefMesh:
    """A reference polyMesh built from the same flat arrays."""

    def __init__(self, m: Mesh):
        R = ref()
        self.m = m
        self.h = R.cfr_mesh_new(len(m.points), np.ascontiguousarray(m.points, np.float64),
                                m.nFaces, m.faceStart, m.faceVerts, m.nInternal, m.owner,
                                m.neighbour if m.nInternal else np.zeros(1, np.int64),
                                len(m.patchStart), m.patchStart, m.patchSize)
        if not self.h:
            raise ValueError("reference polyMesh construction failed")

    def __del__(self):
        try:
            ref().cfr_mesh_free(C.c_void_p(self.h))
        except Exception:
            pass

    def geometry(self):
        m = self.m
        nF, nC = m.nFaces, m.nCells
        g = dict(Sf=np.zeros((nF, 3)), Cf=np.zeros((nF, 3)), magSf=np.zeros(nF), C=np.zeros((nC, 3)),
                 V=np.zeros(nC), dvec=np.zeros((nF, 3)), invDist=np.zeros(nF))
        ref().cfr_geometry(C.c_void_p(self.h), *[_ptr(g[k]) for k in
                                                 ("Sf", "Cf", "magSf", "C", "V", "dvec", "invDist")])
        return g

    def owner_weights(self):
        w = np.zeros(max(1, self.m.nInternal))
        ref().cfr_owner_weights(C.c_void_p(self.h), _ptr(w))
        return w[:self.m.nInternal]

    def grad(self, data, ptype, pval, corrected=True):
        out = np.zeros((self.m.nCells, 3))
        ref().cfr_grad(C.c_void_p(self.h), _ptr(np.ascontiguousarray(data, np.float64)),
                       _ptr(np.ascontiguousarray(ptype, np.int32), C.c_int32),
                       _ptr(np.ascontiguousarray(pval, np.float64)), C.c_int(1 if corrected else 0),
                       _ptr(out))
        return out

    def laplacian_face(self, gammaFace, ptype, pval, lapCorrected, gradCorrected, sign, extraRhs,
                       tol, maxIter, nonOrth, x0, u, relaxation=1.0):
        n = self.m.nCells
        x = np.array(x0, np.float64)
        diag, rhs, offU = np.zeros(n), np.zeros(n), np.zeros(n)
        it = C.c_int(0)
        conv = ref().cfr_laplacian_face_relaxed(
            C.c_void_p(self.h), _ptr(np.ascontiguousarray(gammaFace, np.float64)),
            _ptr(np.ascontiguousarray(ptype, np.int32), C.c_int32),
            _ptr(np.ascontiguousarray(pval, np.float64)), C.c_int(int(lapCorrected)),
            C.c_int(int(gradCorrected)), C.c_int(sign),
            _ptr(None if extraRhs is None else np.ascontiguousarray(extraRhs, np.float64)),
            C.c_double(tol), C.c_int(maxIter), C.c_int(nonOrth), _ptr(x), _ptr(diag), _ptr(rhs),
            _ptr(np.ascontiguousarray(u, np.float64)), _ptr(offU), C.byref(it), C.c_double(relaxation))
        return dict(x=x, diag=diag, rhs=rhs, offU=offU, iterations=it.value, converged=bool(conv))

    def flux_ops(self, U, p, gradP, rAU, Utype, Uval, ptype, pval, lapCorrected):
        m = self.m
        phi, rAUf, dv, ce = np.zeros(m.nFaces), np.zeros(m.nFaces), np.zeros(m.nCells), np.zeros(2)
        ref().cfr_flux_ops(C.c_void_p(self.h), _ptr(np.ascontiguousarray(U, np.float64)),
                           _ptr(np.ascontiguousarray(p, np.float64)),
                           _ptr(np.ascontiguousarray(gradP, np.float64)),
                           _ptr(np.ascontiguousarray(rAU, np.float64)),
                           _ptr(np.ascontiguousarray(Utype, np.int32), C.c_int32),
                           _ptr(np.ascontiguousarray(Uval, np.float64)),
                           _ptr(np.ascontiguousarray(ptype, np.int32), C.c_int32),
                           _ptr(np.ascontiguousarray(pval, np.float64)), C.c_int(int(lapCorrected)),
                           _ptr(phi), _ptr(rAUf), _ptr(dv), _ptr(ce))
        return dict(phi=phi, rAUf=rAUf, divPhi=dv, contErrBefore=ce[0], contErrAfter=ce[1])

    def simple(self, Utype, Uval, ptype, pval, nu, schemes, ctl_vec, U, p, phi):
        res = np.zeros(4)
        ref().cfr_simple(C.c_void_p(self.h), _ptr(np.ascontiguousarray(Utype, np.int32), C.c_int32),
                         _ptr(np.ascontiguousarray(Uval, np.float64)),
                         _ptr(np.ascontiguousarray(ptype, np.int32), C.c_int32),
                         _ptr(np.ascontiguousarray(pval, np.float64)), C.c_double(nu),
                         C.c_int(schemes[0]), C.c_int(schemes[1]), C.c_int(schemes[2]),
                         _ptr(np.ascontiguousarray(ctl_vec, np.float64)), _ptr(U), _ptr(p), _ptr(phi),
                         _ptr(res))
        return res

    def piso_step(self, Utype, Uval, ptype, pval, nu, schemes, ctl_vec, U, p, phi):
        res = np.zeros(5)
        ref().cfr_piso_step(C.c_void_p(self.h), _ptr(np.ascontiguousarray(Utype, np.int32), C.c_int32),
                            _ptr(np.ascontiguousarray(Uval, np.float64)),
                            _ptr(np.ascontiguousarray(ptype, np.int32), C.c_int32),
                            _ptr(np.ascontiguousarray(pval, np.float64)), C.c_double(nu),
                            C.c_int(schemes[0]), C.c_int(schemes[1]), C.c_int(schemes[2]),
                            _ptr(np.ascontiguousarray(ctl_vec, np.float64)), _ptr(U), _ptr(p),
                            _ptr(phi), _ptr(res))
        return res


def ref_blockmesh(nx, ny, nz, mn=(0, 0, 0), mx=(1, 1, 1)):
    nP = (nx + 1) * (ny + 1) * (nz + 1)
    nC = nx * ny * nz
    nF = 3 * nC + nx * ny + ny * nz + nz * nx
    pts = np.zeros((nP, 3))
    fv = np.zeros(4 * nF, np.int64)
    own = np.zeros(nF, np.int64)
    nb = np.zeros(nF, np.int64)
    ps = np.zeros(6, np.int64)
    pz = np.zeros(6, np.int64)
    nI = C.c_int64(0)
    R = ref()
    npatch = R.cfr_blockmesh(C.c_int64(nx), C.c_int64(ny), C.c_int64(nz),
                             _ptr(np.asarray(mn, np.float64)), _ptr(np.asarray(mx, np.float64)),
                             _ptr(pts), _ptr(fv, C.c_int64), _ptr(own, C.c_int64), _ptr(nb, C.c_int64),
                             _ptr(ps, C.c_int64), _ptr(pz, C.c_int64), C.byref(nI))
    return Mesh(pts, np.arange(0, 4 * nF + 1, 4, dtype=np.int64), fv, own, nb[:nI.value].copy(),
                ps[:npatch].copy(), pz[:npatch].copy(), nC)
