"""ctypes binding of the native library (include/cfx.h, include/cfx_host.h).

The library is required: importing this module on a machine where the
library has not been built raises ImportError, and every compute call goes
to the HIP kernels -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libchuanflow_amd.so")
INCLUDE_DIR = os.path.join(os.path.dirname(HERE), "include")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"chuanflow_amd native library not built: {LIB_PATH} "
                      "(run `python -m chuanflow_amd.build`)")

_L = C.CDLL(LIB_PATH)

CFX_OK, CFX_E_INVALID, CFX_E_HIP, CFX_E_STATE, CFX_E_NUMERIC, CFX_E_INTERNAL = range(6)
PRECOND_JACOBI, PRECOND_DIC, PRECOND_LEARNED, PRECOND_MG = 0, 1, 2, 3
FP64, FP32_IR, FP32_MG = 0, 1, 2
SOLVER_CG, SOLVER_BICGSTAB = 0, 1

_vp = C.c_void_p
_i64 = C.c_int64
_dp = C.POINTER(C.c_double)


class Report(C.Structure):
    """cfx_report: SolverReport plus FP64 true residuals and timing."""
    _fields_ = [("iterations", C.c_int), ("initialResidual", C.c_double),
                ("finalResidual", C.c_double), ("converged", C.c_int), ("breakdown", C.c_int),
                ("l2Residual", C.c_double), ("trueResidual", C.c_double),
                ("outerIterations", C.c_int), ("precondFallbacks", C.c_int), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Schemes(C.Structure):
    """cfx_schemes: div 0 upwind / 1 linear / 2 blend; laplacian and grad corrected flags."""
    _fields_ = [("div", C.c_int), ("laplacianCorrected", C.c_int), ("gradCorrected", C.c_int),
                ("blendCoeff", C.c_double)]


class SimpleControls(C.Structure):
    _fields_ = [("maxIterations", C.c_int), ("continuityTol", C.c_double), ("pcorrTol", C.c_double),
                ("momentumTol", C.c_double), ("linearTol", C.c_double), ("linearMaxIter", C.c_int),
                ("nonOrthCorrectors", C.c_int), ("relaxU", C.c_double), ("relaxP", C.c_double),
                ("pRelaxStart", C.c_double), ("bodyForce", C.c_double * 3), ("pPrecision", C.c_int),
                ("uPrecision", C.c_int)]


class SimpleResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("finalContinuityError", C.c_double),
                ("finalMomentumResidual", C.c_double), ("converged", C.c_int), ("diverged", C.c_int),
                ("ms", C.c_double)]


class PisoControls(C.Structure):
    _fields_ = [("dt", C.c_double), ("correctors", C.c_int), ("continuityTol", C.c_double),
                ("linearTol", C.c_double), ("linearMaxIter", C.c_int), ("nonOrthCorrectors", C.c_int),
                ("bodyForce", C.c_double * 3), ("pPrecision", C.c_int), ("projectNeumann", C.c_int),
                ("uPrecision", C.c_int)]


class PisoResult(C.Structure):
    _fields_ = [("correctorsUsed", C.c_int), ("predictedContinuityError", C.c_double),
                ("continuityError", C.c_double), ("pressureCorrectionMax", C.c_double), ("converged", C.c_int),
                ("ms", C.c_double)]


class MgParams(C.Structure):
    """cfx_mg_params: finest-level a (pre), s (scale), c (post); coarse-level constants; hierarchy caps."""
    _fields_ = [("omegaPre", C.c_float), ("scale", C.c_float), ("omegaPost", C.c_float),
                ("restrictScale", C.c_float), ("coarseOmega", C.c_float), ("coarseScale", C.c_float), ("coarseSweeps", C.c_int),
                ("coarsestOmega", C.c_float), ("maxLevels", C.c_int), ("coarseMax", C.c_int),
                ("kcycleDepth", C.c_int), ("fusion", C.c_int), ("smoothFine", C.c_int), ("smoothCoarse", C.c_int),
                ("smoothDeep", C.c_int)]


class TrainOptions(C.Structure):
    """cfx_train_options: Adam steps, error samples per step, smoothing cycles, learning rate, seed."""
    _fields_ = [("steps", C.c_int), ("samples", C.c_int), ("smooth", C.c_int), ("lr", C.c_float),
                ("seed", C.c_uint64)]


class BatchSchedule(C.Structure):
    """cfx_batch_schedule; unnamed fields take the library's defaults (three-launch form)."""
    _fields_ = [("vecElems", C.c_int), ("wgCap", C.c_int), ("graph", C.c_int), ("pollEvery", C.c_int),
                ("persistent", C.c_int), ("claim", C.c_int), ("wgs", C.c_int)]
    _defaults = (4, 2048, 1, 16, 0, 2, 0)

    def __init__(self, *args, **kw):
        vals = dict(zip([f for f, _ in self._fields_], self._defaults))
        vals.update(zip([f for f, _ in self._fields_], args))
        vals.update(kw)
        super().__init__(**vals)

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("maxIterations", C.c_int), ("used", BatchSchedule)]


class CfxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cfx error {code}: {msg}")
        self.code = code


def _sig(name, restype, *args):
    f = getattr(_L, name)
    f.restype = restype
    f.argtypes = list(args)
    return f


# signatures -------------------------------------------------------------
_sig("cfx_create", C.c_int, C.c_int, C.POINTER(_vp))
_sig("cfx_destroy", C.c_int, _vp)
_sig("cfx_last_error", C.c_char_p, _vp)
_sig("cfx_synchronize", C.c_int, _vp)
_sig("cfx_device_info", C.c_int, _vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_int))
_sig("cfx_malloc", C.c_int, _vp, C.c_size_t, C.POINTER(_vp))
_sig("cfx_free", C.c_int, _vp, _vp)
_sig("cfx_memcpy_htod", C.c_int, _vp, _vp, _vp, C.c_size_t)
_sig("cfx_memcpy_dtoh", C.c_int, _vp, _vp, _vp, C.c_size_t)
_sig("cfx_memcpy_dtod", C.c_int, _vp, _vp, _vp, C.c_size_t)
_sig("cfx_zero", C.c_int, _vp, _vp, C.c_size_t)
_sig("cfx_mesh_create", C.c_int, _vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp,
     _vp, C.POINTER(_vp))
_sig("cfx_mesh_destroy", C.c_int, _vp)
_sig("cfx_mesh_sizes", C.c_int, _vp, _vp)
_sig("cfx_mesh_pattern", C.c_int, _vp, _vp, _vp)
_sig("cfx_owner_weights", C.c_int, _vp, _vp)
_sig("cfx_grad", C.c_int, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp)
_sig("cfx_harmonic_gamma_uniform", C.c_int, _vp, C.c_double, _vp)
_sig("cfx_interpolate_to_faces", C.c_int, _vp, _vp, _vp)
_sig("cfx_update_phi_internal", C.c_int, _vp, _vp, _vp)
_sig("cfx_correct_boundary_flux", C.c_int, _vp, _vp, _vp, _vp, _vp)
_sig("cfx_correct_fluxes", C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp)
_sig("cfx_div_phi", C.c_int, _vp, _vp, _vp)
_sig("cfx_continuity_error", C.c_int, _vp, _vp, _dp)
_sig("cfx_matrix_create", C.c_int, _vp, C.c_int, C.POINTER(_vp))
_sig("cfx_matrix_destroy", C.c_int, _vp)
_sig("cfx_matrix_reset", C.c_int, _vp)
_sig("cfx_assemble_laplacian", C.c_int, _vp, _vp, _vp, _vp, C.c_int)
_sig("cfx_assemble_div", C.c_int, _vp, _vp, C.c_int, C.c_double, _vp, _vp, C.c_int)
_sig("cfx_assemble_ddt", C.c_int, _vp, _vp, C.c_double, C.c_int)
_sig("cfx_matrix_add_rhs", C.c_int, _vp, C.c_int, _vp)
_sig("cfx_matrix_get", C.c_int, _vp, _vp, _vp, _vp, _vp)
_sig("cfx_matrix_device_ptrs", C.c_int, _vp, C.POINTER(_vp), C.POINTER(_vp))
_sig("cfx_nonorth_correct", C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp)
_sig("cfx_matrix_solve", C.c_int, _vp, _vp, C.c_int, C.c_double, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int,
     C.c_int, C.c_int, C.c_int, C.POINTER(Report))
_sig("cfx_matrix_set_relaxation", C.c_int, _vp, C.c_double)
_sig("cfx_matrix_set_inner_tol", C.c_int, _vp, C.c_double)
_sig("cfx_matrix_explicit_operator", C.c_int, _vp, _vp, C.c_int, _vp)
_sig("cfx_matrix_spmv", C.c_int, _vp, C.c_int, _vp, _vp)
_sig("cfx_csr_create", C.c_int, _vp, _i64, _vp, _vp, _vp, C.POINTER(_vp))
_sig("cfx_csr_destroy", C.c_int, _vp)
_sig("cfx_csr_set_values", C.c_int, _vp, _vp)
_sig("cfx_csr_set_values_device", C.c_int, _vp, _vp)
_sig("cfx_csr_spmv", C.c_int, _vp, _vp, _vp)
_sig("cfx_csr_solve", C.c_int, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
     C.POINTER(Report))
_sig("cfx_csr_solve_device", C.c_int, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
     C.POINTER(Report))
_sig("cfx_csr_residual_history", C.c_int, _vp, _vp, C.c_int)
_sig("cfx_pcg_solve_noalloc", C.c_int, _vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int,
     C.POINTER(Report))
_sig("cfx_csr_set_inner_tol", C.c_int, _vp, C.c_double)
_sig("cfx_mg_default_params", C.c_int, C.POINTER(MgParams))
_sig("cfx_set_graphs", C.c_int, _vp, C.c_int)
_sig("cfx_stream_bandwidth", C.c_int, _vp, C.c_size_t, C.c_int, _dp, _dp)
_sig("cfx_set_component_batching", C.c_int, _vp, C.c_int)
_sig("cfx_batch_solve", C.c_int, _vp, C.c_int, _vp, _vp, _vp, C.c_double, C.c_int, _vp, _vp, _vp,
     C.POINTER(BatchStats))
_sig("cfx_batch_autotune", C.c_int, _vp, C.c_int, _vp, _vp, _vp, C.c_int, C.POINTER(BatchSchedule), _vp, _vp,
     C.c_int)
_sig("cfx_mpnn_num_params", C.c_int)
_sig("cfx_mpnn_create", C.c_int, _vp, C.c_uint64, C.c_int, C.POINTER(_vp))
_sig("cfx_mpnn_destroy", C.c_int, _vp)
_sig("cfx_mpnn_get_params", C.c_int, _vp, _vp)
_sig("cfx_mpnn_set_params", C.c_int, _vp, _vp)
_sig("cfx_mpnn_forward_csr", C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int)
_sig("cfx_mpnn_backward", C.c_int, _vp, _vp, _vp, _vp)
_sig("cfx_mpnn_adam", C.c_int, _vp, _vp, C.c_float)
_sig("cfx_mpnn_train", C.c_int, _vp, _vp, _vp, _vp, C.c_int, C.POINTER(TrainOptions), _vp)
_sig("cfx_mpnn_evaluate", C.c_int, _vp, _vp, _vp, _vp, C.POINTER(TrainOptions), C.c_int, C.c_int, _dp)
_sig("cfx_csr_mg_loss_grad", C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _dp)
_sig("cfx_csr_set_mpnn", C.c_int, _vp, _vp, _vp)
_sig("cfx_matrix_set_mpnn", C.c_int, _vp, _vp)
_sig("cfx_mesh_set_mpnn", C.c_int, _vp, _vp)
_sig("cfx_mesh_set_mg", C.c_int, _vp, C.POINTER(MgParams))
_sig("cfx_mesh_set_pressure_dump", C.c_int, _vp, _vp, _vp, C.c_int, C.c_int)
_sig("cfx_mesh_pressure_dump_count", C.c_int, _vp, C.POINTER(C.c_int))
_sig("cfx_csr_set_mg", C.c_int, _vp, C.POINTER(MgParams))
_sig("cfx_csr_set_mg_weights", C.c_int, _vp, _vp, _vp, _vp, _vp)
_sig("cfx_matrix_set_mg", C.c_int, _vp, C.POINTER(MgParams))
_sig("cfx_matrix_set_mg_weights", C.c_int, _vp, _vp, _vp, _vp, _vp)
_sig("cfx_csr_mg_info", C.c_int, _vp, C.POINTER(C.c_int), _vp, _vp, C.c_int)
_sig("cfx_csr_mg_tail_level", C.c_int, _vp, C.POINTER(C.c_int))
_sig("cfx_csr_mg_slices", C.c_int, _vp, C.POINTER(C.c_int64))
_sig("cfx_csr_mg_level", C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp)
_sig("cfx_csr_mg_apply", C.c_int, _vp, _vp, _vp)
_sig("cfx_csr_dic_info", C.c_int, _vp, C.POINTER(C.c_int), _vp)
_sig("cfx_csr_dic_apply", C.c_int, _vp, _vp, _vp)
_sig("cfx_linearise", C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp)
_sig("cfx_simple_solve", C.c_int, _vp, C.c_double, C.POINTER(Schemes), C.POINTER(SimpleControls), _vp, _vp, _vp, _vp,
     _vp, _vp, _vp, C.POINTER(SimpleResult), _vp)
_sig("cfx_piso_step", C.c_int, _vp, C.c_double, C.POINTER(Schemes), C.POINTER(PisoControls), _vp, _vp, _vp, _vp,
     _vp, _vp, _vp, C.POINTER(PisoResult))
_sig("cfx_csr_time_spmv", C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, _dp)
_sig("cfx_matrix_time_spmv", C.c_int, _vp, C.c_int, _vp, _vp, C.c_int, C.c_int, _dp)
_sig("cfx_matrix_time_fcg_spmv", C.c_int, _vp, C.c_int, C.POINTER(C.c_double))
_sig("cfx_set_fcg_probe", C.c_int, _vp, C.c_int)
_sig("cfx_matrix_fcg_probe", C.c_int, _vp, C.POINTER(C.c_double), C.POINTER(C.c_int))
_sig("cfx_matrix_smooth_probe", C.c_int, _vp, C.POINTER(C.c_double), C.POINTER(C.c_int))
_sig("cfx_matrix_time_smooth", C.c_int, _vp, C.c_int, C.POINTER(C.c_double))
# host mesh
_sig("cfx_hmesh_box", C.c_int, _i64, _i64, _i64, _vp, _vp, C.c_double, C.c_uint64, C.c_int, C.POINTER(_vp))
_sig("cfx_hmesh_channel", C.c_int, _i64, _i64, _i64, _vp, _vp, C.c_double, C.c_uint64, C.POINTER(_vp))
_sig("cfx_detail_create", C.c_int, _vp, _vp, C.c_double, C.c_double, C.c_char_p, C.POINTER(_vp))
_sig("cfx_detail_destroy", C.c_int, _vp)
_sig("cfx_detail_loss_at", C.c_int, _vp, C.c_double, _vp, _vp)
_sig("cfx_detail_loss_at_many", C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, _vp)
_sig("cfx_detail_info", C.c_int, _vp, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp, C.POINTER(_i64),
     C.POINTER(_i64))
_sig("cfx_detail_inlet", C.c_int, _vp, _vp, _vp)
_sig("cfx_detail_fields", C.c_int, _vp, _vp, _vp, _vp)
_sig("cfx_detail_mesh", _vp, _vp)
_sig("cfx_hmesh_obstacle_channel", C.c_int, _i64, _i64, _i64, _vp, _vp, C.c_double, C.c_uint64, C.c_double,
     C.c_double, C.c_double, C.POINTER(_vp))
_sig("cfx_hmesh_from_arrays", C.c_int, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp,
     C.POINTER(_vp))
_sig("cfx_hmesh_free", C.c_int, _vp)
_sig("cfx_hmesh_sizes", C.c_int, _vp, _vp)
_sig("cfx_hmesh_get", C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("cfx_hmesh_geometry", C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("cfx_hmesh_shuffle", C.c_int, _vp, C.c_uint64)
_sig("cfx_hmesh_renumber_rcm", C.c_int, _vp, _vp)
_sig("cfx_hmesh_bandwidth", C.c_int, _vp, C.POINTER(_i64))
_sig("cfx_hmesh_last_error", C.c_char_p)
_sig("cfx_hmesh_read_openfoam", C.c_int, C.c_char_p, C.c_char_p, C.POINTER(_vp))
_sig("cfx_hmesh_read_gmsh", C.c_int, C.c_char_p, C.c_char_p, C.POINTER(_vp))
_sig("cfx_hmesh_patch", C.c_int, _vp, _i64, C.c_char_p, _i64, C.c_char_p, _i64)
_sig("cfx_vtu_write", C.c_int, _vp, C.c_char_p, C.c_int, _vp, _vp, _vp)


def lib():
    return _L


def declared_symbols():
    """Function names declared in include/*.h."""
    import re
    names = []
    for h in sorted(os.listdir(INCLUDE_DIR)):
        if h.endswith(".h"):
            txt = open(os.path.join(INCLUDE_DIR, h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"\b(cfx_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def _p(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, _vp):
        return a
    return a.ctypes.data_as(_vp)


def _host_check(rc):
    if rc != CFX_OK:
        raise CfxError(rc, _L.cfx_hmesh_last_error().decode())


# --------------------------------------------------------------- context
class Context:
    """A HIP context (one device, one stream)."""

    def __init__(self, device=0):
        h = _vp()
        rc = _L.cfx_create(device, C.byref(h))
        if rc != CFX_OK:
            raise CfxError(rc, f"cannot create a context on device {device}")
        self.h = h

    def check(self, rc):
        if rc != CFX_OK:
            raise CfxError(rc, _L.cfx_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            _L.cfx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self.check(_L.cfx_synchronize(self.h))

    def set_graphs(self, enable: bool):
        """Flexible CG: replay one captured iteration per step (True) or launch kernel by kernel."""
        self.check(_L.cfx_set_graphs(self.h, 1 if enable else 0))

    def set_fcg_probe(self, enable: bool):
        """Measurement mode: multigrid solves time each fine-level k_fcg_A launch with HIP events."""
        self.check(_L.cfx_set_fcg_probe(self.h, 1 if enable else 0))

    def stream_bandwidth(self, nbytes=1 << 30, reps=20):
        """(read GB/s, copy GB/s read + written) of plain streams over nbytes: the roofline yardstick."""
        rd, cp = C.c_double(0), C.c_double(0)
        self.check(_L.cfx_stream_bandwidth(self.h, nbytes, reps, C.byref(rd), C.byref(cp)))
        return rd.value, cp.value

    def set_component_batching(self, enable: bool):
        """FP64 BiCGStab of vector equations: all components in shared launches (True) or one by one."""
        self.check(_L.cfx_set_component_batching(self.h, 1 if enable else 0))

    def device_info(self):
        buf = C.create_string_buffer(64)
        cus = C.c_int(0)
        self.check(_L.cfx_device_info(self.h, buf, 64, C.byref(cus)))
        return buf.value.decode(), cus.value

    def empty(self, shape, dtype=np.float64):
        return DeviceArray(self, shape, dtype)

    def zeros(self, shape, dtype=np.float64):
        a = DeviceArray(self, shape, dtype)
        self.check(_L.cfx_zero(self.h, a.ptr, a.nbytes))
        return a

    def array(self, host, dtype=None):
        host = np.ascontiguousarray(host, dtype=dtype)
        a = DeviceArray(self, host.shape, host.dtype)
        self.check(_L.cfx_memcpy_htod(self.h, a.ptr, _p(host), a.nbytes))
        return a


class DeviceArray:
    """Device buffer owned by a Context."""

    def __init__(self, ctx: Context, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        ctx.check(_L.cfx_malloc(ctx.h, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p

    def get(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx.check(_L.cfx_memcpy_dtoh(self.ctx.h, _p(out), self.ptr, self.nbytes))
        return out

    def set(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.nbytes == self.nbytes
        self.ctx.check(_L.cfx_memcpy_htod(self.ctx.h, self.ptr, _p(host), self.nbytes))

    def copy_from(self, other: "DeviceArray"):
        assert other.nbytes == self.nbytes
        self.ctx.check(_L.cfx_memcpy_dtod(self.ctx.h, self.ptr, other.ptr, self.nbytes))

    def offset(self, elems):
        """raw device pointer elems into the buffer"""
        return _vp(self.ptr.value + int(elems) * self.dtype.itemsize)

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                _L.cfx_free(self.ctx.h, self.ptr)
        except Exception:
            pass


# ------------------------------------------------------------ host mesh
def _types_spec(patch_types):
    if not patch_types:
        return b""
    return ";".join(f"{k}={v}" for k, v in patch_types.items()).encode()


class HostMesh:
    """The library's host polyMesh (BlockMesh, jitter, Kuhn tets, RCM)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def box(cls, nx, ny, nz, mn=(0, 0, 0), mx=(1, 1, 1), jitter=0.0, seed=0, tets=False):
        h = _vp()
        _host_check(_L.cfx_hmesh_box(nx, ny, nz, _p(np.asarray(mn, np.float64)), _p(np.asarray(mx, np.float64)),
                                     jitter, seed, int(tets), C.byref(h)))
        return cls(h)

    @classmethod
    def channel(cls, nx, ny, nz, mn, mx, jitter=0.0, seed=0):
        """Channel block with patches inlet (x min), outlet (x max) and four walls."""
        h = _vp()
        _host_check(_L.cfx_hmesh_channel(nx, ny, nz, _p(np.asarray(mn, np.float64)), _p(np.asarray(mx, np.float64)),
                                         jitter, seed, C.byref(h)))
        return cls(h)

    @classmethod
    def obstacle_channel(cls, nx, ny, nz, mn, mx, jitter, seed, cx, cy, radius):
        """Channel box minus a z-parallel cylinder of cells (7th patch "obstacle")."""
        h = _vp()
        _host_check(_L.cfx_hmesh_obstacle_channel(nx, ny, nz, _p(np.asarray(mn, np.float64)),
                                                  _p(np.asarray(mx, np.float64)), jitter, seed, cx, cy, radius,
                                                  C.byref(h)))
        return cls(h)

    @classmethod
    def read_openfoam(cls, directory, patch_types=None):
        """OpenFOAM ASCII polyMesh directory (or case root); patch_types: {name: type} overrides."""
        h = _vp()
        _host_check(_L.cfx_hmesh_read_openfoam(str(directory).encode(), _types_spec(patch_types), C.byref(h)))
        return cls(h)

    @classmethod
    def read_gmsh(cls, path, patch_types=None):
        """Gmsh ASCII .msh 2.2 / 4.1; patch_types: {name: type} (default wall)."""
        h = _vp()
        _host_check(_L.cfx_hmesh_read_gmsh(str(path).encode(), _types_spec(patch_types), C.byref(h)))
        return cls(h)

    def patches(self):
        """[(name, type)] in patch order."""
        out = []
        for i in range(self.sizes()["nPatches"]):
            name, typ = C.create_string_buffer(512), C.create_string_buffer(512)
            _host_check(_L.cfx_hmesh_patch(self.h, i, name, 512, typ, 512))
            out.append((name.value.decode(), typ.value.decode()))
        return out

    def write_vtu(self, path, fields=None):
        """fields: {name: array [nCells] or [nCells, 3]} written as cell data."""
        fields = fields or {}
        keep = [np.ascontiguousarray(v, np.float64) for v in fields.values()]
        names = (C.c_char_p * max(1, len(keep)))(*[k.encode() for k in fields.keys()])
        comps = (C.c_int * max(1, len(keep)))(*[1 if a.ndim == 1 else a.shape[1] for a in keep])
        data = (_vp * max(1, len(keep)))(*[a.ctypes.data_as(_vp) for a in keep])
        _host_check(_L.cfx_vtu_write(self.h, str(path).encode(), len(keep), C.cast(names, _vp), C.cast(comps, _vp),
                                     C.cast(data, _vp)))

    @classmethod
    def from_arrays(cls, points, faceStart, faceVerts, owner, neighbour, patchStart, patchSize):
        h = _vp()
        pts = np.ascontiguousarray(points, np.float64)
        arrs = [np.ascontiguousarray(a, np.int64) for a in (faceStart, faceVerts, owner, neighbour, patchStart,
                                                             patchSize)]
        _host_check(_L.cfx_hmesh_from_arrays(len(pts), _p(pts), len(arrs[2]), _p(arrs[0]), _p(arrs[1]),
                                             len(arrs[3]), _p(arrs[2]), _p(arrs[3]), len(arrs[4]), _p(arrs[4]),
                                             _p(arrs[5]), C.byref(h)))
        return cls(h)

    def __del__(self):
        try:
            if self.h:
                _L.cfx_hmesh_free(self.h)
        except Exception:
            pass

    def sizes(self):
        s = np.zeros(6, np.int64)
        _host_check(_L.cfx_hmesh_sizes(self.h, _p(s)))
        return dict(nPoints=int(s[0]), nFaces=int(s[1]), nInternal=int(s[2]), nCells=int(s[3]),
                    nPatches=int(s[4]), nFaceVerts=int(s[5]))

    def arrays(self):
        s = self.sizes()
        a = dict(points=np.zeros((s["nPoints"], 3)), faceStart=np.zeros(s["nFaces"] + 1, np.int64),
                 faceVerts=np.zeros(s["nFaceVerts"], np.int64), owner=np.zeros(s["nFaces"], np.int64),
                 neighbour=np.zeros(max(1, s["nInternal"]), np.int64),
                 patchStart=np.zeros(s["nPatches"], np.int64), patchSize=np.zeros(s["nPatches"], np.int64))
        _host_check(_L.cfx_hmesh_get(self.h, *[_p(a[k]) for k in ("points", "faceStart", "faceVerts", "owner",
                                                                   "neighbour", "patchStart", "patchSize")]))
        a["neighbour"] = a["neighbour"][:s["nInternal"]]
        a["nCells"] = s["nCells"]
        return a

    def geometry(self):
        s = self.sizes()
        nF, nC = s["nFaces"], s["nCells"]
        g = dict(Sf=np.zeros((nF, 3)), Cf=np.zeros((nF, 3)), magSf=np.zeros(nF), C=np.zeros((nC, 3)),
                 V=np.zeros(nC), dvec=np.zeros((nF, 3)), invDist=np.zeros(nF))
        _host_check(_L.cfx_hmesh_geometry(self.h, *[_p(g[k]) for k in ("Sf", "Cf", "magSf", "C", "V", "dvec",
                                                                        "invDist")]))
        return g

    def shuffle(self, seed):
        _host_check(_L.cfx_hmesh_shuffle(self.h, seed))

    def renumber_rcm(self):
        newId = np.zeros(self.sizes()["nCells"], np.int64)
        _host_check(_L.cfx_hmesh_renumber_rcm(self.h, _p(newId)))
        return newId

    def bandwidth(self):
        bw = _i64(0)
        _host_check(_L.cfx_hmesh_bandwidth(self.h, C.byref(bw)))
        return bw.value


# ------------------------------------------------------------ device mesh
class Mesh:
    """Device mesh: face streams, gathers, geometry coefficients, CSR pattern."""

    def __init__(self, ctx: Context, arrays: dict, geom: dict):
        self.ctx = ctx
        self.nCells = int(arrays["nCells"])
        self.nInternal = len(arrays["neighbour"])
        self.nFaces = len(arrays["owner"])
        self.nB = self.nFaces - self.nInternal
        self.patchStart = np.asarray(arrays["patchStart"], np.int64)
        self.patchSize = np.asarray(arrays["patchSize"], np.int64)
        own = np.ascontiguousarray(arrays["owner"], np.int64)
        nb = np.ascontiguousarray(arrays["neighbour"], np.int64)
        if len(nb) == 0:
            nb = np.zeros(1, np.int64)
        g = {k: np.ascontiguousarray(v, np.float64) for k, v in geom.items()}
        h = _vp()
        ctx.check(_L.cfx_mesh_create(ctx.h, self.nCells, self.nInternal, self.nFaces, _p(own), _p(nb), _p(g["Sf"]),
                                     _p(g["Cf"]), _p(g["C"]), _p(g["V"]), _p(g["dvec"]), _p(g["magSf"]),
                                     len(self.patchStart), _p(self.patchStart), _p(self.patchSize), C.byref(h)))
        self.h = h
        self.V = g["V"]

    @classmethod
    def borrowed(cls, ctx, handle, n_cells, n_faces, n_internal):
        """A view of a device mesh owned elsewhere (not destroyed by this object)."""
        m = cls.__new__(cls)
        m.ctx, m.h, m._borrowed = ctx, handle, True
        m.nCells, m.nFaces, m.nInternal = n_cells, n_faces, n_internal
        m.nB = n_faces - n_internal
        return m

    def set_mg(self, **kw):
        """Multigrid constants of the drivers' CFX_FP32_MG pressure solves."""
        p = mg_default_params()
        for k, v in kw.items():
            if not hasattr(p, k):
                raise KeyError(k)
            setattr(p, k, v)
        self.ctx.check(_L.cfx_mesh_set_mg(self.h, C.byref(p)))

    def set_mpnn(self, net):
        self._net = net
        self.ctx.check(_L.cfx_mesh_set_mpnn(self.h, net.h if net is not None else None))

    def set_pressure_dump(self, val, rhs, every, max_count):
        """Copy every `every`-th pressure system into device arrays val [max][nnz], rhs [max][n]."""
        self._dump = (val, rhs)
        self.ctx.check(_L.cfx_mesh_set_pressure_dump(self.h, _p(val) if val is not None else None,
                                                     _p(rhs) if rhs is not None else None, every, max_count))

    def pressure_dump_count(self):
        c = C.c_int(0)
        self.ctx.check(_L.cfx_mesh_pressure_dump_count(self.h, C.byref(c)))
        return c.value

    @classmethod
    def from_host(cls, ctx, hm: HostMesh):
        return cls(ctx, hm.arrays(), hm.geometry())

    def __del__(self):
        try:
            if self.h and self.ctx.h and not getattr(self, "_borrowed", False):
                _L.cfx_mesh_destroy(self.h)
        except Exception:
            pass

    def sizes(self):
        s = np.zeros(5, np.int64)
        self.ctx.check(_L.cfx_mesh_sizes(self.h, _p(s)))
        return dict(nCells=int(s[0]), nInternal=int(s[1]), nFaces=int(s[2]), nnz=int(s[3]), nRowBlocks=int(s[4]))

    def pattern(self):
        s = self.sizes()
        rp = np.zeros(s["nCells"] + 1, np.int32)
        cl = np.zeros(s["nnz"], np.int32)
        self.ctx.check(_L.cfx_mesh_pattern(self.h, _p(rp), _p(cl)))
        return rp, cl

    # field ops (device arrays in/out)
    def grad(self, x, linc, link, ncomp=1, comp=0, corrected=True, quirk=True, out=None):
        out = out if out is not None else self.ctx.empty((3, self.nCells))
        self.ctx.check(_L.cfx_grad(self.h, _p(x), _p(linc), _p(link), ncomp, comp, int(corrected), int(quirk),
                                   _p(out)))
        return out

    def harmonic_gamma_uniform(self, gamma, out=None):
        out = out if out is not None else self.ctx.empty(self.nFaces)
        self.ctx.check(_L.cfx_harmonic_gamma_uniform(self.h, gamma, _p(out)))
        return out

    def interpolate_to_faces(self, cell, out=None):
        out = out if out is not None else self.ctx.empty(self.nFaces)
        self.ctx.check(_L.cfx_interpolate_to_faces(self.h, _p(cell), _p(out)))
        return out

    def update_phi_internal(self, U3, phi):
        self.ctx.check(_L.cfx_update_phi_internal(self.h, _p(U3), _p(phi)))

    def correct_boundary_flux(self, U3, linc3, link3, phi):
        self.ctx.check(_L.cfx_correct_boundary_flux(self.h, _p(U3), _p(linc3), _p(link3), _p(phi)))

    def correct_fluxes(self, p, gradP3, rAUf, linc, link, corrected, phi):
        self.ctx.check(_L.cfx_correct_fluxes(self.h, _p(p), _p(gradP3), _p(rAUf), _p(linc), _p(link),
                                             int(corrected), _p(phi)))

    def div_phi(self, phi, out=None):
        out = out if out is not None else self.ctx.empty(self.nCells)
        self.ctx.check(_L.cfx_div_phi(self.h, _p(phi), _p(out)))
        return out

    def continuity_error(self, phi):
        e = C.c_double(0)
        self.ctx.check(_L.cfx_continuity_error(self.h, _p(phi), C.byref(e)))
        return e.value

    def linearise(self, ncomp, types, values):
        """device (lin_c, lin_k) [nB][ncomp] of per-boundary-face BC types/values"""
        t = np.ascontiguousarray(types, np.int32)
        v = np.ascontiguousarray(values, np.float64).reshape(self.nB, ncomp)
        lc, lk = self.ctx.empty((max(1, self.nB), ncomp)), self.ctx.empty((max(1, self.nB), ncomp))
        self.ctx.check(_L.cfx_linearise(self.h, ncomp, _p(t), _p(v), _p(lc), _p(lk)))
        return lc, lk

    def simple(self, nu, schemes: "Schemes", ctl: "SimpleControls", Ubc, pbc, U, p, phi, hist=False):
        """solveIncompressibleSimple on device fields U [3][n], p [n], phi [nFaces]"""
        ut, uv = np.ascontiguousarray(Ubc[0], np.int32), np.ascontiguousarray(Ubc[1], np.float64)
        pt, pv = np.ascontiguousarray(pbc[0], np.int32), np.ascontiguousarray(pbc[1], np.float64)
        res = SimpleResult()
        h = np.zeros((ctl.maxIterations, 4)) if hist else None
        self.ctx.check(_L.cfx_simple_solve(self.h, nu, C.byref(schemes), C.byref(ctl), _p(ut), _p(uv), _p(pt),
                                           _p(pv), _p(U), _p(p), _p(phi), C.byref(res), _p(h)))
        if hist:
            return res, h[:res.iterations]
        return res

    def piso_step(self, nu, schemes: "Schemes", ctl: "PisoControls", Ubc, pbc, U, p, phi):
        ut, uv = np.ascontiguousarray(Ubc[0], np.int32), np.ascontiguousarray(Ubc[1], np.float64)
        pt, pv = np.ascontiguousarray(pbc[0], np.int32), np.ascontiguousarray(pbc[1], np.float64)
        res = PisoResult()
        self.ctx.check(_L.cfx_piso_step(self.h, nu, C.byref(schemes), C.byref(ctl), _p(ut), _p(uv), _p(pt), _p(pv),
                                        _p(U), _p(p), _p(phi), C.byref(res)))
        return res

    def nonorth_correct(self, gammaFace, x, linc, link, rhs, ncomp=1, comp=0, gradCorrected=True, sign=-1):
        self.ctx.check(_L.cfx_nonorth_correct(self.h, _p(gammaFace), _p(x), _p(linc), _p(link), ncomp, comp,
                                              int(gradCorrected), sign, _p(rhs)))


class Matrix:
    """Device fvMatrix on the mesh's fixed pattern."""

    def __init__(self, mesh: Mesh, ncomp=1):
        self.mesh, self.ctx, self.ncomp = mesh, mesh.ctx, ncomp
        h = _vp()
        self.ctx.check(_L.cfx_matrix_create(mesh.h, ncomp, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                _L.cfx_matrix_destroy(self.h)
        except Exception:
            pass

    def set_mg(self, **kw):
        """Multigrid constants for precond=PRECOND_MG / PRECOND_LEARNED (cfx_mg_params fields)."""
        p = mg_default_params()
        for k, v in kw.items():
            if not hasattr(p, k):
                raise KeyError(k)
            setattr(p, k, v)
        self.mesh.ctx.check(_L.cfx_matrix_set_mg(self.h, C.byref(p)))
        return p

    def set_mpnn(self, net):
        self._net = net
        self.mesh.ctx.check(_L.cfx_matrix_set_mpnn(self.h, net.h if net is not None else None))

    def set_mg_weights(self, pre=None, scale=None, post=None, restrict=None):
        self._mgw = (pre, scale, post, restrict)
        self.mesh.ctx.check(_L.cfx_matrix_set_mg_weights(self.h, *[_p(a) if a is not None else None
                                                                   for a in self._mgw]))

    def reset(self):
        self.ctx.check(_L.cfx_matrix_reset(self.h))

    def laplacian(self, gammaFace, linc, link, sign=-1):
        self.ctx.check(_L.cfx_assemble_laplacian(self.h, _p(gammaFace), _p(linc), _p(link), sign))

    def div(self, phi, linc, link, scheme=0, blend=0.9, sign=1):
        self.ctx.check(_L.cfx_assemble_div(self.h, _p(phi), scheme, blend, _p(linc), _p(link), sign))

    def ddt(self, old, dt, sign=1):
        self.ctx.check(_L.cfx_assemble_ddt(self.h, _p(old), dt, sign))

    def add_rhs(self, comp, v):
        self.ctx.check(_L.cfx_matrix_add_rhs(self.h, comp, _p(v)))

    def get(self):
        n, nI = self.mesh.nCells, self.mesh.nInternal
        d = np.zeros((self.ncomp, n))
        r = np.zeros((self.ncomp, n))
        u = np.zeros(max(1, nI))
        lo = np.zeros(max(1, nI))
        self.ctx.check(_L.cfx_matrix_get(self.h, _p(d), _p(r), _p(u), _p(lo)))
        return d, r, u[:nI], lo[:nI]

    def solve(self, x, solver=SOLVER_CG, tol=1e-8, maxIter=2000, nonOrth=2, gammaFace=None, linc=None, link=None,
              sign=-1, gradCorrected=True, precond=PRECOND_JACOBI, precision=FP64):
        rep = Report()
        self.ctx.check(_L.cfx_matrix_solve(self.h, _p(x), solver, tol, maxIter, nonOrth, _p(gammaFace), _p(linc),
                                           _p(link), sign, int(gradCorrected), precond, precision, C.byref(rep)))
        return rep

    def time_spmv(self, comp, x, y, reps=20, precision=FP64):
        ms = C.c_double(0)
        self.ctx.check(_L.cfx_matrix_time_spmv(self.h, comp, _p(x), _p(y), reps, precision, C.byref(ms)))
        return ms.value

    def fcg_probe(self):
        """(average ms, launches) of the fine-level k_fcg_A in the last probed multigrid solve."""
        ms, cnt = C.c_double(0), C.c_int(0)
        self.ctx.check(_L.cfx_matrix_fcg_probe(self.h, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def smooth_probe(self):
        """(average ms, launches) of the multigrid's finest-level post-smoothing sweep in the last probed
        flexible-CG solve."""
        ms, cnt = C.c_double(0), C.c_int(0)
        self.ctx.check(_L.cfx_matrix_smooth_probe(self.h, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def time_smooth(self, reps=20):
        """Average ms of the finest-level post-smoothing sweep over reps back-to-back launches."""
        ms = C.c_double(0)
        self.ctx.check(_L.cfx_matrix_time_smooth(self.h, reps, C.byref(ms)))
        return ms.value

    def set_relaxation(self, relaxation):
        """Schemes::relaxation of this fvMatrix: applied after the final pass of solve()."""
        self.ctx.check(_L.cfx_matrix_set_relaxation(self.h, relaxation))

    def set_inner_tol(self, t):
        """Reliable-update factor of this matrix's FP32-refined solves (0: each solver's default)."""
        self.ctx.check(_L.cfx_matrix_set_inner_tol(self.h, t))

    def time_fcg_spmv(self, reps=20):
        ms = C.c_double(0)
        self.ctx.check(_L.cfx_matrix_time_fcg_spmv(self.h, reps, C.byref(ms)))
        return ms.value

    def explicit_operator(self, x, comp, out=None):
        out = out if out is not None else self.ctx.empty(self.mesh.nCells)
        self.ctx.check(_L.cfx_matrix_explicit_operator(self.h, _p(x), comp, _p(out)))
        return out

    def spmv(self, comp, x, out=None):
        out = out if out is not None else self.ctx.empty(self.mesh.nCells)
        self.ctx.check(_L.cfx_matrix_spmv(self.h, comp, _p(x), _p(out)))
        return out


class Csr:
    """Device CSR matrix for the linear-solver seam (solveCG / solveBiCGStab)."""

    def __init__(self, ctx: Context, rowStart, cols, vals):
        self.ctx = ctx
        rs = np.ascontiguousarray(rowStart, np.int64)
        cl = np.ascontiguousarray(cols, np.int64)
        vv = np.ascontiguousarray(vals, np.float64)
        self.n = len(rs) - 1
        h = _vp()
        ctx.check(_L.cfx_csr_create(ctx.h, self.n, _p(rs), _p(cl), _p(vv), C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                _L.cfx_csr_destroy(self.h)
        except Exception:
            pass

    def set_values_device(self, d_vals):
        """Values from a device array or raw device pointer (nnz FP64)."""
        self.ctx.check(_L.cfx_csr_set_values_device(self.h, _p(d_vals)))

    def set_values(self, vals):
        self.ctx.check(_L.cfx_csr_set_values(self.h, _p(np.ascontiguousarray(vals, np.float64))))

    def spmv(self, x, out=None):
        out = out if out is not None else self.ctx.empty(self.n)
        self.ctx.check(_L.cfx_csr_spmv(self.h, _p(x), _p(out)))
        return out

    def solve(self, b, x0=None, tol=1e-8, maxIter=2000, solver=SOLVER_CG, precond=PRECOND_JACOBI, precision=FP64):
        """Host b / x (the reference solveCG signature); returns (x, Report)."""
        x = np.zeros(self.n) if x0 is None else np.array(x0, np.float64)
        rep = Report()
        self.ctx.check(_L.cfx_csr_solve(self.h, _p(np.ascontiguousarray(b, np.float64)), _p(x), tol, maxIter, solver,
                                        precond, precision, C.byref(rep)))
        return x, rep

    def solve_device(self, b, x, tol=1e-8, maxIter=2000, solver=SOLVER_CG, precond=PRECOND_JACOBI, precision=FP64):
        rep = Report()
        self.ctx.check(_L.cfx_csr_solve_device(self.h, _p(b), _p(x), tol, maxIter, solver, precond, precision,
                                               C.byref(rep)))
        return rep

    def residual_history(self, maxLen):
        h = np.zeros(maxLen)
        self.ctx.check(_L.cfx_csr_residual_history(self.h, _p(h), maxLen))
        return h

    def time_spmv(self, x, y, reps=20, precision=FP64):
        ms = C.c_double(0)
        self.ctx.check(_L.cfx_csr_time_spmv(self.h, _p(x), _p(y), reps, precision, C.byref(ms)))
        return ms.value

    def set_inner_tol(self, t):
        self.ctx.check(_L.cfx_csr_set_inner_tol(self.h, t))

    def set_mg(self, **kw):
        """Multigrid constants (cfx_mg_params fields); unspecified fields keep the library defaults."""
        p = mg_default_params()
        for k, v in kw.items():
            if not hasattr(p, k):
                raise KeyError(k)
            setattr(p, k, v)
        self.ctx.check(_L.cfx_csr_set_mg(self.h, C.byref(p)))
        return p

    def set_mg_weights(self, pre=None, scale=None, post=None, restrict=None):
        """Per-row finest-level a, s, c, t: device FP32 arrays (kept referenced here) or None."""
        self._mgw = (pre, scale, post, restrict)
        self.ctx.check(_L.cfx_csr_set_mg_weights(self.h, *[_p(a) if a is not None else None for a in self._mgw]))

    def mg_info(self):
        lv = C.c_int(0)
        rows = np.zeros(16, np.int64)
        nnz = np.zeros(16, np.int64)
        self.ctx.check(_L.cfx_csr_mg_info(self.h, C.byref(lv), _p(rows), _p(nnz), 16))
        return rows[:lv.value].tolist(), nnz[:lv.value].tolist()

    def mg_tail_level(self):
        lv = C.c_int(0)
        self.ctx.check(_L.cfx_csr_mg_tail_level(self.h, C.byref(lv)))
        return lv.value

    def mg_slices(self):
        """Slots of the finest level's sliced layout (0: the pattern does not allow it)."""
        n = C.c_int64(0)
        self.ctx.check(_L.cfx_csr_mg_slices(self.h, C.byref(n)))
        return n.value

    def mg_level(self, l):
        """(rowPtr, col, FP64 values, agg or None) of hierarchy level l."""
        rows, nnz = self.mg_info()
        n, z = rows[l], nnz[l]
        rp = np.zeros(n + 1, np.int32)
        cl = np.zeros(z, np.int32)
        vv = np.zeros(z)
        agg = np.zeros(n, np.int32) if l + 1 < len(rows) else None
        self.ctx.check(_L.cfx_csr_mg_level(self.h, l, _p(rp), _p(cl), _p(vv), _p(agg) if agg is not None else None))
        return rp, cl, vv, agg

    def mg_apply(self, r, z):
        self.ctx.check(_L.cfx_csr_mg_apply(self.h, _p(r), _p(z)))

    def dic_info(self):
        """(level count, FP64 DIC factor D) of the current values."""
        lv = C.c_int(0)
        d = np.zeros(self.n)
        self.ctx.check(_L.cfx_csr_dic_info(self.h, C.byref(lv), _p(d)))
        return lv.value, d

    def dic_apply(self, r, z):
        self.ctx.check(_L.cfx_csr_dic_apply(self.h, _p(r), _p(z)))

    def mg_loss_grad(self, e, a, t):
        """(L, dL/da, dL/dt) of the cycle-energy objective for error sample e (device FP32 arrays)."""
        ga, gt = self.ctx.zeros(self.n, np.float32), self.ctx.zeros(self.n, np.float32)
        loss = C.c_double(0)
        self.ctx.check(_L.cfx_csr_mg_loss_grad(self.h, _p(e), _p(a), _p(t), _p(ga), _p(gt), C.byref(loss)))
        return loss.value, ga, gt

    def set_mpnn(self, net, geo=None):
        """Learned solves take their finest-level weights from `net` (None detaches)."""
        self._net = (net, geo)
        self.ctx.check(_L.cfx_csr_set_mpnn(self.h, net.h if net is not None else None,
                                           _p(geo) if geo is not None else None))


def mg_default_params():
    p = MgParams()
    if _L.cfx_mg_default_params(C.byref(p)):
        raise CfxError(-1, "cfx_mg_default_params")
    return p


class Mpnn:
    """Message-passing network giving the finest-level multigrid weights (cfx_mpnn)."""

    nparams = _L.cfx_mpnn_num_params()

    def __init__(self, ctx: Context, seed=1, fp16=True):
        self.ctx = ctx
        h = _vp()
        ctx.check(_L.cfx_mpnn_create(ctx.h, seed, 1 if fp16 else 0, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                _L.cfx_mpnn_destroy(self.h)
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        self.ctx.check(_L.cfx_mpnn_get_params(self.h, _p(p)))
        return p

    def set_params(self, p):
        self.ctx.check(_L.cfx_mpnn_set_params(self.h, _p(np.ascontiguousarray(p, np.float32))))

    def forward(self, csr, geo=None, r0=None, train=False):
        a, t = self.ctx.empty(csr.n, np.float32), self.ctx.empty(csr.n, np.float32)
        self.ctx.check(_L.cfx_mpnn_forward_csr(self.h, csr.h, _p(geo) if geo is not None else None,
                                               _p(r0) if r0 is not None else None, _p(a), _p(t),
                                               1 if train else 0))
        return a, t

    def backward(self, ga, gt):
        g = self.ctx.zeros(self.nparams, np.float32)
        self.ctx.check(_L.cfx_mpnn_backward(self.h, _p(ga), _p(gt), _p(g)))
        return g

    def adam(self, grad, lr):
        self.ctx.check(_L.cfx_mpnn_adam(self.h, _p(grad), lr))

    @staticmethod
    def _opts(steps=100, samples=2, smooth=4, lr=1e-3, seed=1):
        return TrainOptions(steps, samples, smooth, lr, seed)

    def train(self, systems, geos=None, r0s=None, **kw):
        o = self._opts(**kw)
        n = len(systems)
        S = (_vp * n)(*[s.h for s in systems])
        G = (_vp * n)(*[_p(g) if g is not None else None for g in geos]) if geos is not None else None
        R = (_vp * n)(*[_p(r) if r is not None else None for r in r0s]) if r0s is not None else None
        hist = np.zeros(max(o.steps, 1))
        self.ctx.check(_L.cfx_mpnn_train(self.h, C.cast(S, _vp), C.cast(G, _vp) if G is not None else None,
                                         C.cast(R, _vp) if R is not None else None, n, C.byref(o), _p(hist)))
        return hist[:o.steps]

    def evaluate(self, csr, geo=None, r0=None, samples=4, with_network=True, **kw):
        o = self._opts(**kw)
        loss = C.c_double(0)
        self.ctx.check(_L.cfx_mpnn_evaluate(self.h, csr.h, _p(geo) if geo is not None else None,
                                            _p(r0) if r0 is not None else None, C.byref(o), samples,
                                            1 if with_network else 0, C.byref(loss)))
        return loss.value


NOALLOC_MAX_ROWS = 524288


def pcg_solve_noalloc(ctx: Context, d_rowPtr, d_col, d_val, d_b, d_x, tol=1e-8, maxIter=2000):
    """cfx_pcg_solve_noalloc: FP64 Jacobi-PCG on a device CSR system (int32 d_rowPtr[n+1], d_col[nnz],
    FP64 d_val[nnz], d_b[n], d_x[n] = initial guess and result) with the library's static scratch."""
    n = int(np.prod(d_b.shape))
    nnz = int(np.prod(d_col.shape))
    rep = Report()
    ctx.check(_L.cfx_pcg_solve_noalloc(ctx.h, n, nnz, _p(d_rowPtr), _p(d_col), _p(d_val), _p(d_b), _p(d_x), tol,
                                       maxIter, C.byref(rep)))
    return rep


def _ptr_array(items):
    arr = (_vp * len(items))(*[_p(a) if not isinstance(a, _vp) else a for a in items])
    return arr


def batch_solve(ctx, systems, bs, xs, tol=1e-8, maxIter=5000, schedule=None):
    """Jacobi-PCG on many CSR systems in one batched launch sequence.
    Returns (reports, finishMs, stats)."""
    K = len(systems)
    S = (_vp * K)(*[s.h for s in systems])
    B = _ptr_array(bs)
    X = _ptr_array(xs)
    reps = (Report * K)()
    fin = np.zeros(K)
    st = BatchStats()
    sched = C.byref(schedule) if schedule is not None else None
    ctx.check(_L.cfx_batch_solve(ctx.h, K, C.cast(S, _vp), C.cast(B, _vp), C.cast(X, _vp), tol, maxIter, sched,
                                 C.cast(reps, _vp), _p(fin), C.byref(st)))
    return list(reps), fin, st


def batch_autotune(ctx, systems, bs, xs, iters=20):
    K = len(systems)
    S = (_vp * K)(*[s.h for s in systems])
    B = _ptr_array(bs)
    X = _ptr_array(xs)
    best = BatchSchedule()
    cand = (BatchSchedule * 64)()
    ms = np.zeros(64)
    ctx.check(_L.cfx_batch_autotune(ctx.h, K, C.cast(S, _vp), C.cast(B, _vp), C.cast(X, _vp), iters, C.byref(best),
                                    C.cast(cand, _vp), _p(ms), 64))
    n = int((ms > 0).sum())
    return best, [(cand[i].as_dict(), ms[i]) for i in range(n)]


class DetailMeasure(C.Structure):
    _fields_ = [("cells", C.c_int64), ("sweeps", C.c_int), ("converged", C.c_int), ("imposedQ", C.c_double),
                ("inletArea", C.c_double), ("plugSpeed", C.c_double), ("dpPa", C.c_double), ("fluxIn", C.c_double),
                ("fluxOut", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Detail:
    """A 3D detail region (DetailBlock): inlet profile, GPU SIMPLE, section-mean pressure drop."""

    def __init__(self, ctx: Context, hm: HostMesh, nu=1e-3, rho=1.0, shape="flat"):
        self.ctx, self.hm = ctx, hm
        h = _vp()
        _host_check(_L.cfx_detail_create(ctx.h, hm.h, nu, rho, shape.encode(), C.byref(h)))
        self.h = h
        a = hm.arrays()
        self.nCells, self.nFaces, self.nInternal = int(a["nCells"]), len(a["owner"]), len(a["neighbour"])

    def __del__(self):
        try:
            if self.h:
                _L.cfx_detail_destroy(self.h)
        except Exception:
            pass

    def loss_at(self, q, ctl: "SimpleControls"):
        m = DetailMeasure()
        _host_check(_L.cfx_detail_loss_at(self.h, q, C.byref(ctl), C.byref(m)))
        return m.dpPa, m.as_dict()

    def info(self):
        ax, fs, area, ns, ni = C.c_int(0), C.c_int(0), C.c_double(0), C.c_int64(0), C.c_int64(0)
        _host_check(_L.cfx_detail_info(self.h, C.byref(ax), C.byref(fs), C.byref(area), C.byref(ns), C.byref(ni)))
        return {"axis": ax.value, "flowSign": fs.value, "inletArea": area.value, "sections": ns.value,
                "inletFaces": ni.value}

    def inlet(self):
        k = self.info()["inletFaces"]
        f, w = np.zeros(k, np.int64), np.zeros(k)
        _host_check(_L.cfx_detail_inlet(self.h, _p(f), _p(w)))
        return f, w

    def fields(self):
        U, p, phi = np.zeros(3 * self.nCells), np.zeros(self.nCells), np.zeros(self.nFaces)
        _host_check(_L.cfx_detail_fields(self.h, _p(U), _p(p), _p(phi)))
        return U.reshape(3, self.nCells), p, phi

    def mesh(self):
        return Mesh.borrowed(self.ctx, _L.cfx_detail_mesh(self.h), self.nCells, self.nFaces, self.nInternal)


def detail_loss_at_many(blocks, qs, ctl: "SimpleControls", threads=1):
    """lossAt of every block (Detail objects; with threads > 1 each needs its own Context):
    returns (dp list, measure dicts)."""
    n = len(blocks)
    H = (_vp * n)(*[b.h for b in blocks])
    q = np.ascontiguousarray(qs, np.float64)
    m = (DetailMeasure * n)()
    _host_check(_L.cfx_detail_loss_at_many(n, C.cast(H, _vp), _p(q), C.byref(ctl), int(threads), C.cast(m, _vp)))
    return [x.dpPa for x in m], [x.as_dict() for x in m]
