"""ctypes binding of libphyflow_amd.so (the C ABI declared in
include/phyflow_amd.h).

The library is loaded lazily on first use.  If it is missing, or if the
GPU is not usable, calls raise RuntimeError: there is no CPU fallback.
ctypes.CDLL releases the GIL for the duration of each call.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphyflow_amd.so")

PHF_OK = 0
PREC_F64 = 0
PREC_F32C = 1
REMESH_M4 = 0
REMESH_TUTTY = 1

STAGES = ("upload", "tree", "p2m", "m2m", "m2l", "l2l", "p2p", "download",
          "remesh", "panel", "tree_far", "total")

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_ptr = ctypes.c_void_p

_lock = threading.Lock()
_lib = None


class PhfError(RuntimeError):
    """A libphyflow_amd call returned an error status."""

    def __init__(self, code, msg):
        super().__init__("libphyflow_amd error %d: %s" % (code, msg))
        self.code = code


class FmmOpts(ctypes.Structure):
    _fields_ = [("order", _i32), ("leaf", _i32), ("prec", _i32), ("ws", _i32)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", _i32), ("converged", _i32), ("lu_ms", _dbl), ("refine_ms", _dbl),
                ("total_ms", _dbl)]


class PanelsTiming(ctypes.Structure):
    _fields_ = [("influence_ms", _dbl), ("normal_ms", _dbl), ("lu_ms", _dbl),
                ("inverse_ms", _dbl), ("total_ms", _dbl)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepParams(ctypes.Structure):
    _fields_ = [("dt", _dbl), ("vinf_x", _dbl), ("vinf_y", _dbl), ("panel_update", _i32),
                ("remesh", _i32), ("remesh_sigma", _dbl), ("remesh_overlap", _dbl),
                ("remesh_x0", _dbl), ("remesh_y0", _dbl), ("remesh_c", _dbl),
                ("pc_enabled", _i32), ("remesh_kernel", _i32), ("pc_w_threshold", _dbl),
                ("pc_total_threshold", _dbl)]


class StepInfo(ctypes.Structure):
    _fields_ = [("n_before", _i64), ("n_after", _i64), ("circ_before", _dbl), ("circ_after", _dbl),
                ("impulse_x_before", _dbl), ("impulse_x_after", _dbl),
                ("impulse_y_before", _dbl), ("impulse_y_after", _dbl),
                ("m2_before", _dbl), ("m2_after", _dbl), ("solve_iterations", _i32),
                ("pad", _i32), ("ms_convection", _dbl), ("ms_remesh", _dbl), ("ms_total", _dbl)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class BatchParams(ctypes.Structure):
    _fields_ = [("prec", _i32), ("order", _i32), ("leaf", _i32), ("ws", _i32), ("k", _dbl)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("ms", _dbl), ("graph", _i32), ("flags", _i32), ("total_points", _i64),
                ("levels_max", _i32), ("pad", _i32), ("remeshed", _i64), ("spilled", _i32),
                ("pad2", _i32)]


class BatchRemeshParams(ctypes.Structure):
    _fields_ = [("dt", _dbl), ("vinf_x", _dbl), ("vinf_y", _dbl), ("overlap", _dbl),
                ("x0h", _dbl), ("y0h", _dbl), ("c", _dbl), ("cell_factor", _dbl)]


class RequestResult(ctypes.Structure):
    _fields_ = [("status", _i32), ("pad", _i32), ("latency_ms", _dbl), ("queue_ms", _dbl),
                ("info", StepInfo)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", _i64), ("failed", _i64), ("p50_ms", _dbl), ("p99_ms", _dbl),
                ("max_ms", _dbl), ("span_s", _dbl), ("hbm_used", _i64), ("hbm_budget", _i64),
                ("slots", _i32), ("pad", _i32)]


METHOD_FMM = 0
METHOD_DIRECT = 1
N_STAGES = len(STAGES)


class VelParams(ctypes.Structure):
    _fields_ = [("method", _i32), ("prec", _i32), ("order", _i32), ("leaf", _i32),
                ("ws", _i32), ("ndirect", _i32), ("sigma", _dbl), ("k", _dbl),
                ("tol", _dbl), ("cutoff", _dbl), ("leaf_side", _dbl)]


class EvalInfo(ctypes.Structure):
    _fields_ = [("L", _i32), ("p", _i32), ("nnear", _i32), ("flags", _i32), ("D", _dbl),
                ("p2p_pairs", _i64), ("m2l_links", _i64), ("ms", _dbl * N_STAGES)]

    def as_dict(self):
        return {"L": self.L, "p": self.p, "nnear": self.nnear, "D": self.D,
                "p2p_pairs": self.p2p_pairs, "m2l_links": self.m2l_links,
                "ms": {k: self.ms[i] for i, k in enumerate(STAGES) if self.ms[i] >= 0}}


_SIGS = {
    "phf_last_error": ([], ctypes.c_char_p),
    "phf_init": ([], ctypes.c_int),
    "phf_device_info": ([ctypes.c_char_p, _i64, ctypes.POINTER(_i32), ctypes.POINTER(_i64)], ctypes.c_int),
    "phf_set_timing": ([_i32], ctypes.c_int),
    "phf_synchronize": ([], ctypes.c_int),
    "phf_last_stage_times": ([_ptr], ctypes.c_int),
    "phf_vv2par_gauss": ([_ptr, _ptr, _ptr, _i64, _ptr, _ptr, _i64, _dbl, _i32, _ptr, _ptr], ctypes.c_int),
    "phf_vv2par_gauss_noalloc": ([_ptr, _ptr, _ptr, _i64, _ptr, _ptr, _i64, _dbl, _i32, _ptr, _ptr,
                                  _ptr], ctypes.c_int),
    "phf_ww2par_gauss": ([_ptr, _ptr, _ptr, _i64, _ptr, _ptr, _i64, _dbl, _i32, _ptr], ctypes.c_int),
    "phf_vv2par_cutoff": ([_ptr, _ptr, _ptr, _i64, _ptr, _ptr, _i64, _dbl, _i32, _ptr, _ptr], ctypes.c_int),
    "phf_ww2par_cutoff": ([_ptr, _ptr, _ptr, _i64, _ptr, _ptr, _i64, _dbl, _i32, _ptr], ctypes.c_int),
    "phf_fmm2d": ([_ptr, _ptr, _ptr, _i64, _ptr, _ptr, _i64, _i32, _dbl, _dbl, _dbl,
                   ctypes.POINTER(FmmOpts), _ptr, _ptr], ctypes.c_int),
    "phf_system_create": ([_ptr, _ptr, _ptr, _i64, ctypes.POINTER(_ptr)], ctypes.c_int),
    "phf_system_destroy": ([_ptr], ctypes.c_int),
    "phf_system_size": ([_ptr, ctypes.POINTER(_i64)], ctypes.c_int),
    "phf_system_get": ([_ptr, _ptr, _ptr, _ptr], ctypes.c_int),
    "phf_system_set": ([_ptr, _ptr, _ptr, _ptr, _i64], ctypes.c_int),
    "phf_system_velocity": ([_ptr, ctypes.POINTER(VelParams), _ptr, _ptr, ctypes.POINTER(EvalInfo)],
                            ctypes.c_int),
    "phf_system_get_velocity": ([_ptr, _ptr, _ptr], ctypes.c_int),
    "phf_m4kernel_1d": ([_ptr, _i64, _dbl, _ptr], ctypes.c_int),
    "phf_regrid_m4": ([_ptr, _ptr, _ptr, _i64, _dbl, _dbl, _dbl, _dbl, _dbl, ctypes.POINTER(_ptr)],
                      ctypes.c_int),
    "phf_diffusion_tutty": ([_ptr, _ptr, _ptr, _i64, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl,
                             ctypes.POINTER(_ptr)], ctypes.c_int),
    "phf_blobset_size": ([_ptr], _i64),
    "phf_population_control": ([_ptr, _ptr, _ptr, _i64, _dbl, _dbl, ctypes.POINTER(_ptr)],
                               ctypes.c_int),
    "phf_blobset_copy": ([_ptr, _ptr, _ptr, _ptr, _ptr, _ptr], ctypes.c_int),
    "phf_blobset_free": ([_ptr], None),
    "phf_panel_influence": ([_ptr] * 8 + [_i64, _ptr], ctypes.c_int),
    "phf_panel_induced_velocity": ([_ptr, _i64, _ptr, _ptr, _i64] + [_ptr] * 8, ctypes.c_int),
    "phf_lu_factor_f32": ([_ptr, _i64, _ptr, _ptr, ctypes.POINTER(_i32)], ctypes.c_int),
    "phf_chol_factor_f32": ([_ptr, _i64, _ptr, ctypes.POINTER(_i32)], ctypes.c_int),
    "phf_dense_solve": ([_ptr, _i64, _i64, _ptr, _ptr, ctypes.POINTER(SolveInfo)], ctypes.c_int),
    "phf_panels_create": ([_ptr, _ptr, _ptr, _ptr, _i64, _i32, _dbl, ctypes.POINTER(_ptr)],
                          ctypes.c_int),
    "phf_panels_destroy": ([_ptr], ctypes.c_int),
    "phf_panels_info": ([_ptr, ctypes.POINTER(_i64), _ptr, _ptr, _ptr], ctypes.c_int),
    "phf_panels_update": ([_ptr, _ptr, _ptr, _ptr, _ptr, ctypes.POINTER(PanelsTiming)],
                          ctypes.c_int),
    "phf_panels_update_async": ([_ptr, _ptr, _ptr, _ptr, _ptr], ctypes.c_int),
    "phf_panels_solve": ([_ptr, _ptr, _ptr, _ptr, ctypes.POINTER(_i32)], ctypes.c_int),
    "phf_system_step_rk2": ([_ptr, ctypes.POINTER(VelParams), _ptr, ctypes.POINTER(StepParams),
                             ctypes.POINTER(StepInfo)], ctypes.c_int),
    "phf_batch_create": ([_ptr, _ptr, _i32, ctypes.POINTER(BatchParams), ctypes.POINTER(_ptr)],
                         ctypes.c_int),
    "phf_batch_destroy": ([_ptr], ctypes.c_int),
    "phf_batch_set_system": ([_ptr, _i32, _ptr, _ptr, _ptr], ctypes.c_int),
    "phf_batch_velocity": ([_ptr, _i32, ctypes.POINTER(BatchInfo)], ctypes.c_int),
    "phf_batch_get_velocity": ([_ptr, _i32, _ptr, _ptr], ctypes.c_int),
    "phf_batch_set_remesh": ([_ptr, ctypes.POINTER(BatchRemeshParams)], ctypes.c_int),
    "phf_batch_step": ([_ptr, _i32, ctypes.POINTER(BatchInfo)], ctypes.c_int),
    "phf_batch_get_remeshed": ([_ptr, _i32, ctypes.POINTER(_i64), _ptr, _ptr, _ptr],
                               ctypes.c_int),
    "phf_server_create": ([_i32, _i64, ctypes.POINTER(_ptr)], ctypes.c_int),
    "phf_server_destroy": ([_ptr], ctypes.c_int),
    "phf_server_add_system": ([_ptr, _ptr, ctypes.POINTER(_i64)], ctypes.c_int),
    "phf_server_remove_system": ([_ptr, _i64], ctypes.c_int),
    "phf_server_submit": ([_ptr, _i64, ctypes.POINTER(VelParams), ctypes.POINTER(StepParams),
                           ctypes.POINTER(_i64)], ctypes.c_int),
    "phf_server_wait": ([_ptr, _i64, _dbl, ctypes.POINTER(RequestResult)], ctypes.c_int),
    "phf_server_drain": ([_ptr], ctypes.c_int),
    "phf_server_stats": ([_ptr, _i32, ctypes.POINTER(ServerStats)], ctypes.c_int),
    "phf_tree_build": ([_ptr, _ptr, _i64, _i32, _ptr, _ptr, _ptr, ctypes.POINTER(_dbl),
                        ctypes.POINTER(_dbl), ctypes.POINTER(_dbl)], ctypes.c_int),
}


def exported_symbols():
    """Names of the C entry points this binding declares."""
    return sorted(_SIGS)


def lib():
    """The loaded library (raises RuntimeError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    "libphyflow_amd.so is not built (expected at %s); run "
                    "__graft_entry__.build() or python phyflow2.0_amd/build.py" % LIB_PATH)
            L = ctypes.CDLL(LIB_PATH)
            for name, (args, res) in _SIGS.items():
                fn = getattr(L, name, None)
                if fn is None:
                    continue
                fn.argtypes = args
                fn.restype = res
            _lib = L
    return _lib


def check(rc):
    if rc != PHF_OK:
        msg = lib().phf_last_error()
        raise PhfError(rc, msg.decode() if msg else "")
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))


# ---------------------------------------------------------------- helpers --
def f64(a, name="array"):
    """Contiguous float64 1-D view/copy of a (the reference reads raw
    float64 buffers without checking layout; this shim enforces it)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 1:
        a = a.ravel()
    return a


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else ctypes.c_void_p(0)


def device_info():
    name = ctypes.create_string_buffer(256)
    cu = _i32()
    hbm = _i64()
    call("phf_device_info", name, 256, ctypes.byref(cu), ctypes.byref(hbm))
    return {"name": name.value.decode(), "compute_units": cu.value, "hbm_bytes": hbm.value}


def set_timing(on):
    call("phf_set_timing", 1 if on else 0)


def last_stage_times():
    ms = np.empty(len(STAGES))
    call("phf_last_stage_times", ptr(ms))
    return {k: float(v) for k, v in zip(STAGES, ms) if v >= 0}
