"""ctypes binding of ``libfastoverlap_hip.so`` (the C ABI in
``include/fastoverlap_hip.h``).

The library is the only compute path: if it is missing or cannot be loaded
every entry point raises ``FastOverlapError`` -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FASTOVERLAP_HIP_LIB", os.path.join(_HERE, "libfastoverlap_hip.so"))

FO_OK = 0
ERRORS = {1: "invalid argument", 2: "HIP error", 3: "out of memory", 4: "unsupported",
          5: "state", 6: "not found", 7: "cancelled"}

# spherical flags / stages
SPH_NO_INVERT, SPH_KEEP_GRID, SPH_NO_CENTER, SPH_GENERAL, SPH_DFT, SPH_AUTOTUNE = \
    1, 2, 4, 8, 16, 32
SPH_BETA_GEMM = 64
SPH_EXACT_PLANES = 128
STAGE_PREP, STAGE_COEFF, STAGE_GRID, STAGE_ROTOPT, STAGE_FINISH, STAGE_ALL = 1, 2, 4, 8, 16, 31
# periodic flags / stages
PER_KEEP_GRID, PER_INPUT_COEFFS, PER_KEEP_COEFFS, PER_KEEP_FFT = 2, 4, 8, 16
PSTAGE_SF, PSTAGE_GRID, PSTAGE_REFINE, PSTAGE_ALL = 1, 2, 4, 7


class FastOverlapError(RuntimeError):
    """Raised for any non-zero status of the C ABI (or a missing library)."""


_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)
_pf = C.POINTER(C.c_float)
_vp = C.c_void_p


class SphBatchDesc(C.Structure):
    _fields_ = [("npairs", C.c_int), ("grid_n", C.c_int), ("natoms", _pi), ("L", _pi),
                ("scale", _pd), ("ngroups", _pi), ("group_sizes", _pi),
                ("group_atoms", _pi), ("flags", C.c_int)]


class PerBatchDesc(C.Structure):
    _fields_ = [("npairs", C.c_int), ("nk", C.c_int), ("F", C.c_int), ("natoms", _pi),
                ("box", _pd), ("scale", _pd), ("ngroups", _pi), ("group_sizes", _pi),
                ("group_atoms", _pi), ("niter", C.c_int), ("flags", C.c_int)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("natoms", C.c_int), ("pos1", _pd), ("pos2", _pd),
                ("ngroups", C.c_int), ("group_sizes", _pi), ("group_atoms", _pi),
                ("L", C.c_int), ("grid_n", C.c_int), ("invert", C.c_int),
                ("box", C.c_double * 3), ("nk", C.c_int), ("F", C.c_int),
                ("scale", C.c_double)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("dist", C.c_double), ("R", C.c_double * 3),
                ("rmat", C.c_double * 9), ("inverted", C.c_int), ("disp", C.c_double * 3),
                ("ind", C.c_int * 6), ("latency_ms", C.c_double), ("queue_ms", C.c_double),
                ("batch_size", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch", C.c_int), ("streams", C.c_int), ("max_wait_ms", C.c_double),
                ("hbm_budget", C.c_ulonglong)]


_PROTOS = {
    "fo_version": (C.c_int, []),
    "fo_last_error": (C.c_char_p, []),
    "fo_device_count": (C.c_int, [_pi]),
    "fo_context_create": (C.c_int, [C.c_int, C.c_ulonglong, C.POINTER(_vp)]),
    "fo_context_destroy": (C.c_int, [_vp]),
    "fo_context_memory": (C.c_int, [_vp, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
    "fo_sph_batch_create": (C.c_int, [_vp, C.POINTER(SphBatchDesc), C.POINTER(_vp)]),
    "fo_sph_batch_create_harmonic": (C.c_int, [_vp, C.POINTER(SphBatchDesc), C.c_int,
                                               C.c_double, C.c_int, C.POINTER(_vp)]),
    "fo_sph_batch_coef_norm2": (C.c_int, [_vp, _pd]),
    "fo_sph_batch_harmonic_coeffs": (C.c_int, [_vp, C.c_int, _pd, _pd]),
    "fo_sph_batch_destroy": (C.c_int, [_vp]),
    "fo_sph_batch_footprint": (C.c_int, [_vp, C.POINTER(C.c_ulonglong)]),
    "fo_sph_batch_set_positions": (C.c_int, [_vp, _pd, _pd]),
    "fo_sph_batch_set_positions_device": (C.c_int, [_vp, _vp, _vp]),
    "fo_sph_batch_set_coeffs": (C.c_int, [_vp, _pd, _pd]),
    "fo_sph_batch_set_start": (C.c_int, [_vp, _pd]),
    "fo_sph_batch_set_rotation": (C.c_int, [_vp, _pd]),
    "fo_sph_batch_run": (C.c_int, [_vp, C.c_int, C.c_int]),
    "fo_sph_batch_sync": (C.c_int, [_vp]),
    "fo_sph_batch_results": (C.c_int, [_vp, _pd, _pd, _pd, _pi, _pd, _pd, _pi]),
    "fo_sph_batch_peaks": (C.c_int, [_vp, _pi, _pd, _pd, _pd, _pd, _pi, _pi]),
    "fo_sph_batch_opt": (C.c_int, [_vp, _pd, _pd, _pd, _pi, _pi]),
    "fo_sph_batch_scales": (C.c_int, [_vp, _pd]),
    "fo_sph_batch_coeffs": (C.c_int, [_vp, _pd]),
    "fo_sph_batch_grid": (C.c_int, [_vp, C.c_int, _pf]),
    "fo_sph_batch_norient": (C.c_int, [_vp]),
    "fo_sph_batch_timing": (C.c_int, [_vp, C.c_int, _pf]),
    "fo_sph_prepare": (C.c_int, [_vp, C.c_int]),
    "fo_sph_align_noalloc": (C.c_int, [_vp, C.POINTER(SphBatchDesc), _vp, _vp, _vp, _vp, _vp,
                                       _vp, _vp]),
    "fo_sph_batch_find_peaks": (C.c_int, [_vp, C.c_int, C.c_int, _pd, _pi, _pf]),
    "fo_sph_batch_start": (C.c_int, [_vp, _pd]),
    "fo_sph_batch_tuning": (C.c_int, [_vp, _pi, C.c_int]),
    "fo_sph_batch_precision": (C.c_int, [_vp, _pd, _pd, _pd]),
    "fo_sph_harm": (C.c_int, [_vp, C.c_int, C.c_int, _pd, _pd, _pd]),
    "fo_sph_energy": (C.c_int, [_vp, C.c_int, _pd, _pd, _pd, _pd]),
    "fo_sph_wigner_D": (C.c_int, [_vp, C.c_int, _pd, _pd, _pd]),
    "fo_sph_wigner_table": (C.c_int, [_vp, C.c_int, _pd]),
    "fo_soft_forward": (C.c_int, [_vp, C.c_int, _pd, _pd]),
    "fo_lap_align": (C.c_int, [_vp, C.c_int, _pd, _pd, C.c_int, _pi, _pi, C.c_int, _pd, _pi,
                               _pd]),
    "fo_per_batch_create": (C.c_int, [_vp, C.POINTER(PerBatchDesc), C.POINTER(_vp)]),
    "fo_per_batch_destroy": (C.c_int, [_vp]),
    "fo_per_batch_footprint": (C.c_int, [_vp, C.POINTER(C.c_ulonglong)]),
    "fo_per_batch_set_positions": (C.c_int, [_vp, _pd, _pd]),
    "fo_per_batch_set_positions_device": (C.c_int, [_vp, _vp, _vp]),
    "fo_per_batch_set_coeffs": (C.c_int, [_vp, _pd, _pd]),
    "fo_per_batch_set_disp": (C.c_int, [_vp, _pd]),
    "fo_per_batch_run": (C.c_int, [_vp, C.c_int, C.c_int]),
    "fo_per_batch_sync": (C.c_int, [_vp]),
    "fo_per_batch_results": (C.c_int, [_vp, _pd, _pd, _pd, _pi, _pd]),
    "fo_per_batch_peaks": (C.c_int, [_vp, _pi, _pd, _pd, _pd, _pd, _pi, _pi]),
    "fo_per_batch_coeffs": (C.c_int, [_vp, _pd, _pd]),
    "fo_per_batch_product": (C.c_int, [_vp, _pd]),
    "fo_per_batch_grid": (C.c_int, [_vp, C.c_int, _pf]),
    "fo_per_batch_timing": (C.c_int, [_vp, C.c_int, _pf]),
    "fo_per_batch_fft": (C.c_int, [_vp, C.c_int, _pf]),
    "fo_per_batch_find_peaks": (C.c_int, [_vp, C.c_int, C.c_int, _pd, _pi, _pf]),
    "fo_per_batch_set_coeffs_indexed": (C.c_int, [_vp, _pd, C.c_int, _pi, _pi]),
    "fo_server_create": (C.c_int, [_vp, C.POINTER(ServerConfig), C.POINTER(_vp)]),
    "fo_server_destroy": (C.c_int, [_vp]),
    "fo_server_submit": (C.c_int, [_vp, C.POINTER(Request), C.POINTER(C.c_longlong)]),
    "fo_server_wait": (C.c_int, [_vp, C.c_longlong, C.c_double, C.POINTER(Result), _pd, _pi]),
    "fo_server_stats": (C.c_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                  C.POINTER(C.c_ulonglong)]),
    "fo_f90_cluster_align": (C.c_int, [_pd, _pd, C.c_int, C.c_int, _pd, C.c_int, C.c_int, _pi,
                                       _pi, _pd, _pd, _pd, _pi, _pi]),
    "fo_f90_bulk_align": (C.c_int, [_pd, _pd, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_int, C.c_int, C.c_int, _pi, _pi, _pd, _pd,
                                    _pi, _pd]),
    "fo_oh_operation": (C.c_int, [C.c_int, _pd]),
    "fo_f90_cluster_align_harm": (C.c_int, [_pd, _pd, C.c_int, C.c_int, C.c_int, C.c_double,
                                            _pd, C.c_int, C.c_int, _pi, _pi, _pd, _pd, _pd, _pi,
                                            _pi]),
    "fo_f90_harm_overlap_matrices": (C.c_int, [_pd, C.c_int, C.c_int, C.c_int, C.c_int,
                                               C.c_double, C.c_double, C.c_int, _pi, _pi, _pd,
                                               _pd]),
    "fo_f90_bulk_defaults": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, _pd, _pi,
                                       _pi]),
    "fo_f90_bulk_align_group": (C.c_int, [_pd, C.c_int, _pd, C.c_int, C.c_int, C.c_double,
                                          C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_int, _pi, _pi, _pd, _pd]),
}

_lib = None
_lib_lock = threading.Lock()


def load(path: str | None = None):
    """Load the shared library (once) and bind every prototype."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise FastOverlapError(
                f"{p} not found: build it with `python -m fastoverlap_amd.build` "
                "(the HIP library is the only compute path)")
        lib = C.CDLL(p)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def exported_symbols():
    return list(_PROTOS)


def check(rc: int):
    if rc != FO_OK:
        msg = load().fo_last_error()
        msg = msg.decode() if msg else ""
        raise FastOverlapError(f"{ERRORS.get(rc, rc)}: {msg}")


def call(name: str, *args):
    check(getattr(load(), name)(*args))


# ---- numpy <-> ctypes -------------------------------------------------------

def dptr(a):
    return None if a is None else a.ctypes.data_as(_pd)


def iptr(a):
    return None if a is None else a.ctypes.data_as(_pi)


def fptr(a):
    return None if a is None else a.ctypes.data_as(_pf)


def as_f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---- default context ---------------------------------------------------------

class Context:
    """A device context (tables are cached per context)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = _vp()
        call("fo_context_create", device, hbm_budget, C.byref(h))
        self.handle = h
        self.device = device

    def memory(self):
        u, b = C.c_ulonglong(), C.c_ulonglong()
        call("fo_context_memory", self.handle, C.byref(u), C.byref(b))
        return u.value, b.value

    def close(self):
        if self.handle:
            load().fo_context_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass


_default = None
_default_lock = threading.Lock()


def default_context() -> Context:
    global _default
    with _default_lock:
        if _default is None:
            _default = Context(int(os.environ.get("FASTOVERLAP_DEVICE", "0")))
        return _default
