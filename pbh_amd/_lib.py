"""ctypes binding of `_pm3d_hip.so` (declarations in include/pm3d_hip.h).

Loaded once, from next to this module, as the reference loads its native kernels
(pbhgr/pbhgr/cosmo_ev.py:31-47: `Path(__file__).with_name("_kernels.so")`, explicit argtypes and
restype).  Unlike the reference there is no pure-Python fallback: if the library is missing or
does not load, every entry point raises.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_void_p
from pathlib import Path

import numpy as np

SO_PATH = Path(__file__).with_name("_pm3d_hip.so")

PM3D_FP64 = 0
PM3D_MIXED = 1

ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "EUNSUPPORTED", -5: "ESTATE"}


class Pm3dCfg(ctypes.Structure):
    _fields_ = [("N", c_int32), ("Np", c_int32), ("L", c_double), ("smooth", c_double),
                ("precision", c_int32), ("frac_bits", c_int32)]


class Pm3dStepScalars(ctypes.Structure):
    _fields_ = [("a", c_double), ("da", c_double), ("c1", c_double), ("den1", c_double),
                ("den_d", c_double), ("den2", c_double), ("s_old", c_double), ("s_new", c_double),
                ("sm2", c_double), ("need_forces", c_int32), ("pad", c_int32)]


class Pm3dPeakIC(ctypes.Structure):
    _fields_ = [("mu", c_double), ("sx", c_double), ("sy", c_double), ("sz", c_double),
                ("noise_rms", c_double), ("noise_kcut", c_double), ("coef", c_double), ("seed", ctypes.c_uint64)]


class Pm3dSchedEntry(ctypes.Structure):
    _fields_ = [("c1p", c_double), ("denp", c_double), ("c1", c_double), ("den1", c_double),
                ("da", c_double), ("den_d", c_double), ("den2", c_double), ("s", c_double), ("sm2", c_double)]


class Pm3dMemberStatus(ctypes.Structure):
    _fields_ = [("done", c_int32), ("nsteps", c_int32), ("jmin", c_double)]


PD = POINTER(c_double)
PI64 = POINTER(c_int64)

# name -> (restype, argtypes); every symbol of include/pm3d_hip.h
SIGNATURES = {
    "pm3d_create": (c_int, [POINTER(Pm3dCfg), POINTER(c_void_p)]),
    "pm3d_destroy": (c_int, [c_void_p]),
    "pm3d_last_error": (c_char_p, [c_void_p]),
    "pm3d_synchronize": (c_int, [c_void_p]),
    "pm3d_device_bytes": (c_int64, [c_void_p]),
    "pm3d_set_state": (c_int, [c_void_p, PD, PD, c_int64]),
    "pm3d_get_state": (c_int, [c_void_p, PD, PD]),
    "pm3d_cic_index": (c_int, [c_void_p, PD, c_int64, PI64, PD]),
    "pm3d_deposit": (c_int, [c_void_p, PD, c_int64, PD]),
    "pm3d_deposit_noalloc": (c_int, [c_void_p, c_int64, c_int32, c_double, c_void_p, c_void_p]),
    "pm3d_interp": (c_int, [c_void_p, PD, PD, c_int64, PD]),
    "pm3d_forces": (c_int, [c_void_p, c_double, c_double, c_double, PD, PD]),
    "pm3d_step": (c_int, [c_void_p, POINTER(Pm3dStepScalars), PD, PD]),
    "pm3d_deposit_serial": (c_int64, [c_void_p]),
    "pm3d_delta_max_at": (c_int, [c_void_p, c_int64, PD]),
    "pm3d_get_delta": (c_int, [c_void_p, PD]),
    "pm3d_rfftn": (c_int, [c_void_p, PD, PD]),
    "pm3d_irfftn": (c_int, [c_void_p, PD, PD]),
    "pm3d_poisson_gradient": (c_int, [c_void_p, PD, c_double, c_double, PD]),
    "pm3d_min_jacobian": (c_int, [c_void_p, PD]),
    "pm3d_set_timing": (c_int, [c_void_p, c_int]),
    "pm3d_get_timings": (c_int, [c_void_p, PI64, PD, c_int]),
    "pm3d_kernel_names": (c_char_p, []),
    "pm3d_totals": (c_int, [c_void_p, PD, PD, PD]),
    "pm3d_stream_probe": (c_int, [c_int64, c_int, PD, PD]),
    "pm3d_fft_tiles": (c_int, [c_void_p, POINTER(c_int32), c_int]),
    "pm3d_deposit_tiles": (c_int, [c_void_p, POINTER(c_int32)]),
    "pm3d_deposit_trials": (c_int, [c_void_p, POINTER(c_int64)]),
    "pm3d_frac_bits": (c_int, [c_void_p]),
    "pm3d_device_info": (c_int, [c_void_p, c_char_p, c_int, c_char_p, c_int, PI64]),
    "pm3d_set_analysis_sets": (c_int, [c_void_p, POINTER(c_int32), PI64, c_int]),
    "pm3d_analyse": (c_int, [c_void_p, PD, c_int, PD, PD, PI64, PD]),
    "pm3d_make_peak_ic": (c_int, [c_void_p, POINTER(Pm3dPeakIC)]),
    "pm3d_copy_state": (c_int, [c_void_p, c_void_p]),
    "pm3d_set_schedule": (c_int, [c_void_p, POINTER(Pm3dSchedEntry), c_int]),
    "pm3d_schedule_step": (c_int, [c_void_p]),
    "pm3d_schedule_status": (c_int, [c_void_p, POINTER(c_int), PD]),
    "pm3d_schedule_end": (c_int, [c_void_p, POINTER(c_int)]),
    "pm3d_batch_create": (c_int, [POINTER(c_void_p)]),
    "pm3d_batch_destroy": (c_int, [c_void_p]),
    "pm3d_batch_last_error": (c_char_p, [c_void_p]),
    "pm3d_batch_set": (c_int, [c_void_p, POINTER(c_void_p), c_int]),
    "pm3d_batch_run": (c_int, [c_void_p, c_int]),
    "pm3d_batch_set_steps": (c_int, [c_void_p, c_int]),
    "pm3d_batch_synchronize": (c_int, [c_void_p]),
    "pm3d_batch_status": (c_int, [c_void_p, POINTER(Pm3dMemberStatus), c_int]),
}

_lib = None


def load():
    """Load the library (once).  Raises if it is missing: there is no CPU path."""
    global _lib
    if _lib is not None:
        return _lib
    if not SO_PATH.exists():
        raise RuntimeError(f"{SO_PATH.name} is not built; run `python -m pbh_amd.build`")
    lib = ctypes.CDLL(str(SO_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class Pm3dError(RuntimeError):
    pass


def check(rc: int, handle=None) -> int:
    """Raise on a negative return code; pass counts (>= 0) through."""
    if rc < 0:
        msg = load().pm3d_last_error(handle)
        raise Pm3dError(f"pm3d {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")
    return rc


def dptr(a: np.ndarray):
    return a.ctypes.data_as(PD)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(PI64)


def as_f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)
