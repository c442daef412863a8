"""ctypes bindings of the C ABI in include/remc.h (libremc.so).

The library is loaded from this package directory; a missing or unloadable
library raises immediately -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libremc.so")

NTYPES = 20


class remc_forcefield(ctypes.Structure):
    _fields_ = [("table", ctypes.c_float * (NTYPES * NTYPES)),
                ("lambda_", ctypes.c_double), ("e0", ctypes.c_double), ("xi", ctypes.c_double),
                ("lj_cf", ctypes.c_double), ("dh_cf", ctypes.c_double), ("kbt_cf", ctypes.c_double),
                ("crowder_sigma", ctypes.c_double), ("crowder_cutoff", ctypes.c_double)]


_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)


class remc_system_desc(ctypes.Structure):
    _fields_ = [("n_beads", ctypes.c_int), ("n_bodies", ctypes.c_int),
                ("body_start", _ip), ("type", _ip), ("charge", _fp), ("sigma", _fp),
                ("crowder", ctypes.POINTER(ctypes.c_ubyte)), ("box", ctypes.c_float),
                ("x", _fp), ("y", _fp), ("z", _fp), ("rx", _fp), ("ry", _fp), ("rz", _fp),
                ("cx", _fp), ("cy", _fp), ("cz", _fp), ("quat", _dp)]


class remc_counters(ctypes.Structure):
    _fields_ = [("accept", ctypes.c_uint64), ("accept_boltz", ctypes.c_uint64),
                ("reject", ctypes.c_uint64), ("steps", ctypes.c_uint64)]


class remc_step_record(ctypes.Structure):
    _fields_ = [("body", ctypes.c_int32), ("move", ctypes.c_int32), ("accepted", ctypes.c_int32),
                ("ticks", ctypes.c_int32), ("delta", ctypes.c_double)]


RECORD_DTYPE = np.dtype([("body", np.int32), ("move", np.int32), ("accepted", np.int32),
                         ("ticks", np.int32), ("delta", np.float64)])

# every entry point declared in include/remc.h: name -> (restype, argtypes)
_vp = ctypes.c_void_p
SIGNATURES = {
    "remc_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "remc_ctx_destroy": (ctypes.c_int, [_vp]),
    "remc_last_error": (ctypes.c_char_p, []),
    "remc_set_forcefield": (ctypes.c_int, [_vp, ctypes.POINTER(remc_forcefield)]),
    "remc_set_seed": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "remc_system_add": (ctypes.c_int, [_vp, ctypes.POINTER(remc_system_desc), _ip]),
    "remc_replicas_add": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _fp, _dp, _ip]),
    "remc_get_sizes": (ctypes.c_int, [_vp, _ip, _ip, _ip]),
    "remc_energy_plain": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "remc_energy_all": (ctypes.c_int, [_vp, _dp]),
    "remc_energy_all_noalloc": (ctypes.c_int, [_vp, _dp]),
    "remc_refresh_energies": (ctypes.c_int, [_vp]),
    "remc_energy_noncrowder_all": (ctypes.c_int, [_vp, _dp]),
    "remc_mc_steps": (ctypes.c_int, [_vp, ctypes.c_int]),
    "remc_mc_steps_plain": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "remc_mc_steps_fb": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "remc_exchange": (ctypes.c_int, [_vp, _ip]),
    "remc_run": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "remc_set_run_refresh": (ctypes.c_int, [_vp, ctypes.c_int]),
    "remc_set_recording": (ctypes.c_int, [_vp, ctypes.c_int]),
    "remc_get_records": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(remc_step_record), ctypes.c_int]),
    "remc_get_energies": (ctypes.c_int, [_vp, _dp]),
    "remc_get_temperatures": (ctypes.c_int, [_vp, _dp]),
    "remc_get_counters": (ctypes.c_int, [_vp, ctypes.POINTER(remc_counters)]),
    "remc_get_ladder": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "remc_get_state": (ctypes.c_int, [_vp, ctypes.c_int] + [_fp] * 9 + [_dp]),
    "remc_set_state": (ctypes.c_int, [_vp, ctypes.c_int] + [_fp] * 9 + [_dp]),
    "remc_set_energies": (ctypes.c_int, [_vp, _dp]),
    "remc_set_counters": (ctypes.c_int, [_vp, ctypes.POINTER(remc_counters)]),
    "remc_set_ladder": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "remc_get_exchange_state": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64)]),
    "remc_set_exchange_state": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64]),
    "remc_drms": (ctypes.c_int, [_vp, ctypes.c_int, _fp, _fp, _fp, _dp]),
    "remc_radial_distribution": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_double, ctypes.c_double, _dp]),
    "remc_centre_histogram": (ctypes.c_int, [_vp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.POINTER(ctypes.c_uint64)]),
    "remc_adjoint": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "remc_adjoint_all": (ctypes.c_int, [_vp, _dp, _dp]),
    "remc_timing_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "remc_timing_get": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_int64)]),
    "remc_synchronize": (ctypes.c_int, [_vp]),
    "remc_set_mc_block": (ctypes.c_int, [_vp, ctypes.c_int]),
    "remc_get_fb_block": (ctypes.c_int, [_vp, _ip]),
    "remc_get_mc_block": (ctypes.c_int, [_vp, _ip]),
    "remc_set_mc_coop": (ctypes.c_int, [_vp, ctypes.c_int]),
    "remc_get_mc_coop": (ctypes.c_int, [_vp, _ip]),
    "remc_set_mc_tuning": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "remc_set_mc_window": (ctypes.c_int, [_vp, ctypes.c_int]),
}

_lib = None


def load(path=LIB_PATH):
    """Load libremc.so (cached).  Raises if it is missing or lacks a symbol."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("native library %s is not built; run `python -m remcdockingsimulator_amd.build`" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError if the symbol is missing
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


class RemcError(RuntimeError):
    pass


def check(lib, rc, what):
    if rc != 0:
        raise RemcError("%s failed (%d): %s" % (what, rc, lib.remc_last_error().decode()))


def fptr(a):
    return a.ctypes.data_as(_fp)


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)
