"""ctypes binding of libpimdtunnel.so (the C ABI declared in include/rp_hip.h).

The library is loaded from the package's in-tree ``lib/`` directory.  There is no
CPU fallback: if the library or a HIP device is missing, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libpimdtunnel.so")

RP_OK, RP_ERR_ARG, RP_ERR_HIP, RP_ERR_NAN, RP_ERR_NOISE, RP_ERR_STATE, RP_ERR_CONVERGE = range(7)

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
vp = C.c_void_p


class RpError(RuntimeError):
    """A non-zero status from the library (the reference would `stop` here)."""

    def __init__(self, code, msg):
        super().__init__(f"rp status {code}: {msg}")
        self.code = code


class PolymerDesc(C.Structure):
    _fields_ = [("n", C.c_int), ("betan", C.c_double), ("xi", C.c_double), ("a", dp),
                ("b", dp), ("dbdl", dp), ("stream", C.c_uint64)]


# Host-potential callbacks (rp_host_potforce_fn / rp_host_hessian_fn of rp_hip.h)
HOST_POTFORCE = C.CFUNCTYPE(C.c_int, vp, C.c_int, dp, dp, dp)
HOST_HESSIAN = C.CFUNCTYPE(C.c_int, vp, C.c_int, dp, dp)

_SIGS = {
    "rp_init": [C.c_int],
    "rp_finalize": [],
    "rp_last_error": [C.c_char_p, C.c_int],
    "rp_version": [],
    "rp_pes_create": [C.c_int, C.c_int, C.c_int, dp, C.c_int, ip, C.c_int, C.POINTER(vp)],
    "rp_pes_create_host": [C.c_int, C.c_int, HOST_POTFORCE, HOST_HESSIAN, vp, C.POINTER(vp)],
    "rp_pes_destroy": [vp],
    "rp_pes_set_v0": [vp, C.c_double],
    "rp_pes_potforce": [vp, C.c_int, dp, dp, dp],
    "rp_pes_potforce_noalloc": [vp, C.c_int, dp, dp, dp],
    "rp_pes_hessian": [vp, C.c_int, dp, dp],
    "rp_pes_param_grad": [vp, C.c_int, dp, dp, ip],
    "rp_nm_setup": [vp, C.c_int, C.c_int, C.c_int, dp, C.c_double, C.c_double, C.c_double,
                    C.c_double, C.c_int, dp, dp, C.POINTER(vp)],
    "rp_chain_destroy": [vp],
    "rp_chain_set_ends": [vp, dp, dp],
    "rp_chain_set_state": [vp, dp, dp],
    "rp_chain_get_state": [vp, dp, dp],
    "rp_chain_get_nm": [vp, dp, dp, dp, dp],
    "rp_chain_set_rng": [vp, C.c_uint64, C.c_uint64, C.c_int64],
    "rp_chain_get_cursor": [vp, i64p],
    "rp_chain_set_kick_cursor": [vp, C.c_int64],
    "rp_chain_get_kick_cursor": [vp, i64p],
    "rp_nm_transform": [vp, C.c_int, C.c_int, ip, dp, dp],
    "rp_step_v": [vp, C.c_double],
    "rp_step_nm": [vp, C.c_double],
    "rp_step_langevin": [vp, dp],
    "rp_time_step_pile": [vp, dp],
    "rp_time_step_nm": [vp],
    "rp_andersen_resample": [vp, dp],
    "rp_chain_set_path": [vp, C.c_int, dp, dp],
    "rp_init_path": [vp, C.c_double, dp],
    "rp_propagate_pile": [vp, C.c_int64, C.c_int64, C.c_double, dp, C.c_double, dp, C.c_int64,
                          dp, i64p],
    "rp_propagate_nm": [vp, C.c_int64, C.c_int64, C.c_int64, dp, dp, C.c_int64, i32p,
                        C.c_int64, dp],
    "rp_batch_create": [vp, dp, C.c_int, C.POINTER(PolymerDesc), C.c_double, C.c_double,
                        C.c_double, C.c_int, C.c_int64, C.c_double, C.c_int, C.c_uint64,
                        C.POINTER(vp)],
    "rp_batch_destroy": [vp],
    "rp_batch_set_path": [vp, C.c_int, dp, dp],
    "rp_batch_init_path": [vp, dp],
    "rp_batch_set_state": [vp, C.c_int, dp, dp],
    "rp_batch_get_state": [vp, C.c_int, dp, dp],
    "rp_batch_run": [vp, C.c_int64, C.c_int64, dp, C.c_int64],
    "rp_batch_set_kicks": [vp, i32p, C.c_int64],
    "rp_batch_get_cursor": [vp, i64p],
    "rp_batch_get_kick_cursor": [vp, i64p],
    "rp_batch_set_cursor": [vp, i64p, i64p],
    "rp_batch_families": [vp, ip, ip, ip],
    "rp_batch_family_shape": [vp, C.c_int, ip, ip, ip],
    "rp_batch_ti_readout": [vp, dp, i64p, i64p],
    "rp_batch_set_launch": [vp, C.c_int, C.c_int],
    "rp_batch_set_graphs": [vp, C.c_int],
    "rp_batch_autotune": [vp, C.c_int, ip, ip],
    "rp_batch_last_kernel_ms": [vp, dp, i64p],
    "rp_ring_create": [vp, dp, C.c_int, ip, dp, dp, dp, C.c_int, C.POINTER(vp)],
    "rp_ring_destroy": [vp],
    "rp_ring_set_x": [vp, dp],
    "rp_ring_get_x": [vp, dp],
    "rp_ring_um": [vp, dp, dp],
    "rp_ring_hvp": [vp, dp, dp],
    "rp_ring_set_x_device": [vp, vp],
    "rp_ring_um_device": [vp, dp, vp],
    "rp_ring_hvp_device": [vp, vp, vp],
    "rp_ring_hessian_band": [vp, C.c_int, C.c_int, dp],
    "rp_ring_logdet": [vp, C.c_int, C.c_int, dp, ip, dp],
    "rp_ring_lowest_eigs": [vp, C.c_int, C.c_int, dp],
    "rp_ring_set_preconditioner": [vp, C.c_double],
    "rp_instanton": [vp, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp, ip, ip],
    "rp_ring_param_grad": [vp, dp, ip],
}

_lib = None


def lib():
    """The loaded library; raises if it is not built (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def exported_symbols():
    L = lib()
    return {name for name in _SIGS if getattr(L, name, None) is not None}


def last_error():
    buf = C.create_string_buffer(1024)
    lib().rp_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc):
    if rc != RP_OK:
        raise RpError(rc, last_error())
    return rc


_initialised = False


def init(device=0):
    global _initialised
    if not _initialised:
        check(lib().rp_init(device))
        _initialised = True


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def ptr(a):
    if a is None:
        return None
    if not a.flags.c_contiguous:  # the library reads plain memory: a strided view would alias
        raise ValueError("ptr: array is not contiguous")
    return a.ctypes.data_as(dp)


def iptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(ip)
