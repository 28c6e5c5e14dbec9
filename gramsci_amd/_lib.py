"""ctypes binding of the C ABI in include/gramsci_hip.h (libgramsci_hip.so).

The shared library is built in-tree (gramsci_amd/libgramsci_hip.so, see
gramsci_amd/csrc/Makefile).  There is no fallback: if the library is missing
or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GS_LIBRARY selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("GS_LIBRARY") or os.path.join(HERE, "libgramsci_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "gramsci_hip.h")

GS_OK, GS_EINVAL, GS_ENOMEM, GS_EHIP, GS_ENODEV, GS_ELIMIT = 0, -1, -2, -3, -4, -5
GS_FP64, GS_FP32_GUARD = 0, 1


class GsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gramsci_hip error {code}: {msg}")
        self.code = code


class gs_params(ctypes.Structure):
    _fields_ = [("rmin", ctypes.c_double), ("rmax", ctypes.c_double),
                ("nbins", ctypes.c_int32), ("nmu", ctypes.c_int32), ("log_bins", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("cell_div", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class gs_stats(ctypes.Structure):
    _fields_ = [("ms_grid", ctypes.c_double), ("ms_count", ctypes.c_double), ("ms_kernel_avg", ctypes.c_double),
                ("n_units", ctypes.c_uint64), ("n_binned", ctypes.c_uint64), ("n_candidates", ctypes.c_uint64),
                ("n_recheck", ctypes.c_uint64), ("n_overflow", ctypes.c_uint64), ("n_evaluated", ctypes.c_uint64),
                ("cell_div", ctypes.c_int32), ("reserved", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib = None


def exported_names():
    """Function names declared in include/gramsci_hip.h."""
    import re
    with open(HEADER_PATH) as f:
        text = f.read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(gs_\w+)\s*\(", text, flags=re.M)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    # torch ships its own HIP runtime with the same soname (libamdhip64.so.7);
    # importing torch first makes this library bind to that one instance, so
    # device pointers can be shared with torch tensors in one process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built (run `make -C gramsci_amd/csrc` or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp, dp, i32p, u64p = ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_int32), P(ctypes.c_uint64)
    pp, sp = P(gs_params), P(gs_stats)
    L.gs_last_error.restype = ctypes.c_char_p
    L.gs_version.restype = ctypes.c_char_p
    L.gs_create.argtypes = [P(vp), ctypes.c_int]
    L.gs_destroy.argtypes = [vp]
    L.gs_set_catalogue.argtypes = [vp, dp, dp, i32p, ctypes.c_int64, ctypes.c_int64]
    L.gs_set_catalogue_device.argtypes = [vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int64]
    L.gs_count_2pcf.argtypes = [vp, pp, u64p, dp, dp, dp, sp]
    L.gs_count_3pcf.argtypes = [vp, pp, u64p, dp, dp, dp, sp]
    L.gs_count_4pcf.argtypes = [vp, pp, u64p, dp, dp, dp, dp, sp]
    L.gs_count_3pcf_equi.argtypes = [vp, pp, u64p, dp, dp, dp, sp]
    L.gs_count_4pcf_tetra.argtypes = [vp, pp, u64p, dp, dp, dp, dp, sp]
    L.gs_pcf2.argtypes = [dp, dp, i32p, ctypes.c_int64, ctypes.c_int64, pp, u64p, dp, dp, sp]
    L.gs_pcf3.argtypes = [dp, dp, i32p, ctypes.c_int64, ctypes.c_int64, pp, u64p, dp, dp, sp]
    L.gs_pcf4.argtypes = [dp, dp, i32p, ctypes.c_int64, ctypes.c_int64, pp, u64p, dp, dp, dp, dp, sp]
    L.gs_bin_thresholds.argtypes = [pp, dp]
    L.gs_adjoint_2pcf.argtypes = [vp, pp, dp, dp, dp, sp]
    i64p = P(ctypes.c_int64)
    L.gs_set_catalogue_batched.argtypes = [vp, dp, dp, i32p, i64p, i64p, ctypes.c_int]
    L.gs_decompose.argtypes = [vp, P(ctypes.c_float), i32p, ctypes.c_int64, ctypes.c_float, ctypes.c_int, i64p, i32p,
                               i32p, ctypes.c_int64, P(ctypes.c_float)]
    L.gs_count_2pcf_batched.argtypes = [vp, pp, u64p, dp, dp, dp, sp]
    L.gs_count_3pcf_batched.argtypes = [vp, pp, u64p, dp, dp, dp, sp]
    L.gs_adjoint_2pcf_device.argtypes = [vp, pp, dp, dp, vp, sp]
    L.gs_count_2pcf_noalloc.argtypes = [vp, pp, vp, vp, vp, ctypes.c_int64, ctypes.c_int64, u64p, dp, dp, dp, sp]
    for name in exported_names():
        getattr(L, name)  # every declared symbol must resolve
        if name not in ("gs_last_error", "gs_version"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc):
    if rc != GS_OK:
        raise GsError(rc, lib().gs_last_error().decode(errors="replace"))


def ptr(a, t=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


def make_params(rmin, rmax, nbins, nmu=1, log_bins=False, precision=GS_FP32_GUARD, cell_div=0) -> gs_params:
    return gs_params(float(rmin), float(rmax), int(nbins), int(nmu), int(bool(log_bins)), int(precision),
                     int(cell_div), 0)


def bin_thresholds(p: gs_params) -> np.ndarray:
    t = np.zeros(p.nbins + 2)
    check(lib().gs_bin_thresholds(ctypes.byref(p), ptr(t)))
    return t
