"""CPU-side checks of the C ABI: the library loads, exports every symbol the
header declares, and its host-only pieces (exact bin thresholds, argument
validation) agree with the oracle.  No GPU is touched."""
import ctypes
import math

import numpy as np
import pytest

import oracle as O
from gramsci_amd import _lib as L


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = L.exported_names()
    assert {"gs_create", "gs_count_2pcf", "gs_count_3pcf", "gs_count_4pcf", "gs_bin_thresholds"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n
    assert lib.gs_version().startswith(b"gramsci_hip")


def test_calibration_library_loads():
    # the VALU issue-rate calibration kernel that bench.py times beside the pair kernel
    import os
    so = os.path.join(os.path.dirname(L.__file__), "libgs_calib.so")
    lib = ctypes.CDLL(so)
    assert lib.gs_calib_threads() == 256 and lib.gs_calib_fma_per_iter() == 128
    f = ctypes.c_float(0.0)
    # argument checks refuse before any GPU call: an output buffer smaller than the grid
    assert lib.gs_calib_fma32(4, 1, ctypes.c_void_p(1), ctypes.c_int64(1023), ctypes.byref(f)) == -1
    assert lib.gs_calib_fma32(0, 1, ctypes.c_void_p(1), ctypes.c_int64(1 << 20), ctypes.byref(f)) == -1


def _bin_from_thresholds(t, s):
    nb = len(t) - 2
    if not (s > t[0] and s < t[nb + 1]):
        return 0
    b = int(np.searchsorted(t[1:nb + 1], s, side="right"))   # #{k: t[k] <= s}
    return b + 1   # 1-based; nb+1 = overflow


@pytest.mark.parametrize("rmin,rmax,nbins,log", [
    (10.0, 50.0, 10, False), (1.0, 150.0, 30, True), (0.0, 5.0, 5, False), (1.0, 8.0, 3, True),
    (2.5, 37.3, 17, False), (0.1, 99.9, 127, True), (4.0, 12.0, 5, False)])
def test_thresholds_match_reference_bin_formula(rmin, rmax, nbins, log):
    p = L.make_params(rmin, rmax, nbins, 1, log)
    t = L.bin_thresholds(p)
    assert np.all(np.diff(t[1:nbins + 1]) >= 0)
    op = O.Params(rmin, rmax, nbins, 1, log)
    rng = np.random.default_rng(nbins)
    origin = np.zeros(3)
    # pairs (0,0,0)-(x,0,0): s = x*x exactly as the reference forms it; probe
    # x around every threshold and at random
    xs = []
    for k in range(nbins + 2):
        r = math.sqrt(t[k])
        for d in range(-40, 41):
            xs.append(r + d * math.ulp(max(r, 1e-300)))
    xs += list(rng.uniform(0, rmax * 1.1, 2000))
    for x in xs:
        if x < 0:
            continue
        s = x * x
        assert _bin_from_thresholds(t, s) == O.pair_rbin(op, origin, np.array([x, 0.0, 0.0])), (x, s)


def test_thresholds_three_axis_pairs():
    # s formed from three components, left to right, as in the reference
    p = L.make_params(1.0, 30.0, 8, 1, True)
    t = L.bin_thresholds(p)
    op = O.Params(1.0, 30.0, 8, 1, True)
    rng = np.random.default_rng(0)
    a = rng.uniform(0, 100, (20000, 3))
    b = a + rng.normal(0, 12, (20000, 3))
    for x, y in zip(a, b):
        d = y - x
        s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
        assert _bin_from_thresholds(t, s) == O.pair_rbin(op, x, y)


@pytest.mark.parametrize("args", [
    (5.0, 1.0, 3, 1, False),     # rmax < rmin
    (0.0, 1.0, 3, 1, True),      # log with rmin = 0
    (0.0, 1.0, 0, 1, False),     # nbins = 0
    (0.0, 1.0, 128, 1, False),   # nbins > 127
    (0.0, 1.0, 3, 0, False),     # nmu = 0
    (-1.0, 1.0, 3, 1, False),    # rmin < 0
])
def test_invalid_params(args):
    p = L.make_params(*args)
    with pytest.raises(L.GsError) as e:
        L.bin_thresholds(p)
    assert e.value.code == L.GS_EINVAL


def test_null_arguments_rejected():
    lib = L.lib()
    assert lib.gs_create(None, 0) == L.GS_EINVAL
    assert lib.gs_set_catalogue(None, None, None, None, 1, 0) == L.GS_EINVAL
    assert lib.gs_count_2pcf(None, None, None, None, None, None, None) == L.GS_EINVAL
    assert lib.gs_destroy(None) == L.GS_OK
