"""Loading and parsing of the golden fixtures in tests/golden/ (see make_golden.py)."""
from __future__ import annotations

import json
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return {c["name"]: c for c in json.load(f)}


def load_input(name):
    z = np.load(os.path.join(GOLDEN, name, "input.npz"))
    d = z["xyz_d"]
    r = z["xyz_r"] if "xyz_r" in z else np.zeros((0, 3))
    wd = z["w_d"] if "w_d" in z else None
    wr = z["w_r"] if "w_r" in z else None
    return d, r, wd, wr


def read_text(name, run):
    with open(os.path.join(GOLDEN, name, f"{run}.txt")) as f:
        return f.read()


def parse_rows(text):
    """Numeric rows of a reference output file (header lines skipped)."""
    rows = []
    for line in text.splitlines():
        toks = line.split()
        try:
            rows.append([float(t) for t in toks])
        except ValueError:
            continue
    return np.array(rows)


def e147_ulp(v):
    """One unit in the last place of Fortran E14.7 output of v (0.ddddddd E+xx)."""
    if v == 0 or not math.isfinite(v):
        return 0.0
    e = math.floor(math.log10(abs(v))) + 1
    return 10.0 ** (e - 7)


def close_e147(printed, value, ulps=1.0, cancel=1e-12):
    """printed (parsed from E14.7) equals value to within `ulps` units of the 7th
    digit, or to within cancel * max|value| (the floor set by the reference's own
    FP64 summation of terms of both signs, e.g. 1e-27 printed for an exact 0)."""
    printed = np.asarray(printed, dtype=float)
    value = np.asarray(value, dtype=float)
    fin = np.isfinite(value)
    floor = cancel * (np.abs(value[fin]).max() if fin.any() else 0.0)
    ok = np.ones(printed.shape, bool)
    for idx in np.ndindex(printed.shape):
        p, v = printed[idx], value[idx]
        if math.isnan(p) or math.isinf(p):
            ok[idx] = (math.isnan(p) and math.isnan(v)) or (p == v)
            continue
        tol = ulps * max(e147_ulp(p), e147_ulp(v)) * 1.0000001
        ok[idx] = abs(p - v) <= max(tol, floor)
    return ok
