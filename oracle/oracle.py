"""ctypes view of the CPU restatement in gs_oracle.c, plus helpers to run the
reference binary built by oracle/Makefile.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_gs.so")
REF_BIN = os.path.join(HERE, "_ref", "gramsci")
REF_BIN_FAST = os.path.join(HERE, "_ref", "gramsci_fast")


class _Params(ctypes.Structure):
    _fields_ = [("rmin", ctypes.c_double), ("rmax", ctypes.c_double),
                ("nbins", ctypes.c_int32), ("nmu", ctypes.c_int32), ("log_bins", ctypes.c_int32)]


@dataclass
class Params:
    rmin: float
    rmax: float
    nbins: int
    nmu: int = 1
    log_bins: bool = False

    def c(self) -> _Params:
        return _Params(self.rmin, self.rmax, self.nbins, self.nmu, int(bool(self.log_bins)))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", HERE, "liboracle_gs.so"])
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        dp, i32p, u64p = P(ctypes.c_double), P(ctypes.c_int32), P(ctypes.c_uint64)
        pp = P(_Params)
        L.or_radial_bins.argtypes = [pp, dp]
        L.or_normalise.argtypes = [dp, ctypes.c_int64, ctypes.c_int64, dp]
        L.or_pair_rbin.argtypes = [pp, dp, dp]
        L.or_pair_mubin.argtypes = [pp, dp, dp]
        L.or_2pcf.argtypes = [dp, dp, i32p, ctypes.c_int64, ctypes.c_int64, pp, u64p, dp, dp, dp, u64p]
        L.or_3pcf.argtypes = [dp, dp, i32p, ctypes.c_int64, ctypes.c_int64, pp, u64p, dp, dp, u64p]
        L.or_4pcf.argtypes = [dp, dp, i32p, ctypes.c_int64, ctypes.c_int64, pp, u64p, dp, dp, dp, dp, u64p]
        L.or_normal_bin.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_3pcf_equi.argtypes = [dp, dp, i32p, ctypes.c_int64, ctypes.c_int64, pp, u64p, dp, dp]
        L.or_4pcf_tetra.argtypes = [dp, dp, i32p, ctypes.c_int64, ctypes.c_int64, pp, u64p, dp, dp, dp, dp]
        L.or_adjoint_2pcf.argtypes = [dp, dp, i32p, ctypes.c_int64, ctypes.c_int64, pp, dp, dp, dp]
        for f in ("or_2pcf", "or_3pcf", "or_4pcf", "or_adjoint_2pcf", "or_pair_rbin", "or_pair_mubin",
                  "or_n3cfg", "or_n4cfg", "or_bintable3", "or_bintable4"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


def _prep(xyz, w, buf):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    assert xyz.ndim == 2 and xyz.shape[1] == 3
    w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    buf = None if buf is None else np.ascontiguousarray(buf, dtype=np.int32)
    return xyz, w, buf


def radial_bins(p: Params) -> np.ndarray:
    e = np.zeros(p.nbins + 1)
    lib().or_radial_bins(ctypes.byref(p.c()), _ptr(e, ctypes.c_double))
    return e


def normalise(w, n_data, n_rand):
    wn = np.zeros(n_data + n_rand)
    w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    lib().or_normalise(_ptr(w, ctypes.c_double), n_data, n_rand, _ptr(wn, ctypes.c_double))
    return wn


def pair_rbin(p: Params, xi, xj) -> int:
    a = np.ascontiguousarray(xi, dtype=np.float64)
    b = np.ascontiguousarray(xj, dtype=np.float64)
    return lib().or_pair_rbin(ctypes.byref(p.c()), _ptr(a, ctypes.c_double), _ptr(b, ctypes.c_double))


def pair_mubin(p: Params, xi, xj) -> int:
    a = np.ascontiguousarray(xi, dtype=np.float64)
    b = np.ascontiguousarray(xj, dtype=np.float64)
    return lib().or_pair_mubin(ctypes.byref(p.c()), _ptr(a, ctypes.c_double), _ptr(b, ctypes.c_double))


def n3cfg(nbins):
    return lib().or_n3cfg(nbins)


def n4cfg(nbins):
    return lib().or_n4cfg(nbins)


def pcf2(xyz, w, n_data, p: Params, buf=None):
    """Returns dict(cnt[3,nb,nm] u64, wsum[3,nb,nm], n2[nb,nm], n3[nb,nm], overflow)."""
    xyz, w, buf = _prep(xyz, w, buf)
    n = xyz.shape[0]
    cnt = np.zeros((3, p.nbins, p.nmu), np.uint64)
    wsum = np.zeros((3, p.nbins, p.nmu))
    n2 = np.zeros((p.nbins, p.nmu))
    n3 = np.zeros((p.nbins, p.nmu))
    ov = ctypes.c_uint64(0)
    rc = lib().or_2pcf(_ptr(xyz, ctypes.c_double), _ptr(w, ctypes.c_double), _ptr(buf, ctypes.c_int32),
                       n_data, n - n_data, ctypes.byref(p.c()), _ptr(cnt, ctypes.c_uint64),
                       _ptr(wsum, ctypes.c_double), _ptr(n2, ctypes.c_double), _ptr(n3, ctypes.c_double),
                       ctypes.byref(ov))
    if rc:
        raise ValueError(f"or_2pcf failed: {rc}")
    return dict(cnt=cnt, wsum=wsum, n2=n2, n3=n3, overflow=int(ov.value))


def pcf3(xyz, w, n_data, p: Params, buf=None):
    """Returns dict(cnt u64, n2, n3, n_cand): cnt[ncfg,4], n2/n3[ncfg] for
    nmu == 1; cnt[ncfg,nmu,4], n2/n3[ncfg,nmu] for the RSD 3PCF (nmu >= 2)."""
    xyz, w, buf = _prep(xyz, w, buf)
    n = xyz.shape[0]
    nc = n3cfg(p.nbins)
    cnt = np.zeros((nc, p.nmu, 4), np.uint64)
    n2 = np.zeros((nc, p.nmu))
    n3 = np.zeros((nc, p.nmu))
    cand = ctypes.c_uint64(0)
    rc = lib().or_3pcf(_ptr(xyz, ctypes.c_double), _ptr(w, ctypes.c_double), _ptr(buf, ctypes.c_int32),
                       n_data, n - n_data, ctypes.byref(p.c()), _ptr(cnt, ctypes.c_uint64),
                       _ptr(n2, ctypes.c_double), _ptr(n3, ctypes.c_double), ctypes.byref(cand))
    if rc:
        raise ValueError(f"or_3pcf failed: {rc}")
    if p.nmu == 1:
        cnt, n2, n3 = cnt[:, 0], n2[:, 0], n3[:, 0]
    return dict(cnt=cnt, n2=n2, n3=n3, n_cand=int(cand.value))


def normal_bin(nmu, mu1, mu2) -> int:
    """find_normal: 1-based mu bin of a triplet from its two apex-side mu bins."""
    return int(lib().or_normal_bin(nmu, mu1, mu2))


def pcf3_equi(xyz, w, n_data, p: Params, buf=None):
    """Equilateral 3PCF: dict(cnt[nbins,nmu,4] u64, n2[nbins,nmu], n3[nbins,nmu])."""
    xyz, w, buf = _prep(xyz, w, buf)
    n = xyz.shape[0]
    cnt = np.zeros((p.nbins, p.nmu, 4), np.uint64)
    n2 = np.zeros((p.nbins, p.nmu))
    n3 = np.zeros((p.nbins, p.nmu))
    rc = lib().or_3pcf_equi(_ptr(xyz, ctypes.c_double), _ptr(w, ctypes.c_double), _ptr(buf, ctypes.c_int32),
                            n_data, n - n_data, ctypes.byref(p.c()), _ptr(cnt, ctypes.c_uint64),
                            _ptr(n2, ctypes.c_double), _ptr(n3, ctypes.c_double))
    if rc:
        raise ValueError(f"or_3pcf_equi failed: {rc}")
    return dict(cnt=cnt, n2=n2, n3=n3)


def pcf4_tetra(xyz, w, n_data, p: Params, buf=None):
    """Regular-tetrahedron 4PCF: dict(cnt[nbins,16] u64, n2[nbins], n3[nbins],
    disc_n[6,nbins], disc_r[6,nbins])."""
    xyz, w, buf = _prep(xyz, w, buf)
    n = xyz.shape[0]
    nb = p.nbins
    cnt = np.zeros((nb, 16), np.uint64)
    n2, n3 = np.zeros(nb), np.zeros(nb)
    dn, dr = np.zeros((6, nb)), np.zeros((6, nb))
    rc = lib().or_4pcf_tetra(_ptr(xyz, ctypes.c_double), _ptr(w, ctypes.c_double), _ptr(buf, ctypes.c_int32),
                             n_data, n - n_data, ctypes.byref(p.c()), _ptr(cnt, ctypes.c_uint64),
                             _ptr(n2, ctypes.c_double), _ptr(n3, ctypes.c_double), _ptr(dn, ctypes.c_double),
                             _ptr(dr, ctypes.c_double))
    if rc:
        raise ValueError(f"or_4pcf_tetra failed: {rc}")
    return dict(cnt=cnt, n2=n2, n3=n3, disc_n=dn, disc_r=dr)


def pcf4(xyz, w, n_data, p: Params, buf=None):
    """Returns dict(cnt[ncfg,nb,nb,16], n2/n3[ncfg,nb,nb], disc_n/disc_r[6,ncfg,nb,nb], n_quads)."""
    xyz, w, buf = _prep(xyz, w, buf)
    n = xyz.shape[0]
    nc, nb = n4cfg(p.nbins), p.nbins
    cnt = np.zeros((nc, nb, nb, 16), np.uint64)
    n2 = np.zeros((nc, nb, nb))
    n3 = np.zeros((nc, nb, nb))
    dn = np.zeros((6, nc, nb, nb))
    dr = np.zeros((6, nc, nb, nb))
    q = ctypes.c_uint64(0)
    rc = lib().or_4pcf(_ptr(xyz, ctypes.c_double), _ptr(w, ctypes.c_double), _ptr(buf, ctypes.c_int32),
                       n_data, n - n_data, ctypes.byref(p.c()), _ptr(cnt, ctypes.c_uint64),
                       _ptr(n2, ctypes.c_double), _ptr(n3, ctypes.c_double),
                       _ptr(dn, ctypes.c_double), _ptr(dr, ctypes.c_double), ctypes.byref(q))
    if rc:
        raise ValueError(f"or_4pcf failed: {rc}")
    return dict(cnt=cnt, n2=n2, n3=n3, disc_n=dn, disc_r=dr, n_quads=int(q.value))


def adjoint2(xyz, w, n_data, p: Params, gN2, gN3, buf=None):
    """dL/dw (raw weights) for dL/dN2 = gN2, dL/dN3 = gN3 of shape (nbins, nmu)."""
    xyz, w, buf = _prep(xyz, w, buf)
    n = xyz.shape[0]
    g2 = np.ascontiguousarray(gN2, dtype=np.float64).reshape(p.nbins, p.nmu)
    g3 = np.ascontiguousarray(gN3, dtype=np.float64).reshape(p.nbins, p.nmu)
    out = np.zeros(n)
    rc = lib().or_adjoint_2pcf(_ptr(xyz, ctypes.c_double), _ptr(w, ctypes.c_double), _ptr(buf, ctypes.c_int32),
                               n_data, n - n_data, ctypes.byref(p.c()), _ptr(g2, ctypes.c_double),
                               _ptr(g3, ctypes.c_double), _ptr(out, ctypes.c_double))
    if rc:
        raise ValueError(f"or_adjoint_2pcf failed: {rc}")
    return out


# ------------------------------------------------------------------ reference binary

def write_catalogue(path, xyz, w=None, extra=None):
    """ASCII x y z w [extra] with 17 significant digits (exact round trip)."""
    xyz = np.asarray(xyz, dtype=np.float64)
    w = np.ones(len(xyz)) if w is None else np.asarray(w, dtype=np.float64)
    cols = [xyz[:, 0], xyz[:, 1], xyz[:, 2], w]
    fmt = ["%.17g"] * 4
    if extra is not None:
        cols.append(np.asarray(extra))
        fmt.append("%d")
    np.savetxt(path, np.column_stack(cols), fmt=fmt)


def run_reference(args, cwd, binary=REF_BIN, threads=None, timeout=3600):
    """Runs the reference program with the given argument list; returns (stdout, wall_s)."""
    import time
    env = dict(os.environ)
    if threads is not None:
        env["OMP_NUM_THREADS"] = str(threads)
    t0 = time.perf_counter()
    r = subprocess.run([binary] + list(args), cwd=cwd, env=env, capture_output=True, text=True,
                       timeout=timeout)
    dt = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(f"reference failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
    return r.stdout, dt


def reference_table(xyz_d, w_d, xyz_r, w_r, p: Params, mode="-2pcf", threads=None, binary=REF_BIN):
    """Runs the reference on the catalogues and returns the numeric rows of its output file."""
    with tempfile.TemporaryDirectory() as td:
        write_catalogue(os.path.join(td, "d.txt"), xyz_d, w_d)
        args = ["-gal", "d.txt"]
        if xyz_r is not None and len(xyz_r):
            write_catalogue(os.path.join(td, "r.txt"), xyz_r, w_r)
            args += ["-ran", "r.txt"]
        args += ["-rmin", repr(float(p.rmin)), "-rmax", repr(float(p.rmax)), "-nbins", str(p.nbins),
                 "-nmu", str(p.nmu), "-out", "o.txt"]
        if p.log_bins:
            args.append("-log")
        args.append(mode)
        run_reference(args, td, binary=binary, threads=threads)
        with open(os.path.join(td, "o.txt")) as f:
            text = f.read()
    return text


def decompose(xyz, region, rmax, nreg=None):
    """The reference's domain decomposition tool (domain_decomposition.f90:35-74)
    in single precision, as the tool's default reals: for labels r = 1..nreg,
    box = min/max of the points labelled r (:38-45), grown by rmax (:48-53);
    members = points strictly inside the grown box (:62-64), in input order;
    buffer = 0 inside the closed box, else 1 (:67-70).  An empty label has box
    (+huge, -huge) and no member.  Returns (off[nreg+1], idx, buf, box[nreg,6])
    with box rows xmin, ymin, zmin, xmax, ymax, zmax."""
    x = np.asarray(xyz, np.float32)
    region = np.asarray(region)
    if nreg is None:
        nreg = int(region.max()) if len(region) else 0
    r32 = np.float32(rmax)
    off, idx, buf, boxes = [0], [], [], []
    big = np.finfo(np.float32).max
    for r in range(1, nreg + 1):
        m = region == r
        if m.any():
            lo, hi = x[m].min(0), x[m].max(0)
        else:
            lo, hi = np.full(3, big, np.float32), np.full(3, -big, np.float32)
        lo2, hi2 = lo - r32, hi + r32            # float32 arithmetic
        inside = np.all((x > lo2) & (x < hi2), axis=1)
        core = np.all((x >= lo) & (x <= hi), axis=1)
        k = np.nonzero(inside)[0]
        idx.append(k)
        buf.append(np.where(core[k], 0, 1))
        off.append(off[-1] + len(k))
        boxes.append(np.concatenate([lo, hi]))
    cat = (lambda a, t: np.concatenate(a).astype(t) if a else np.zeros(0, t))
    return (np.array(off, np.int64), cat(idx, np.int32), cat(buf, np.int32),
            np.array(boxes, np.float32).reshape(-1, 6))
