"""Python host side of the engine: a thin object over the C ABI.

    eng = Engine()                          # context on HIP device 0 (gfx950)
    eng.set_catalogue(xyz, w, n_data)       # numpy (host) or torch.cuda tensors (device)
    r = eng.pcf2(rmin, rmax, nbins, nmu=10, log_bins=False, precision="fp32")
    r["cnt"], r["n2"], r["n3"], r["xi"], r["stats"]

The counting path is always the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

PRECISION = {"fp64": L.GS_FP64, "fp32": L.GS_FP32_GUARD, "fp32_guard": L.GS_FP32_GUARD}


def n3cfg(nbins: int) -> int:
    return nbins * (nbins + 1) * (nbins + 2) // 6


def n4cfg(nbins: int) -> int:
    return nbins * (nbins + 1) * (nbins + 2) * (nbins + 3) // 24


def _xi(n2, n3):
    with np.errstate(divide="ignore", invalid="ignore"):
        return n2 / n3


class Engine:
    def __init__(self, device: int = 0):
        self._lib = L.lib()
        self._ctx = ctypes.c_void_p()
        L.check(self._lib.gs_create(ctypes.byref(self._ctx), int(device)))
        self.n_data = 0
        self.n_rand = 0
        self.generation = 0     # incremented by every catalogue upload

    def close(self):
        if self._ctx:
            self._lib.gs_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ catalogue
    def set_catalogue(self, xyz, w=None, n_data=None, buf=None):
        """xyz (n,3) float64 with data first then randoms; w raw weights (n,) or None;
        buf (n,) int32 with 1 = never an apex, or None.  numpy arrays are copied
        from host memory; torch CUDA tensors are copied device to device."""
        try:
            import torch
            is_dev = isinstance(xyz, torch.Tensor) and xyz.is_cuda
        except ImportError:  # pragma: no cover
            is_dev = False
        n = int(xyz.shape[0])
        nd = n if n_data is None else int(n_data)
        if nd < 0 or nd > n:
            raise ValueError("n_data out of range")
        if is_dev:
            import torch
            xyz = xyz.to(torch.float64).contiguous()
            w = None if w is None else w.to(torch.float64).contiguous()
            buf = None if buf is None else buf.to(torch.int32).contiguous()
            torch.cuda.current_stream().synchronize()
            L.check(self._lib.gs_set_catalogue_device(
                self._ctx, ctypes.c_void_p(xyz.data_ptr()),
                None if w is None else ctypes.c_void_p(w.data_ptr()),
                None if buf is None else ctypes.c_void_p(buf.data_ptr()), nd, n - nd))
        else:
            xyz = np.ascontiguousarray(xyz, dtype=np.float64)
            if xyz.ndim != 2 or xyz.shape[1] != 3:
                raise ValueError("xyz must be (n, 3)")
            w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
            buf = None if buf is None else np.ascontiguousarray(buf, dtype=np.int32)
            if w is not None and w.shape != (n,):
                raise ValueError("w must be (n,)")
            if buf is not None and buf.shape != (n,):
                raise ValueError("buf must be (n,)")
            L.check(self._lib.gs_set_catalogue(self._ctx, L.ptr(xyz), L.ptr(w), L.ptr(buf, ctypes.c_int32),
                                               nd, n - nd))
        self.n_data, self.n_rand = nd, n - nd
        self.generation += 1

    def set_catalogue_batched(self, xyz, w, buf, off, ndata):
        """Independent regions: region r is points off[r]:off[r+1] (data first,
        ndata[r] of them, then randoms); buf flags never-apex shell points."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
        buf = None if buf is None else np.ascontiguousarray(buf, dtype=np.int32)
        off = np.ascontiguousarray(off, dtype=np.int64)
        ndata = np.ascontiguousarray(ndata, dtype=np.int64)
        L.check(self._lib.gs_set_catalogue_batched(self._ctx, L.ptr(xyz), L.ptr(w), L.ptr(buf, ctypes.c_int32),
                                                   L.ptr(off, ctypes.c_int64), L.ptr(ndata, ctypes.c_int64),
                                                   len(ndata)))
        self.nreg = len(ndata)
        self.generation += 1

    def decompose(self, xyz, region, rmax, nreg=None):
        """Domain decomposition on the GPU (gs_decompose): labels 1..nreg (default
        max(region)) -> dict(off[nreg+1], idx, buf, box[nreg, 6]); region r's
        members are idx[off[r]:off[r+1]] with buffer flags buf[...].  Single
        precision, as the reference's tool."""
        x = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        reg = np.ascontiguousarray(region, dtype=np.int32)
        if nreg is None:
            nreg = int(reg.max()) if len(reg) else 0
        if nreg < 1:
            raise ValueError("no region label >= 1")
        off = np.zeros(nreg + 1, np.int64)
        box = np.zeros((nreg, 6), np.float32)
        fp = ctypes.c_float
        L.check(self._lib.gs_decompose(self._ctx, L.ptr(x, fp), L.ptr(reg, ctypes.c_int32), len(x), fp(rmax), nreg,
                                       L.ptr(off, ctypes.c_int64), None, None, 0, L.ptr(box, fp)))
        idx = np.zeros(max(int(off[-1]), 1), np.int32)
        buf = np.zeros(max(int(off[-1]), 1), np.int32)
        L.check(self._lib.gs_decompose(self._ctx, L.ptr(x, fp), L.ptr(reg, ctypes.c_int32), len(x), fp(rmax), nreg,
                                       L.ptr(off, ctypes.c_int64), L.ptr(idx, ctypes.c_int32),
                                       L.ptr(buf, ctypes.c_int32), int(off[-1]), None))
        return dict(off=off, idx=idx[:off[-1]], buf=buf[:off[-1]], box=box)

    def set_catalogue_regions(self, xyz, w, region, rmax, nreg=None):
        """Decomposes the catalogue on the GPU and sets the regions as a batched
        catalogue: region r holds its members as the tool writes them to
        "<r>.loadnodes" (single-precision coordinates and weights), every member
        counted as data, with the tool's buffer flags.  pcf2_batched then gives
        per region what the -cut run of that file computes with those flags
        (its wsum class 0 is the un-normalised N2).  Returns the decomposition."""
        d = self.decompose(xyz, region, rmax, nreg)
        x = np.asarray(xyz, np.float32).reshape(-1, 3).astype(np.float64)[d["idx"]]
        ws = None if w is None else np.asarray(w, np.float32).astype(np.float64)[d["idx"]]
        self.set_catalogue_batched(x, ws, d["buf"], d["off"], np.diff(d["off"]))
        return d

    def pcf2_batched(self, rmin, rmax, nbins, nmu=1, log_bins=False, precision="fp32", cell_div=0):
        p = self.params(rmin, rmax, nbins, nmu, log_bins, precision, cell_div)
        R = self.nreg
        cnt = np.zeros((R, 3, nbins, nmu), np.uint64)
        wsum = np.zeros((R, 3, nbins, nmu))
        n2 = np.zeros((R, nbins, nmu))
        n3 = np.zeros((R, nbins, nmu))
        st = L.gs_stats()
        L.check(self._lib.gs_count_2pcf_batched(self._ctx, ctypes.byref(p), L.ptr(cnt, ctypes.c_uint64),
                                                L.ptr(wsum), L.ptr(n2), L.ptr(n3), ctypes.byref(st)))
        return dict(cnt=cnt, wsum=wsum, n2=n2, n3=n3, xi=_xi(n2, n3), stats=st.as_dict())

    def pcf3_batched(self, rmin, rmax, nbins, nmu=1, log_bins=False, precision="fp32", cell_div=0):
        """Per-region 3PCF of the batched catalogue: cnt[R, ncfg, nmu, 4], n2/n3[R, ncfg, nmu]."""
        p = self.params(rmin, rmax, nbins, nmu, log_bins, precision, cell_div)
        R, nc = self.nreg, n3cfg(nbins)
        cnt = np.zeros((R, nc, nmu, 4), np.uint64)
        wsum = np.zeros((R, nc, nmu, 4))
        n2 = np.zeros((R, nc, nmu))
        n3 = np.zeros((R, nc, nmu))
        st = L.gs_stats()
        L.check(self._lib.gs_count_3pcf_batched(self._ctx, ctypes.byref(p), L.ptr(cnt, ctypes.c_uint64),
                                                L.ptr(wsum), L.ptr(n2), L.ptr(n3), ctypes.byref(st)))
        return dict(cnt=cnt, wsum=wsum, n2=n2, n3=n3, zeta=_xi(n2, n3), stats=st.as_dict())

    # ------------------------------------------------------------ counts
    @staticmethod
    def params(rmin, rmax, nbins, nmu=1, log_bins=False, precision="fp32", cell_div=0):
        return L.make_params(rmin, rmax, nbins, nmu, log_bins, PRECISION[precision], cell_div)

    def pcf2(self, rmin, rmax, nbins, nmu=1, log_bins=False, precision="fp32", cell_div=0):
        p = self.params(rmin, rmax, nbins, nmu, log_bins, precision, cell_div)
        cnt = np.zeros((3, nbins, nmu), np.uint64)
        wsum = np.zeros((3, nbins, nmu))
        n2 = np.zeros((nbins, nmu))
        n3 = np.zeros((nbins, nmu))
        st = L.gs_stats()
        L.check(self._lib.gs_count_2pcf(self._ctx, ctypes.byref(p), L.ptr(cnt, ctypes.c_uint64), L.ptr(wsum),
                                        L.ptr(n2), L.ptr(n3), ctypes.byref(st)))
        return dict(cnt=cnt, wsum=wsum, n2=n2, n3=n3, xi=_xi(n2, n3), stats=st.as_dict())

    def pcf2_noalloc(self, xyz, w, n_data, rmin, rmax, nbins, nmu=1, log_bins=False, precision="fp32",
                     cell_div=0, buf=None):
        """gs_count_2pcf_noalloc: the 2PCF of a catalogue in device memory (torch CUDA tensors:
        xyz (n, 3) float64, w (n,) float64 or None, buf (n,) int32 or None) with no device
        allocation; the context's own catalogue is not touched."""
        import torch
        xyz = xyz.to(torch.float64).contiguous()
        w = None if w is None else w.to(torch.float64).contiguous()
        buf = None if buf is None else buf.to(torch.int32).contiguous()
        n = int(xyz.shape[0])
        p = self.params(rmin, rmax, nbins, nmu, log_bins, precision, cell_div)
        cnt = np.zeros((3, nbins, nmu), np.uint64)
        wsum = np.zeros((3, nbins, nmu))
        n2 = np.zeros((nbins, nmu))
        n3 = np.zeros((nbins, nmu))
        st = L.gs_stats()
        torch.cuda.current_stream().synchronize()
        L.check(self._lib.gs_count_2pcf_noalloc(
            self._ctx, ctypes.byref(p), ctypes.c_void_p(xyz.data_ptr()),
            None if w is None else ctypes.c_void_p(w.data_ptr()),
            None if buf is None else ctypes.c_void_p(buf.data_ptr()), int(n_data), n - int(n_data),
            L.ptr(cnt, ctypes.c_uint64), L.ptr(wsum), L.ptr(n2), L.ptr(n3), ctypes.byref(st)))
        return dict(cnt=cnt, wsum=wsum, n2=n2, n3=n3, xi=_xi(n2, n3), stats=st.as_dict())

    def adjoint2(self, dl_dn2, dl_dn3, rmin, rmax, nbins, nmu=1, log_bins=False, precision="fp32", cell_div=0,
                 out=None):
        """dL/dw for the raw weights of the current catalogue, given dL/dN2 and
        dL/dN3 of shape (nbins, nmu).  Returns a numpy array, or fills `out`
        (a float64 torch CUDA tensor of n elements) on the device."""
        p = self.params(rmin, rmax, nbins, nmu, log_bins, precision, cell_div)
        g2 = np.ascontiguousarray(dl_dn2, dtype=np.float64).reshape(nbins, nmu)
        g3 = np.ascontiguousarray(dl_dn3, dtype=np.float64).reshape(nbins, nmu)
        st = L.gs_stats()
        if out is not None:
            import torch
            if out.dtype != torch.float64 or not out.is_cuda or out.numel() != self.n_data + self.n_rand:
                raise ValueError("out must be a float64 CUDA tensor with one entry per point")
            L.check(self._lib.gs_adjoint_2pcf_device(self._ctx, ctypes.byref(p), L.ptr(g2), L.ptr(g3),
                                                     ctypes.c_void_p(out.data_ptr()), ctypes.byref(st)))
            return out
        res = np.zeros(self.n_data + self.n_rand)
        L.check(self._lib.gs_adjoint_2pcf(self._ctx, ctypes.byref(p), L.ptr(g2), L.ptr(g3), L.ptr(res),
                                          ctypes.byref(st)))
        self.last_stats = st.as_dict()
        return res

    def pcf3(self, rmin, rmax, nbins, nmu=1, log_bins=False, precision="fp32", cell_div=0):
        """3PCF over all configurations; nmu >= 2 is the RSD 3PCF.  Arrays are
        [ncfg, 4] / [ncfg] for nmu == 1 and [ncfg, nmu, 4] / [ncfg, nmu] otherwise."""
        return self._pcf3(self._lib.gs_count_3pcf, n3cfg(nbins), rmin, rmax, nbins, nmu, log_bins, precision,
                          cell_div)

    def pcf3_equi(self, rmin, rmax, nbins, nmu=1, log_bins=False, precision="fp32", cell_div=0):
        """Equilateral 3PCF: arrays [nbins, nmu, 4] / [nbins, nmu] (the radial bin replaces the configuration)."""
        r = self._pcf3(self._lib.gs_count_3pcf_equi, nbins, rmin, rmax, nbins, nmu, log_bins, precision, cell_div,
                       squeeze=False)
        return r

    def _pcf3(self, fn, nc, rmin, rmax, nbins, nmu, log_bins, precision, cell_div, squeeze=True):
        p = self.params(rmin, rmax, nbins, nmu, log_bins, precision, cell_div)
        cnt = np.zeros((nc, nmu, 4), np.uint64)
        wsum = np.zeros((nc, nmu, 4))
        n2 = np.zeros((nc, nmu))
        n3 = np.zeros((nc, nmu))
        st = L.gs_stats()
        L.check(fn(self._ctx, ctypes.byref(p), L.ptr(cnt, ctypes.c_uint64), L.ptr(wsum), L.ptr(n2), L.ptr(n3),
                   ctypes.byref(st)))
        if squeeze and nmu == 1:
            cnt, wsum, n2, n3 = cnt[:, 0], wsum[:, 0], n2[:, 0], n3[:, 0]
        return dict(cnt=cnt, wsum=wsum, n2=n2, n3=n3, zeta=_xi(n2, n3), stats=st.as_dict())

    def pcf4(self, rmin, rmax, nbins, log_bins=False, precision="fp64", cell_div=0):
        p = self.params(rmin, rmax, nbins, 1, log_bins, precision, cell_div)
        nc = n4cfg(nbins)
        cnt = np.zeros((nc, nbins, nbins, 16), np.uint64)
        n2 = np.zeros((nc, nbins, nbins))
        n3 = np.zeros((nc, nbins, nbins))
        dn = np.zeros((6, nc, nbins, nbins))
        dr = np.zeros((6, nc, nbins, nbins))
        st = L.gs_stats()
        L.check(self._lib.gs_count_4pcf(self._ctx, ctypes.byref(p), L.ptr(cnt, ctypes.c_uint64), L.ptr(n2),
                                        L.ptr(n3), L.ptr(dn), L.ptr(dr), ctypes.byref(st)))
        return dict(cnt=cnt, n2=n2, n3=n3, disc_n=dn, disc_r=dr, eta=_xi(n2, n3), stats=st.as_dict())

    def pcf4_tetra(self, rmin, rmax, nbins, log_bins=False, precision="fp64", cell_div=0):
        """Regular-tetrahedron 4PCF: cnt[nbins, 16], n2/n3[nbins], disc_n/disc_r[6, nbins]."""
        p = self.params(rmin, rmax, nbins, 1, log_bins, precision, cell_div)
        cnt = np.zeros((nbins, 16), np.uint64)
        n2 = np.zeros(nbins)
        n3 = np.zeros(nbins)
        dn = np.zeros((6, nbins))
        dr = np.zeros((6, nbins))
        st = L.gs_stats()
        L.check(self._lib.gs_count_4pcf_tetra(self._ctx, ctypes.byref(p), L.ptr(cnt, ctypes.c_uint64), L.ptr(n2),
                                              L.ptr(n3), L.ptr(dn), L.ptr(dr), ctypes.byref(st)))
        return dict(cnt=cnt, n2=n2, n3=n3, disc_n=dn, disc_r=dr, eta=_xi(n2, n3), stats=st.as_dict())
