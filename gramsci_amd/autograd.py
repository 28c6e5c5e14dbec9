"""torch.autograd wrapper of the weighted Landy-Szalay xi(r, mu).

    xi = weighted_xi(w, xyz, n_data, rmin, rmax, nbins, nmu, log_bins)

w is a float64 CUDA tensor of raw per-point weights (data first, then
randoms), xyz a float64 CUDA tensor (n, 3).  The forward pass is the engine's
2PCF count (N2, N3 exactly as the reference normalises and sums them); the
backward pass is the hand-written adjoint kernel (gs_adjoint_2pcf_device),
so gradients flow to anything w was computed from (e.g. w = exp(t @ theta)).
"""
from __future__ import annotations

import torch

from .engine import Engine

_engines = {}


def _engine(device_index: int) -> Engine:
    if device_index not in _engines:
        _engines[device_index] = Engine(device_index)
    return _engines[device_index]


class WeightedXi(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, xyz, n_data, params, buf):
        eng = _engine(w.device.index or 0)
        wd = w.detach().to(torch.float64).contiguous()
        eng.set_catalogue(xyz, wd, n_data, buf)
        r = eng.pcf2(**params)
        n2 = torch.from_numpy(r["n2"]).to(w.device)
        n3 = torch.from_numpy(r["n3"]).to(w.device)
        ctx.save_for_backward(wd, n2, n3)
        ctx.eng, ctx.xyz, ctx.n_data, ctx.params, ctx.buf = eng, xyz, n_data, params, buf
        ctx.generation = eng.generation
        return n2 / n3

    @staticmethod
    def backward(ctx, gxi):
        wd, n2, n3 = ctx.saved_tensors
        eng = ctx.eng
        g = gxi.detach().to(torch.float64)
        # dxi/dN2 = 1/N3, dxi/dN3 = -N2/N3^2.  A bin whose incoming gradient is exactly zero
        # (e.g. masked out of the loss) contributes nothing, even where N3 = 0 makes its own
        # derivatives non-finite; any other non-finite value is passed on, so the adjoint
        # returns NaN gradients as for a loss that depends on an undefined xi.
        zero = g == 0
        g2 = torch.where(zero, torch.zeros_like(g), g / n3).cpu().numpy()
        g3 = torch.where(zero, torch.zeros_like(g), -g * n2 / (n3 * n3)).cpu().numpy()
        # the context may hold another catalogue since the forward pass
        if eng.generation != ctx.generation:
            eng.set_catalogue(ctx.xyz, wd, ctx.n_data, ctx.buf)
            ctx.generation = eng.generation
        out = torch.empty_like(wd)
        eng.adjoint2(g2, g3, out=out, **ctx.params)
        return out, None, None, None, None


def weighted_xi(w, xyz, n_data, rmin, rmax, nbins, nmu=1, log_bins=False, precision="fp32", cell_div=0, buf=None):
    params = dict(rmin=rmin, rmax=rmax, nbins=nbins, nmu=nmu, log_bins=log_bins, precision=precision,
                  cell_div=cell_div)
    return WeightedXi.apply(w, xyz, int(n_data), params, buf)
