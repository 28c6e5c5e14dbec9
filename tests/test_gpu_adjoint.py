"""The hand-written 2PCF adjoint (gs_adjoint_2pcf) and its torch.autograd
wrapper, against the oracle's brute-force CPU gradient (itself pinned by
finite differences in test_oracle_adjoint.py) and against FP64 central
finite differences of the GPU forward pass.

Tolerances: dL/dw within 1e-10 of max|dL/dw| of the oracle (the GPU sums each
point's pair terms in FP64 in a different order); the scale invariance
sum_data w dL/dw = 0 to 1e-12 of sum |w dL/dw|; dL/dtheta within 1e-6
relative of the finite-difference estimate.
"""
import numpy as np
import pytest

import oracle as O
from gramsci_amd import catalog as C
from gramsci_amd.engine import Engine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _subset_c5(n_gal=1250, n_ran=3750, seed=0):
    """A 5,000-point subset at the density of the differentiable configuration
    (2M points in 1000^3): side (5000 / 2e-3)^(1/3) ~ 136."""
    side = (float(n_gal + n_ran) / 2e-3) ** (1.0 / 3.0)
    d = C.clustered(n_gal, side, 500 + seed, members=5, sigma=2.0)
    r = C.uniform(n_ran, side, 510 + seed)
    fields = C.grf_templates(8, 32, side, 50.0, 520 + seed)
    t = C.sample_trilinear(fields, d, side)
    theta = np.random.default_rng(530 + seed).normal(0, 0.05, 8)
    return d, r, t, theta, side


def _loss_grads(n2, n3, target):
    xi = n2 / n3
    g = 2.0 * (xi - target)
    return float(np.sum((xi - target) ** 2)), g / n3, -g * n2 / n3 ** 2


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_adjoint_matches_oracle(eng, prec):
    d, r, t, theta, side = _subset_c5()
    xyz = np.concatenate([d, r])
    w = np.concatenate([np.exp(t @ theta), np.ones(len(r))])
    p = O.Params(0.0, 120.0, 20, 10, False)
    ref = O.pcf2(xyz, w, len(d), p)
    target = np.zeros((20, 10))
    _, g2, g3 = _loss_grads(ref["n2"], ref["n3"], target)
    g_ref = O.adjoint2(xyz, w, len(d), p, g2, g3)
    eng.set_catalogue(xyz, w, len(d))
    g_gpu = eng.adjoint2(g2, g3, p.rmin, p.rmax, p.nbins, p.nmu, False, prec)
    err = np.max(np.abs(g_gpu - g_ref)) / np.max(np.abs(g_ref))
    print(f"{prec}: max |dL/dw - oracle| / max|dL/dw| = {err:.3g}", flush=True)
    assert err < 1e-10
    s = np.abs(w * g_gpu).sum()
    assert abs(np.dot(w[:len(d)], g_gpu[:len(d)])) < 1e-12 * s
    assert abs(np.dot(w[len(d):], g_gpu[len(d):])) < 1e-12 * s


def test_adjoint_buffer_and_log_bins(eng):
    rng = np.random.default_rng(4)
    d = C.clustered(2000, 80.0, 41, members=5)
    r = C.uniform(4000, 80.0, 42)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(2000, 0.3, 43), C.lognormal_weights(4000, 0.1, 44)])
    buf = (rng.random(6000) < 0.25).astype(np.int32)
    p = O.Params(1.0, 25.0, 8, 5, True)
    ref = O.pcf2(xyz, w, 2000, p, buf=buf)
    _, g2, g3 = _loss_grads(ref["n2"], ref["n3"], np.full((8, 5), 0.1))
    g_ref = O.adjoint2(xyz, w, 2000, p, g2, g3, buf=buf)
    eng.set_catalogue(xyz, w, 2000, buf)
    g_gpu = eng.adjoint2(g2, g3, p.rmin, p.rmax, p.nbins, p.nmu, True, "fp32")
    assert np.max(np.abs(g_gpu - g_ref)) <= 1e-10 * np.max(np.abs(g_ref))


def test_autograd_theta_vs_finite_differences():
    import torch
    from gramsci_amd.autograd import weighted_xi
    d, r, t, theta0, side = _subset_c5(seed=1)
    xyz = torch.from_numpy(np.concatenate([d, r])).cuda()
    tt = torch.from_numpy(t).cuda()
    ones = torch.ones(len(r), dtype=torch.float64, device="cuda")
    p = dict(rmin=0.0, rmax=120.0, nbins=20, nmu=10)
    target = torch.zeros((20, 10), dtype=torch.float64, device="cuda")

    def loss_of(theta):
        w = torch.cat([torch.exp(tt @ theta), ones])
        xi = weighted_xi(w, xyz, len(d), **p)
        return ((xi - target) ** 2).sum()

    theta = torch.from_numpy(theta0).cuda().requires_grad_(True)
    L = loss_of(theta)
    L.backward()
    grad = theta.grad.detach().cpu().numpy()
    fd = np.zeros(8)
    with torch.no_grad():
        for k in range(8):
            h = 1e-4
            tp = theta.detach().clone()
            tm = theta.detach().clone()
            tp[k] += h
            tm[k] -= h
            fd[k] = (loss_of(tp).item() - loss_of(tm).item()) / (2 * h)
    print("dL/dtheta", grad, "fd", fd, flush=True)
    assert np.all(np.abs(grad - fd) <= 1e-6 * np.max(np.abs(fd)) + 1e-9 * np.abs(fd).max())


def test_adjoint_fp32_reproducible_and_matches_fp64_at_scale(eng):
    # 120k points at the density of the differentiable configuration, with buffer points: the FP32
    # half-stencil kernel (each pair once, terms to both ends, exact fixed-point sums) gives the same
    # bits on every call and agrees with the full-stencil FP64 kernel
    side = (120_000 / 2e-3) ** (1.0 / 3.0)
    d = C.clustered(30_000, side, 61, members=5, sigma=2.0)
    r = C.uniform(90_000, side, 62)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(30_000, 0.3, 63), C.lognormal_weights(90_000, 0.05, 64)])
    buf = (np.random.default_rng(65).random(len(xyz)) < 0.1).astype(np.int32)
    eng.set_catalogue(xyz, w, len(d), buf)
    rng = np.random.default_rng(66)
    g2 = rng.normal(0, 1, (20, 10))
    g3 = rng.normal(0, 1, (20, 10)) * 1e-3
    args = (0.0, 120.0, 20, 10, False)
    a = eng.adjoint2(g2, g3, *args, "fp32")
    b = eng.adjoint2(g2, g3, *args, "fp32")
    c = eng.adjoint2(g2, g3, *args, "fp64")
    assert np.array_equal(a, b)
    err = np.max(np.abs(a - c)) / np.max(np.abs(c))
    print(f"fp32 half-stencil vs fp64: {err:.3g}", flush=True)
    assert err < 1e-10


def test_adjoint_nonfinite_gradient_propagates(eng):
    # a non-finite dL/dN entry takes the FP64-summing kernel in both precisions: NaN propagates as in
    # the forward arithmetic instead of being rounded into the fixed-point sums
    d, r, t, theta, side = _subset_c5(seed=2)
    xyz = np.concatenate([d, r])
    w = np.concatenate([np.exp(t @ theta), np.ones(len(r))])
    eng.set_catalogue(xyz, w, len(d))
    g2 = np.ones((20, 10))
    g2[3, 4] = np.nan
    for prec in ("fp32", "fp64"):
        g = eng.adjoint2(g2, np.zeros((20, 10)), 0.0, 120.0, 20, 10, False, prec)
        assert np.isnan(g).any(), prec


def test_adjoint_tiny_and_empty_catalogues(eng):
    # 0 to 3 points, with and without randoms, in both precisions: the oracle's gradient (an empty
    # catalogue gives an empty one)
    pts = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [2.0, 0.5, 0.0]])
    g2 = np.arange(8.0).reshape(4, 2)
    g3 = np.ones((4, 2))
    p = O.Params(0.5, 3.0, 4, 2, False)
    for n, nd in ((0, 0), (1, 1), (2, 1), (3, 2), (3, 3)):
        xyz = pts[:n]
        w = np.linspace(1.0, 2.0, n)
        eng.set_catalogue(xyz, w, nd)
        ref = O.adjoint2(xyz, w, nd, p, g2, g3) if n else np.zeros(0)
        for prec in ("fp32", "fp64"):
            g = eng.adjoint2(g2, g3, p.rmin, p.rmax, p.nbins, p.nmu, False, prec)
            assert g.shape == (n,), (n, prec)
            assert np.all(np.abs(g - ref) <= 1e-12 * max(np.abs(ref).max() if n else 0.0, 1.0)), (n, nd, prec, g, ref)


def test_adjoint_full_size_properties(eng):
    # the differentiable configuration at its real size (500k Neyman-Scott galaxies with template
    # weights exp(theta . t) + 1.5M randoms in 1000^3, 0 < r < 120, 20 bins, nmu = 10): the FP32
    # half-stencil gradient is bit-identical from call to call, agrees with the full-stencil FP64
    # kernel, and satisfies the scale invariance sum over each catalogue of w dL/dw = 0
    d = C.neyman_scott(250000, 12500, 20, 2.0, 1000.0, 501)
    r = C.uniform(1500000, 1000.0, 502)
    fields = C.grf_templates(8, 128, 1000.0, 50.0, 503)
    t = C.sample_trilinear(fields, d, 1000.0)
    theta = np.random.default_rng(504).normal(0, 0.05, 8)
    xyz = np.concatenate([d, r])
    w = np.concatenate([np.exp(t @ theta), np.ones(len(r))])
    eng.set_catalogue(xyz, w, len(d))
    rng = np.random.default_rng(505)
    g2 = rng.normal(0, 1, (20, 10))
    g3 = rng.normal(0, 1, (20, 10)) * 1e-2
    args = (0.0, 120.0, 20, 10, False)
    a = eng.adjoint2(g2, g3, *args, "fp32")
    b = eng.adjoint2(g2, g3, *args, "fp32")
    c = eng.adjoint2(g2, g3, *args, "fp64")
    err = np.max(np.abs(a - c)) / np.max(np.abs(c))
    print(f"2M-point adjoint: fp32 vs fp64 {err:.3g}", flush=True)
    assert np.array_equal(a, b)
    assert err < 1e-10
    s = np.abs(w * a).sum()
    assert abs(np.dot(w[:len(d)], a[:len(d)])) < 1e-12 * s
    assert abs(np.dot(w[len(d):], a[len(d):])) < 1e-12 * s


@pytest.mark.parametrize("log,buffered", [(False, False), (True, True)])
def test_adjoint_isotropic(eng, log, buffered):
    # nmu = 1 (no mu binning): the isotropic variants of both adjoint kernels against the oracle
    rng = np.random.default_rng(7)
    d = C.clustered(2500, 90.0, 71, members=5)
    r = C.uniform(5000, 90.0, 72)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(2500, 0.3, 73), C.lognormal_weights(5000, 0.1, 74)])
    buf = (rng.random(len(xyz)) < 0.2).astype(np.int32) if buffered else None
    p = O.Params(1.0 if log else 0.0, 20.0, 8, 1, log)
    ref = O.pcf2(xyz, w, 2500, p, buf=buf)
    _, g2, g3 = _loss_grads(ref["n2"], ref["n3"], np.full((8, 1), 0.05))
    g_ref = O.adjoint2(xyz, w, 2500, p, g2, g3, buf=buf)
    eng.set_catalogue(xyz, w, 2500, buf)
    for prec in ("fp32", "fp64"):
        g = eng.adjoint2(g2, g3, p.rmin, p.rmax, p.nbins, 1, log, prec)
        err = np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref))
        print(f"isotropic adjoint {prec} log={log} buffered={buffered}: {err:.3g}", flush=True)
        assert err < 1e-10


def test_adjoint_linearity_and_extreme_scales(eng):
    # dL/dw is linear in (dL/dN2, dL/dN3): scaling them by s scales the gradient by s, also where the
    # fixed-point scale of the FP32 kernel would leave 2^-1000 .. 2^1000 (s = 1e-300 takes the
    # FP64-summing kernel) or sits near its far end (s = 1e290)
    d, r, t, theta, side = _subset_c5(n_gal=800, n_ran=2400, seed=3)
    xyz = np.concatenate([d, r])
    w = np.concatenate([np.exp(t @ theta), np.ones(len(r))])
    eng.set_catalogue(xyz, w, len(d))
    rng = np.random.default_rng(9)
    g2 = rng.normal(0, 1, (20, 10))
    g3 = rng.normal(0, 1, (20, 10))
    args = (0.0, 120.0, 20, 10, False, "fp32")
    base = eng.adjoint2(g2, g3, *args)
    for s in (1e-300, 1e290, -3.0):
        g = eng.adjoint2(s * g2, s * g3, *args) / s
        err = np.max(np.abs(g - base)) / np.max(np.abs(base))
        print(f"scale {s:g}: {err:.3g}", flush=True)
        assert err < 1e-12, s


def _empty_rr_bin_case():
    # randoms on a lattice of spacing 2.5 have no pair closer than 2.5, so bin 0 = [0, 2) has
    # RR = 0 (N3 = 0: xi infinite or undefined there) while the clustered data fill it
    side = 60.0
    d = C.clustered(1500, side, 71, members=5, sigma=1.5)
    r = C.lattice(24, 2.5, 0.0)
    w = np.concatenate([C.lognormal_weights(len(d), 0.2, 72), np.ones(len(r))])
    return d, r, w, dict(rmin=0.0, rmax=20.0, nbins=10, nmu=4)


def test_autograd_masked_empty_bin_gives_finite_gradient():
    import torch
    from gramsci_amd.autograd import weighted_xi
    d, r, w0, p = _empty_rr_bin_case()
    xyz_np = np.concatenate([d, r])
    xyz = torch.from_numpy(xyz_np).cuda()
    w = torch.from_numpy(w0).cuda().requires_grad_(True)
    xi = weighted_xi(w, xyz, len(d), **p)
    ok = torch.isfinite(xi)
    assert not bool(ok[0].any()) and bool(ok[1:].all())
    target = torch.full_like(xi, 0.5)
    # mask the input of the loss (not its output, whose backward would form 0 * inf), so the
    # masked bins receive an incoming gradient of exactly zero
    xi_masked = torch.where(ok, xi, target)
    loss = ((xi_masked - target) ** 2).sum()
    loss.backward()
    g = w.grad.detach().cpu().numpy()
    assert np.isfinite(g).all()
    # the same gradient from the oracle with the masked bin's dL/dN set to zero
    P = O.Params(p["rmin"], p["rmax"], p["nbins"], p["nmu"], False)
    ref = O.pcf2(xyz_np, w0, len(d), P)
    x = ref["n2"] / np.where(ref["n3"] == 0, 1.0, ref["n3"])
    gx = np.where(ref["n3"] != 0, 2.0 * (x - 0.5), 0.0)
    g2 = np.where(ref["n3"] != 0, gx / np.where(ref["n3"] == 0, 1.0, ref["n3"]), 0.0)
    g3 = np.where(ref["n3"] != 0, -gx * ref["n2"] / np.where(ref["n3"] == 0, 1.0, ref["n3"]) ** 2, 0.0)
    g_ref = O.adjoint2(xyz_np, w0, len(d), P, g2, g3)
    assert np.max(np.abs(g - g_ref)) <= 1e-10 * np.max(np.abs(g_ref))


def test_autograd_unmasked_empty_bin_gives_nonfinite_gradient():
    # a loss that depends on an infinite xi has no finite gradient: the wrapper passes the
    # non-finite dL/dN on (no rounding to a finite number), and the adjoint propagates it
    import torch
    from gramsci_amd.autograd import weighted_xi
    d, r, w0, p = _empty_rr_bin_case()
    xyz = torch.from_numpy(np.concatenate([d, r])).cuda()
    w = torch.from_numpy(w0).cuda().requires_grad_(True)
    xi = weighted_xi(w, xyz, len(d), **p)
    loss = ((xi - 0.5) ** 2).sum()
    loss.backward()
    g = w.grad.detach().cpu().numpy()
    assert not np.isfinite(g).all()
