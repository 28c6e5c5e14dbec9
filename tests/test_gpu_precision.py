"""Per-bin precision of the weighted sums against the FP64 oracle.

The engine sums weight products as exact fixed-point integers: each point's
weight is scaled by a power of two of its class (data or random) that keeps
the class's largest weight below 2^26, each pair product is rounded to an
integer below 2^52, and the integers are added exactly.  A bin's class sum
S_b therefore carries at most n_b * 2^-53 * max_product_class of rounding
error, i.e. a relative error of at most 2^-53 * max_product_class /
mean_product_b: set by the spread of weights within a class, not by the ratio
of data to random weights.  The oracle sums the FP64 products in long double
(relative error below ~1.2e-16 per bin).

Asserted here, per bin with no floor (only empty bins are skipped):
  * wsum of every class (DD, DR, RR) within 1e-13 relative of the oracle's;
  * N3 = RR / Wr^2 within 1e-13 relative;
and reported (printed, for the measurement notes): the largest per-bin
relative error of wsum per class, of N2, N3 and xi, and of xi scaled by its
condition number (|DD/Wd^2| + |DR/(Wd Wr)| + |RR/Wr^2|) / |N2|, which is what
the cancellation in N2 = DD/Wd^2 - DR/(Wd Wr) + RR/Wr^2 amplifies.

Workloads: a 150^3 sub-box of the (r, mu) configuration at its density
(clustered galaxies with lognormal(0, 0.2) weights, unit-weight randoms,
1 < r < 150, 30 log bins, nmu = 20); the 5,000 + 15,000-point subset of the
differentiable configuration (w = exp(theta . t) from 8 Gaussian random field
templates, 0 < r < 120, 20 bins, nmu = 10); and a catalogue whose random
weights are 1e-4 of the data weights (the classes' ranges differ by 10^8 in
their products).
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


def _rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    nz = b != 0
    return float(np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz]))) if nz.any() else 0.0


def _c2_subbox():
    # the (r, mu) configuration's recipe at its density (4M points in 1000^3), cut to 150^3
    side = 150.0
    n_gal = int(round(1_000_000 * (side / 1000.0) ** 3))
    d = C.neyman_scott(n_gal // 2, n_gal // 40, 20, 2.0, side, 211)
    r = C.uniform(3 * n_gal, side, 212)
    w = np.concatenate([C.lognormal_weights(len(d), 0.2, 213), np.ones(len(r))])
    return np.concatenate([d, r]), w, len(d), O.Params(1.0, 150.0, 30, 20, True)


def _c5_subset():
    side = (20_000 / 2e-3) ** (1.0 / 3.0)
    d = C.clustered(5000, side, 501, members=5, sigma=2.0)
    r = C.uniform(15000, side, 502)
    t = C.sample_trilinear(C.grf_templates(8, 32, side, 50.0, 503), d, side)
    theta = np.random.default_rng(504).normal(0, 0.05, 8)
    w = np.concatenate([np.exp(t @ theta), np.ones(len(r))])
    return np.concatenate([d, r]), w, len(d), O.Params(0.0, 120.0, 20, 10, False)


def _split_classes():
    d = C.clustered(4000, 100.0, 26)
    r = C.uniform(8000, 100.0, 27)
    w = np.concatenate([C.lognormal_weights(4000, 0.3, 28), 1e-4 * C.lognormal_weights(8000, 0.3, 29)])
    return np.concatenate([d, r]), w, 4000, O.Params(1.0, 25.0, 10, 5, True)


CASES = {"c2_subbox": _c2_subbox, "c5_subset": _c5_subset, "split_classes": _split_classes}


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_weighted_sums_per_bin(eng, case, prec):
    xyz, w, nd, p = CASES[case]()
    ref = O.pcf2(xyz, w, nd, p)
    eng.set_catalogue(xyz, w, nd)
    out = eng.pcf2(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, prec)
    assert np.array_equal(out["cnt"], ref["cnt"])
    err = {f"wsum_{c}": _rel(out["wsum"][k], ref["wsum"][k]) for k, c in enumerate(("DD", "DR", "RR"))}
    err["N2"] = _rel(out["n2"], ref["n2"])
    err["N3"] = _rel(out["n3"], ref["n3"])
    with np.errstate(divide="ignore", invalid="ignore"):
        xi_o, xi_r = out["n2"] / out["n3"], ref["n2"] / ref["n3"]
        ok = np.isfinite(xi_r) & (ref["n2"] != 0)
        err["xi"] = _rel(xi_o[ok], xi_r[ok])
        # condition number of N2 = DD/Wd^2 - DR/(Wd Wr) + RR/Wr^2 per bin
        wd, wr = w[:nd].sum(), w[nd:].sum()
        terms = (np.abs(ref["wsum"][0]) / wd ** 2 + np.abs(ref["wsum"][1]) / (wd * wr)
                 + np.abs(ref["wsum"][2]) / wr ** 2)
        kappa = terms / np.abs(ref["n2"])
        err["xi/kappa"] = float(np.max(np.abs(xi_o[ok] - xi_r[ok]) / np.abs(xi_r[ok]) / kappa[ok]))
        err["max_kappa"] = float(np.max(kappa[ok]))
    print(f"per-bin relative error, {case} {prec}: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()),
          flush=True)
    for c in ("DD", "DR", "RR"):
        assert err[f"wsum_{c}"] <= 1e-13, (c, err)
    assert err["N3"] <= 1e-13, err
    # N2 and xi cancel; their per-bin error stays within 1e-13 of the terms' sum (the condition
    # number times the relative error of the class sums)
    assert err["xi/kappa"] <= 1e-13, err
