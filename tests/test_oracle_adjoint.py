"""Pins the oracle's 2PCF adjoint (or_adjoint_2pcf) independently of its
code: central finite differences of the oracle's forward N2/N3, and the exact
scale invariance of the normalised weights (sum over data of w dL/dw = 0, and
the same over randoms)."""
import numpy as np
import pytest

import oracle as O
from gramsci_amd import catalog as C


def _loss_and_grads(xyz, w, nd, p, target, buf=None):
    r = O.pcf2(xyz, w, nd, p, buf=buf)
    xi = r["n2"] / r["n3"]
    L = float(np.sum((xi - target) ** 2))
    gxi = 2.0 * (xi - target)
    gN2 = gxi / r["n3"]
    gN3 = -gxi * r["n2"] / r["n3"] ** 2
    return L, gN2, gN3


@pytest.mark.parametrize("nmu,log,with_buf", [(1, False, False), (4, True, False), (3, False, True)])
def test_adjoint_matches_finite_differences(nmu, log, with_buf):
    d = C.clustered(300, 30.0, 5, members=5)
    r = C.uniform(600, 30.0, 6)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(300, 0.3, 7), C.lognormal_weights(600, 0.2, 8)])
    buf = (np.random.default_rng(1).random(900) < 0.2).astype(np.int32) if with_buf else None
    p = O.Params(1.0, 8.0, 5, nmu, log)
    target = np.full((5, nmu), 0.05)
    L0, gN2, gN3 = _loss_and_grads(xyz, w, 300, p, target, buf)
    grad = O.adjoint2(xyz, w, 300, p, gN2, gN3, buf)
    rng = np.random.default_rng(2)
    for k in list(rng.choice(300, 5, replace=False)) + list(300 + rng.choice(600, 5, replace=False)):
        h = 1e-5 * w[k]
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (_loss_and_grads(xyz, wp, 300, p, target, buf)[0] - _loss_and_grads(xyz, wm, 300, p, target, buf)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(grad).max()), (k, fd, grad[k])


def test_adjoint_scale_invariance():
    d = C.clustered(400, 40.0, 9, members=5)
    r = C.uniform(800, 40.0, 10)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(400, 0.3, 11), np.ones(800)])
    p = O.Params(1.0, 10.0, 6, 2, False)
    _, gN2, gN3 = _loss_and_grads(xyz, w, 400, p, np.zeros((6, 2)))
    g = O.adjoint2(xyz, w, 400, p, gN2, gN3)
    scale = np.abs(w * g).sum()
    assert abs(np.dot(w[:400], g[:400])) <= 1e-12 * scale
    assert abs(np.dot(w[400:], g[400:])) <= 1e-12 * scale
