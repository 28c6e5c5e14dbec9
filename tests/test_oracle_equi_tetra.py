"""Pins the oracle's equilateral 3PCF (-equi) and regular-tetrahedron 4PCF
(-tetra) against the reference program's own output files (tests/golden/,
written by make_golden.py) and against brute-force enumerations that use only
the oracle's pair predicate.
"""
import itertools

import numpy as np
import pytest

import oracle as O
from golden_util import close_e147, load_input, manifest, parse_rows, read_text

MAN = manifest()
CASES_E = [n for n, c in MAN.items() if c["mode"] == "-equi"]
CASES_T = [n for n, c in MAN.items() if c["mode"] == "-tetra"]


def _params(c):
    return O.Params(**c["params"])


def _cat(name):
    d, r, wd, wr = load_input(name)
    xyz = np.concatenate([d, r]) if len(r) else d
    w = None
    if wd is not None:
        w = np.concatenate([wd, wr if wr is not None else np.ones(len(r))])
    return xyz, w, len(d), len(r)


def _bins(p, xyz):
    n = len(xyz)
    b = np.zeros((n, n), int)
    for i in range(n):
        for j in range(n):
            if i != j:
                b[i, j] = O.pair_rbin(p, xyz[i], xyz[j])
    return b


# ------------------------------------------------------------------ -equi

@pytest.mark.parametrize("name", CASES_E)
def test_equi_matches_reference_output(name):
    c = MAN[name]
    p = _params(c)
    xyz, w, nd, nr = _cat(name)
    run = "DR" if "DR" in c["runs"] else "D"
    rows = parse_rows(read_text(name, run))
    res = O.pcf3_equi(xyz, w, nd, p)
    nb, nm = p.nbins, p.nmu
    assert rows.shape == (nb * nm, 7)
    edges = O.radial_bins(p)
    assert close_e147(rows[:, 0], np.repeat(edges[:-1], nm)).all()
    assert close_e147(rows[:, 1], np.repeat(edges[1:], nm)).all()
    mu = np.tile(np.arange(nm), nb)   # (k-1)/mu_scale/2 = (k-1)/nmu (gramsci.F90:451)
    assert close_e147(rows[:, 2], mu / nm).all() and close_e147(rows[:, 3], (mu + 1) / nm).all()
    assert close_e147(rows[:, 4], res["n2"].reshape(-1)).all()
    assert close_e147(rows[:, 5], res["n3"].reshape(-1)).all()


@pytest.mark.parametrize("name", [n for n in CASES_E if not MAN[n]["weighted"]])
def test_equi_exact_counts_from_reference(name):
    c = MAN[name]
    p = _params(c)
    xyz, _, nd, nr = _cat(name)
    cnt = O.pcf3_equi(xyz, None, nd, p)["cnt"].astype(np.int64).reshape(-1, 4)
    checks = {"D": (nd, cnt[:, 0]), "R": (nr, cnt[:, 3]), "T": (nd + nr, cnt.sum(1))}
    for run, (n, expect) in checks.items():
        if run not in c["runs"]:
            continue
        rows = parse_rows(read_text(name, run))
        got = np.rint(rows[:, 4] * float(n) ** 3).astype(np.int64)
        assert np.array_equal(got, expect), (run, got, expect)


def test_equi_is_the_3pcf_diagonal():
    # without RSD the equilateral counts are the 3PCF's (b, b, b) configurations
    name = "equi"
    p = _params(MAN[name])
    xyz, _, nd, nr = _cat(name)
    e = O.pcf3_equi(xyz, None, nd, p)
    t = O.pcf3(xyz, None, nd, p)
    diag = [O.n3cfg(p.nbins) - O.n3cfg(p.nbins - b) for b in range(p.nbins)]   # cfg index of (b, b, b)
    assert np.array_equal(e["cnt"][:, 0, :], t["cnt"][diag])
    assert np.allclose(e["n2"][:, 0], t["n2"][diag], rtol=1e-13, atol=0)


def test_equi_bruteforce_tiny():
    rng = np.random.default_rng(5)
    xyz = rng.uniform(0, 16, (70, 3))
    p = O.Params(2.0, 8.0, 3)
    b = _bins(p, xyz)
    expect = np.zeros(3, np.int64)
    for i, j, k in itertools.combinations(range(len(xyz)), 3):
        if b[i, j] and b[i, j] == b[j, k] == b[i, k]:
            expect[b[i, j] - 1] += 3         # once per apex
    got = O.pcf3_equi(xyz, None, len(xyz), p)["cnt"][:, 0, 0].astype(np.int64)
    assert np.array_equal(got, expect)


# ----------------------------------------------------------------- -tetra

def _tetra_disc(dn, dr):
    with np.errstate(divide="ignore", invalid="ignore"):
        q = dn / dr
        return q[0] * q[4] + q[2] * q[3] + q[1] * q[5]   # gramsci.F90:652-654


@pytest.mark.parametrize("name", CASES_T)
def test_tetra_matches_reference_output(name):
    c = MAN[name]
    p = _params(c)
    xyz, w, nd, nr = _cat(name)
    run = "DR" if "DR" in c["runs"] else "D"
    rows = parse_rows(read_text(name, run))
    res = O.pcf4_tetra(xyz, w, nd, p)
    assert rows.shape == (p.nbins, 8)
    edges = O.radial_bins(p)
    assert close_e147(rows[:, 0], edges[:-1]).all() and close_e147(rows[:, 1], edges[1:]).all()
    assert close_e147(rows[:, 4], res["n2"]).all()
    assert close_e147(rows[:, 5], res["n3"]).all()
    disc = _tetra_disc(res["disc_n"], res["disc_r"])
    ok = close_e147(rows[:, 7], disc, ulps=2.0)
    ok |= np.isnan(disc) & ~np.isfinite(rows[:, 7])
    assert ok.all(), (rows[:, 7], disc)


@pytest.mark.parametrize("name", [n for n in CASES_T if not MAN[n]["weighted"]])
def test_tetra_exact_counts_from_reference(name):
    c = MAN[name]
    p = _params(c)
    d, r, _, _ = load_input(name)
    cats = {"D": d, "R": r, "T": np.concatenate([d, r]) if len(r) else d}
    for run, cat in cats.items():
        if run not in c["runs"]:
            continue
        expect = O.pcf4_tetra(cat, None, len(cat), p)["cnt"][:, 0].astype(np.int64)
        rows = parse_rows(read_text(name, run))
        got = np.rint(rows[:, 4] * float(len(cat)) ** 4).astype(np.int64)
        assert np.array_equal(got, expect), run
    if "T" in c["runs"]:
        xyz, _, nd, nr = _cat(name)
        tot = O.pcf4_tetra(xyz, None, nd, p)["cnt"].sum(-1)
        rows = parse_rows(read_text(name, "T"))
        assert np.array_equal(np.rint(rows[:, 4] * float(nd + nr) ** 4).astype(np.int64), tot)


def test_tetra_bruteforce_tiny():
    # every 4-point set with all six sides in one bin is counted once per apex
    rng = np.random.default_rng(7)
    xyz = rng.uniform(0, 9, (40, 3))
    p = O.Params(1.5, 6.0, 2)
    b = _bins(p, xyz)
    expect = np.zeros(2, np.int64)
    for q in itertools.combinations(range(len(xyz)), 4):
        e = {b[u, v] for u, v in itertools.combinations(q, 2)}
        if len(e) == 1 and 0 not in e:
            expect[e.pop() - 1] += 4
    got = O.pcf4_tetra(xyz, None, len(xyz), p)["cnt"][:, 0].astype(np.int64)
    assert np.array_equal(got, expect)
    assert expect.sum() > 0


def test_tetra_three_neighbour_apex_counted():
    # unlike the perimeter 4PCF, an apex with exactly three neighbours forms a
    # tetrahedron: a regular tetrahedron of side 3 gives 4 apex-ordered counts
    xyz = np.array([[1.0, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) * (3.0 / np.sqrt(8.0)) + 10.0
    p = O.Params(1.0, 5.0, 2)
    assert O.pcf4_tetra(xyz, None, 4, p)["cnt"].sum() == 4
    assert O.pcf4(xyz, None, 4, p)["cnt"].sum() == 0
