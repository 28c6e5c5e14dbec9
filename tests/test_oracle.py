"""Pins the CPU oracle (oracle/gs_oracle.c) before anything is compared with it.

Checks, per golden fixture written by the reference program itself:
  * N2, N3, xi and the bin-edge columns agree with the printed reference output
    to the last printed digit (+-1 unit: the reference's thread-order ties);
  * exact ordered DD / RR / total counts, recovered from the reference's
    D-only, R-only and D+R runs (N2 * N^k rounds to an integer), equal the
    oracle's integer counts per bin;
and, independently of the reference:
  * lattice pair counts per bin against the closed-form lattice enumeration;
  * total pair counts against scipy.spatial.cKDTree.count_neighbors;
  * 3PCF apex counts are multiples of 3, DD is even, and 2PCF/3PCF counts are
    invariant under a permutation of the input.
"""
import itertools
import math
import os

import numpy as np
import pytest

import oracle as O
from golden_util import close_e147, load_input, manifest, parse_rows, read_text

MAN = manifest()
CASES_2 = [n for n, c in MAN.items() if c["mode"] == "-2pcf"]
CASES_3 = [n for n, c in MAN.items() if c["mode"] == "-3pcf"]
CASES_4 = [n for n, c in MAN.items() if c["mode"] == "-4pcf"]


def _params(c):
    return O.Params(**c["params"])


def _cat(name):
    d, r, wd, wr = load_input(name)
    xyz = np.concatenate([d, r]) if len(r) else d
    w = None
    if wd is not None:
        w = np.concatenate([wd, wr if wr is not None else np.ones(len(r))])
    return xyz, w, len(d), len(r)


# ----------------------------------------------------------------------- 2PCF

@pytest.mark.parametrize("name", CASES_2)
def test_2pcf_matches_reference_output(name):
    c = MAN[name]
    p = _params(c)
    xyz, w, nd, nr = _cat(name)
    run = "DR" if "DR" in c["runs"] else "D"
    rows = parse_rows(read_text(name, run))
    res = O.pcf2(xyz, w, nd, p)
    assert rows.shape == (p.nbins * p.nmu, 7)
    edges = O.radial_bins(p)
    mu_scale = 0.5 * p.nmu
    k = np.arange(1, p.nmu + 1)
    exp_cols = np.column_stack([
        np.repeat(edges[:-1], p.nmu), np.repeat(edges[1:], p.nmu),
        np.tile((k - 1.0) / mu_scale - 1.0, p.nbins), np.tile(k / mu_scale - 1.0, p.nbins),
        res["n2"].ravel(), res["n3"].ravel(),
        np.where(res["n3"].ravel() != 0, res["n2"].ravel() / np.where(res["n3"].ravel() == 0, 1, res["n3"].ravel()),
                 np.nan)])
    ok = close_e147(rows[:, :6], exp_cols[:, :6])
    assert ok.all(), np.argwhere(~ok)[:5]
    if nr:
        ok = close_e147(rows[:, 6], exp_cols[:, 6])
        assert ok.all()
    assert res["overflow"] == 0


@pytest.mark.parametrize("name", [n for n in CASES_2 if len(MAN[n]["runs"]) > 1 or MAN[n]["runs"] == ["D"]])
def test_2pcf_exact_counts_from_reference(name):
    c = MAN[name]
    p = _params(c)
    xyz, w, nd, nr = _cat(name)
    res = O.pcf2(xyz, None, nd, p)
    cnt = res["cnt"].astype(np.int64)
    checks = {"D": (nd, cnt[0]), "R": (nr, cnt[2]), "T": (nd + nr, cnt.sum(0))}
    done = 0
    for run, (n, expect) in checks.items():
        if run not in c["runs"]:
            continue
        rows = parse_rows(read_text(name, run))
        got = np.rint(rows[:, 4] * float(n) * float(n)).astype(np.int64).reshape(p.nbins, p.nmu)
        assert np.array_equal(got, expect), (run, got, expect)
        done += 1
    assert done >= 1
    if "D" in c["runs"]:
        assert (cnt[0] % 1 == 0).all() and cnt[0].sum() % 2 == 0  # ordered DD pairs come in twos


def test_2pcf_dr_split_is_consistent():
    name = "c1_lin"
    p = _params(MAN[name])
    xyz, _, nd, nr = _cat(name)
    res = O.pcf2(xyz, None, nd, p)
    rows = {run: parse_rows(read_text(name, run)) for run in ("D", "R", "T")}
    dd = np.rint(rows["D"][:, 4] * nd * nd)
    rr = np.rint(rows["R"][:, 4] * nr * nr)
    tt = np.rint(rows["T"][:, 4] * (nd + nr) ** 2)
    assert np.array_equal(tt - dd - rr, res["cnt"][1, :, 0].astype(float))


def _lattice_pairs_by_k(n_side, kmax):
    """Ordered pairs of an n_side^3 lattice at squared distance k, by enumeration of offsets."""
    out = np.zeros(kmax + 1, np.int64)
    r = int(math.isqrt(kmax)) + 1
    for a, b, c in itertools.product(range(-r, r + 1), repeat=3):
        k = a * a + b * b + c * c
        if 0 < k <= kmax:
            out[k] += (n_side - abs(a)) * (n_side - abs(b)) * (n_side - abs(c))
    return out


def test_lattice_closed_form():
    p = _params(MAN["lattice_lin"])
    xyz, _, nd, _ = _cat("lattice_lin")
    res = O.pcf2(xyz, None, nd, p)
    by_k = _lattice_pairs_by_k(12, 16)
    # rmin=1 < d < rmax=4, d = sqrt(k): bin edges at d = 2, 3 fall on k = 4, 9 exactly
    expect = np.zeros(3, np.int64)
    for k in range(2, 16):
        expect[int(math.floor(math.sqrt(k) - 1.0)) + 1 - 1] += by_k[k]
    assert np.array_equal(res["cnt"][0, :, 0].astype(np.int64), expect)


def test_total_pairs_match_cKDTree():
    from scipy.spatial import cKDTree
    name = "rmu_lin"
    p = _params(MAN[name])
    xyz, _, nd, _ = _cat(name)
    t = cKDTree(xyz)
    hi = t.count_neighbors(t, np.nextafter(p.rmax, 0))
    lo = t.count_neighbors(t, p.rmin)
    res = O.pcf2(xyz, None, nd, p)
    assert int(res["cnt"].sum()) == hi - lo


def test_2pcf_permutation_invariant():
    name = "rmu_log"
    p = _params(MAN[name])
    xyz, _, nd, nr = _cat(name)
    a = O.pcf2(xyz, None, nd, p)["cnt"]
    rng = np.random.default_rng(5)
    pd, pr = rng.permutation(nd), nd + rng.permutation(nr)
    b = O.pcf2(xyz[np.concatenate([pd, pr])], None, nd, p)["cnt"]
    assert np.array_equal(a, b)


def test_mu_bins_mirror():
    # theta(i->j) = -theta(j->i) exactly, so the ordered-pair mu histogram of DD is symmetric
    name = "rmu_lin"
    p = _params(MAN[name])
    xyz, _, nd, _ = _cat(name)
    c = O.pcf2(xyz[:nd], None, nd, p)["cnt"][0].astype(np.int64)
    assert np.array_equal(c, c[:, ::-1])


# ----------------------------------------------------------------------- 3PCF

@pytest.mark.parametrize("name", CASES_3)
def test_3pcf_matches_reference_output(name):
    c = MAN[name]
    p = _params(c)
    xyz, w, nd, nr = _cat(name)
    run = "DR" if "DR" in c["runs"] else "D"
    rows = parse_rows(read_text(name, run))
    res = O.pcf3(xyz, w, nd, p)
    nc, nm = O.n3cfg(p.nbins), p.nmu
    rsd = nm > 1
    assert rows.shape == ((nc * nm, 11) if rsd else (nc, 9))
    edges = O.radial_bins(p)
    cfgs = [(i, j, k) for i in range(p.nbins) for j in range(i, p.nbins) for k in range(j, p.nbins)]
    ecols = np.array([[edges[i], edges[i + 1], edges[j], edges[j + 1], edges[k], edges[k + 1]] for i, j, k in cfgs])
    ecols = np.repeat(ecols, nm, axis=0)
    assert close_e147(rows[:, :6], ecols).all()
    col = 6
    if rsd:   # mu columns (l-1)/nmu, l/nmu (gramsci.F90:532)
        mu = np.tile(np.arange(nm), nc)
        assert close_e147(rows[:, 6], mu / nm).all() and close_e147(rows[:, 7], (mu + 1) / nm).all()
        col = 8
    assert close_e147(rows[:, col], res["n2"].reshape(-1)).all()
    assert close_e147(rows[:, col + 1], res["n3"].reshape(-1)).all()
    cnt = res["cnt"].reshape(nc, nm, 4)
    assert (cnt.sum((1, 2)) % 3 == 0).all()     # each triangle once per apex


@pytest.mark.parametrize("name", [n for n in CASES_3 if not MAN[n]["weighted"]])
def test_3pcf_exact_counts_from_reference(name):
    c = MAN[name]
    p = _params(c)
    xyz, _, nd, nr = _cat(name)
    cnt = O.pcf3(xyz, None, nd, p)["cnt"].astype(np.int64).reshape(-1, 4)
    checks = {"D": (nd, cnt[:, 0]), "R": (nr, cnt[:, 3]), "T": (nd + nr, cnt.sum(1))}
    col = 8 if p.nmu > 1 else 6
    for run, (n, expect) in checks.items():
        if run not in c["runs"]:
            continue
        rows = parse_rows(read_text(name, run))
        got = np.rint(rows[:, col] * float(n) ** 3).astype(np.int64)
        assert np.array_equal(got, expect), (run, got, expect)


def test_3pcf_rsd_sums_to_isotropic():
    # the RSD split only redistributes the isotropic triplets over mu bins
    name = "p3_rsd"
    c = MAN[name]
    p = _params(c)
    xyz, _, nd, nr = _cat(name)
    a = O.pcf3(xyz, None, nd, p)["cnt"].sum(1)
    b = O.pcf3(xyz, None, nd, O.Params(p.rmin, p.rmax, p.nbins, 1, p.log_bins))["cnt"]
    assert np.array_equal(a, b)


def test_normal_bin_special_values():
    # find_normal(mu1, mu2): the radicand 0.75 - a^2 - b^2 + ab of the bin centres
    # a, b; at a = b = 0 (odd nmu, central bins) it is 0.75 and
    # 1.1547 * sqrt(0.75) = 0.99999... < 1, so the top bin nmu is reached;
    # a negative radicand (a = -b near +-1) lands in bin 1.
    for nmu in (3, 5, 7, 127):
        mid = (nmu + 1) // 2
        assert O.normal_bin(nmu, mid, mid) == nmu
        assert O.normal_bin(nmu, 1, nmu) == 1
        assert O.normal_bin(nmu, nmu, 1) == 1
    for nmu in range(2, 40):
        for a in range(1, nmu + 1):
            for b in range(1, nmu + 1):
                assert 1 <= O.normal_bin(nmu, a, b) <= nmu


def test_3pcf_permutation_invariant():
    name = "p3"
    p = _params(MAN[name])
    xyz, _, nd, nr = _cat(name)
    a = O.pcf3(xyz, None, nd, p)["cnt"]
    rng = np.random.default_rng(9)
    perm = np.concatenate([rng.permutation(nd), nd + rng.permutation(nr)])
    b = O.pcf3(xyz[perm], None, nd, p)["cnt"]
    assert np.array_equal(a, b)


def test_3pcf_bruteforce_tiny():
    # independent O(N^3) enumeration of triangles on a tiny catalogue, using only
    # the oracle's pair predicate: each triangle with all three sides kept is
    # counted once per apex, i.e. 3 times, in the config of its sorted sides.
    rng = np.random.default_rng(3)
    xyz = rng.uniform(0, 20, (70, 3))
    p = O.Params(2.0, 9.0, 3)
    n = len(xyz)
    b = np.zeros((n, n), int)
    for i in range(n):
        for j in range(n):
            if i != j:
                b[i, j] = O.pair_rbin(p, xyz[i], xyz[j])
    expect = np.zeros(O.n3cfg(3), np.int64)
    cfg = {}
    q = 0
    for i in range(1, 4):
        for j in range(i, 4):
            for k in range(j, 4):
                q += 1
                cfg[(i, j, k)] = q
    for i, j, k in itertools.combinations(range(n), 3):
        if b[i, j] and b[j, k] and b[i, k]:
            expect[cfg[tuple(sorted((b[i, j], b[j, k], b[i, k])))] - 1] += 3
    got = O.pcf3(xyz, None, n, p)["cnt"][:, 0].astype(np.int64)
    assert np.array_equal(got, expect)


# ----------------------------------------------------------------------- 4PCF

def _disc_products(dn, dr):
    with np.errstate(divide="ignore", invalid="ignore"):
        q = dn / dr
    return np.stack([q[0] * q[4], q[2] * q[3], q[1] * q[5]], axis=-1)


@pytest.mark.parametrize("name", CASES_4)
def test_4pcf_matches_reference_output(name):
    c = MAN[name]
    p = _params(c)
    xyz, w, nd, nr = _cat(name)
    rows = parse_rows(read_text(name, "DR"))
    res = O.pcf4(xyz, w, nd, p)
    nc, nb = O.n4cfg(p.nbins), p.nbins
    assert rows.shape == (nc * nb * nb, 14)
    n2 = res["n2"].reshape(-1)
    n3 = res["n3"].reshape(-1)
    assert close_e147(rows[:, 8], n2).all()
    assert close_e147(rows[:, 9], n3).all()
    disc = _disc_products(res["disc_n"].reshape(6, -1), res["disc_r"].reshape(6, -1))
    # the disc products of ratios amplify the reference's own FP64 summation order:
    # allow 2 units of the last printed digit
    ok = close_e147(rows[:, 11:14], disc, ulps=2.0)
    # 0 * inf: where a disc sum is exactly 0 in the oracle the reference holds a
    # summation residue of either sign, so its product prints as +-Inf or NaN
    ok |= np.isnan(disc) & ~np.isfinite(rows[:, 11:14])
    assert ok.all(), np.argwhere(~ok)[:5]


@pytest.mark.parametrize("name", [n for n in CASES_4 if not MAN[n]["weighted"]])
def test_4pcf_exact_counts_from_reference(name):
    # each reference run is matched by the oracle on the same catalogue: an apex
    # with exactly 3 neighbours is skipped (gramsci.F90:675), so an R-only run is
    # not the all-random pattern of the D+R run.
    c = MAN[name]
    p = _params(c)
    d, r, _, _ = load_input(name)
    cats = {"D": d, "R": r, "T": np.concatenate([d, r])}
    for run, cat in cats.items():
        if run not in c["runs"]:
            continue
        expect = O.pcf4(cat, None, len(cat), p)["cnt"][..., 0].astype(np.int64).reshape(-1)
        rows = parse_rows(read_text(name, run))
        got = np.rint(rows[:, 8] * float(len(cat)) ** 4).astype(np.int64)
        assert np.array_equal(got, expect), run
    # and the D+R pattern split adds up to the T run
    xyz, _, nd, nr = _cat(name)
    tot = O.pcf4(xyz, None, nd, p)["cnt"].sum(-1).reshape(-1)
    rows = parse_rows(read_text(name, "T"))
    assert np.array_equal(np.rint(rows[:, 8] * float(nd + nr) ** 4).astype(np.int64), tot)


def test_4pcf_three_neighbour_apex_skipped():
    # an apex with exactly three linked neighbours would form one quad, but the
    # reference skips apexes with nn <= 3 (gramsci.F90:675)
    xyz = np.array([[0.0, 0, 0], [3.0, 0, 0], [0, 3.0, 0], [0, 0, 3.0]])
    p = O.Params(1.0, 5.0, 2)
    assert O.pcf4(xyz, None, 4, p)["cnt"].sum() == 0
    xyz5 = np.vstack([xyz, [[2.0, 2.0, 2.0]]])
    assert O.pcf4(xyz5, None, 5, p)["cnt"].sum() > 0


def test_4pcf_order_dependence():
    # the perimeter rule depends on the id order of the neighbours: a permutation
    # changes the binned quads (a property of the reference kept on purpose)
    name = "p4"
    p = _params(MAN[name])
    xyz, _, nd, nr = _cat(name)
    a = O.pcf4(xyz, None, nd, p)["cnt"].sum()
    rng = np.random.default_rng(1)
    perm = np.concatenate([rng.permutation(nd), nd + rng.permutation(nr)])
    b = O.pcf4(xyz[perm], None, nd, p)["cnt"].sum()
    assert a != b


def test_invalid_params_rejected():
    xyz = np.zeros((4, 3))
    with pytest.raises(ValueError):
        O.pcf2(xyz, None, 4, O.Params(5.0, 1.0, 3))
    with pytest.raises(ValueError):
        O.pcf2(xyz, None, 4, O.Params(0.0, 1.0, 3, log_bins=True))
    with pytest.raises(ValueError):
        O.pcf2(xyz, None, 4, O.Params(0.0, 1.0, 200))


def ray_catalogue(seed=5):
    """Points on rays from the origin (pairs with theta = +-1 exactly on the axes) mixed with
    uniform points; at nmu = 127 the reference's int8 store of floor(t)+1 wraps for theta >= 1."""
    rng = np.random.default_rng(seed)
    k = np.arange(1, 161, dtype=float)
    rays = [np.stack([0.5 * k, 0 * k, 0 * k], 1), np.stack([0 * k, 0.75 * k, 0 * k], 1),
            np.stack([0 * k, 0 * k, -0.625 * k], 1), np.stack([0.5 * k, 0.5 * k, 0 * k], 1),
            np.stack([-0.25 * k, 0.25 * k, 0.25 * k], 1)]
    return np.concatenate(rays + [rng.uniform(-30.0, 30.0, (400, 3))])


@pytest.mark.skipif(not os.path.exists(O.REF_BIN), reason="reference not built (make -C oracle ref)")
def test_mu_int8_wrap_matches_reference():
    xyz = ray_catalogue()
    p = O.Params(0.5, 12.0, 4, 127, False)
    res = O.pcf2(xyz, None, len(xyz), p)
    rows = parse_rows(O.reference_table(xyz, None, None, None, p))
    got = np.rint(rows[:, 4] * float(len(xyz)) ** 2).astype(np.int64).reshape(p.nbins, p.nmu)
    assert np.array_equal(got, res["cnt"][0].astype(np.int64))
    # the wrap is exercised: theta = 1 pairs land in mu bin 1, not bin 127
    assert res["cnt"][0][:, 0].sum() > res["cnt"][0][:, -1].sum() > 0
