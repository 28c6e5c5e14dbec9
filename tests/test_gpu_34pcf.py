"""3PCF and 4PCF parity of the HIP path against the oracle and the
reference's output files.

Bar: triplet / quad counts per configuration and class bit-exact (both
precisions for the 3PCF; the 4PCF is evaluated in FP64); N2/N3 within 1e-12
relative of the oracle, with an absolute floor of 1e-12 * max|N| for cells
where a signed sum cancels; the 4PCF disc sums (FP64 atomic sums, whose order
varies) within 1e-12 relative with a floor of 1e-9 * max|disc|.
"""
import numpy as np
import pytest

import oracle as O
from golden_util import close_e147, load_input, manifest, parse_rows, read_text
from gramsci_amd import catalog as C
from gramsci_amd.engine import Engine

pytestmark = pytest.mark.gpu

MAN = manifest()
REL = 1e-12


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _cat(name):
    d, r, wd, wr = load_input(name)
    xyz = np.concatenate([d, r]) if len(r) else d
    w = None
    if wd is not None:
        w = np.concatenate([wd, wr if wr is not None else np.ones(len(r))])
    return xyz, w, len(d)


def _close(a, b, rel=REL):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    fin = np.isfinite(b)
    floor = rel * max(np.abs(b[fin]).max() if fin.any() else 0.0, 1e-300)
    return np.all(np.abs(a[fin] - b[fin]) <= np.maximum(rel * np.abs(b[fin]), floor))


def _close_disc(a, b):
    # the 4PCF disc sums are FP64 atomic sums (their order varies from run to run) of signed terms; a
    # cell whose terms cancel carries their rounding, about n * 2^-53 of their size, so the floor is
    # 1e-9 of the largest cell
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    fin = np.isfinite(b)
    floor = 1e-9 * max(np.abs(b[fin]).max() if fin.any() else 0.0, 1e-300)
    return np.all(np.abs(a[fin] - b[fin]) <= np.maximum(REL * np.abs(b[fin]), floor))


def check3(eng, xyz, w, nd, p, buf=None, precisions=("fp64", "fp32")):
    ref = O.pcf3(xyz, w, nd, p, buf=buf)
    outs = []
    for prec in precisions:
        eng.set_catalogue(xyz, w, nd, buf)
        r = eng.pcf3(p.rmin, p.rmax, p.nbins, nmu=p.nmu, log_bins=p.log_bins, precision=prec)
        print(f"  3pcf {prec}: units={r['stats']['n_units']} binned={r['stats']['n_binned']} "
              f"recheck={r['stats']['n_recheck']}", flush=True)
        assert np.array_equal(r["cnt"], ref["cnt"]), (prec, np.argwhere(r["cnt"] != ref["cnt"])[:5])
        assert r["stats"]["n_units"] == ref["n_cand"]
        # RSD visits every apex's pairs; the isotropic modes each triangle once (shorter lists)
        ev = r["stats"]["n_evaluated"]
        assert ev == ref["n_cand"] if p.nmu > 1 else ev <= ref["n_cand"]
        assert _close(r["n2"], ref["n2"]) and _close(r["n3"], ref["n3"])
        outs.append(r)
    return ref, outs


def check4(eng, xyz, w, nd, p, buf=None):
    ref = O.pcf4(xyz, w, nd, p, buf=buf)
    eng.set_catalogue(xyz, w, nd, buf)
    r = eng.pcf4(p.rmin, p.rmax, p.nbins, p.log_bins)
    print(f"  4pcf: units={r['stats']['n_units']} binned={r['stats']['n_binned']}", flush=True)
    assert np.array_equal(r["cnt"], ref["cnt"])
    assert r["stats"]["n_units"] == ref["n_quads"]
    assert _close(r["n2"], ref["n2"]) and _close(r["n3"], ref["n3"])
    assert _close_disc(r["disc_n"], ref["disc_n"]) and _close_disc(r["disc_r"], ref["disc_r"])
    return ref, r


@pytest.mark.parametrize("name", [n for n, c in MAN.items() if c["mode"] == "-3pcf"])
def test_3pcf_golden(eng, name):
    c = MAN[name]
    p = O.Params(**c["params"])
    xyz, w, nd = _cat(name)
    print(f"case {name}", flush=True)
    ref, outs = check3(eng, xyz, w, nd, p)
    rows = parse_rows(read_text(name, "DR" if "DR" in c["runs"] else "D"))
    col = 8 if p.nmu > 1 else 6      # RSD rows carry two mu columns (gramsci.F90:531-532)
    for r in outs:
        assert close_e147(rows[:, col], r["n2"].reshape(-1)).all()
        assert close_e147(rows[:, col + 1], r["n3"].reshape(-1)).all()


@pytest.mark.parametrize("name", [n for n, c in MAN.items() if c["mode"] == "-4pcf"])
def test_4pcf_golden(eng, name):
    c = MAN[name]
    p = O.Params(**c["params"])
    xyz, w, nd = _cat(name)
    print(f"case {name}", flush=True)
    ref, r = check4(eng, xyz, w, nd, p)
    rows = parse_rows(read_text(name, "DR"))
    assert close_e147(rows[:, 8], r["n2"].reshape(-1)).all()
    assert close_e147(rows[:, 9], r["n3"].reshape(-1)).all()


def test_3pcf_exact_counts_vs_reference_runs(eng):
    name = "p3"
    p = O.Params(**MAN[name]["params"])
    d, r, _, _ = load_input(name)
    for run, cat in (("D", d), ("R", r), ("T", np.concatenate([d, r]))):
        eng.set_catalogue(cat, None, len(cat))
        out = eng.pcf3(p.rmin, p.rmax, p.nbins, precision="fp32")
        rows = parse_rows(read_text(name, run))
        got = np.rint(rows[:, 6] * float(len(cat)) ** 3).astype(np.int64)
        assert np.array_equal(out["cnt"][:, 0].astype(np.int64), got), run


def test_4pcf_exact_counts_vs_reference_runs(eng):
    name = "p4"
    p = O.Params(**MAN[name]["params"])
    d, r, _, _ = load_input(name)
    for run, cat in (("D", d), ("R", r), ("T", np.concatenate([d, r]))):
        eng.set_catalogue(cat, None, len(cat))
        out = eng.pcf4(p.rmin, p.rmax, p.nbins)
        rows = parse_rows(read_text(name, run))
        got = np.rint(rows[:, 8] * float(len(cat)) ** 4).astype(np.int64)
        assert np.array_equal(out["cnt"][..., 0].reshape(-1).astype(np.int64), got), run


def test_3pcf_clustered_log_weighted_buffer(eng):
    rng = np.random.default_rng(7)
    d = C.clustered(6000, 120.0, 71)
    r = C.uniform(12000, 120.0, 72)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(6000, 0.2, 73), C.lognormal_weights(12000, 0.1, 74)])
    buf = (rng.random(len(xyz)) < 0.25).astype(np.int32)
    check3(eng, xyz, w, 6000, O.Params(2.0, 14.0, 6, 1, True), buf=buf)
    check3(eng, xyz, None, 6000, O.Params(3.0, 12.0, 4, 1, False))


@pytest.mark.parametrize("sx", ["2", "8"])
def test_apex_x_cell_subdivisions(eng, monkeypatch, sx):
    # the triplet and quad gathers on a grid with x cells of h / sx (GS_T3_CELL_SX): counts, sums and
    # N2/N3 match the oracle (RSD and weighted 3PCF with buffer points, 4PCF with weights and buffers)
    monkeypatch.setenv("GS_T3_CELL_SX", sx)
    rng = np.random.default_rng(11)
    d = C.clustered(5000, 110.0, 111)
    r = C.uniform(10000, 110.0, 112)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(5000, 0.2, 113), np.ones(10000)])
    buf = (rng.random(len(xyz)) < 0.2).astype(np.int32)
    print(f"sx={sx}", flush=True)
    check3(eng, xyz, w, 5000, O.Params(2.0, 13.0, 5, 3, True), buf=buf)
    check3(eng, xyz, None, 5000, O.Params(3.0, 12.0, 4, 1, False))
    check4(eng, xyz[:6000], w[:6000], 5000, O.Params(3.0, 8.0, 3, 1, False), buf=buf[:6000])


@pytest.mark.parametrize("nmu", [1, 4])
def test_3pcf_two_pass_overflow(eng, monkeypatch, nmu):
    # a first-pass neighbour list of 128 entries (about the mean neighbour count here): the
    # apexes with more neighbours go to the second pass, and every apex is counted once
    rng = np.random.default_rng(11)
    d = C.clustered(6000, 120.0, 75)
    r = C.uniform(12000, 120.0, 76)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(6000, 0.2, 77), np.ones(12000)])
    buf = (rng.random(len(xyz)) < 0.1).astype(np.int32)
    p = O.Params(2.0, 14.0, 5, nmu, False)
    eng.set_catalogue(xyz, w, 6000, buf)
    base = eng.pcf3(p.rmin, p.rmax, p.nbins, nmu=nmu, precision="fp32")
    monkeypatch.setenv("GS_T3_KMAX", "128")
    ref, outs = check3(eng, xyz, w, 6000, p, buf=buf)
    for o in outs:
        assert np.array_equal(o["cnt"], base["cnt"]) and np.array_equal(o["wsum"], base["wsum"])
        assert o["stats"]["n_units"] == base["stats"]["n_units"]
        # every apex's candidates and overflow pairs are counted in exactly one pass
        assert o["stats"]["n_candidates"] == base["stats"]["n_candidates"]
        assert o["stats"]["n_overflow"] == base["stats"]["n_overflow"]


@pytest.mark.parametrize("wg", ["64", "128", "256"])
@pytest.mark.parametrize("nmu,weighted", [(1, True), (3, False)])
def test_3pcf_work_group_sizes(eng, monkeypatch, wg, nmu, weighted):
    # one apex per work-group of 64, 128 or 256 threads (256 unless GS_T3_WG says otherwise):
    # counts equal the oracle's, sums and statistics equal across sizes
    rng = np.random.default_rng(13)
    d = C.clustered(3000, 90.0, 85)
    r = C.uniform(6000, 90.0, 86)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(3000, 0.2, 87), np.ones(6000)]) if weighted else None
    buf = (rng.random(len(xyz)) < 0.1).astype(np.int32)
    p = O.Params(2.0, 12.0, 4, nmu, False)
    monkeypatch.setenv("GS_T3_WG", "256")
    eng.set_catalogue(xyz, w, 3000, buf)
    base = eng.pcf3(p.rmin, p.rmax, p.nbins, nmu=nmu, precision="fp32")
    monkeypatch.setenv("GS_T3_WG", wg)
    ref, outs = check3(eng, xyz, w, 3000, p, buf=buf)
    for o in outs:
        assert np.array_equal(o["cnt"], base["cnt"]) and np.array_equal(o["wsum"], base["wsum"])
        assert o["stats"]["n_candidates"] == base["stats"]["n_candidates"]


def test_3pcf_config3_full_size_properties(eng):
    # the 3PCF configuration at its real size (200k Neyman-Scott galaxies + 400k uniform randoms in
    # 500^3, 10 < r < 30, 5 bins; ~7e10 apex-pair candidates): FP32 + guard counts equal FP64's at two
    # cell sizes, and every (configuration, class) count is a multiple of 3 (each triangle is counted
    # once from each of its three apexes)
    d = C.neyman_scott(100000, 5000, 20, 2.0, 500.0, 301)
    r = C.uniform(400000, 500.0, 302)
    xyz = np.concatenate([d, r])
    eng.set_catalogue(xyz, None, len(d))
    a = eng.pcf3(10.0, 30.0, 5, precision="fp32")
    b = eng.pcf3(10.0, 30.0, 5, precision="fp64")
    c = eng.pcf3(10.0, 30.0, 5, precision="fp32", cell_div=2)
    print("C3 full size: units", a["stats"]["n_units"], "binned", a["stats"]["n_binned"], flush=True)
    assert np.array_equal(a["cnt"], b["cnt"]) and np.array_equal(a["cnt"], c["cnt"])
    assert a["stats"]["n_units"] == b["stats"]["n_units"] == c["stats"]["n_units"]
    assert np.all(a["cnt"] % 3 == 0)
    assert a["stats"]["n_units"] > 5e10


def test_4pcf_config3_full_size_properties(eng, monkeypatch):
    # the 4PCF configuration at its real size (a seeded 50k-galaxy + 100k-random subset of the 3PCF
    # catalogue, 10 < r < 20, 3 bins): counts and statistics equal at two cell sizes and with a
    # forced 64-entry first-pass neighbour list (most apexes then go to the second pass)
    d = C.neyman_scott(100000, 5000, 20, 2.0, 500.0, 301)
    r = C.uniform(400000, 500.0, 302)
    rng = np.random.default_rng(303)
    d = d[np.sort(rng.choice(len(d), 50000, replace=False))]
    r = r[np.sort(rng.choice(len(r), 100000, replace=False))]
    xyz = np.concatenate([d, r])
    eng.set_catalogue(xyz, None, len(d))
    a = eng.pcf4(10.0, 20.0, 3)
    b = eng.pcf4(10.0, 20.0, 3, cell_div=2)
    monkeypatch.setenv("GS_T4_KMAX", "64")
    c = eng.pcf4(10.0, 20.0, 3)
    print("4PCF full size: units", a["stats"]["n_units"], "binned", a["stats"]["n_binned"], flush=True)
    for o in (b, c):
        assert np.array_equal(o["cnt"], a["cnt"])
        assert o["stats"]["n_units"] == a["stats"]["n_units"] and o["stats"]["n_binned"] == a["stats"]["n_binned"]
    assert a["stats"]["n_binned"] > 0
    # and the oracle on the same 150k points (about 15 s of one CPU core): every quad count per
    # (configuration, side bins, random pattern), the evaluated quads, N2 and N3
    ref = O.pcf4(xyz, None, len(d), O.Params(10.0, 20.0, 3, 1, False))
    assert np.array_equal(a["cnt"], ref["cnt"])
    assert a["stats"]["n_units"] == ref["n_quads"]
    assert _close(a["n2"], ref["n2"]) and _close(a["n3"], ref["n3"])


@pytest.mark.parametrize("batch", ["7", "16"])
@pytest.mark.parametrize("kmax1", ["", "64"])
def test_apex_batches(eng, monkeypatch, batch, kmax1):
    # apexes fetched from the work queue 7 or 16 at a time (the automatic size is 1 for catalogues
    # this small), alone or with a 64-entry first-pass neighbour list, so that apexes in the middle
    # of a batch go to the second pass: RSD weighted 3PCF, equilateral 3PCF and 4PCF give the
    # counts, sums and statistics of single-apex fetches in one pass, and the 3PCF and 4PCF counts
    # equal the oracle's
    rng = np.random.default_rng(17)
    d = C.clustered(4000, 100.0, 91)
    r = C.uniform(8000, 100.0, 92)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(4000, 0.2, 93), np.ones(8000)])
    buf = (rng.random(len(xyz)) < 0.1).astype(np.int32)
    eng.set_catalogue(xyz, w, 4000, buf)
    monkeypatch.setenv("GS_APEX_BATCH", "1")
    monkeypatch.setenv("GS_T3_KMAX", "")
    monkeypatch.setenv("GS_T4_KMAX", "")
    base3 = eng.pcf3(2.0, 12.0, 4, nmu=3, precision="fp32")
    base_e = eng.pcf3_equi(2.0, 12.0, 4, precision="fp32")
    base4 = eng.pcf4(2.0, 8.0, 3)
    monkeypatch.setenv("GS_APEX_BATCH", batch)
    monkeypatch.setenv("GS_T3_KMAX", kmax1)
    monkeypatch.setenv("GS_T4_KMAX", kmax1)
    o3 = eng.pcf3(2.0, 12.0, 4, nmu=3, precision="fp32")
    oe = eng.pcf3_equi(2.0, 12.0, 4, precision="fp32")
    o4 = eng.pcf4(2.0, 8.0, 3)
    if kmax1:
        # mean neighbour counts here: ~87 within r < 12 (most 3PCF apexes overflow 64 entries),
        # ~26 within r < 8 (the 4PCF's clustered apexes overflow)
        ref4 = O.pcf4(xyz, w, 4000, O.Params(2.0, 8.0, 3, 1, False), buf=buf)
        assert np.array_equal(o4["cnt"], ref4["cnt"]) and o4["stats"]["n_units"] == ref4["n_quads"]
    for o, b in ((o3, base3), (oe, base_e), (o4, base4)):
        assert np.array_equal(o["cnt"], b["cnt"])
        assert o["stats"]["n_units"] == b["stats"]["n_units"]
        assert o["stats"]["n_candidates"] == b["stats"]["n_candidates"]
    # 3PCF weighted sums are exact fixed point; the 4PCF's weighted products are FP64 sums
    assert np.array_equal(o3["wsum"], base3["wsum"]) and _close(o4["n2"], base4["n2"])
    ref = O.pcf3(xyz, w, 4000, O.Params(2.0, 12.0, 4, 3, False), buf=buf)
    assert np.array_equal(o3["cnt"], ref["cnt"])


def test_3pcf_config3_density_slab(eng):
    # a 150^3 slab at the density of the 3PCF configuration (600k points in 500^3)
    d = C.clustered(5400, 150.0, 81)
    r = C.uniform(10800, 150.0, 82)
    xyz = np.concatenate([d, r])
    ref, outs = check3(eng, xyz, None, 5400, O.Params(10.0, 30.0, 5, 1, False))
    assert ref["n_cand"] > 1e8


def test_4pcf_weighted_buffer_and_order(eng):
    rng = np.random.default_rng(9)
    d = C.clustered(3000, 70.0, 91)
    r = C.uniform(3000, 70.0, 92)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(3000, 0.2, 93), np.ones(3000)])
    buf = (rng.random(len(xyz)) < 0.2).astype(np.int32)
    check4(eng, xyz, w, 3000, O.Params(3.0, 9.0, 3, 1, False), buf=buf)
    # the result depends on the input order (ascending original index), as in the reference
    perm = np.concatenate([rng.permutation(3000), 3000 + rng.permutation(3000)])
    check4(eng, xyz[perm], None, 3000, O.Params(3.0, 9.0, 3, 1, False))


@pytest.mark.parametrize("tetra", [False, True])
def test_4pcf_two_pass_overflow(eng, monkeypatch, tetra):
    # a 64-entry first-pass list: apexes with more neighbours (about half here) are counted by
    # the second pass; counts equal the oracle's and the single-pass run's
    d = C.clustered(3000, 60.0, 97)
    r = C.uniform(3000, 60.0, 98)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(3000, 0.2, 99), np.ones(3000)])
    p = O.Params(2.0, 7.0, 3, 1, False)
    eng.set_catalogue(xyz, w, 3000)
    fn = eng.pcf4_tetra if tetra else eng.pcf4
    base = fn(p.rmin, p.rmax, p.nbins, p.log_bins)
    monkeypatch.setenv("GS_T4_KMAX", "64")
    two = fn(p.rmin, p.rmax, p.nbins, p.log_bins)
    assert np.array_equal(two["cnt"], base["cnt"])
    assert two["stats"]["n_units"] == base["stats"]["n_units"]
    assert two["stats"]["n_candidates"] == base["stats"]["n_candidates"]
    assert _close(two["n2"], base["n2"]) and _close(two["n3"], base["n3"])
    if not tetra:
        ref = O.pcf4(xyz, w, 3000, p)
        assert np.array_equal(two["cnt"], ref["cnt"]) and two["stats"]["n_units"] == ref["n_quads"]
        assert _close_disc(two["disc_n"], ref["disc_n"]) and _close_disc(two["disc_r"], ref["disc_r"])


def test_4pcf_lattice_and_log(eng):
    lat = C.lattice(9, 1.0, 0.0)
    check4(eng, lat, None, len(lat), O.Params(0.5, 2.5, 2, 1, False))
    d = C.clustered(2500, 50.0, 95)
    check4(eng, d, None, len(d), O.Params(1.0, 8.0, 4, 1, True))


def test_4pcf_rejects_mu_bins_and_3pcf_large_neighbour_sets(eng):
    from gramsci_amd._lib import GsError
    d = C.uniform(2000, 50.0, 1)
    eng.set_catalogue(d, None, len(d))
    p = Engine.params(1.0, 5.0, 3, nmu=4)
    import ctypes
    from gramsci_amd import _lib as L
    # the 4PCF has no mu binning; rejected before any output is written
    rc = L.lib().gs_count_4pcf(eng._ctx, ctypes.byref(p), None, None, None, None, None, None)
    assert rc == L.GS_EINVAL
    rc = L.lib().gs_count_4pcf_tetra(eng._ctx, ctypes.byref(p), None, None, None, None, None, None)
    assert rc == L.GS_EINVAL
    del GsError


@pytest.mark.parametrize("prec,nmu,weighted,equi", [("fp32", 1, True, False), ("fp64", 1, False, False),
                                                    ("fp32", 3, True, False), ("fp32", 1, True, True)])
def test_3pcf_device_memory_neighbour_lists(eng, monkeypatch, prec, nmu, weighted, equi):
    # every LDS list cut to 64 entries (about 90 neighbours per apex here): the apexes with more
    # neighbours run in the last pass with their lists in device memory; counts equal the oracle's,
    # and counts, sums and statistics equal those of the normal two-pass run
    rng = np.random.default_rng(31)
    d = C.clustered(3000, 90.0, 85)
    r = C.uniform(6000, 90.0, 86)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(3000, 0.2, 87), np.ones(6000)]) if weighted else None
    buf = (rng.random(len(xyz)) < 0.1).astype(np.int32)
    p = O.Params(2.0, 12.0, 4, nmu, False)
    eng.set_catalogue(xyz, w, 3000, buf)
    run = (lambda: eng.pcf3_equi(p.rmin, p.rmax, p.nbins, nmu=nmu, precision=prec)) if equi else \
        (lambda: eng.pcf3(p.rmin, p.rmax, p.nbins, nmu=nmu, precision=prec))
    base = run()
    monkeypatch.setenv("GS_T3_KMAX_LDS", "64")
    out = run()
    assert np.array_equal(out["cnt"], base["cnt"]) and np.array_equal(out["wsum"], base["wsum"])
    for k in ("n_units", "n_binned", "n_candidates", "n_overflow"):
        assert out["stats"][k] == base["stats"][k], k
    ref = (O.pcf3_equi if equi else O.pcf3)(xyz, w, 3000, p, buf=buf)
    assert np.array_equal(out["cnt"].reshape(ref["cnt"].shape), ref["cnt"])
    assert _close(out["n2"].reshape(-1), ref["n2"].reshape(-1))


def _esym(w, k):
    """Elementary symmetric polynomial e_k of positive weights (k <= 3), from power sums."""
    p1, p2, p3 = (np.sum(w ** j) for j in (1, 2, 3))
    return [1.0, p1, (p1 * p1 - p2) / 2.0, (p1 ** 3 - 3.0 * p1 * p2 + 2.0 * p3) / 6.0][k]


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_3pcf_more_than_4096_neighbours(eng, prec):
    # 4300 points in a 10^3 box with rmax = 20 > its diagonal: every point is every other's
    # neighbour (k = 4299, more than any LDS list holds).  With one bin, every triangle is counted
    # once from each apex: class c (c randoms) holds 3 C(nd, 3-c) C(nr, c) triangles and the weight
    # sum 3 e_{3-c}(w_data) e_c(w_rand)
    from math import comb
    nd, nr = 2000, 2300
    xyz = C.uniform(nd + nr, 10.0, 41)
    w = C.lognormal_weights(nd + nr, 0.2, 42)
    eng.set_catalogue(xyz, w, nd)
    out = eng.pcf3(1e-3, 20.0, 1, precision=prec)
    cnt = out["cnt"].reshape(-1)
    assert [int(x) for x in cnt] == [3 * comb(nd, 3 - c) * comb(nr, c) for c in range(4)]
    ws = out["wsum"].reshape(-1)
    for c in range(4):
        exp = 3.0 * _esym(w[:nd], 3 - c) * _esym(w[nd:], c)
        assert abs(ws[c] - exp) <= 1e-12 * exp, (c, ws[c], exp)
    assert out["stats"]["n_units"] == (nd + nr) * comb(nd + nr - 1, 2)
    # with mu bins, the RSD 3PCF summed over mu is the isotropic count
    rsd = eng.pcf3(1e-3, 20.0, 1, nmu=4, precision=prec)
    assert np.array_equal(rsd["cnt"].sum(axis=1).reshape(-1), cnt)


def test_3pcf_rsd_exact_counts_vs_reference_runs(eng):
    name = "p3_rsd"
    p = O.Params(**MAN[name]["params"])
    d, r, _, _ = load_input(name)
    for run, cat in (("D", d), ("R", r), ("T", np.concatenate([d, r]))):
        eng.set_catalogue(cat, None, len(cat))
        out = eng.pcf3(p.rmin, p.rmax, p.nbins, nmu=p.nmu, precision="fp32")
        rows = parse_rows(read_text(name, run))
        got = np.rint(rows[:, 8] * float(len(cat)) ** 3).astype(np.int64)
        assert np.array_equal(out["cnt"][..., 0].reshape(-1).astype(np.int64), got), run


@pytest.mark.parametrize("weighted", [False, True])
def test_3pcf_rsd_global_histogram(eng, weighted):
    # 12 bins x 10 mu bins: the histogram does not fit in LDS and goes to global atomics
    xyz = C.clustered(6000, 120.0, 77, members=10, sigma=2.0)
    nd = 3000
    w = C.lognormal_weights(len(xyz), 0.3, 78) if weighted else None
    p = O.Params(3.0, 15.0, 12, 10)
    print(f"case rsd global hist weighted={weighted}", flush=True)
    check3(eng, xyz, w, nd, p)


@pytest.mark.parametrize("tetra,weighted", [(False, True), (False, False), (True, True)])
def test_4pcf_device_memory_neighbour_lists(eng, monkeypatch, tetra, weighted):
    # every LDS list cut to 64 entries: the apexes with more neighbours run in the last pass with
    # their lists in device memory and the neighbour-neighbour bins evaluated on the fly (no link
    # matrix); counts equal the oracle's and the normal run's, sums within 1e-12
    rng = np.random.default_rng(33)
    d = C.clustered(3000, 70.0, 91)
    r = C.uniform(3000, 70.0, 92)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(3000, 0.2, 93), np.ones(3000)]) if weighted else None
    buf = (rng.random(len(xyz)) < 0.1).astype(np.int32)
    p = O.Params(3.0, 10.0, 3 if not tetra else 2, 1, False)
    eng.set_catalogue(xyz, w, 3000, buf)
    run = (lambda: eng.pcf4_tetra(p.rmin, p.rmax, p.nbins)) if tetra else (lambda: eng.pcf4(p.rmin, p.rmax, p.nbins))
    base = run()
    monkeypatch.setenv("GS_T4_KMAX_LDS", "64")
    monkeypatch.setenv("GS_T4_KMAX", "64")
    out = run()
    assert np.array_equal(out["cnt"], base["cnt"])
    for k in ("n_units", "n_binned", "n_candidates", "n_overflow"):
        assert out["stats"][k] == base["stats"][k], k
    assert _close(out["n2"], base["n2"]) and _close_disc(out["disc_n"], base["disc_n"])
    ref = (O.pcf4_tetra if tetra else O.pcf4)(xyz, w, 3000, p, buf=buf)
    assert np.array_equal(out["cnt"], ref["cnt"])
    assert _close(out["n2"], ref["n2"]) and _close(out["n3"], ref["n3"])
    assert _close_disc(out["disc_n"], ref["disc_n"]) and _close_disc(out["disc_r"], ref["disc_r"])


@pytest.mark.parametrize("tetra", [False, True])
def test_4pcf_more_than_256_neighbours(eng, tetra):
    # 300 points in a 10^3 box with rmax = 20 > its diagonal: k = 299 for every apex, more than an
    # LDS list holds.  With one bin every quad is counted from each apex, and the random pattern
    # follows from the id order (randoms after data, so among an apex's three neighbours in id
    # order the randoms come last): pattern = r_i | (c >= 3) << 1 | (c >= 2) << 2 | (c >= 1) << 3
    # for c randoms among them
    from math import comb
    nd, nr = 150, 150
    xyz = C.uniform(nd + nr, 10.0, 51)
    eng.set_catalogue(xyz, None, nd)
    out = eng.pcf4_tetra(1e-3, 20.0, 1) if tetra else eng.pcf4(1e-3, 20.0, 1)
    exp = np.zeros(16, np.int64)
    for ri, n_i in ((0, nd), (1, nr)):
        for c in range(4):
            pat = ri | (int(c >= 3) << 1) | (int(c >= 2) << 2) | (int(c >= 1) << 3)
            exp[pat] += n_i * comb(nd - (1 - ri), 3 - c) * comb(nr - ri, c)
    assert np.array_equal(out["cnt"].reshape(-1, 16).sum(0).astype(np.int64), exp)
    assert int(out["cnt"].sum()) == (nd + nr) * comb(nd + nr - 1, 3)
