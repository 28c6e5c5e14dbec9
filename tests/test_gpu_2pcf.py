"""2PCF parity of the HIP path (through the C ABI) against the oracle and the
reference's own output files.

Bar: integer counts bit-exact for every bin and class, in both precisions
(FP64 and FP32 + FP64 guard band); N2 and N3 within 1e-12 relative of the
oracle (absolute floor 1e-12 * max|N|, the cancellation floor of a signed sum);
the printed reference columns reproduced to their last digit.
"""
import numpy as np
import pytest

import oracle as O
from golden_util import close_e147, load_input, manifest, parse_rows, read_text
from gramsci_amd import catalog as C
from gramsci_amd.engine import Engine

pytestmark = pytest.mark.gpu

MAN = manifest()
CASES_2 = [n for n, c in MAN.items() if c["mode"] == "-2pcf"]
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
    floor = rel * max(np.abs(b).max(), 1e-300)
    return np.all(np.abs(a - b) <= np.maximum(rel * np.abs(b), floor))


def _check(eng, xyz, w, nd, p, buf=None, precisions=("fp64", "fp32"), cell_divs=(0,)):
    ref = O.pcf2(xyz, w, nd, p, buf=buf)
    out = []
    for prec in precisions:
        for cd in cell_divs:
            eng.set_catalogue(xyz, w, nd, buf)
            r = eng.pcf2(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, prec, cd)
            print(f"  {prec} cell_div={cd}: binned={r['stats']['n_binned']} recheck={r['stats']['n_recheck']} "
                  f"cand={r['stats']['n_candidates']}", flush=True)
            assert np.array_equal(r["cnt"], ref["cnt"]), (prec, cd, np.argwhere(r["cnt"] != ref["cnt"])[:5])
            assert r["stats"]["n_overflow"] == ref["overflow"]
            if w is not None:
                assert _close(r["wsum"], ref["wsum"], 1e-14)
            assert _close(r["n2"], ref["n2"]), (prec, np.max(np.abs(r["n2"] - ref["n2"])))
            assert _close(r["n3"], ref["n3"])
            out.append(r)
    return ref, out


@pytest.mark.parametrize("name", CASES_2)
def test_golden_fixtures(eng, name):
    c = MAN[name]
    p = O.Params(**c["params"])
    xyz, w, nd = _cat(name)
    print(f"case {name}", flush=True)
    ref, outs = _check(eng, xyz, w, nd, p)
    run = "DR" if "DR" in c["runs"] else "D"
    rows = parse_rows(read_text(name, run))
    for r in outs:
        assert close_e147(rows[:, 4], r["n2"].ravel()).all()
        assert close_e147(rows[:, 5], r["n3"].ravel()).all()


@pytest.mark.parametrize("name", [n for n in CASES_2 if "D" in MAN[n]["runs"]])
def test_exact_counts_vs_reference_runs(eng, name):
    # D-only reference run: N2 * Nd^2 is the exact ordered DD count per bin
    c = MAN[name]
    p = O.Params(**c["params"])
    d, _, _, _ = load_input(name)
    eng.set_catalogue(d, None, len(d))
    rows = parse_rows(read_text(name, "D"))
    dd = np.rint(rows[:, 4] * float(len(d)) ** 2).astype(np.int64)
    for prec in ("fp64", "fp32"):
        r = eng.pcf2(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, prec)
        assert np.array_equal(r["cnt"][0].ravel().astype(np.int64), dd)


def test_config1_full_size(eng):
    # the plain CPU configuration at its real size: 20k + 40k uniform in [0,500)^3
    d = C.uniform(20000, 500.0, 101)
    r = C.uniform(40000, 500.0, 102)
    xyz = np.concatenate([d, r])
    p = O.Params(10.0, 50.0, 10, 1, False)
    ref, outs = _check(eng, xyz, None, 20000, p, cell_divs=(0, 1, 3))
    assert int(ref["cnt"].sum()) > 1e7


def test_config2_full_size_properties(eng):
    # the (r, mu) configuration at its real size (1M Neyman-Scott galaxies with lognormal weights +
    # 3M uniform randoms in 1000^3, 1 < r < 150, 30 log bins, nmu = 20), where the oracle would take
    # days: the FP32 + guard path gives the FP64 path's counts and fixed-point sums bit for bit, at two
    # cell sizes, and every class's per-r-bin total is even (each pair is counted in both orientations
    # with the same exact distance)
    d = C.neyman_scott(500_000, 25_000, 20, 2.0, 1000.0, 201)
    r = C.uniform(3_000_000, 1000.0, 202)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(len(d), 0.2, 203), np.ones(len(r))])
    eng.set_catalogue(xyz, w, len(d))
    args = (1.0, 150.0, 30, 20, True)
    f32 = eng.pcf2(*args, "fp32")
    f64 = eng.pcf2(*args, "fp64")
    f32m = eng.pcf2(*args, "fp32", cell_div=5)
    print("C2 full size: binned", int(f32["cnt"].sum()), "rechecks", f32["stats"]["n_recheck"], flush=True)
    assert np.array_equal(f32["cnt"], f64["cnt"]) and np.array_equal(f32["wsum"], f64["wsum"])
    assert np.array_equal(f32m["cnt"], f32["cnt"]) and np.array_equal(f32m["wsum"], f32["wsum"])
    assert np.all(f32["cnt"].sum(axis=2) % 2 == 0)
    assert int(f32["cnt"].sum()) > 1e11


def test_rmu_log_weighted_clustered(eng):
    d = C.clustered(30000, 300.0, 7)
    r = C.uniform(60000, 300.0, 8)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(30000, 0.2, 9), np.ones(60000)])
    p = O.Params(1.0, 40.0, 12, 10, True)
    _check(eng, xyz, w, 30000, p, cell_divs=(0, 2))


def test_buffer_points(eng):
    # buffer points are neighbours but never apexes (gramsci.F90:363)
    rng = np.random.default_rng(4)
    d = C.clustered(8000, 150.0, 44)
    r = C.uniform(16000, 150.0, 45)
    xyz = np.concatenate([d, r])
    buf = (rng.random(len(xyz)) < 0.3).astype(np.int32)
    p = O.Params(2.0, 25.0, 6, 4, False)
    _check(eng, xyz, None, 8000, p, buf=buf)


@pytest.mark.parametrize("sx", [1, 2, 16, 64])
def test_x_cell_subdivisions(eng, monkeypatch, sx):
    # the grid's x cells are h / sx (GS_CELL_SX): counts, weighted sums and N2/N3 match the oracle for
    # every subdivision (weighted, clustered, with buffer points, a lattice with exact edge distances),
    # and finer x cells never examine more candidate pairs
    monkeypatch.setenv("GS_CELL_SX", str(sx))
    rng = np.random.default_rng(46)
    d = C.clustered(6000, 120.0, 47)
    r = C.uniform(12000, 120.0, 48)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(6000, 0.3, 49), np.ones(12000)])
    buf = (rng.random(len(xyz)) < 0.2).astype(np.int32)
    p = O.Params(1.0, 20.0, 8, 5, True)
    print(f"sx={sx}", flush=True)
    _, outs = _check(eng, xyz, w, 6000, p, buf=buf, cell_divs=(0, 3))
    monkeypatch.setenv("GS_CELL_SX", "1")
    eng.set_catalogue(xyz, w, 6000, buf)
    base = eng.pcf2(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, "fp32")
    assert outs[2]["stats"]["n_candidates"] <= base["stats"]["n_candidates"]
    monkeypatch.setenv("GS_CELL_SX", str(sx))
    _check(eng, C.lattice(11, 1.0, -5.0), None, 11 ** 3, O.Params(1.0, 4.0, 3, 4, False))


@pytest.mark.parametrize("dual", ["0", "1"])
def test_two_apexes_per_lane(eng, monkeypatch, dual):
    # the weighted FP32 kernel with and without its two-apexes-per-lane schedule (GS_PAIR_DUAL forces
    # it; unset, it runs where it costs no resident waves): counts, N2/N3 and weighted sums match the
    # oracle, with buffer points, ragged rows, (r, mu) log bins and plain linear bins from r = 0, and the
    # FP32 sums equal the FP64 path's bit for bit
    monkeypatch.setenv("GS_PAIR_DUAL", dual)
    rng = np.random.default_rng(61)
    d = C.clustered(7000, 130.0, 62)
    r = C.uniform(14000, 130.0, 63)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(7000, 0.3, 64), np.ones(14000)])
    buf = (rng.random(len(xyz)) < 0.15).astype(np.int32)
    print(f"dual={dual}", flush=True)
    for p, b, cds in ((O.Params(1.0, 25.0, 9, 6, True), buf, (0, 2)),
                      (O.Params(1.0, 25.0, 9, 6, True), None, (0, 3)),
                      (O.Params(0.0, 20.0, 7, 1, False), buf, (0, 3))):
        _, outs = _check(eng, xyz, w, 7000, p, buf=b, cell_divs=cds)
        assert np.array_equal(outs[0]["wsum"], outs[len(cds)]["wsum"])


def test_lattice_edges_and_origin(eng):
    # exact-edge distances and mu values on a lattice that contains the origin
    # (mid = 0 pairs give theta = NaN in the reference)
    lat = C.lattice(11, 1.0, -5.0)
    for p in (O.Params(1.0, 4.0, 3, 4, False), O.Params(1.0, 8.0, 3, 6, True), O.Params(0.0, 3.0, 3, 2, False)):
        _check(eng, lat, None, len(lat), p)


def test_duplicates_and_tiny_catalogues(eng):
    d = C.uniform(500, 30.0, 3)
    d = np.concatenate([d, d[:50], d[:5]])
    _check(eng, d, None, len(d), O.Params(0.0, 6.0, 4, 3, False))
    two = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])
    _check(eng, two, None, 1, O.Params(0.5, 2.0, 2, 2, False))
    one = np.array([[0.0, 0.0, 0.0]])
    _check(eng, one, None, 1, O.Params(0.5, 2.0, 2, 1, False))


def test_device_pointer_input(eng):
    import torch
    d = C.clustered(5000, 100.0, 12)
    r = C.uniform(10000, 100.0, 13)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(5000, 0.2, 14), C.lognormal_weights(10000, 0.1, 15)])
    p = O.Params(2.0, 20.0, 8, 5, True)
    ref = O.pcf2(xyz, w, 5000, p)
    eng.set_catalogue(torch.from_numpy(xyz).cuda(), torch.from_numpy(w).cuda(), 5000)
    r = eng.pcf2(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, "fp32")
    assert np.array_equal(r["cnt"], ref["cnt"])
    assert _close(r["n2"], ref["n2"]) and _close(r["n3"], ref["n3"])
    # the weight sums of a device upload (GPU reduction) and of a host upload (host loop) are the
    # same exact fixed-point sum, so the normalised results agree to the last bit
    eng.set_catalogue(xyz, w, 5000)
    h = eng.pcf2(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, "fp32")
    assert np.array_equal(h["n2"], r["n2"]) and np.array_equal(h["n3"], r["n3"])


def test_device_weights_wide_range_and_errors(eng):
    import torch
    d = C.clustered(4000, 100.0, 16)
    r = C.uniform(8000, 100.0, 17)
    xyz = np.concatenate([d, r])
    rng = np.random.default_rng(18)
    w = 10.0 ** rng.uniform(-12, 6, len(xyz))        # 18 decades of weights
    w[::97] = 0.0
    p = O.Params(2.0, 20.0, 8, 3, False)
    ref = O.pcf2(xyz, w, 4000, p)
    eng.set_catalogue(torch.from_numpy(xyz).cuda(), torch.from_numpy(w).cuda(), 4000)
    a = eng.pcf2(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, "fp32")
    assert np.array_equal(a["cnt"], ref["cnt"])
    assert _close(a["n2"], ref["n2"]) and _close(a["n3"], ref["n3"])
    eng.set_catalogue(xyz, w, 4000)
    b = eng.pcf2(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, "fp32")
    assert np.array_equal(a["n2"], b["n2"])
    for bad in (np.nan, -1.0, np.inf):
        wb = w.copy()
        wb[1234] = bad
        with pytest.raises(Exception, match=r"w\[1234\]"):
            eng.set_catalogue(torch.from_numpy(xyz).cuda(), torch.from_numpy(wb).cuda(), 4000)
    eng.set_catalogue(torch.from_numpy(xyz).cuda(), torch.from_numpy(w).cuda(), 4000)   # usable again
    c = eng.pcf2(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, "fp32")
    assert np.array_equal(a["n2"], c["n2"])


def test_fp32_equals_fp64_large(eng):
    # a density-matched slab of the (r, mu) GPU configuration: the two
    # precisions must give identical counts and identical fixed-point sums
    d = C.clustered(200000, 400.0, 201)
    r = C.uniform(600000, 400.0, 202)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(200000, 0.2, 203), np.ones(600000)])
    eng.set_catalogue(xyz, w, 200000)
    a = eng.pcf2(1.0, 60.0, 30, 20, True, "fp64")
    b = eng.pcf2(1.0, 60.0, 30, 20, True, "fp32")
    print("binned", a["stats"]["n_binned"], "rechecks", b["stats"]["n_recheck"], flush=True)
    assert np.array_equal(a["cnt"], b["cnt"])
    assert np.array_equal(a["wsum"], b["wsum"])
    assert np.array_equal(a["n2"], b["n2"])
    # mu mirror symmetry of the DD and RR histograms holds exactly
    assert np.array_equal(a["cnt"][0], a["cnt"][0][:, ::-1])
    assert b["stats"]["n_recheck"] < 1e-3 * b["stats"]["n_binned"]


def test_repeat_is_deterministic(eng):
    d = C.clustered(20000, 200.0, 301)
    w = C.lognormal_weights(20000, 0.3, 302)
    eng.set_catalogue(d, w, 20000)
    a = eng.pcf2(1.0, 30.0, 10, 6, True, "fp32")
    b = eng.pcf2(1.0, 30.0, 10, 6, True, "fp32", cell_div=2)
    assert np.array_equal(a["cnt"], b["cnt"])
    assert np.array_equal(a["wsum"], b["wsum"])   # fixed point: order independent


def test_mu_int8_wrap_at_127_mu_bins(eng):
    # pairs on rays from the origin have theta = +-1; at nmu = 127 the reference's int8 store of
    # floor(t)+1 wraps and sends them to mu bin 1 (the oracle follows it; pinned against the
    # reference binary in test_oracle.py), which the FP32 guard must hand to the FP64 recheck
    from test_oracle import ray_catalogue
    xyz = ray_catalogue()
    nd = len(xyz) // 2
    w = C.lognormal_weights(len(xyz), 0.2, 6)
    for wt in (None, w):
        ref, outs = _check(eng, xyz, wt, nd, O.Params(0.5, 12.0, 4, 127, False))
        assert ref["cnt"][0][:, 0].sum() > 0


@pytest.mark.parametrize("shift", [5.0e4, -2.0e5, -60.0])
def test_catalogue_far_from_or_around_origin(eng, shift):
    # (r, mu) counts with the line of sight from the origin: a box 50000 / 200000 units away (large
    # absolute coordinates, small mu error bound) and one straddling the origin (|mid| near 0)
    d = C.clustered(4000, 120.0, 31) + shift
    r = C.uniform(8000, 120.0, 32) + shift
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(4000, 0.2, 33), np.ones(8000)])
    ref, outs = _check(eng, xyz, w, 4000, O.Params(1.0, 25.0, 8, 12, True))
    assert ref["cnt"].sum() > 0
