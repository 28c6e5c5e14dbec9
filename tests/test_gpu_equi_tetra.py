"""Equilateral 3PCF (-equi) and regular-tetrahedron 4PCF (-tetra) parity of
the HIP path against the oracle and the reference's output files.

Bar: counts per bin (and mu bin) and class bit-exact in both precisions
(the tetrahedron kernel is FP64); N2/N3 within 1e-12 relative of the oracle,
with an absolute floor of 1e-12 * max|N| for cancelling cells; the disc sums
(FP64 atomic sums, whose order varies) with a floor of 1e-9 * max|disc|.
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


def check_equi(eng, xyz, w, nd, p, buf=None):
    ref = O.pcf3_equi(xyz, w, nd, p, buf=buf)
    outs = []
    for prec in ("fp64", "fp32"):
        eng.set_catalogue(xyz, w, nd, buf)
        r = eng.pcf3_equi(p.rmin, p.rmax, p.nbins, nmu=p.nmu, log_bins=p.log_bins, precision=prec)
        print(f"  equi {prec}: binned={r['stats']['n_binned']} recheck={r['stats']['n_recheck']}", flush=True)
        assert np.array_equal(r["cnt"], ref["cnt"]), (prec, np.argwhere(r["cnt"] != ref["cnt"])[:5])
        assert _close(r["n2"], ref["n2"]) and _close(r["n3"], ref["n3"])
        outs.append(r)
    return ref, outs


def check_tetra(eng, xyz, w, nd, p, buf=None):
    ref = O.pcf4_tetra(xyz, w, nd, p, buf=buf)
    eng.set_catalogue(xyz, w, nd, buf)
    r = eng.pcf4_tetra(p.rmin, p.rmax, p.nbins, p.log_bins)
    print(f"  tetra: binned={r['stats']['n_binned']}", flush=True)
    assert np.array_equal(r["cnt"], ref["cnt"])
    assert _close(r["n2"], ref["n2"]) and _close(r["n3"], ref["n3"])
    assert _close_disc(r["disc_n"], ref["disc_n"]) and _close_disc(r["disc_r"], ref["disc_r"])
    return ref, r


@pytest.mark.parametrize("name", [n for n, c in MAN.items() if c["mode"] == "-equi"])
def test_equi_golden(eng, name):
    c = MAN[name]
    p = O.Params(**c["params"])
    xyz, w, nd = _cat(name)
    print(f"case {name}", flush=True)
    ref, outs = check_equi(eng, xyz, w, nd, p)
    rows = parse_rows(read_text(name, "DR" if "DR" in c["runs"] else "D"))
    for r in outs:
        assert close_e147(rows[:, 4], r["n2"].reshape(-1)).all()
        assert close_e147(rows[:, 5], r["n3"].reshape(-1)).all()


@pytest.mark.parametrize("name", [n for n, c in MAN.items() if c["mode"] == "-tetra"])
def test_tetra_golden(eng, name):
    c = MAN[name]
    p = O.Params(**c["params"])
    xyz, w, nd = _cat(name)
    print(f"case {name}", flush=True)
    ref, r = check_tetra(eng, xyz, w, nd, p)
    rows = parse_rows(read_text(name, "DR" if "DR" in c["runs"] else "D"))
    assert close_e147(rows[:, 4], r["n2"]).all()
    assert close_e147(rows[:, 5], r["n3"]).all()


def test_equi_exact_counts_vs_reference_runs(eng):
    name = "equi"
    p = O.Params(**MAN[name]["params"])
    d, r, _, _ = load_input(name)
    for run, cat in (("D", d), ("R", r), ("T", np.concatenate([d, r]))):
        eng.set_catalogue(cat, None, len(cat))
        out = eng.pcf3_equi(p.rmin, p.rmax, p.nbins, precision="fp32")
        rows = parse_rows(read_text(name, run))
        got = np.rint(rows[:, 4] * float(len(cat)) ** 3).astype(np.int64)
        assert np.array_equal(out["cnt"][..., 0].reshape(-1).astype(np.int64), got), run


def test_tetra_exact_counts_vs_reference_runs(eng):
    name = "tetra"
    p = O.Params(**MAN[name]["params"])
    d, r, _, _ = load_input(name)
    for run, cat in (("D", d), ("R", r), ("T", np.concatenate([d, r]))):
        eng.set_catalogue(cat, None, len(cat))
        out = eng.pcf4_tetra(p.rmin, p.rmax, p.nbins)
        rows = parse_rows(read_text(name, run))
        got = np.rint(rows[:, 4] * float(len(cat)) ** 4).astype(np.int64)
        assert np.array_equal(out["cnt"][:, 0].astype(np.int64), got), run


def test_equi_rsd_log_weighted_buffer(eng):
    xyz = C.clustered(5000, 100.0, 81, members=10, sigma=2.0)
    nd = 2500
    w = C.lognormal_weights(len(xyz), 0.3, 82)
    buf = (xyz[:, 0] < 10.0).astype(np.int32)
    print("case equi rsd log weighted buffer", flush=True)
    check_equi(eng, xyz, w, nd, O.Params(2.0, 12.0, 5, 7, True), buf=buf)


def test_equi_matches_3pcf_diagonal(eng):
    xyz = C.clustered(4000, 90.0, 83, members=10, sigma=2.0)
    eng.set_catalogue(xyz, None, 2000)
    e = eng.pcf3_equi(3.0, 12.0, 4, precision="fp32")
    t = eng.pcf3(3.0, 12.0, 4, precision="fp32")
    diag = [O.n3cfg(4) - O.n3cfg(4 - b) for b in range(4)]
    assert np.array_equal(e["cnt"][:, 0, :], t["cnt"][diag])


def test_tetra_weighted_log_buffer(eng):
    xyz = C.clustered(4000, 60.0, 84, members=10, sigma=1.5)
    nd = 2000
    w = C.lognormal_weights(len(xyz), 0.3, 85)
    buf = (xyz[:, 1] > 50.0).astype(np.int32)
    print("case tetra weighted log buffer", flush=True)
    check_tetra(eng, xyz, w, nd, O.Params(1.0, 6.0, 3, 1, True), buf=buf)


def test_tetra_three_neighbour_apex(eng):
    xyz = np.array([[1.0, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) * (3.0 / np.sqrt(8.0)) + 10.0
    eng.set_catalogue(xyz, None, 4)
    assert eng.pcf4_tetra(1.0, 5.0, 2)["cnt"].sum() == 4
