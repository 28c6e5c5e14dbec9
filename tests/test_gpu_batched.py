"""Batched regions (one persistent hipGraph-captured launch over many uneven
sub-volumes with buffer shells) against the oracle per region and against
one plain launch per region.  Integer counts bit-exact; N2/N3 within 1e-12."""
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


def _close(a, b, rel=1e-12):
    fin = np.isfinite(b)
    floor = rel * max(np.abs(b[fin]).max() if fin.any() else 0.0, 1e-300)
    return np.all(np.abs(a[fin] - b[fin]) <= np.maximum(rel * np.abs(b[fin]), floor))


@pytest.fixture(scope="module")
def batch():
    xyz, buf, off, ndata = C.regions(7, 3, rmax=15.0, median=1500, sigma=0.8, n_min=300, n_max=6000)
    return xyz, buf, off, ndata


@pytest.mark.parametrize("prec,nmu,log", [("fp32", 5, False), ("fp64", 5, False), ("fp32", 1, True)])
def test_batched_matches_oracle_per_region(eng, batch, prec, nmu, log):
    xyz, buf, off, ndata = batch
    p = O.Params(1.0, 15.0, 6, nmu, log)
    eng.set_catalogue_batched(xyz, None, buf, off, ndata)
    out = eng.pcf2_batched(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, prec, cell_div=3)
    print("batched", out["stats"], flush=True)
    for r in range(len(ndata)):
        sl = slice(off[r], off[r + 1])
        ref = O.pcf2(xyz[sl], None, int(ndata[r]), p, buf=buf[sl])
        assert np.array_equal(out["cnt"][r], ref["cnt"]), r
        assert _close(out["n2"][r], ref["n2"]) and _close(out["n3"][r], ref["n3"])


def test_batched_equals_separate_launches_and_replays(eng, batch):
    xyz, buf, off, ndata = batch
    rng = np.random.default_rng(0)
    w = C.lognormal_weights(len(xyz), 0.2, 17)
    args = (1.0, 15.0, 8, 4, True, "fp32")
    sep = []
    for r in range(len(ndata)):
        sl = slice(off[r], off[r + 1])
        eng.set_catalogue(xyz[sl], w[sl], int(ndata[r]), buf[sl])
        sep.append(eng.pcf2(*args))
    eng.set_catalogue_batched(xyz, w, buf, off, ndata)
    a = eng.pcf2_batched(*args)                     # autotuned cell size
    # replays of the captured graph, with unrelated eager work in between
    junk = rng.uniform(0, 10, (500, 3))
    for rep in range(3):
        b = eng.pcf2_batched(*args)
        assert np.array_equal(a["cnt"], b["cnt"]) and np.array_equal(a["wsum"], b["wsum"]), rep
    for r in range(len(ndata)):
        assert np.array_equal(a["cnt"][r], sep[r]["cnt"]), r
        assert np.array_equal(a["wsum"][r], sep[r]["wsum"]), r
        assert np.array_equal(a["n2"][r], sep[r]["n2"]), r
    # a different cell size gives the same answer
    c = eng.pcf2_batched(*args[:-1], args[-1], cell_div=5)
    assert np.array_equal(a["cnt"], c["cnt"]) and np.array_equal(a["wsum"], c["wsum"])
    del junk


@pytest.mark.parametrize("prec,nmu,weighted", [("fp32", 1, False), ("fp64", 1, True), ("fp32", 3, True)])
def test_batched_3pcf_matches_oracle_per_region(eng, batch, prec, nmu, weighted):
    xyz, buf, off, ndata = batch
    w = C.lognormal_weights(len(xyz), 0.2, 19) if weighted else None
    p = O.Params(2.0, 8.0, 4, nmu, False)
    eng.set_catalogue_batched(xyz, w, buf, off, ndata)
    out = eng.pcf3_batched(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, prec, cell_div=2)
    print("batched 3pcf", out["stats"], flush=True)
    for r in range(len(ndata)):
        sl = slice(off[r], off[r + 1])
        ref = O.pcf3(xyz[sl], None if w is None else w[sl], int(ndata[r]), p, buf=buf[sl])
        assert np.array_equal(out["cnt"][r].reshape(ref["cnt"].shape), ref["cnt"]), r
        assert _close(out["n2"][r].reshape(-1), ref["n2"].reshape(-1))
        assert _close(out["n3"][r].reshape(-1), ref["n3"].reshape(-1))


def test_batched_3pcf_equals_separate_launches_and_replays(eng, batch):
    xyz, buf, off, ndata = batch
    w = C.lognormal_weights(len(xyz), 0.2, 23)
    args = (2.0, 8.0, 4, 1, False, "fp32")
    sep = []
    for r in range(len(ndata)):
        sl = slice(off[r], off[r + 1])
        eng.set_catalogue(xyz[sl], w[sl], int(ndata[r]), buf[sl])
        sep.append(eng.pcf3(*args[:3], nmu=1, precision="fp32"))
    eng.set_catalogue_batched(xyz, w, buf, off, ndata)
    a = eng.pcf3_batched(*args)                     # autotuned cell size
    for rep in range(3):
        b = eng.pcf3_batched(*args)
        assert np.array_equal(a["cnt"], b["cnt"]) and np.array_equal(a["wsum"], b["wsum"]), rep
    for r in range(len(ndata)):
        assert np.array_equal(a["cnt"][r][:, 0], sep[r]["cnt"]), r
        assert np.array_equal(a["wsum"][r][:, 0], sep[r]["wsum"]), r


def test_batched_3pcf_two_pass_overflow(eng, batch, monkeypatch):
    # a 96-entry first-pass neighbour list: most apexes go to the captured second pass,
    # whose apex count is read on the device; results equal the oracle's per region
    xyz, buf, off, ndata = batch
    w = C.lognormal_weights(len(xyz), 0.2, 29)
    p = O.Params(2.0, 8.0, 4, 1, False)
    monkeypatch.setenv("GS_T3_KMAX", "96")
    eng.set_catalogue_batched(xyz, w, buf, off, ndata)        # fresh plans read the override
    out = eng.pcf3_batched(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, "fp32", cell_div=2)
    again = eng.pcf3_batched(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, "fp32", cell_div=2)   # replay
    assert np.array_equal(out["cnt"], again["cnt"]) and np.array_equal(out["wsum"], again["wsum"])
    for r in range(len(ndata)):
        sl = slice(off[r], off[r + 1])
        ref = O.pcf3(xyz[sl], w[sl], int(ndata[r]), p, buf=buf[sl])
        assert np.array_equal(out["cnt"][r].reshape(ref["cnt"].shape), ref["cnt"]), r
        assert _close(out["n2"][r].reshape(-1), ref["n2"].reshape(-1))


def test_batched_3pcf_device_memory_lists(eng, batch, monkeypatch):
    # every LDS list cut to 64 entries: apexes with more neighbours (rmax = 12) are queued by the
    # captured passes and counted after each replay with their lists in device memory; results
    # equal the oracle's per region, and a replay gives the same counts and sums
    xyz, buf, off, ndata = batch
    w = C.lognormal_weights(len(xyz), 0.2, 31)
    p = O.Params(2.0, 12.0, 3, 1, False)
    monkeypatch.setenv("GS_T3_KMAX_LDS", "64")
    monkeypatch.setenv("GS_T3_KMAX", "64")
    eng.set_catalogue_batched(xyz, w, buf, off, ndata)        # fresh plans read the overrides
    out = eng.pcf3_batched(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, "fp32", cell_div=2)
    again = eng.pcf3_batched(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins, "fp32", cell_div=2)
    assert np.array_equal(out["cnt"], again["cnt"]) and np.array_equal(out["wsum"], again["wsum"])
    for r in range(len(ndata)):
        sl = slice(off[r], off[r + 1])
        ref = O.pcf3(xyz[sl], w[sl], int(ndata[r]), p, buf=buf[sl])
        assert np.array_equal(out["cnt"][r].reshape(ref["cnt"].shape), ref["cnt"]), r
        assert _close(out["n2"][r].reshape(-1), ref["n2"].reshape(-1))


def test_batch_survives_single_catalogue_uploads(eng, batch):
    # the batch keeps its own copy of the catalogue: uploading a single catalogue on the
    # same context (smaller or larger) changes neither replays of captured pipelines nor
    # pipelines captured afterwards
    xyz, buf, off, ndata = batch
    eng.set_catalogue_batched(xyz, None, buf, off, ndata)
    a2 = eng.pcf2_batched(1.0, 15.0, 6, 3, False, "fp32", cell_div=3)
    small = C.uniform(500, 30.0, 5)
    eng.set_catalogue(small, None, len(small))
    b2 = eng.pcf2_batched(1.0, 15.0, 6, 3, False, "fp32", cell_div=3)        # replay
    big = C.uniform(len(xyz) * 2, 400.0, 6)
    eng.set_catalogue(big, None, len(big))
    a3 = eng.pcf3_batched(2.0, 8.0, 4, 1, False, "fp32", cell_div=2)          # new capture
    assert np.array_equal(a2["cnt"], b2["cnt"])
    for r in range(len(ndata)):
        sl = slice(off[r], off[r + 1])
        ref = O.pcf3(xyz[sl], None, int(ndata[r]), O.Params(2.0, 8.0, 4), buf=buf[sl])
        assert np.array_equal(a3["cnt"][r][:, 0], ref["cnt"]), r


def test_batched_full_size_equals_separate_launches(eng):
    # the 64-region configuration at its real size (28M points with their rmax buffer shells):
    # one captured launch over all regions gives, region by region, the counts of one plain launch
    # per region, for the 2PCF (5 < r < 100, 20 bins, nmu = 10) and the 3PCF (5 < r < 25, 4 bins)
    xyz, buf, off, ndata = C.regions(64, 401, rmax=100.0)
    eng.set_catalogue_batched(xyz, None, buf, off, ndata)
    p2 = (5.0, 100.0, 20, 10, False, "fp32")
    p3 = (5.0, 25.0, 4, 1, False, "fp32")
    a2 = eng.pcf2_batched(*p2)
    a3 = eng.pcf3_batched(*p3)
    m2, m3 = a2["stats"]["cell_div"], a3["stats"]["cell_div"]
    print("64 regions:", len(xyz), "points, 2PCF binned", a2["stats"]["n_binned"], "m", m2,
          "3PCF binned", a3["stats"]["n_binned"], "m", m3, flush=True)
    for k in range(len(ndata)):
        sl = slice(off[k], off[k + 1])
        eng.set_catalogue(xyz[sl], None, int(ndata[k]), buf[sl])
        assert np.array_equal(eng.pcf2(*p2[:5], p2[5], cell_div=m2)["cnt"], a2["cnt"][k]), k
        assert np.array_equal(eng.pcf3(*p3[:3], nmu=1, precision="fp32", cell_div=m3)["cnt"], a3["cnt"][k][:, 0]), k
