"""GPU domain decomposition (gs_decompose) and its two consumers:

  * host/domain_decomposition, the drop-in for the reference's tool: its
    "<r>.loadnodes" files must be byte-identical to the reference tool's;
  * the batched launch fed with the regions (Engine.set_catalogue_regions):
    per-region counts and un-normalised sums equal the oracle's 2PCF of each
    region's member catalogue, and over a slab partition the regions' counts add
    up to the whole catalogue's.

Bar: membership, order, buffer flags and boxes identical to the restatement of
the tool (oracle.decompose); counts bit-exact; sums within 1e-12 relative.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from gramsci_amd import catalog as C
from gramsci_amd.engine import Engine
from test_oracle_decompose import REF_TOOL, make_case, write_allpoints

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "host", "domain_decomposition")
HOST = os.path.join(ROOT, "host", "gramsci")


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("snap", [False, True])
def test_decompose_matches_restatement(eng, snap):
    xyz, w, gal, region = make_case(21 + snap, n=60000, snap=snap)
    for rmax in (0.0, 7.5, 20.0):
        print(f"case snap={snap} rmax={rmax}", flush=True)
        d = eng.decompose(xyz, region, rmax)
        off, idx, buf, box = O.decompose(xyz, region, rmax)
        assert np.array_equal(d["off"], off)
        assert np.array_equal(d["idx"], idx) and np.array_equal(d["buf"], buf)
        assert np.array_equal(d["box"], box)


def test_decompose_sizes_and_errors(eng):
    xyz, w, gal, region = make_case(31, n=5000)
    with pytest.raises(Exception):
        eng.decompose(xyz, np.zeros(len(xyz), np.int32), 10.0)        # no label >= 1
    d = eng.decompose(xyz[:1], np.ones(1, np.int32), 10.0)             # one point: its own region
    assert list(d["off"]) == [0, 1] and list(d["buf"]) == [0]


@pytest.mark.skipif(not (os.path.exists(TOOL) and os.path.exists(REF_TOOL)), reason="tools not built")
def test_tool_files_identical_to_reference(tmp_path):
    xyz, w, gal, region = make_case(41, n=50000, snap=True)
    for d in ("ref", "gpu"):
        (tmp_path / d).mkdir()
        write_allpoints(tmp_path / d / "allpoints", xyz, w, gal, region)
    a = subprocess.run([REF_TOOL, "15.0"], cwd=tmp_path / "ref", capture_output=True, text=True, timeout=300)
    b = subprocess.run([TOOL, "15.0"], cwd=tmp_path / "gpu", capture_output=True, text=True, timeout=300)
    assert a.returncode == 0 and b.returncode == 0, b.stdout + b.stderr
    files = sorted(p.name for p in (tmp_path / "ref").glob("*.loadnodes"))
    assert files == sorted(p.name for p in (tmp_path / "gpu").glob("*.loadnodes"))
    assert len(files) == 7
    for f in files:
        assert (tmp_path / "ref" / f).read_bytes() == (tmp_path / "gpu" / f).read_bytes(), f


@pytest.mark.skipif(not (os.path.exists(TOOL) and os.path.exists(HOST)), reason="tools not built")
def test_cut_workflow_end_to_end(tmp_path):
    # GPU tool -> host -cut on every region file.  The reference's own -cut run
    # cannot serve as the check here: built with flang it stops with a
    # segmentation fault before its first output (num_rand is never set in -cut
    # mode, gramsci.F90:1304-1317).  The check is the oracle on the file's
    # contents with -cut semantics (gramsci.F90:1342-1362): every point is data,
    # weights are used as read, and the 5th column -- the gal column of the
    # tool's 6-column files -- is the buffer flag.
    xyz, w, gal, region = make_case(51, n=40000)
    write_allpoints(tmp_path / "allpoints", xyz, w, gal, region)
    assert subprocess.run([TOOL, "12.0"], cwd=tmp_path, capture_output=True, timeout=300).returncode == 0
    p = O.Params(2.0, 12.0, 5, 1)
    for k in (1, 2, 3):
        rows = np.loadtxt(tmp_path / f"{k}.loadnodes")
        r = subprocess.run([HOST, "-cut", f"{k}.loadnodes", "-rmin", "2.0", "-rmax", "12.0", "-nbins", "5", "-2pcf"],
                           cwd=tmp_path, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stdout + r.stderr
        from golden_util import parse_rows
        out = parse_rows((tmp_path / f"{k}.loadnodes").read_text())
        ref = O.pcf2(rows[:, :3], None, len(rows), p, buf=rows[:, 4].astype(np.int32))
        wsum = O.pcf2(rows[:, :3], rows[:, 3], len(rows), p, buf=rows[:, 4].astype(np.int32))["wsum"][0, :, 0]
        assert out.shape == (5, 7)
        # N2 of -cut = the un-normalised sum of w_i w_j over the kept ordered pairs, printed from the
        # engine's exact weight sums: equal to the oracle's FP64 sum to the last printed digit
        from golden_util import close_e147
        assert close_e147(out[:, 4], wsum).all(), (out[:, 4], wsum)
        assert ref["cnt"][0].sum() > 0


def test_regions_feed_the_batched_launch(eng):
    xyz, w, gal, region = make_case(61, n=60000)
    p = O.Params(2.0, 12.0, 6, 4, True)
    d = eng.set_catalogue_regions(xyz, w, region, p.rmax)
    res = eng.pcf2_batched(p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins)
    x32 = xyz.astype(np.float32).astype(np.float64)
    w32 = w.astype(np.float32).astype(np.float64)
    for k in range(len(d["off"]) - 1):
        sel = d["idx"][d["off"][k]:d["off"][k + 1]]
        if len(sel) == 0:
            assert res["cnt"][k].sum() == 0
            continue
        ref = O.pcf2(x32[sel], w32[sel], len(sel), p, buf=d["buf"][d["off"][k]:d["off"][k + 1]])
        assert np.array_equal(res["cnt"][k], ref["cnt"]), k
        np.testing.assert_allclose(res["wsum"][k][0], ref["wsum"][0], rtol=1e-12, atol=0)


def test_slab_partition_sums_to_whole(eng):
    # four x-slabs with gaps wider than nothing: every point is a core point of
    # exactly one region and all its neighbours lie in that region's grown box
    xyz = C.clustered(80000, 200.0, 71, members=10, sigma=2.0)
    region = 1 + np.minimum((xyz[:, 0] // 50.0).astype(np.int32), 3)
    p = O.Params(1.0, 10.0, 5, 1)
    eng.set_catalogue_regions(xyz, None, region, p.rmax)
    per = eng.pcf2_batched(p.rmin, p.rmax, p.nbins)["cnt"][:, 0].sum(0)
    eng.set_catalogue(xyz.astype(np.float32).astype(np.float64), None, len(xyz))
    whole = eng.pcf2(p.rmin, p.rmax, p.nbins)["cnt"][0]
    assert np.array_equal(per, whole)
