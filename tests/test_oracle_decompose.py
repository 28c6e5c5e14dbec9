"""Pins the numpy restatement of the reference's domain decomposition tool
(oracle.decompose) against the tool itself, compiled from its own source by
`make -C oracle ref` into oracle/_ref/domain_decomposition: every
"<r>.loadnodes" file it writes must list exactly the restatement's members,
in the same order, with the same coordinates, weights, gal flags and buffer
flags.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

REF_TOOL = os.path.join(os.path.dirname(O.REF_BIN), "domain_decomposition")

pytestmark = pytest.mark.skipif(not os.path.exists(REF_TOOL), reason="reference tool not built (make -C oracle ref)")


def write_allpoints(path, xyz, w, gal, region):
    with open(path, "w") as f:
        for (x, y, z), wi, g, r in zip(xyz, w, gal, region):
            f.write("%.17g %.17g %.17g %.17g %d 0.5 %d\n" % (x, y, z, wi, g, r))


def make_case(seed, n=40000, snap=False):
    # n >= 40000 keeps every array of the tool above the allocator's mmap
    # threshold, so its never-read extra element is a zero point, which lies
    # outside every grown box here (all coordinates are >= 100).
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(100.0, 400.0, (n, 3))
    if snap:
        xyz = np.round(xyz * 2.0) / 2.0       # many points exactly on box faces
    w = np.exp(rng.normal(0.0, 0.3, n))
    gal = (rng.random(n) < 0.3).astype(int)
    region = 1 + (xyz[:, 0] > 200).astype(int) + (xyz[:, 0] > 300).astype(int) + 3 * (xyz[:, 1] > 250).astype(int)
    region[rng.random(n) < 0.01] = 0          # unlabelled points: members of boxes only
    region[(region == 5) | (region == 6)] = 7   # labels 5 and 6 empty, 7 used
    return xyz, w, gal, region


def read_loadnodes(path):
    rows = [line.split() for line in open(path)]
    if not rows:
        return np.zeros((0, 3), np.float32), np.zeros(0, np.float32), np.zeros(0, int), np.zeros(0, int)
    a = np.array(rows)
    return (a[:, :3].astype(np.float32), a[:, 3].astype(np.float32), a[:, 4].astype(int), a[:, 5].astype(int))


@pytest.mark.parametrize("snap", [False, True])
def test_decompose_matches_reference_tool(tmp_path, snap):
    xyz, w, gal, region = make_case(11 + snap, snap=snap)
    write_allpoints(tmp_path / "allpoints", xyz, w, gal, region)
    rmax = 20.0
    r = subprocess.run([REF_TOOL, repr(rmax)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    off, idx, buf, box = O.decompose(xyz, region, rmax)
    assert len(off) == 8
    x32, w32 = xyz.astype(np.float32), w.astype(np.float32)
    for k in range(1, 8):
        px, pw, pg, pb = read_loadnodes(tmp_path / f"{k}.loadnodes")
        sel = idx[off[k - 1]:off[k]]
        assert len(px) == len(sel), k
        assert np.array_equal(px, x32[sel]) and np.array_equal(pw, w32[sel])
        assert np.array_equal(pg, gal[sel]) and np.array_equal(pb, buf[off[k - 1]:off[k]])
    assert off[5] == off[4] and off[6] == off[5]           # empty labels 5 and 6
    if snap:
        assert buf.sum() > 0 and (buf == 0).sum() > 0
