"""gs_count_2pcf_noalloc (no device allocation, scratch in a file-scope device
array) against gs_count_2pcf on the same catalogues: identical counts,
class sums, N2 and N3, on two calls in a row with the second catalogue smaller
than the first (the array keeps the first call's contents), in both
precisions, weighted with mu bins and with buffer points; plus the size
limit, refused on the host before any GPU work."""
import numpy as np
import pytest

from gramsci_amd import catalog as C
from gramsci_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _case(nd, nr, seed, box, weighted=True, buffered=False):
    d = C.clustered(nd, box, seed)
    r = C.uniform(nr, box, seed + 1)
    xyz = np.concatenate([d, r])
    w = np.concatenate([C.lognormal_weights(nd, 0.2, seed + 2), C.lognormal_weights(nr, 0.1, seed + 3)]) \
        if weighted else None
    buf = (np.random.default_rng(seed + 4).random(nd + nr) < 0.1).astype(np.int32) if buffered else None
    return xyz, w, buf


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("args,buffered", [((1.0, 30.0, 12, 6, True), False), ((2.0, 25.0, 8, 1, False), True)])
def test_noalloc_equals_count_2pcf(prec, args, buffered):
    import torch
    eng = Engine(0)
    try:
        for nd, nr, seed, box in ((60000, 120000, 11, 220.0), (7000, 9000, 21, 90.0)):   # second call smaller
            xyz, w, buf = _case(nd, nr, seed, box, buffered=buffered)
            eng.set_catalogue(xyz, w, nd, buf)
            ref = eng.pcf2(*args, precision=prec)
            dxyz = torch.from_numpy(xyz).cuda()
            dw = torch.from_numpy(w).cuda()
            dbuf = None if buf is None else torch.from_numpy(buf).cuda()
            out = eng.pcf2_noalloc(dxyz, dw, nd, *args, precision=prec, buf=dbuf)
            print(f"{prec} {args} n={nd + nr}: binned {out['stats']['n_binned']} vs {ref['stats']['n_binned']}",
                  flush=True)
            assert out["stats"]["n_binned"] > 0
            assert np.array_equal(out["cnt"], ref["cnt"])
            assert np.array_equal(out["wsum"], ref["wsum"])
            assert np.array_equal(out["n2"], ref["n2"]) and np.array_equal(out["n3"], ref["n3"])
            # (the FP32 recheck count is left out: which end of a pair is the apex follows the atomic
            # order of points inside a cell, so it varies by a few between any two calls)
            for k in ("n_candidates", "n_overflow"):
                assert out["stats"][k] == ref["stats"][k], k
    finally:
        eng.close()


def test_noalloc_unweighted_and_limit():
    import torch
    from gramsci_amd import _lib as L
    eng = Engine(0)
    try:
        xyz, _, _ = _case(3000, 6000, 31, 80.0, weighted=False)
        eng.set_catalogue(xyz, None, 3000)
        ref = eng.pcf2(1.0, 20.0, 6, 4, False)
        out = eng.pcf2_noalloc(torch.from_numpy(xyz).cuda(), None, 3000, 1.0, 20.0, 6, 4, False)
        assert np.array_equal(out["cnt"], ref["cnt"]) and np.array_equal(out["n2"], ref["n2"])
        big = torch.zeros((600_000, 3), dtype=torch.float64, device="cuda")
        with pytest.raises(L.GsError) as e:
            eng.pcf2_noalloc(big, None, 600_000, 1.0, 20.0, 6)
        assert e.value.code == L.GS_ELIMIT
    finally:
        eng.close()
