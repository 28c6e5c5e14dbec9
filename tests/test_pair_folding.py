"""CPU check of the folded neighbour-pair schedules of the triplet kernel
(gramsci_amd/csrc/triplet3.hip, step 2).  Folded row r (0 <= r < K/2) holds
the K pairs (r, r+1+c) for c < K-1-r and (K-2-r, c) for K-1-r <= c < K; for
even K the last folded row is row r alone (K/2 pairs, its other K/2 positions
unused).  A work-group of n threads (256, or 64 / 128) is n / 64 waves.

Rows (LDS lists where they keep >= 7/8 of the lanes busy, and every
device-memory list): wave w takes folded rows w G, (w + n/64) G, ... with
G = 64 // K rows per pass for K <= 64 (lane l: row l // K, column l % K), else
one row per pass with lane l at columns l, l + 64, ...

Flat (LDS lists otherwise): the pair index pi = r K + c, 0 <= pi < K(K-1)/2,
starts at the thread index and advances by n: (r, c) by (n // K, n % K) with at
most one carry.

Every unordered pair u < v of K neighbours must be visited exactly once (each
apex pair is one triangle candidate of the reference's k1 < k2 loops,
gramsci.F90:479-489)."""
import pytest

NTHR = 256


def by_rows(K):
    """The kernel's choice for an LDS list of K neighbours."""
    if K > 64:
        return K * 8 >= -(-K // 64) * 64 * 7
    G = 64 // K if K > 0 else 1
    return G * K * 8 >= 64 * 7


def flat_schedule(K, nthr=NTHR):
    """(wave, step, lane, u, v) for every visit of the flat schedule."""
    P = K * (K - 1) // 2
    if P == 0:
        return
    dq, dr = divmod(nthr, K)
    for tid in range(nthr):
        wv, ln = divmod(tid, 64)
        r, c = divmod(tid, K)
        step = 0
        pi = tid
        while pi < P:
            assert r * K + c == pi          # the incremental (r, c) is the index's
            first = c < K - 1 - r
            yield wv, step, ln, (r if first else K - 2 - r), (r + 1 + c if first else c)
            c += dr
            carry = c >= K
            c -= K if carry else 0
            r += dq + (1 if carry else 0)
            pi += nthr
            step += 1


def row_schedule(K, nthr=NTHR, packed=True):
    """(wave, pass row, lane, u, v) for every visit of the row schedule (packed: G rows per pass for
    K <= 64, as for LDS lists; unpacked: one row per pass, as for device-memory lists)."""
    R = K // 2
    nw = nthr // 64
    G = (64 // K if 0 < K <= 64 else 1) if packed else 1
    for wv in range(nw):
        for ln in range(64):
            rr = ln // K if (packed and 0 < K <= 64) else 0
            c0 = ln - rr * K if (packed and 0 < K <= 64) else ln
            if rr >= G:
                continue
            for r0 in range(wv * G, R, nw * G):
                r = r0 + rr
                if r >= R:
                    continue
                t = K - 1 - r
                length = t if 2 * r + 2 == K else K
                for c in range(c0, length, 64):
                    first = c < t
                    yield wv, r0, ln, (r if first else K - 2 - r), (r + 1 + c if first else c)


def unpacked_rows(K, nthr=NTHR):
    return row_schedule(K, nthr, packed=False)


SIZES = list(range(0, 40)) + [63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513, 768, 1000, 4096]


@pytest.mark.parametrize("nthr", [64, 128, 256])
@pytest.mark.parametrize("K", SIZES)
@pytest.mark.parametrize("sched", [flat_schedule, row_schedule, unpacked_rows])
def test_every_pair_once(sched, K, nthr):
    got = [(u, v) for _, _, _, u, v in sched(K, nthr)]
    assert all(0 <= u < v < K for u, v in got)
    assert len(got) == K * (K - 1) // 2
    assert len(set(got)) == len(got)


@pytest.mark.parametrize("K", [20, 40, 63, 70, 100, 490])
@pytest.mark.parametrize("sched", [flat_schedule, row_schedule])
def test_neighbouring_lanes_read_neighbouring_entries(sched, K):
    # in a wave's step, lanes l and l+1 share u and have v one apart, except where a folded row's
    # two segments meet or a row ends (at most 2 (64 // K + 1) + 2 breaks per step)
    steps = {}
    for wv, st, ln, u, v in sched(K):
        steps.setdefault((wv, st), {})[ln] = (u, v)
    for lanes in steps.values():
        ls = sorted(lanes)
        same = sum(lanes[l][0] == lanes[l + 1][0] and lanes[l + 1][1] == lanes[l][1] + 1
                   for l in ls if l + 1 in lanes)
        assert same >= len(ls) - 2 * (64 // K + 1) - 2


@pytest.mark.parametrize("K", [33, 40, 64, 65, 100, 120, 300, 490, 1000])
def test_chosen_schedule_keeps_lanes_busy(K):
    # busy lanes per wave step of the schedule the kernel picks for K, counted from the schedule:
    # at least 7/8 for rows (by the choice), and for the flat schedule all but the last step full.
    # (Below about 2 nw rows a step of either schedule covers the whole apex, and utilisation is moot.)
    sched = row_schedule if by_rows(K) else flat_schedule
    slots = {}
    for wv, st, ln, _, _ in sched(K):
        slots[(wv, st)] = slots.get((wv, st), 0) + 1
    busy = K * (K - 1) // 2
    full = 64 * len(slots)
    if sched is flat_schedule:
        assert full - busy < 4 * 64          # only each wave's last step can be partial
    else:
        # rows: per row of K pairs at most 64 - K % 64 idle lanes (the even-K last row aside)
        assert busy / full >= 7 / 8 * (K - 1) / K - 0.03
