"""CPU check of the folded neighbour-pair schedule of the triplet kernel
(gramsci_amd/csrc/triplet3.hip, step 2): n threads (256, or 64 / 128 for short
neighbour lists), thread t starting at pair index t and advancing by n; pair index pi = r*K + c maps to
(u, v) = (r, r+1+c) if c < K-1-r, else (K-2-r, c).  Every unordered pair
u < v of K neighbours must be visited exactly once (each apex pair is one
triangle candidate of the reference's k1 < k2 loops, gramsci.F90:479-489)."""
import pytest

NTHR = 256


def schedule(K, nthr=NTHR):
    P = K * (K - 1) // 2
    Kc = max(K, 1)
    dq = nthr // Kc
    dr = nthr - dq * Kc
    for tid in range(nthr):
        r, c = divmod(tid, Kc)
        pi = tid
        while pi < P:
            first = c < K - 1 - r
            u = r if first else K - 2 - r
            v = r + 1 + c if first else c
            yield u, v
            pi += nthr
            c += dr
            carry = c >= Kc
            c -= Kc if carry else 0
            r += dq + (1 if carry else 0)


@pytest.mark.parametrize("nthr", [64, 128, 256])
@pytest.mark.parametrize("K", list(range(0, 40)) + [63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513, 768, 1000,
                                                    4096])
def test_every_pair_once(K, nthr):
    # the kernel runs 256 threads per apex, or 64 / 128 for short neighbour lists
    got = list(schedule(K, nthr))
    assert all(0 <= u < v < K for u, v in got)
    assert len(got) == K * (K - 1) // 2
    assert len(set(got)) == len(got)


def test_neighbouring_lanes_read_neighbouring_entries():
    # lanes t and t+1 of a step share u and have v one apart, except where a folded row ends
    K = 700
    seq = {}
    for tid in range(NTHR):
        r, c = divmod(tid, K)
        first = c < K - 1 - r
        seq[tid] = (r if first else K - 2 - r, r + 1 + c if first else c)
    same_u = sum(seq[t][0] == seq[t + 1][0] and seq[t + 1][1] == seq[t][1] + 1 for t in range(NTHR - 1))
    assert same_u >= NTHR - 3
