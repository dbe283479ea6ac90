"""GPU k-best suboptimal alignments (locarna --kbest / --better,
Aligner::suboptimal) against the oracle, whose restricted top levels are
pinned by enumeration (tests/test_oracle_kbest.py): the same alignments in
the same order, bit for bit (scores, edges, restrictions)."""
import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_pair

pytestmark = pytest.mark.gpu


def _compare(pp_pairs, kw, kbest, threshold=-1000000):
    po = oracle.locarna_params(**kw)
    pa = api.locarna_params(**kw)
    P = [api.Pair(a, b, pa) for a, b in pp_pairs]
    B = api.Batch(P)
    for k, (a, b) in enumerate(pp_pairs):
        print("kbest pair", k, kw, kbest, flush=True)
        got = B.suboptimal(k, kbest, threshold)
        want = oracle.suboptimal(a, b, kbest, threshold, po)
        assert len(got) == len(want), k
        for g, w in zip(got, want):
            assert g[0] == w[0] and g[1] == w[1] and g[2] == w[2], (k, g[:3], w)
    return B


@pytest.mark.parametrize("local", [1, 0], ids=["local", "global"])
def test_kbest_tiny(local):
    kw = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, sequ_local=local, indel=-150, indel_opening=-200)
    B = _compare([tiny_pair(4100 + s, 6, 7) for s in range(10)], kw, 6)
    B.destroy()


@pytest.mark.parametrize("local", [1, 0], ids=["local", "global"])
def test_kbest_workload_pairs(local):
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=3)
    B = _compare([(pps[a], pps[b]) for a, b in pairs], dict(max_bp_span=span, sequ_local=local), 5)
    B.destroy()


def test_kbest_threshold():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=1)
    a, b = pps[pairs[0][0]], pps[pairs[0][1]]
    kw = dict(max_bp_span=span, sequ_local=1)
    first = oracle.suboptimal(a, b, 3, params=oracle.locarna_params(**kw))
    thr = first[-1][0] - 1
    B = _compare([(a, b)], kw, -1, thr)
    B.destroy()
