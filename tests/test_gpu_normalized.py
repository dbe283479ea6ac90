"""Penalized and normalized alignment on the GPU (lcg_batch_run_penalized /
lcg_batch_run_normalized; locarna --penalized / --normalized): scores,
alignments and structures equal the oracle's (itself pinned against
exhaustive enumeration and, for normalized alignment, against Dinkelbach
iterations of pinned penalized alignments in test_oracle_bruteforce.py).
Calls the kernels through the C ABI."""
import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_pair

pytestmark = pytest.mark.gpu

TINY = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, indel=-150, indel_opening=-200)


def _batch(pp, kw):
    return api.Batch([api.Pair(a, b, api.locarna_params(**kw)) for a, b in pp])


def _same(B, k, o, label):
    edges, sA, sB = B.alignment(k)
    assert B.score(k) == o.score, (label, k, B.score(k), o.score)
    assert edges == o.edges and sA == o.structA and sB == o.structB, (label, k)


@pytest.mark.parametrize("local", [1, 0])
@pytest.mark.parametrize("P", [25, 80, -40])
def test_penalized(local, P):
    kw = dict(TINY, sequ_local=local)
    pp = [tiny_pair(5100 + s) for s in range(12)]
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=10)
    for label, pairs_, kw_ in (("tiny", pp, kw), ("allpairs64", [(pps[a], pps[b]) for a, b in pairs],
                                                   dict(max_bp_span=span, sequ_local=local))):
        B = _batch(pairs_, kw_)
        try:
            B.run_penalized(P)
            for k, (a, b) in enumerate(pairs_):
                _same(B, k, oracle.align_variant(a, b, oracle.locarna_params(**kw_), penalized=P), (label, P))
        finally:
            B.destroy()


@pytest.mark.parametrize("L", [0, 6, 30])
def test_normalized_tiny(L):
    pp = [tiny_pair(5300 + s) for s in range(12)]
    kw = dict(TINY, sequ_local=1)
    B = _batch(pp, kw)
    try:
        it, ms = B.run_normalized(L)
        print("normalized L", L, "iterations", it, flush=True)
        for k, (a, b) in enumerate(pp):
            _same(B, k, oracle.align_variant(a, b, oracle.locarna_params(**kw), normalized=L), ("tiny", L))
        # a plain run afterwards reports plain scores again
        B.run()
        for k, (a, b) in enumerate(pp):
            assert B.score(k) == oracle.align(a, b, oracle.locarna_params(**kw)).score
    finally:
        B.destroy()


def test_normalized_workloads():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=24)
    pp = [(pps[a], pps[b]) for a, b in pairs]
    pps2, pairs2, span2 = synth.workload("long500", limit_pairs=1)
    for label, pp_, kw in (("allpairs64", pp, dict(max_bp_span=span, sequ_local=1)),
                           ("long500", [(pps2[a], pps2[b]) for a, b in pairs2], dict(max_bp_span=span2, sequ_local=1))):
        B = _batch(pp_, kw)
        try:
            it, ms = B.run_normalized(20)
            print(label, "normalized iterations", it, "ms", round(ms, 2), flush=True)
            for k, (a, b) in enumerate(pp_):
                _same(B, k, oracle.align_variant(a, b, oracle.locarna_params(**kw), normalized=20), label)
        finally:
            B.destroy()


def test_normalized_refuses_global():
    B = _batch([tiny_pair(1)], TINY)
    try:
        with pytest.raises(api.LcgError):
            B.run_normalized(5)
    finally:
        B.destroy()
