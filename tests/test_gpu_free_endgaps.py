"""Free end gaps (locarna --free-endgaps) on the GPU: scores, alignments and
structures equal the oracle's (itself pinned against exhaustive enumeration
in test_oracle_bruteforce.py) on tiny pairs and on all-pairs workload pairs
for several end specifications; LocARNA-P refuses them.  Calls the kernels
through the C ABI."""
import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_pair

pytestmark = pytest.mark.gpu

SPECS = ["+---", "-+--", "--+-", "---+", "++++", "+-+-", "-+-+", "++--", "--++"]


def _check(pp, kw, label):
    pa = api.locarna_params(**kw)
    po = oracle.locarna_params(**kw)
    P = [api.Pair(a, b, pa) for a, b in pp]
    B = api.Batch(P)
    try:
        for mode in (0, 2):
            B.run(mode=mode)
            for k, (a, b) in enumerate(pp):
                o = oracle.align(a, b, po)
                edges, sA, sB = B.alignment(k)
                if o.score < -(1 << 60):  # minus infinity (InftyInt range; the library's is below -2^28)
                    assert B.score(k) < -(1 << 28), (label, mode, k, B.score(k))
                else:
                    assert B.score(k) == o.score, (label, mode, k, B.score(k), o.score)
                assert edges == o.edges and sA == o.structA and sB == o.structB, (label, mode, k)
    finally:
        B.destroy()


@pytest.mark.parametrize("spec", SPECS)
def test_tiny_pairs(spec):
    fe = api.free_endgaps(spec)
    pp = [tiny_pair(4400 + s * 13 + fe) for s in range(12)]
    print("free end gaps", spec, "tiny", flush=True)
    _check(pp, dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, indel=-150, indel_opening=-200,
                    free_endgaps=fe), spec)


@pytest.mark.parametrize("spec", ["++++", "+-+-", "-+-+", "---+"])
def test_workload_pairs(spec):
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=40)
    print("free end gaps", spec, "allpairs64 subset", flush=True)
    _check([(pps[a], pps[b]) for a, b in pairs], dict(max_bp_span=span, free_endgaps=api.free_endgaps(spec)),
           spec)


def test_locarna_p_refuses_free_endgaps():
    pps, pairs, span = synth.workload("pf16x150", limit_pairs=1)
    pa = api.locarna_p_params(max_bp_span=span, free_endgaps=1)
    B = api.Batch([api.Pair(pps[pairs[0][0]], pps[pairs[0][1]], pa)])
    try:
        with pytest.raises(api.LcgError):
            B.pf_run()
    finally:
        B.destroy()
