"""GPU structure-local alignment (--struct-local, eight states inside arc
matches) against the oracle, whose optimum is pinned by enumeration
(tests/test_oracle_struct_local.py): scores, alignments (excluded positions
without edges), structures and D bit for bit."""
import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_pair

pytestmark = pytest.mark.gpu


def _compare(pp_pairs, kw):
    po, pa = oracle.locarna_params(**kw), api.locarna_params(**kw)
    ores = [oracle.align(a, b, po) for a, b in pp_pairs]
    B = api.Batch([api.Pair(a, b, pa) for a, b in pp_pairs])
    print("struct-local batch", len(pp_pairs), kw, flush=True)
    B.run(mode=0, trace=True)
    for k, o in enumerate(ores):
        assert B.score(k) == o.score, (k, B.score(k), o.score)
        edges, sA, sB = B.alignment(k)
        assert edges == o.edges, k
        assert (sA, sB) == (o.structA, o.structB), k
        for t, (x, y) in enumerate(zip(B.D(k), o.D)):
            if y < -(1 << 60):
                assert x < -(1 << 28), (k, t)
            else:
                assert x == y, (k, t)
    return B


@pytest.mark.parametrize("exclusion", [0, -60, -400])
@pytest.mark.parametrize("local", [0, 1], ids=["global", "local"])
def test_struct_local_tiny(local, exclusion):
    kw = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, sequ_local=local, indel=-150,
              indel_opening=-200, struct_local=1, exclusion=exclusion)
    _compare([tiny_pair(4500 + s, arcs=(2, 4)) for s in range(24)], kw).destroy()


@pytest.mark.parametrize("local", [0, 1], ids=["global", "local"])
def test_struct_local_workload_pairs(local):
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=10)
    _compare([(pps[a], pps[b]) for a, b in pairs],
             dict(max_bp_span=span, sequ_local=local, struct_local=1, exclusion=-100)).destroy()


def test_struct_local_long_pair_multi_workgroup():
    pps, pairs, span = synth.workload("long500", limit_pairs=1)
    _compare([(pps[a], pps[b]) for a, b in pairs], dict(max_bp_span=span, struct_local=1, exclusion=-50)).destroy()


def test_struct_local_refusals():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=1)
    a, b = pps[pairs[0][0]], pps[pairs[0][1]]
    B = api.Batch([api.Pair(a, b, api.locarna_p_params(max_bp_span=span, struct_local=1))])
    with pytest.raises(api.LcgError):
        B.pf_run()
    B.destroy()
