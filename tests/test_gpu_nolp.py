"""--noLP (no lonely base pairs) on the GPU against the oracle: scores,
alignment edges, structures and the D table bit for bit, on tiny inputs with
stacks of two and three arcs (the oracle pinned by enumeration), on
workload pairs (global and sequence-local) and on a long pair that several
workgroups share; penalized alignment; LocARNA-P refuses --noLP."""
import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_stacked_pair

pytestmark = pytest.mark.gpu


def _compare(batch, k, o, label):
    assert batch.score(k) == o.score, label
    edges, sA, sB = batch.alignment(k)
    assert edges == o.edges, label
    assert sA == o.structA and sB == o.structB, label
    for t, (x, y) in enumerate(zip(batch.D(k), o.D)):
        if y < -(1 << 60):
            assert x < -(1 << 28), (label, t)
        else:
            assert x == y, (label, t)


def _run(pp, kw, label):
    kw = dict(kw, no_lonely_pairs=1)
    ores = [oracle.align(a, b, oracle.locarna_params(**kw)) for a, b in pp]
    batch = api.Batch([api.Pair(a, b, api.locarna_params(**kw)) for a, b in pp])
    print(label, len(pp), "pairs", flush=True)
    batch.run(mode=0, trace=True)
    for k, o in enumerate(ores):
        _compare(batch, k, o, (label, k))
    return batch


@pytest.mark.parametrize("local", [0, 1])
def test_nolp_tiny(local):
    pp = [tiny_stacked_pair(4000 + s) for s in range(40)]
    kw = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, struct_weight=400, tau=100, sequ_local=local)
    _run(pp, kw, "tiny").destroy()


@pytest.mark.parametrize("local", [0, 1])
def test_nolp_workload(local):
    pps, pairs, span = synth.workload("allpairs64")
    pp = [(pps[a], pps[b]) for a, b in pairs[::20]]
    b = _run(pp, dict(max_bp_span=span, sequ_local=local), "allpairs64")
    used = sum(1 for k in range(len(pp)) if "(" in b.alignment(k)[1])
    assert used > 0  # stacks are aligned
    b.destroy()


def test_nolp_long_pair_several_workgroups():
    pps, pairs, span = synth.workload("long500", limit_pairs=1)
    b = _run([(pps[a], pps[b]) for a, b in pairs], dict(max_bp_span=span), "long500")
    assert b.workgroups > 1
    b.destroy()


def test_nolp_penalized():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=6)
    kw = dict(max_bp_span=span, sequ_local=1, no_lonely_pairs=1)
    pp = [(pps[a], pps[b]) for a, b in pairs]
    batch = api.Batch([api.Pair(a, b, api.locarna_params(**kw)) for a, b in pp])
    batch.run_penalized(40)
    for k, (a, b) in enumerate(pp):
        o = oracle.align_variant(a, b, oracle.locarna_params(**kw), penalized=40)
        assert batch.score(k) == o.score and batch.alignment(k)[0] == o.edges, k
    batch.destroy()


def test_nolp_refused_by_pf():
    pps, pairs, span = synth.workload("pf16x150", limit_pairs=1)
    a, b = pairs[0]
    batch = api.Batch([api.Pair(pps[a], pps[b], api.locarna_p_params(max_bp_span=span, no_lonely_pairs=1))])
    with pytest.raises(api.LcgError, match=r"\(-1\)"):
        batch.pf_run()
    batch.destroy()
