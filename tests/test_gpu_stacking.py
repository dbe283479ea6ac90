"""GPU --stacking / --new-stacking (integer path) against the oracle, whose
stacked D-fill is pinned by enumeration (tests/test_oracle_stacking.py):
scores, alignments, structures and D bit for bit."""
import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_stacking_pair

pytestmark = pytest.mark.gpu


def _compare(pp_pairs, kw, shape=None):
    po = oracle.locarna_params(**kw)
    pa = api.locarna_params(**kw)
    ores = [oracle.align(a, b, po) for a, b in pp_pairs]
    P = [api.Pair(a, b, pa) for a, b in pp_pairs]
    B = api.Batch(P, shape=shape)
    print("stacking run", len(P), kw, shape, flush=True)
    B.run(mode=0, trace=True)
    for k, o in enumerate(ores):
        assert B.score(k) == o.score, (k, B.score(k), o.score)
        edges, sA, sB = B.alignment(k)
        assert edges == o.edges, k
        assert (sA, sB) == (o.structA, o.structB), k
        assert len(B.D(k)) == len(o.D), k
        for t, (x, y) in enumerate(zip(B.D(k), o.D)):
            if y < -(1 << 60):  # the oracle's minus infinity (64-bit)
                assert x < -(1 << 28), (k, t)
            else:
                assert x == y, (k, t)
    B.destroy()


@pytest.mark.parametrize("stacking", [1, 2, 3])
@pytest.mark.parametrize("local", [0, 1], ids=["global", "local"])
def test_stacking_tiny(stacking, local):
    kw = dict(min_prob=0.0005, max_bp_span=-1, min_trace_prob=0.0, stacking=stacking, exp_prob=0.1,
              sequ_local=local, indel=-150, indel_opening=-200, tau=50)
    _compare([tiny_stacking_pair(6100 + s) for s in range(24)], kw)


@pytest.mark.parametrize("local", [0, 1], ids=["global", "local"])
def test_nolp_stacking_tiny(local):
    kw = dict(min_prob=0.0005, max_bp_span=-1, min_trace_prob=0.0, stacking=1, exp_prob=0.1, sequ_local=local,
              indel=-150, indel_opening=-200, tau=50, no_lonely_pairs=1)
    _compare([tiny_stacking_pair(6300 + s) for s in range(24)], kw)


def _workload_with_stacking(name, npairs):
    """The named workload's first pairs, their pp files carrying joint
    probabilities of stacked pairs (#STACK section)."""
    seed, count, lmin, lmax, g0, g1, span, pairing = synth.WORKLOADS[name]
    seqs = synth.sequences(seed, count, lmin, lmax, g0, g1)
    pairs = synth.pair_list(count, pairing)[:npairs]
    pp = {x: synth.pp_text("s%03d" % x, seqs[x], synth.with_stacking(synth.base_pair_probs(seqs[x], span), seed=x))
          for x in sorted({x for p in pairs for x in p})}
    return [(pp[a], pp[b]) for a, b in pairs], span


@pytest.mark.parametrize("local", [0, 1], ids=["global", "local"])
def test_stacking_workload_pairs(local):
    pp_pairs, span = _workload_with_stacking("allpairs64", 12)
    _compare(pp_pairs, dict(max_bp_span=span, stacking=1, exp_prob=0.005, sequ_local=local, tau=50))


def test_stacking_long_pair_multi_workgroup():
    """One long pair on several workgroups (the D of inner arc matches from
    other workgroups' rows)."""
    pp_pairs, span = _workload_with_stacking("long500", 1)
    _compare(pp_pairs, dict(max_bp_span=span, stacking=3, exp_prob=0.002, tau=50))


def test_pf_refuses_stacking():
    pp_pairs, span = _workload_with_stacking("allpairs64", 1)
    pa = api.locarna_p_params(max_bp_span=span, stacking=1, exp_prob=0.005)
    B = api.Batch([api.Pair(a, b, pa) for a, b in pp_pairs])
    with pytest.raises(api.LcgError):
        B.pf_run()
    B.destroy()
