"""GPU fragment match probabilities (locarna_p --fragment-match-probs,
AlignerP::compute_fragment_match_prob) against the oracle, whose values are
pinned by enumeration (tests/test_oracle_fragment.py): within 1e-9
absolute plus 1e-7 relative (FP64 sums in another order)."""
import random

import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_pair

pytestmark = pytest.mark.gpu


def _close(x, y):
    return abs(x - y) <= 1e-9 + 1e-7 * abs(y)


def _frags(rng, n, m, count):
    out = []
    for _ in range(count):
        i = rng.randint(1, n)
        j = rng.randint(i, n)
        k = rng.randint(1, m)
        l = rng.randint(k, m)
        out.append((i, j, k, l))
    return out


def _check(pp_pairs, kw, per_pair, seed, shape=None):
    rng = random.Random(seed)
    pa = api.locarna_p_params(**kw)
    po = oracle.locarna_p_params(**kw)
    P = [api.Pair(a, b, pa) for a, b in pp_pairs]
    B = api.Batch(P, shape=shape)
    frags = []
    for k, p in enumerate(P):
        f = [(1, p.lenA, 1, p.lenB)] + _frags(rng, p.lenA, p.lenB, per_pair)
        frags.append(f)
        B.pf_set_fragments(k, f)
    print("fragments", len(P), kw, shape, flush=True)
    B.pf_run()
    for k, (a, b) in enumerate(pp_pairs):
        got = B.pf_fragment_probs(k)
        want = oracle.pf_fragments(a, b, frags[k], po)
        assert len(got) == len(frags[k])
        assert abs(got[0] - 1.0) <= 1e-9, (k, got[0])  # the whole alignment
        for f, x, y in zip(frags[k], got, want):
            assert _close(x, y), (k, f, x, y)
    return B


@pytest.mark.parametrize("variant", ["default", "band", "scale"])
def test_fragments_tiny(variant):
    kw = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0)
    if variant == "band":
        kw.update(min_trace_prob=0.01)
    elif variant == "scale":
        kw.update(pf_scale=0.5)
    B = _check([tiny_pair(8100 + s) for s in range(12)], kw, 8, seed=len(variant))
    B.destroy()


def test_fragments_one_workgroup_per_pair():
    kw = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0)
    B = _check([tiny_pair(8300 + s) for s in range(6)], kw, 6, seed=3, shape={"pf_wg_per_pair": 1})
    B.destroy()


def test_fragments_workload_pairs():
    """pf16x150 pairs (fixed-column group and top-level kernels) and the
    longest grad256 pair (top level on the general kernels)."""
    pps, pairs, span = synth.workload("pf16x150", limit_pairs=2)
    B = _check([(pps[a], pps[b]) for a, b in pairs], dict(max_bp_span=span), 3, seed=11)
    B.destroy()
    pps, pairs, span = synth.workload("grad256")
    a, b = max(pairs, key=lambda ab: len(pps[ab[0]]) + len(pps[ab[1]]))
    B = _check([(pps[a], pps[b])], dict(max_bp_span=span), 2, seed=12)
    B.destroy()


def test_fragment_list_grows_and_clears():
    kw = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0)
    pp = [tiny_pair(8400 + s) for s in range(3)]
    pa = api.locarna_p_params(**kw)
    po = oracle.locarna_p_params(**kw)
    P = [api.Pair(a, b, pa) for a, b in pp]
    B = api.Batch(P)
    B.pf_run()
    assert B.pf_fragment_probs(0) == []
    f1 = [(1, 2, 1, 2)]
    B.pf_set_fragments(0, f1)
    B.pf_run()
    assert _close(B.pf_fragment_probs(0)[0], oracle.pf_fragments(pp[0][0], pp[0][1], f1, po)[0])
    # a longer list than the buffers were laid out for
    f2 = [(1, 2, 1, 2), (2, 3, 1, 3), (1, P[0].lenA, 2, P[0].lenB)]
    B.pf_set_fragments(0, f2)
    B.pf_run()
    for x, y in zip(B.pf_fragment_probs(0), oracle.pf_fragments(pp[0][0], pp[0][1], f2, po)):
        assert _close(x, y)
    B.pf_set_fragments(0, [])
    B.pf_run()
    assert B.pf_fragment_probs(0) == []
    with pytest.raises(api.LcgError):
        B.pf_set_fragments(0, [(0, 1, 1, 1)])
    with pytest.raises(api.LcgError):
        B.pf_set_fragments(0, [(1, P[0].lenA + 1, 1, 1)])
    B.destroy()
