"""Pins the oracle's fragment match probabilities (locarna_p
--fragment-match-probs, AlignerP::compute_fragment_match_prob,
aligner_p.icc:1440-1479) against exhaustive enumeration on tiny inputs
(tests/bruteforce.py: the weight of the structure alignments with a group
passing through the cells (i-1, k-1) and (j, l), gap runs through either
cell split in two)."""
import pytest

import bruteforce as bf
from oracle import oracle
from tiny import tiny_pair


def _params(**kw):
    d = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, tau=50)
    d.update(kw)
    return oracle.locarna_p_params(**d)


def _all_fragments(lenA, lenB, count):
    frags = [(i, j, k, l) for i in range(1, lenA + 1) for j in range(i, lenA + 1)
             for k in range(1, lenB + 1) for l in range(k, lenB + 1)]
    return frags[::max(1, len(frags) // count)]


CASES = [
    ("full", {}),
    ("band", dict(min_trace_prob=0.01)),
    ("scale", dict(pf_scale=0.5)),
    ("cold", dict(temperature_alipf=100, indel=-100, indel_opening=-200)),
]


@pytest.mark.parametrize("label,kw", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("seed", range(5))
def test_fragments_against_enumeration(label, kw, seed):
    ppA, ppB = tiny_pair(5000 + seed * 17 + sum(map(ord, label)))
    params = _params(**kw)
    r = oracle.pf(ppA, ppB, params)
    P = bf.Problem(r.lenA, r.lenB, r.sigma, r.gapA, r.gapB, params.indel_opening, r.arcmatches,
                   r.am_scores, r.min_col, r.max_col)
    frags = _all_fragments(r.lenA, r.lenB, 24)
    got = oracle.pf_fragments(ppA, ppB, frags, params)
    T = float(params.temperature_alipf)
    for f, g, e in zip(frags, got, bf.fragment_probabilities(P, T, frags)):
        assert abs(g - e) <= 1e-12, f


def test_whole_sequences_and_single_cells():
    """The fragment [1..n] x [1..m] is the whole alignment (probability 1);
    a fragment of one base each is the probability that the alignment
    passes through (i-1, k-1) and (i, k) in one group."""
    ppA, ppB = tiny_pair(4242)
    params = _params()
    r = oracle.pf(ppA, ppB, params)
    got = oracle.pf_fragments(ppA, ppB, [(1, r.lenA, 1, r.lenB)], params)
    assert abs(got[0] - 1.0) <= 1e-12


def test_gap_split_term_is_observed():
    """The enumeration without the split rule (a run through a cut counted
    once) disagrees with the oracle on some fragment of these inputs: the
    check sees that term."""
    differs = False
    for seed in range(4):
        ppA, ppB = tiny_pair(7100 + seed)
        params = _params()
        r = oracle.pf(ppA, ppB, params)
        P = bf.Problem(r.lenA, r.lenB, r.sigma, r.gapA, r.gapB, params.indel_opening, r.arcmatches,
                       r.am_scores, r.min_col, r.max_col)
        frags = _all_fragments(r.lenA, r.lenB, 12)
        got = oracle.pf_fragments(ppA, ppB, frags, params)
        T = float(params.temperature_alipf)
        for g, e in zip(got, bf.fragment_probabilities(P, T, frags, split=False)):
            if abs(e - g) > 1e-9:
                differs = True
    assert differs


def test_scale_invariance():
    ppA, ppB = tiny_pair(31337)
    frags = [(2, 4, 1, 3), (1, 3, 2, 5), (3, 5, 3, 5)]
    a = oracle.pf_fragments(ppA, ppB, frags, _params(pf_scale=1.0))
    b = oracle.pf_fragments(ppA, ppB, frags, _params(pf_scale=0.3))
    for x, y in zip(a, b):
        assert abs(x - y) <= 1e-12 * max(1.0, x)


def test_out_of_range_refused():
    ppA, ppB = tiny_pair(1)
    with pytest.raises(RuntimeError):
        oracle.pf_fragments(ppA, ppB, [(0, 2, 1, 2)], _params())
    with pytest.raises(RuntimeError):
        oracle.pf_fragments(ppA, ppB, [(3, 2, 1, 2)], _params())
