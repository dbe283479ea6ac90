"""--noLP (no lonely base pairs) in the oracle's integer DP against
exhaustive enumeration on tiny inputs with stacked arcs: the optimum over
structure alignments whose every matched arc pair has its directly inner or
directly outer arc pair matched too (aligner.cc:612-657, the D-fill of
stacks), global and sequence-local; the traced alignment is one of the
optimal ones (aligner.cc:1030-1078); without stackable arc pairs no arc
match is used."""
import pytest

import bruteforce as bf
from oracle import oracle
from tiny import tiny_stacked_pair


def problem_from(r, alpha, local=False, nolp=True):
    return bf.Problem(r.lenA, r.lenB, r.sigma, r.gapA, r.gapB, alpha, r.arcmatches, r.am_scores, r.min_col,
                      r.max_col, local=local, no_lonely_pairs=nolp)


CASES = [
    ("ribosum", dict(min_trace_prob=0.0)),
    ("strong_structure", dict(min_trace_prob=0.0, struct_weight=400, tau=100)),
    ("matchmismatch", dict(use_ribosum=0, indel=-50, indel_opening=-100, struct_weight=150, min_trace_prob=0.0)),
]


@pytest.mark.parametrize("label,kw", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("seed", range(10))
def test_nolp_global(label, kw, seed):
    ppA, ppB = tiny_stacked_pair(500 + seed * 13 + len(label))
    params = oracle.locarna_params(min_prob=0.01, max_bp_span=-1, no_lonely_pairs=1, **kw)
    r = oracle.align(ppA, ppB, params)
    P = problem_from(r, params.indel_opening)
    best, opt_edges = bf.optimal_global_edges(P)
    assert r.score == best
    assert r.edges in opt_edges
    # the unrestricted optimum is never worse
    free = oracle.align(ppA, ppB, oracle.locarna_params(min_prob=0.01, max_bp_span=-1, **kw))
    assert free.score >= r.score


@pytest.mark.parametrize("seed", range(10))
def test_nolp_local(seed):
    ppA, ppB = tiny_stacked_pair(900 + seed)
    params = oracle.locarna_params(min_prob=0.01, max_bp_span=-1, no_lonely_pairs=1, sequ_local=1,
                                   min_trace_prob=0.0, indel=-150, indel_opening=-200, struct_weight=300)
    r = oracle.align(ppA, ppB, params)
    best, ends = bf.best_local(problem_from(r, params.indel_opening, local=True))
    assert r.score == best


def test_nolp_uses_stacks():
    # over many inputs, noLP optima that use arc matches use them in stacks
    used = 0
    for seed in range(40):
        ppA, ppB = tiny_stacked_pair(3000 + seed)
        params = oracle.locarna_params(min_prob=0.01, max_bp_span=-1, no_lonely_pairs=1, min_trace_prob=0.0,
                                       struct_weight=400, tau=100)
        r = oracle.align(ppA, ppB, params)
        pairsA = [i for i, c in enumerate(r.structA) if c == '(']
        if pairsA:
            used += 1
        for i, c in enumerate(r.structA):
            if c == '(':
                # its partner's neighbours are paired too (a stack)
                assert (i + 1 < len(r.structA) and r.structA[i + 1] == '(') or (i > 0 and r.structA[i - 1] == '(')
    assert used > 5
