"""Pins the oracle's structure-local alignment (locarna --struct-local:
aligner.cc:118-128, 264-565, 574-610, 1012-1025, 1233-1343) by exhaustive
enumeration (tests/bruteforce.py: inside every arc match one excluded block
per sequence, at the exclusion score, no gap costs, no edges)."""
import pytest

import bruteforce as bf
from oracle import oracle
from tiny import tiny_pair


def _problem(r, p, local):
    return bf.Problem(r.lenA, r.lenB, r.sigma, r.gapA, r.gapB, p.indel_opening, r.arcmatches, r.am_scores,
                      r.min_col, r.max_col, local=local, struct_local=True, exclusion=p.exclusion)


@pytest.mark.parametrize("exclusion", [0, -60, -400])
@pytest.mark.parametrize("local", [False, True], ids=["global", "local"])
@pytest.mark.parametrize("seed", range(8))
def test_struct_local_optimum(seed, local, exclusion):
    a, b = tiny_pair(seed + 4000, arcs=(2, 4))
    p = oracle.locarna_params(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, sequ_local=int(local),
                              indel=-150, indel_opening=-200, struct_local=1, exclusion=exclusion)
    r = oracle.align(a, b, p)
    P = _problem(r, p, local)
    best = bf.best_local(P)[0] if local else bf.best_global(P)[0]
    assert r.score == best


@pytest.mark.parametrize("seed", range(6))
def test_struct_local_D_is_the_best_excluded_interior(seed):
    """D(am) = s(am) + the best interior with up to one excluded block per
    sequence, for every arc match."""
    a, b = tiny_pair(seed + 4100, nA=7, nB=7, arcs=(3, 4))
    p = oracle.locarna_params(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, indel=-150, indel_opening=-200,
                              struct_local=1, exclusion=-50)
    r = oracle.align(a, b, p)
    P = _problem(r, p, False)
    memo = {}
    for k, (al, ar, bl, br) in enumerate(r.arcmatches):
        inner = bf.group_alignments(P, al, bl, ar - 1, br - 1, False, memo)
        assert r.D[k] == max(sc for sc, _, _ in inner) + r.am_scores[k], k


def test_struct_local_with_nolp_or_stacking_refused():
    a, b = tiny_pair(4200)
    for kw in (dict(no_lonely_pairs=1), dict(stacking=1)):
        with pytest.raises(RuntimeError):
            oracle.align(a, b, oracle.locarna_params(struct_local=1, **kw))


@pytest.mark.parametrize("seed", range(8))
def test_struct_local_trace_is_an_optimal_alignment(seed):
    """The traced edges (excluded positions have none) are those of an
    optimal structure-local alignment."""
    a, b = tiny_pair(seed + 4300, arcs=(2, 4))
    p = oracle.locarna_params(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, indel=-150, indel_opening=-200,
                              struct_local=1, exclusion=-20)
    r = oracle.align(a, b, p)
    P = _problem(r, p, False)
    alns = bf.group_alignments(P, 0, 0, P.n, P.m, True, {})
    best = max(sc for sc, _, _ in alns)
    optimal = [bf.edges(P, st, 0, 0, True)[0] for sc, st, _ in alns if sc == best]
    assert r.score == best and r.edges in optimal
