"""Pins the oracle's k-best suboptimal alignments (locarna --kbest,
Aligner::suboptimal, aligner.cc:1382-1514): every enumerated alignment is
the optimum of its restriction by exhaustive enumeration
(tests/bruteforce.py best_restricted: score, the traced edges of a global
restriction, the end cell of a local one), the first is the unrestricted
optimum, and the splitting rule (split the longer side at the middle of
the alignment's extent, both halves keep that position) is checked on the
restrictions the oracle reports."""
import pytest

import bruteforce as bf
from oracle import oracle
from tiny import tiny_pair


def _problem(r, params, local):
    return bf.Problem(r.lenA, r.lenB, r.sigma, r.gapA, r.gapB, params.indel_opening, r.arcmatches,
                      r.am_scores, r.min_col, r.max_col, local=local)


@pytest.mark.parametrize("band", [0.0, 0.05], ids=["full", "band"])
@pytest.mark.parametrize("local", [1, 0], ids=["local", "global"])
@pytest.mark.parametrize("seed", range(8))
def test_kbest_against_enumeration(band, local, seed):
    """With a band, a global restriction whose end cell lies outside it has
    no alignment (minus infinity) and ends the enumeration."""
    ppA, ppB = tiny_pair(4100 + seed, 6, 7)
    params = oracle.locarna_params(min_prob=0.01, max_bp_span=-1, min_trace_prob=band, sequ_local=local,
                                   indel=-150, indel_opening=-200)
    res = oracle.suboptimal(ppA, ppB, 6, params=params)
    r = oracle.align(ppA, ppB, params)
    P = _problem(r, params, bool(local))
    assert res and res[0][0] == r.score and res[0][1] == r.edges
    for score, edges, (sA, eA, sB, eB) in res:
        best, bedges, end = bf.best_restricted(P, sA, eA, sB, eB)
        assert score == best, (sA, eA, sB, eB)
        # the traceback checks band membership (aligner.cc:1082-1229): where
        # the restriction's origin lies outside the band, a path through its
        # init row or column is scored but not traced to its start
        if not local and P.valid(sA - 1, sB - 1):
            assert edges == bedges
        else:
            last = [e for e in edges if e[0] > 0 and e[1] > 0][-1]
            assert last == end
    # each restriction after the first is a half of an earlier one: inside
    # it, with three of its four bounds
    rs = [x[2] for x in res]
    for t, (sA, eA, sB, eB) in enumerate(rs[1:], 1):
        assert any(a <= sA and eA <= e and b <= sB and eB <= f and
                   sum((sA == a, eA == e, sB == b, eB == f)) >= 3 for (a, e, b, f) in rs[:t])


def test_kbest_threshold_and_count():
    ppA, ppB = tiny_pair(4242, 7, 7)
    params = oracle.locarna_params(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, sequ_local=1)
    all5 = oracle.suboptimal(ppA, ppB, 5, params=params)
    assert len(all5) <= 5
    assert oracle.suboptimal(ppA, ppB, 2, params=params) == all5[:2]
    thr = all5[0][0] - 1
    better = oracle.suboptimal(ppA, ppB, -1, threshold=thr, params=params)
    assert all(sc > thr for sc, _, _ in better) and better == all5[:len(better)]
