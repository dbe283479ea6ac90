"""Pins the oracle on the edge-case pairs of tests/edge.py against
exhaustive enumeration (tests/bruteforce.py): optimal global and
sequence-local scores and the traced alignment, the LocARNA-P partition
function and its arc-match / base-match probabilities."""
import math

import pytest

import bruteforce as bf
from edge import CASES
from oracle import oracle

KW = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0)
LOCAL = dict(sequ_local=1, indel=-150, indel_opening=-200)


def _problem(r, alpha, local=False):
    return bf.Problem(r.lenA, r.lenB, r.sigma, r.gapA, r.gapB, alpha, r.arcmatches, r.am_scores,
                      r.min_col, r.max_col, local=local)


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("local", [False, True])
def test_edge_optimum(name, local):
    a, b = CASES[name]
    po = oracle.locarna_params(**KW, **(LOCAL if local else {}))
    r = oracle.align(a, b, po)
    P = _problem(r, po.indel_opening, local)
    if local:
        best, ends = bf.best_local(P)
        assert r.score == best
        if best > 0:
            # the traced alignment ends in the first optimal cell in row-major order
            last = [e for e in r.edges if e[0] > 0 and e[1] > 0][-1]
            assert last == min(ends)
        else:
            assert r.edges == []
    else:
        best, edges, arcs, nal = bf.best_global(P)
        assert nal > 0
        assert r.score == best
        assert r.edges == edges


@pytest.mark.parametrize("name", sorted(CASES))
def test_edge_partition_function(name):
    a, b = CASES[name]
    po = oracle.locarna_p_params(**KW, tau=50)
    r = oracle.pf(a, b, po)
    P = _problem(r, po.indel_opening)
    Z, am, full, ref, openings = bf.pf_details(P, float(po.temperature_alipf), po.min_am_prob)
    assert abs(math.log(r.pf) - math.log(Z)) <= 1e-12 * max(1.0, abs(math.log(Z)))
    for k in range(len(am)):
        assert abs(r.am_probs[k] - am[k]) <= 1e-11, k
    # summation order differs from the enumeration's (about 1e-12 for the
    # identical pair, whose Z is large, and for the expected gap openings)
    tol = 1e-11
    W = r.lenB + 1
    for i in range(1, r.lenA + 1):
        for j in range(1, r.lenB + 1):
            assert abs(r.bm_full[i * W + j] - full.get((i, j), 0.0)) <= tol, (i, j)
            assert abs(r.bm_probs[i * W + j] - ref.get((i, j), 0.0)) <= tol, (i, j)
    assert abs(r.openings - openings) <= 1e-10 * max(1.0, openings)
