"""Pins the oracle's LocARNA-P path (inside partition function, arc-match
and base-match probabilities, expected gap openings) against exhaustive
enumeration on tiny inputs (tests/bruteforce.py)."""
import math

import pytest

import bruteforce as bf
from oracle import oracle
from tiny import tiny_pair


def _params(**kw):
    d = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, tau=50)
    d.update(kw)
    return oracle.locarna_p_params(**d)


CASES = [
    ("full", {}),
    ("band", dict(min_trace_prob=0.01)),
    ("scale", dict(pf_scale=0.5)),
    ("cold", dict(temperature_alipf=100, indel=-100, indel_opening=-200)),
    ("noribosum", dict(use_ribosum=0, struct_weight=150, tau=100)),
    ("amthreshold", dict(min_am_prob=0.04)),
]


@pytest.mark.parametrize("label,kw", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("seed", range(8))
def test_pf_against_enumeration(label, kw, seed):
    ppA, ppB = tiny_pair(9000 + seed * 13 + sum(map(ord, label)))
    params = _params(**kw)
    r = oracle.pf(ppA, ppB, params)
    P = bf.Problem(r.lenA, r.lenB, r.sigma, r.gapA, r.gapB, params.indel_opening, r.arcmatches,
                   r.am_scores, r.min_col, r.max_col)
    Z, am, full, ref, openings = bf.pf_details(P, float(params.temperature_alipf), params.min_am_prob)
    # Z carries pf_scale^(1 + #arc matches) factors; compare log Z at pf_scale 1
    if params.pf_scale == 1.0:
        assert abs(math.log(r.pf) - math.log(Z)) <= 1e-12 * max(1.0, abs(math.log(Z)))
    for k in range(len(am)):
        assert abs(r.am_probs[k] - am[k]) <= 1e-12, k
    W = r.lenB + 1
    for i in range(1, r.lenA + 1):
        for j in range(1, r.lenB + 1):
            assert abs(r.bm_full[i * W + j] - full.get((i, j), 0.0)) <= 1e-12, (i, j)
            assert abs(r.bm_probs[i * W + j] - ref.get((i, j), 0.0)) <= 1e-12, (i, j)
    assert abs(r.openings - openings) <= 1e-12 * max(1.0, openings)


def test_include_am_in_bm():
    ppA, ppB = tiny_pair(777)
    params = _params()
    r0 = oracle.pf(ppA, ppB, params)
    r1 = oracle.pf(ppA, ppB, params, include_am_in_bm=True)
    W = r0.lenB + 1
    add = {}
    for k, (al, ar, bl, br) in enumerate(r0.arcmatches):
        add[(al, bl)] = add.get((al, bl), 0.0) + r0.am_probs[k]
        add[(ar, br)] = add.get((ar, br), 0.0) + r0.am_probs[k]
    for i in range(1, r0.lenA + 1):
        for j in range(1, r0.lenB + 1):
            assert abs(r1.bm_probs[i * W + j] - r0.bm_probs[i * W + j] - add.get((i, j), 0.0)) < 1e-15
