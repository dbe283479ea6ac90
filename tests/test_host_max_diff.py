"""--max-diff (band around the diagonal without a reference alignment,
TraceController::constrain_wo_ref, trace_controller.cc:361-390),
--max-diff-am and --max-diff-at-am (ArcMatches::is_valid_arcmatch,
arc_matches.cc:19-48) on the host: band, trace band on top of it, arc
matches and scores equal the oracle's; the options take effect.  Runs on
the CPU."""
import pytest

from locarna_amd import api, synth
from oracle import oracle

CASES = [dict(max_diff=12), dict(max_diff=3), dict(max_diff_am=4), dict(max_diff_at_am=6),
         dict(max_diff=20, max_diff_am=10, max_diff_at_am=15)]


@pytest.mark.parametrize("kw", CASES, ids=[str(c) for c in CASES])
def test_band_and_arcmatches_equal_oracle(kw):
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=6)
    for a, b in pairs:
        pa = api.locarna_params(max_bp_span=span, **kw)
        po = oracle.locarna_params(max_bp_span=span, **kw)
        P = api.Pair(pps[a], pps[b], pa)
        o = oracle.align(pps[a], pps[b], po, trace=False)
        assert P.band() == (o.min_col, o.max_col), (kw, a, b)
        ams, scores = P.arcmatches()
        assert ams == o.arcmatches and scores == o.am_scores, (kw, a, b)


def test_options_take_effect():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=4)
    a, b = pairs[0]
    base = api.Pair(pps[a], pps[b], api.locarna_params(max_bp_span=span, min_trace_prob=0.0))
    band = api.Pair(pps[a], pps[b], api.locarna_params(max_bp_span=span, min_trace_prob=0.0, max_diff=5))
    assert band.band_cells < base.band_cells
    am = api.Pair(pps[a], pps[b], api.locarna_params(max_bp_span=span, max_diff_am=2, max_diff_at_am=3))
    full = api.Pair(pps[a], pps[b], api.locarna_params(max_bp_span=span))
    assert am.num_arcmatches < full.num_arcmatches
    for (al, ar, bl, br) in am.arcmatches()[0]:
        assert abs((ar - al) - (br - bl)) <= 2 and abs(al - bl) <= 3 and abs(ar - br) <= 3


def _max_diff_band(la, lb, delta):
    a = synth.pp_text("a", ("ACGU" * la)[:la], [])
    b = synth.pp_text("b", ("GUCA" * lb)[:lb], [])
    o = oracle.align(a, b, oracle.locarna_params(min_trace_prob=0.0, max_diff=delta), trace=False)
    return o.min_col, o.max_col


@pytest.mark.parametrize("la,lb", [(1, 1), (7, 7), (30, 30), (12, 31), (31, 12), (1, 9), (9, 1), (25, 26)])
@pytest.mark.parametrize("delta", [0, 1, 3, 10, 100])
def test_max_diff_band_geometry(la, lb, delta):
    # Pins the oracle's band (row bounds by integer ceil / floor) by cell
    # membership: (i, j) is in the band iff it lies within a half-width w of
    # the diagonal j = i lb / la, tested as (la + lb) |j la - i lb| <= Y with
    # Y = 2 delta la lb, raised for unequal lengths to (la + lb) max(la, lb)
    # / 2 (rounded down), i.e. w >= half a column and w >= half the slope.
    mn, mx = _max_diff_band(la, lb, delta)
    Y = 2 * delta * la * lb
    if la != lb:
        Y = max(Y, (la + lb) * max(la, lb) // 2)
    for i in range(la + 1):
        got = {j for j in range(lb + 1) if mn[i] <= j <= mx[i]}
        want = {j for j in range(lb + 1) if (la + lb) * abs(j * la - i * lb) <= Y}
        assert got == want, (i, mn[i], mx[i])
        if la == lb:
            # equal lengths: the documented |i - j| <= max-diff
            assert got == {j for j in range(lb + 1) if abs(i - j) <= delta}, i
        # every row non-empty, consecutive rows overlap or touch: a path exists
        assert got
        if i > 0:
            assert mn[i] <= mx[i - 1] + 1 and mn[i - 1] <= mn[i] and mx[i - 1] <= mx[i]
    assert mn[0] == 0 and mx[la] == lb
