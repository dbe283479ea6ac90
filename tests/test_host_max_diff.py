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
