"""Host logic of the arc-match scores text (lcg_pair_prepare_scored /
lcg_pair_write_arcmatch_scores; locarna --read-arcmatch-scores and
--write-arcmatch-scores): writing a pair's arc matches and reading them back
gives the same arc matches, scores, groups and band; probabilities are
scaled and truncated; malformed lines fail with their line number.  Runs on
the CPU."""
import collections

import pytest

from locarna_amd import api, synth


def _pairs(limit=8):
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=limit)
    return [(pps[a], pps[b]) for a, b in pairs], api.locarna_params(max_bp_span=span)


def _matches(P):
    ams, sc = P.arcmatches()
    return collections.Counter(zip(ams, sc))


def test_round_trip():
    pp, pa = _pairs()
    for a, b in pp:
        P = api.Pair(a, b, pa)
        text = P.arcmatch_scores_text()
        assert text.count("\n") == P.num_arcmatches
        Q = api.Pair.from_arcmatch_scores(a, b, text, pa)
        assert Q.band() == P.band()
        assert (Q.num_arcmatches, Q.num_groups, Q.cells) == (P.num_arcmatches, P.num_groups, P.cells)
        assert _matches(Q) == _matches(P)
        # the index order is the text's line order
        assert Q.arcmatch_scores_text() == text


def test_probabilities_scaled_and_truncated():
    pp, pa = _pairs(1)
    a, b = pp[0]
    P = api.Pair(a, b, pa)
    ams, _ = P.arcmatches()
    probs = [0.1 + 0.8 * ((7 * t) % 13) / 13.0 for t in range(len(ams))]
    text = "".join("%d %d %d %d %.6f\n" % (al, ar, bl, br, p) for (al, ar, bl, br), p in zip(ams, probs))
    Q = api.Pair.from_arcmatch_scores(a, b, text, pa, probability_scale=10000)
    got = dict(zip(*Q.arcmatches()))
    for (am, p) in zip(ams, probs):
        assert got[am] == int(float("%.6f" % p) * 10000)


def test_arcs_outside_the_band_are_left_out():
    pp, pa = _pairs(1)
    a, b = pp[0]
    P = api.Pair(a, b, pa)
    mn, mx = P.band()
    # an arc match whose left ends are far outside the band
    al, bl = 1, min(P.lenB - 10, mx[1] + 20)
    extra = "%d %d %d %d 999\n" % (al, al + 5, bl, bl + 5)
    Q = api.Pair.from_arcmatch_scores(a, b, P.arcmatch_scores_text() + extra, pa)
    assert Q.num_arcmatches == P.num_arcmatches


@pytest.mark.parametrize("line,what", [("1 2 x 4 5", "Invalid line 2"), ("9 5 1 6 3", "Invalid line 2"),
                                       ("1 10 1 100000 3", "Invalid line 2"), ("1 3 5 9 2", "j - i < 3")])
def test_bad_lines(line, what):
    pp, pa = _pairs(1)
    a, b = pp[0]
    with pytest.raises(api.LcgError, match=what):
        api.Pair.from_arcmatch_scores(a, b, "1 10 2 12 5\n" + line + "\n", pa)
