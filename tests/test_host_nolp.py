"""--noLP on the host (CPU): the product's groups are the oracle's (the
group (al, bl) holds the stacks whose outer arc pair starts at (al-1,
bl-1), arc_matches.cc:313-355), the arc matches are the same as without
--noLP, and the paths that do not build --noLP groups refuse it."""
import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_stacked_pair


def _pairs(n=6):
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=n)
    return [(pps[a], pps[b]) for a, b in pairs], span


def test_groups_match_oracle():
    pp, span = _pairs()
    pp += [tiny_stacked_pair(70 + s) for s in range(6)]
    for a, b in pp:
        kw = dict(max_bp_span=span, no_lonely_pairs=1)
        P = api.Pair(a, b, api.locarna_params(**kw))
        r = oracle.align(a, b, oracle.locarna_params(**kw), trace=False)
        assert P.num_groups == r.counters["G"]
        plain = api.Pair(a, b, api.locarna_params(max_bp_span=span))
        assert plain.arcmatches() == P.arcmatches()


def test_device_compaction_refuses_nolp():
    pp, span = _pairs(2)
    pa = api.locarna_params(max_bp_span=span, no_lonely_pairs=1)
    with pytest.raises(api.LcgError, match=r"\(-1\)"):
        api.Pair(pp[0][0], pp[0][1], pa, device_compaction=True)
    with pytest.raises(api.LcgError, match=r"\(-1\)"):
        api.prepare_pairs(pp, pa, device_compaction=True, threads=2)
    # host preparation of many pairs takes it
    prep, _ = api.prepare_pairs(pp, pa, threads=2)
    assert len(prep) == 2
