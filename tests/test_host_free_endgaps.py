"""Free end gaps on the host (locarna --free-endgaps): the trace band with
free end gaps equals the oracle's band (its PFGotoh is pinned against
enumeration in test_oracle_band.py), batch preparation bands such pairs on
the host, and the combinations the reference refuses fail.  Runs on the
CPU."""
import pytest

from locarna_amd import api, synth
from oracle import oracle


@pytest.mark.parametrize("spec", ["+-+-", "-+-+", "++++", "+---", "---+"])
def test_band_with_free_endgaps_equals_oracle(spec):
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=6)
    fe = api.free_endgaps(spec)
    for a, b in pairs:
        P = api.Pair(pps[a], pps[b], api.locarna_params(max_bp_span=span, free_endgaps=fe))
        o = oracle.align(pps[a], pps[b], oracle.locarna_params(max_bp_span=span, free_endgaps=fe), trace=False)
        assert P.band() == (o.min_col, o.max_col), (spec, a, b)


def test_free_endgaps_change_the_band():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=6)
    changed = 0
    for a, b in pairs:
        x = api.Pair(pps[a], pps[b], api.locarna_params(max_bp_span=span)).band()
        y = api.Pair(pps[a], pps[b], api.locarna_params(max_bp_span=span, free_endgaps=15)).band()
        changed += x != y
    assert changed > 0


def test_batch_prepare_bands_free_endgaps_on_host():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=5)
    pp = [(pps[a], pps[b]) for a, b in pairs]
    pa = api.locarna_params(max_bp_span=span, free_endgaps=api.free_endgaps("++--"))
    many, st = api.prepare_pairs(pp, pa, device_band=False, threads=2)
    for (a, b), x in zip(pp, many):
        assert x.band() == api.Pair(a, b, pa).band()


def test_rejected_combinations():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=1)
    a, b = pps[pairs[0][0]], pps[pairs[0][1]]
    with pytest.raises(api.LcgError, match="local"):
        api.Pair(a, b, api.locarna_params(sequ_local=1, free_endgaps=1))
    with pytest.raises(ValueError):
        api.free_endgaps("++-")
    assert api.free_endgaps("+--+") == 9
