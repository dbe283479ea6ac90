"""Host logic of lcg_pairs_prepare without the device band: pairs prepared
together on host threads equal pairs prepared one by one (band, arc
matches, groups, DP work), for host and device compaction, and an input
error names its pair.  Runs on the CPU."""
import pytest

from locarna_amd import api, synth


@pytest.mark.parametrize("device_compaction", [False, True])
def test_batch_prepare_equals_single(device_compaction):
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=60)
    pp = [(pps[a], pps[b]) for a, b in pairs]
    pa = api.locarna_params(max_bp_span=span)
    many, st = api.prepare_pairs(pp, pa, device_compaction=device_compaction, threads=4)
    assert (st["band_ms"], st["device_banded"], st["host_banded"]) == (0.0, 0, len(pp))
    for k, (a, b) in enumerate(pp):
        one = api.Pair(a, b, pa, device_compaction=device_compaction)
        x = many[k]
        assert x.band() == one.band(), k
        assert (x.lenA, x.lenB, x.num_arcmatches, x.num_groups, x.cells) == \
            (one.lenA, one.lenB, one.num_arcmatches, one.num_groups, one.cells), k
        if not device_compaction:
            assert x.arcmatches() == one.arcmatches(), k


def test_batch_prepare_reports_bad_pair():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=3)
    pp = [(pps[a], pps[b]) for a, b in pairs]
    pp[1] = (pp[1][0], "not a pp file")
    with pytest.raises(api.LcgError, match="pair 1"):
        api.prepare_pairs(pp, api.locarna_params(max_bp_span=span), threads=2)
