"""--stacking on the host (CPU, no device calls): the pp reader's #STACK
rule (rna_data.cc:1083-1100) and the combinations the library refuses
(device compaction, arc-match score files, LocARNA-P is refused at run
time)."""
import pytest

from locarna_amd import api
from locarna_amd.synth import pp_text

SEQ = "GGGAAAACCC"


def _params(stacking, **kw):
    d = dict(min_trace_prob=0.0, stacking=stacking, exp_prob=0.1, max_bp_span=-1)
    d.update(kw)
    return api.locarna_params(**d)


def test_stack_keyword_required():
    bad = "#PP 2.0\n\na " + SEQ + "\n\n#END\n\n#SECTION BASEPAIRS\n\n2 9 0.8 0.6\n3 8 0.75\n\n#END\n"
    with pytest.raises(api.LcgError, match="stack keyword is missing"):
        api.Pair(bad, bad, _params(1))
    # without stacking the fourth column is not read
    api.Pair(bad, bad, _params(0)).free()


def test_stacking_arcmatch_scores_unchanged():
    """The plain arc-match scores do not depend on the stacking option."""
    pp = pp_text("a", SEQ, [(2, 9, 0.8, 0.6), (3, 8, 0.75)], bpcut=0.0005)
    a = api.Pair(pp, pp, _params(1))
    b = api.Pair(pp, pp, _params(0))
    assert a.arcmatches() == b.arcmatches()
    a.free()
    b.free()


def test_refused_combinations():
    pp = pp_text("a", SEQ, [(2, 9, 0.8, 0.6), (3, 8, 0.75)], bpcut=0.0005)
    with pytest.raises(api.LcgError):
        api.Pair(pp, pp, _params(1), device_compaction=True)
    with pytest.raises(api.LcgError):
        api.Pair.from_arcmatch_scores(pp, pp, "2 9 2 9 100\n", _params(1))
    with pytest.raises(api.LcgError):
        api.Pair(pp, pp, _params(4))  # unknown bit


def test_struct_local_refusals_on_the_host():
    """--struct-local with --noLP or stacking, or with the device
    compaction, is refused before any device work."""
    from tiny import tiny_pair
    a, b = tiny_pair(4600)
    for kw in (dict(no_lonely_pairs=1), dict(stacking=1)):
        with pytest.raises(api.LcgError):
            api.Pair(a, b, api.locarna_params(struct_local=1, **kw))
    with pytest.raises(api.LcgError, match="host"):
        api.Pair(a, b, api.locarna_params(struct_local=1), device_compaction=True)
