"""Score range of the int32 kernels: a pair whose alignment scores could
leave the range the kernels represent exactly (|score| <= 2^27; minus
infinity at -2^29, values below -2^28 read as minus infinity) is refused at
preparation with LCG_E_CAPACITY instead of being narrowed, on every
preparation path: host arc matches, device compaction, an arc-match scores
text and the batched preparation.  The reference computes in 64-bit
InftyInt (infty_int.hh:56-591) and has no such limit.  Runs on the CPU."""
import pytest

from locarna_amd import api, synth

E_CAP = r"\(-4\)"


def _pp(limit=1):
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=limit)
    return [(pps[a], pps[b]) for a, b in pairs], span


def test_ordinary_parameters_accepted():
    pp, span = _pp(4)
    for a, b in pp:
        api.Pair(a, b, api.locarna_params(max_bp_span=span))
        api.Pair(a, b, api.locarna_params(max_bp_span=span), device_compaction=True)


@pytest.mark.parametrize("device_compaction", [False, True])
def test_huge_indel_refused(device_compaction):
    (a, b), = _pp(1)[0]
    span = _pp(1)[1]
    with pytest.raises(api.LcgError, match=E_CAP):
        api.Pair(a, b, api.locarna_params(max_bp_span=span, indel=-10 ** 6), device_compaction=device_compaction)
    # the same magnitude spread over a short pair stays in range
    sa = synth.make_pp("a", "GGGAAACCC", 100)
    api.Pair(sa, sa, api.locarna_params(indel=-10 ** 6), device_compaction=device_compaction)


def test_huge_struct_weight_refused():
    (a, b), = _pp(1)[0]
    # without a trace band, so that the arc matches stay (the band's own
    # partition function sees the struct weight too)
    pa = api.locarna_params(struct_weight=10 ** 7, min_trace_prob=0.0)
    with pytest.raises(api.LcgError, match=E_CAP):
        api.Pair(a, b, pa, device_compaction=True)
    with pytest.raises(api.LcgError, match=E_CAP):
        api.Pair(a, b, pa)
    P = api.Pair(a, b, api.locarna_params(struct_weight=10 ** 6, min_trace_prob=0.0))
    assert P.num_arcmatches > 0 and max(abs(x) for x in P.arcmatches()[1]) > 10 ** 6


def test_arcmatch_scores_text_out_of_range_refused():
    (a, b), = _pp(1)[0]
    span = _pp(1)[1]
    pa = api.locarna_params(max_bp_span=span)
    text = api.Pair(a, b, pa).arcmatch_scores_text()
    lines = text.splitlines()
    f = lines[0].split()
    lines[0] = " ".join(f[:4] + [str(1 << 28)])
    with pytest.raises(api.LcgError, match=E_CAP):
        api.Pair.from_arcmatch_scores(a, b, "\n".join(lines) + "\n", pa)
    lines[0] = " ".join(f[:4] + [str(-(1 << 28))])
    with pytest.raises(api.LcgError, match=E_CAP):
        api.Pair.from_arcmatch_scores(a, b, "\n".join(lines) + "\n", pa)
    # a large score that keeps every alignment within range is accepted
    lines[0] = " ".join(f[:4] + [str(1 << 20)])
    api.Pair.from_arcmatch_scores(a, b, "\n".join(lines) + "\n", pa)


def test_batched_preparation_reports_capacity():
    pp, span = _pp(3)
    with pytest.raises(api.LcgError, match=E_CAP):
        api.prepare_pairs(pp, api.locarna_params(max_bp_span=span, indel=-10 ** 6), threads=2)
    with pytest.raises(api.LcgError, match=E_CAP):
        api.prepare_pairs(pp, api.locarna_params(max_bp_span=span, indel=-10 ** 6), device_compaction=True,
                          threads=2)
