"""Host preprocessing of the product (pp reader, trace band, arc matches,
arc-match scores) against the oracle, bit for bit.  CPU only."""
import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_pair


def _check(ppA, ppB, kw_o, kw_p):
    o = oracle.align(ppA, ppB, oracle.locarna_params(**kw_o), trace=False)
    p = api.Pair(ppA, ppB, api.locarna_params(**kw_p))
    assert (p.lenA, p.lenB) == (o.lenA, o.lenB)
    mn, mx = p.band()
    assert mn == o.min_col and mx == o.max_col
    ams, scores = p.arcmatches()
    assert ams == o.arcmatches
    assert scores == o.am_scores
    return o, p


@pytest.mark.parametrize("seed", range(6))
def test_random_pairs_default_flags(seed):
    seqs = synth.sequences(100 + seed, 2, 60, 140, 0.35, 0.65)
    ppA = synth.make_pp("a", seqs[0], 100)
    ppB = synth.make_pp("b", seqs[1], 100)
    _check(ppA, ppB, {}, {})


@pytest.mark.parametrize("seed", range(4))
def test_random_pairs_locarna_p_flags(seed):
    seqs = synth.sequences(200 + seed, 2, 60, 120, 0.5, 0.5)
    ppA = synth.make_pp("a", seqs[0], 100)
    ppB = synth.make_pp("b", seqs[1], 100)
    kw = dict(tau=50, band_extended=0, min_trace_prob=1e-5)
    _check(ppA, ppB, kw, kw)


@pytest.mark.parametrize("seed", range(8))
def test_tiny_pairs_variants(seed):
    ppA, ppB = tiny_pair(300 + seed)
    kw = dict(min_prob=0.01, max_bp_span=-1, tau=100, use_ribosum=seed % 2,
              min_trace_prob=0.05 * (seed % 3), sequ_local=seed % 2)
    _check(ppA, ppB, kw, kw)


def test_workload_pair120():
    pps, pairs, span = synth.workload("pair120")
    _check(pps[0], pps[1], {}, {})


def test_local_band_has_empty_rows_and_packs():
    """Sequence-local bands contain empty rows (min_col > max_col) at the
    sequence ends; the product must pack them (regression: band cell
    numbering)."""
    pps, pairs, _ = synth.workload("allpairs64", limit_pairs=10)
    a, b = pairs[9]
    o, p = _check(pps[a], pps[b], dict(sequ_local=1), dict(sequ_local=1))
    assert any(mn > mx for mn, mx in zip(o.min_col, o.max_col))
