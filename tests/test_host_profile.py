"""Host preparation of pairs of alignment profiles (mlocarna's progressive
stage) against the oracle: multi-row pp reader, sum-of-pairs sigma table,
gap-frequency gap costs, row-averaged band and riboX term, bit for bit.
CPU only."""
import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_profile_pair


def _check(ppA, ppB, kw):
    o = oracle.align(ppA, ppB, oracle.locarna_params(**kw), trace=False)
    p = api.Pair(ppA, ppB, api.locarna_params(**kw))
    assert (p.lenA, p.lenB) == (o.lenA, o.lenB)
    assert p.band() == (o.min_col, o.max_col)
    ams, scores = p.arcmatches()
    assert ams == o.arcmatches and scores == o.am_scores
    sig, gA, gB, rA, rB = p.score_tables()
    assert sig == o.sigma
    assert gA[1:] == o.gapA[1:] and gB[1:] == o.gapB[1:]
    return p, rA, rB


@pytest.mark.parametrize("seed", range(10))
def test_tiny_profiles(seed):
    a, b = tiny_profile_pair(1000 + seed)
    kw = dict(min_prob=0.01, max_bp_span=-1, tau=100 * (seed % 2), use_ribosum=int(seed % 3 != 2),
              min_trace_prob=0.05 * (seed % 3), sequ_local=seed % 2)
    _check(a, b, kw)


def _merged(workload, x, y, span):
    pps, _, _ = synth.workload(workload, limit_pairs=64)
    p = oracle.locarna_params(max_bp_span=span)
    r = oracle.align(pps[x], pps[y], p)
    return oracle.consensus_pp(pps[x], pps[y], r.edges, p), pps


@pytest.mark.parametrize("local", [0, 1])
def test_consensus_profile_against_sequence_and_profile(local):
    """Profiles as the progressive stage makes them (consensus pp of an
    aligned pair) against a sequence and against another profile."""
    ab, pps = _merged("allpairs64", 1, 0, 100)
    cd, _ = _merged("allpairs64", 3, 2, 100)
    p, rA, rB = _check(ab, pps[5], dict(sequ_local=local))
    assert (rA, rB) == (2, 1)
    p, rA, rB = _check(ab, cd, dict(sequ_local=local, tau=50))
    assert (rA, rB) == (2, 2)


def test_profile_sequence_batch_mix_refused():
    """A batch holds pairs of one sigma kind (the kernels read codes or a
    column table); the refusal comes before any device work."""
    ab, pps = _merged("allpairs64", 1, 0, 100)
    pa = api.locarna_params()
    P = [api.Pair(ab, pps[5], pa), api.Pair(pps[2], pps[3], pa)]
    with pytest.raises(api.LcgError, match="profiles"):
        api.Batch(P)


def test_profile_rows_of_unequal_length_refused():
    bad = synth.pp_text_rows([("a", "ACGUA"), ("b", "ACG")], [])
    with pytest.raises(api.LcgError):
        api.Pair(bad, synth.pp_text_rows([("c", "ACGUA")], []), api.locarna_params())


def test_profiles_need_host_arc_matches():
    ab, pps = _merged("allpairs64", 1, 0, 100)
    with pytest.raises(api.LcgError, match="host"):
        api.Pair(ab, pps[5], api.locarna_params(), device_compaction=True)
