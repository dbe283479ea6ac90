"""mlocarna's pipeline from sequences on the GPU (lcg_mlocarna_run): the
input pp texts equal the folding step's, every all-pairs score equals the
oracle's alignment of those texts, the guide tree equals the oracle's
UPGMA of the scores, and every merge's consensus pp text equals the
oracle's progressive alignment along that tree."""
import pytest

from locarna_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def test_mlocarna_pipeline_matches_oracle():
    seqs = synth.sequences(2, 64, 60, 200, 0.35, 0.65)[:10]
    names = ["seq%02d" % k for k in range(len(seqs))]
    print("mlocarna pipeline", len(seqs), "sequences", flush=True)
    M = api.Mlocarna(names, seqs, max_bp_span=100)
    n = len(seqs)
    F = api.Fold(seqs, max_bp_span=100, min_prob=0.0005)
    pps = [M.pp(k) for k in range(n)]
    assert pps == [F.pp(k, names[k]) for k in range(n)]
    po = oracle.locarna_params(max_bp_span=100)
    S = M.scores()
    pairs = [(a, b) for a in range(n) for b in range(a)]
    ref = oracle.align_many([(pps[a], pps[b]) for a, b in pairs], po, 8)
    for (a, b), sc in zip(pairs, ref):
        assert S[a][b] == sc and S[b][a] == sc, (a, b)
    merges = M.tree()
    assert merges == [tuple(m) for m in oracle.upgma(S)]
    final, steps = oracle.progressive(pps, merges, po)
    assert M.pp(2 * n - 2) == final
    clustal = M.clustal()
    assert clustal.startswith("CLUSTAL W") and all(nm in clustal for nm in names)
    st = M.stats()
    assert st["wall_ms"] >= st["fold_ms"] + st["pairs_ms"] + st["progressive_ms"] - 1.0


def test_mlocarna_refusals():
    with pytest.raises(api.LcgError):
        api.Mlocarna(["only"], ["ACGUACGUAGCUAGCUAGC"])
