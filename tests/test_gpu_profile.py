"""Alignment of alignment profiles and mlocarna's progressive stage on the
GPU against the oracle (whose profile tables, consensus writer and UPGMA
are pinned in tests/test_oracle_profile.py): scores, alignments,
structures, D, consensus pp texts and CLUSTAL rows bit for bit."""
import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import replicate_rows, tiny_pair, tiny_profile_pair

pytestmark = pytest.mark.gpu


def _D_equal(got, want, label):
    assert len(got) == len(want), label
    for t, (x, y) in enumerate(zip(got, want)):
        if y < -(1 << 60):
            assert x < -(1 << 28), (label, t)
        else:
            assert x == y, (label, t)


def _compare(pp_pairs, kw, consensus=True):
    po, pa = oracle.locarna_params(**kw), api.locarna_params(**kw)
    ores = [oracle.align(a, b, po) for a, b in pp_pairs]
    B = api.Batch([api.Pair(a, b, pa) for a, b in pp_pairs])
    print("profile batch", len(pp_pairs), kw, flush=True)
    B.run(mode=0, trace=True)
    for k, o in enumerate(ores):
        assert B.score(k) == o.score, (k, B.score(k), o.score)
        edges, sA, sB = B.alignment(k)
        assert edges == o.edges, k
        assert (sA, sB) == (o.structA, o.structB), k
        _D_equal(B.D(k), o.D, k)
        if consensus and not kw.get("sequ_local"):
            assert B.consensus_pp(k) == oracle.consensus_pp(pp_pairs[k][0], pp_pairs[k][1], o.edges, po), k
    return B


@pytest.mark.parametrize("local", [0, 1], ids=["global", "local"])
def test_tiny_profiles(local):
    kw = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, sequ_local=local, indel=-150,
              indel_opening=-200, tau=50)
    _compare([tiny_profile_pair(2000 + s, rows=(2, 3)) for s in range(24)], kw).destroy()


def test_replicated_rows_align_as_sequences():
    pairs = [tiny_pair(2100 + s, nA=7, nB=8, arcs=(2, 4)) for s in range(12)]
    kw = dict(min_prob=0.01, max_bp_span=-1, tau=100, min_trace_prob=0.01)
    single = _compare(pairs, kw)
    rep = _compare([(replicate_rows(a, 3), replicate_rows(b, 2)) for a, b in pairs], kw, consensus=False)
    for k in range(len(pairs)):
        assert rep.score(k) == single.score(k) and rep.alignment(k) == single.alignment(k)
    single.destroy()
    rep.destroy()


def _merge_texts(names, span):
    pps, _, _ = synth.workload("allpairs64", limit_pairs=2016)
    p = oracle.locarna_params(max_bp_span=span)
    out = []
    for x, y in names:
        r = oracle.align(pps[x], pps[y], p)
        out.append(oracle.consensus_pp(pps[x], pps[y], r.edges, p))
    return out, pps


@pytest.mark.parametrize("local", [0, 1], ids=["global", "local"])
def test_workload_profiles(local):
    """Profiles of two and four rows (consensus pp texts of aligned
    allpairs64 pairs) against sequences and against each other."""
    (ab, cd, ef), pps = _merge_texts([(1, 0), (3, 2), (5, 4)], 100)
    po = oracle.locarna_params(max_bp_span=100)
    r = oracle.align(ab, cd, po)
    abcd = oracle.consensus_pp(ab, cd, r.edges, po)
    pairs = [(ab, pps[7]), (pps[9], cd), (ab, cd), (cd, ef), (abcd, ef), (abcd, pps[11])]
    B = _compare(pairs, dict(max_bp_span=100, sequ_local=local))
    if not local:
        # CLUSTAL: rows of both sides, no score in the header of profiles
        txt = B.clustal(2)
        assert txt.startswith("CLUSTAL W --- LocARNA 2.0.1\n\n")
        block = txt.split("\n\n")[1]  # the first block of 120 columns
        assert [l.split()[0] for l in block.split("\n") if l.strip()] == ["s001", "s000", "s003", "s002"]
    B.destroy()


def test_progressive_matches_oracle():
    """mlocarna's progressive stage over 12 allpairs64 sequences: the
    all-pairs scores on the GPU, the guide tree, every merge's score and
    consensus pp text, and the final CLUSTAL rows."""
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=66)  # sequences 0..11
    n = 12
    texts = [pps[x] for x in range(n)]
    pa, po = api.locarna_params(max_bp_span=span), oracle.locarna_params(max_bp_span=span)
    B = api.Batch([api.Pair(texts[a], texts[b], pa) for a, b in pairs])
    B.run(mode=2)
    S = [[0] * n for _ in range(n)]
    for k, (a, b) in enumerate(pairs):
        S[a][b] = S[b][a] = B.score(k)
    B.destroy()
    want_scores = oracle.align_many([(texts[a], texts[b]) for a, b in pairs], po, nthreads=8)
    assert [S[a][b] for a, b in pairs] == want_scores
    merges = api.guide_tree_upgma(S)
    assert merges == oracle.upgma(S)
    print("progressive", n, merges, flush=True)
    R = api.Progressive(texts, merges, pa)
    final, steps = oracle.progressive(texts, merges, po)
    assert R.scores() == [sc for sc, _ in steps]
    assert R.pp(2 * n - 2) == final
    # every intermediate consensus text equals the oracle's
    text = list(texts)
    for t, (a, b) in enumerate(merges):
        r = oracle.align(text[a], text[b], po)
        text.append(oracle.consensus_pp(text[a], text[b], r.edges, po))
        assert R.pp(n + t) == text[-1], t
    aln = R.clustal()
    rows = [l.split() for l in aln.split("\n")[2:] if l.strip()]
    seqs = {"s%03d" % x: texts[x].split("\n")[2].split()[1] for x in range(n)}
    got = {}
    for name, chunk in rows:
        got[name] = got.get(name, "") + chunk
    assert {k: v.replace("-", "") for k, v in got.items()} == seqs
    st = R.stats()
    assert st["rounds"] >= 1 and st["batches"] >= st["rounds"]
    R.destroy()
