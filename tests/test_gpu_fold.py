"""Base-pair probabilities on the GPU (lcg_fold_run, the input step that
locarna_rnafold_pp performs for mlocarna) against the folding oracle, whose
recursions are pinned by enumeration (tests/test_oracle_fold.py): pair and
stacked-pair probabilities within 1e-10, log Z within 1e-12 relative, and
the pp texts."""
import math
import random

import numpy as np
import pytest

import fold_brute
from locarna_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def _parse(text):
    head, pairs = [], {}
    body = False
    for line in text.split("\n"):
        if body and line and line[0].isdigit():
            f = line.split()
            pairs[(int(f[0]), int(f[1]))] = f[2:]
        else:
            head.append(line)
        if line.startswith("#BPCUT"):
            body = True
    return head, pairs


def _same_pp(got, want, P, cut):
    """Texts equal except where a probability lies within 1e-9 of the cut
    or of a rounding boundary of the 4-digit format."""
    hg, pg = _parse(got)
    hw, pw = _parse(want)
    assert hg == hw
    for key in set(pg) ^ set(pw):
        assert abs(P[key] - cut) < 1e-9, key
    for key in set(pg) & set(pw):
        for a, b in zip(pg[key], pw[key]):
            if a != b:
                assert abs(float(a) - float(b)) <= 1.01e-4 * max(abs(float(a)), abs(float(b))), (key, a, b)
        assert len(pg[key]) == len(pw[key]) or abs(P[key] - cut) < 1e-9


def _check(seqs, span, min_prob=0.0005, stacking=True, names=None):
    names = names or ["s%03d" % k for k in range(len(seqs))]
    print("fold", len(seqs), "sequences, span", span, flush=True)
    F = api.Fold(seqs, max_bp_span=span, min_prob=min_prob, stacking=stacking)
    for k, seq in enumerate(seqs):
        P, P2 = F.probs(k)
        Po, P2o, lz = oracle.fold(seq, span)
        assert np.abs(P - Po).max() <= 1e-10, k
        assert np.abs(P2 - P2o).max() <= 1e-10, k
        assert abs(F.log_z(k) - lz) <= 1e-12 * max(1.0, abs(lz)), k
        _same_pp(F.pp(k, names[k]), oracle.fold_pp(names[k], seq, Po, P2o, min_prob, stacking), Po, min_prob)
    return F


def test_fold_allpairs64_sequences():
    seqs = synth.sequences(2, 64, 60, 200, 0.35, 0.65)
    F = _check(seqs, 100)
    assert F.stats()["device_ms"] > 0


def test_fold_long500_sequences():
    seqs = synth.sequences(3, 8, 500, 500, 0.5, 0.5)[:3]
    _check(seqs, 200, stacking=False)


def test_fold_tiny_matches_enumeration():
    rng = random.Random(11)
    seqs = ["".join(rng.choice("GGCCAU") for _ in range(rng.randint(8, 17))) for _ in range(12)]
    for span in (-1, 9):
        F = api.Fold(seqs, max_bp_span=span)
        for k, seq in enumerate(seqs):
            Z, Pb, P2b = fold_brute.fold(seq, span)
            P, P2 = F.probs(k)
            assert abs(F.log_z(k) - math.log(Z)) <= 1e-12
            for i in range(len(seq) + 1):
                for j in range(len(seq) + 1):
                    assert abs(P[i][j] - Pb.get((i, j), 0.0)) <= 1e-12
                    assert abs(P2[i][j] - P2b.get((i, j), 0.0)) <= 1e-12


def test_fold_edge_cases():
    seqs = ["", "A", "ACGU", "NNNNNNNNNNNN", "gggaaaucccuuu", "GGGAAATCCC", "GGGGAAAACCCC" * 3,
            synth.sequences(4, 1, 150, 150, 0.5, 0.5)[0]]
    _check(seqs, -1)
    _check(seqs, 5, stacking=False)
    F = api.Fold(seqs, max_bp_span=-1)
    assert F.log_z(0) == 0.0 and F.log_z(3) == 0.0
    assert "GGGAAAUCCCUUU" in F.pp(4, "x")
    assert F.probs(5)[0].max() > 0.1  # T read as U: the hairpin forms


def test_fold_scaled_matches_oracle():
    """A forced scale per base changes how values are stored, not the
    probabilities or log Z."""
    seqs = synth.sequences(2, 64, 60, 200, 0.35, 0.65)[:8]
    for g in (-0.45, 0.2):
        print("fold scale", g, flush=True)
        F = api.Fold(seqs, max_bp_span=100, stacking=True, scale_kcal_per_nt=g)
        for k, seq in enumerate(seqs):
            P, P2 = F.probs(k)
            Po, P2o, lz = oracle.fold(seq, 100)
            assert np.abs(P - Po).max() <= 1e-10
            assert np.abs(P2 - P2o).max() <= 1e-10
            assert abs(F.log_z(k) - lz) <= 1e-11 * abs(lz)


def test_fold_beyond_double_range():
    """A perfect 200-pair GC helix (log Z above 709, exp(709) being the
    double range) and two random 2,500-nt sequences fold with the scale
    chosen by bisection."""
    helix = "G" * 200 + "AAAA" + "C" * 200
    rnd = synth.sequences(9, 2, 2500, 2500, 0.5, 0.5)
    print("fold helix and 2 x 2500 nt", flush=True)
    F = api.Fold([helix], max_bp_span=-1)
    assert F.log_z(0) > 709
    P, P2 = F.probs(0)
    n = len(helix)
    assert min(P[i][n + 1 - i] for i in range(5, 195)) > 0.9
    assert P2[10][n - 9] > 0.9
    G = api.Fold(rnd, max_bp_span=150)
    for k in range(2):
        assert G.log_z(k) > 709
        P, P2 = G.probs(k)
        occ = P.sum(axis=1) + P.sum(axis=0)
        assert occ.max() <= 1.0 + 1e-9 and P.min() >= 0
        inner = np.zeros_like(P)
        inner[:-1, 1:] = P[1:, :-1]
        assert (P2 <= np.minimum(P, inner) + 1e-9).all()
    with pytest.raises(api.LcgError, match="range of double"):
        api.Fold([helix], scale_kcal_per_nt=0.5)


def test_fold_pp_feeds_the_aligner():
    seqs = synth.sequences(2, 4, 60, 120, 0.35, 0.65)
    F = api.Fold(seqs, max_bp_span=100, min_prob=0.0005)
    pps = [F.pp(k, "s%d" % k) for k in range(len(seqs))]
    pa, po = api.locarna_params(), oracle.locarna_params()
    pairs = [(pps[0], pps[1]), (pps[2], pps[3])]
    B = api.Batch([api.Pair(a, b, pa) for a, b in pairs])
    B.run(mode=0, trace=True)
    for k, (a, b) in enumerate(pairs):
        assert B.score(k) == oracle.align(a, b, po).score


def test_rnafold_pp_program(tmp_path):
    """The locarna_rnafold_pp replacement reads the "name sequence" line
    mlocarna pipes into it (mlocarna:3172) or a FASTA file and writes the
    same text as lcg_fold_write_pp."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(api.__file__), "_build", "locarna_rnafold_pp")
    seq = synth.sequences(2, 3, 60, 200, 0.35, 0.65)[2]
    out = tmp_path / "s2.pp"
    r = subprocess.run([exe, "-p", "0.0005", "--stacking", "--maxBPspan", "100", "-o", str(out)],
                       input="s2 %s\n#A1 ....\n" % seq, text=True, capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr
    F = api.Fold([seq], max_bp_span=100, min_prob=0.0005, stacking=True)
    assert out.read_text() == F.pp(0, "s2")
    fa = tmp_path / "in.fa"
    fa.write_text(">s2 description\n%s\n%s\n" % (seq[:50], seq[50:]))
    r = subprocess.run([exe, "--min-prob=0.0005", "--maxBPspan=100", "--stacking", str(fa)], text=True,
                       capture_output=True, timeout=120)
    assert r.returncode == 0 and r.stdout == F.pp(0, "s2")
    r = subprocess.run([exe, "--noLP"], input="x ACGU\n", text=True, capture_output=True, timeout=60)
    assert r.returncode != 0 and "not supported" in r.stderr
