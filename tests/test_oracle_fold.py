"""The folding oracle (oracle/fold.cc) against exhaustive enumeration of
secondary structures (tests/fold_brute.py), its probability identities,
and the library's host folding (the seeded input generator, lcg_synth_bpp)
against the oracle.  CPU only."""
import math
import random

import numpy as np
import pytest

import fold_brute
from oracle import oracle


def _compare(seq, span):
    Z, P, P2 = fold_brute.fold(seq, span)
    Po, P2o, lz = oracle.fold(seq, span)
    n = len(seq)
    assert abs(lz - math.log(Z)) <= 1e-12 * max(1.0, abs(lz))
    for i in range(n + 1):
        for j in range(n + 1):
            assert abs(Po[i][j] - P.get((i, j), 0.0)) <= 1e-12, (seq, span, i, j)
            assert abs(P2o[i][j] - P2.get((i, j), 0.0)) <= 1e-12, (seq, span, i, j)


@pytest.mark.parametrize("seed", range(8))
def test_oracle_fold_matches_enumeration(seed):
    rng = random.Random(seed)
    n = rng.randint(8, 18)
    alphabet = rng.choice(["ACGU", "GGCCAU", "GCGCU"])
    seq = "".join(rng.choice(alphabet) for _ in range(n))
    for span in (-1, 7, 11, 15):
        _compare(seq, span)


def test_enumeration_reaches_multiloops():
    # two hairpins closed by an outer helix: GGG..CCC around GCAAAGC GCAAAGC
    seq = "GGGAGCAAAGCAGCAAAGCACCC"
    s = " " + seq
    multi = 0
    for st in fold_brute.structures(s, -1):
        partner = [0] * (len(s) + 1)
        for (i, j) in st:
            partner[i], partner[j] = j, i
        if any(len(fold_brute._children(partner, i, j)) >= 2 for (i, j) in st):
            multi += 1
    assert multi > 0
    _compare(seq, -1)
    _compare(seq.lower().replace("u", "t"), 20)


def test_interior_loop_limit():
    # only G and C pair (no U): the pair (1, 40) around (17, 22) leaves
    # 15 + 18 = 33 unpaired bases, more than 30, and is not an interior loop
    seq = "GG" + "A" * 14 + "GAAAAC" + "A" * 16 + "CC"
    assert len(seq) == 40
    s = " " + seq
    dropped = [st for st in fold_brute.structures(s, -1) if fold_brute.energy(s, st) is None]
    assert dropped
    _compare(seq, -1)


@pytest.fixture(scope="module")
def allpairs_seqs():
    from locarna_amd import synth
    return synth.sequences(2, 64, 60, 200, 0.35, 0.65)


def test_oracle_fold_identities(allpairs_seqs):
    for seq in allpairs_seqs[:4]:
        P, P2, lz = oracle.fold(seq, 100)
        n = len(seq)
        occ = P.sum(axis=1) + P.sum(axis=0)  # probability that position i pairs
        assert occ.max() <= 1.0 + 1e-12
        inner = np.zeros_like(P)
        inner[:-1, 1:] = P[1:, :-1]  # inner[i][j] = P[i+1][j-1]
        assert (P2 <= np.minimum(P, inner) + 1e-12).all()
        for i in range(n + 1):
            for j in range(n + 1):
                if j - i + 1 > 100 or j - i < 4:
                    assert P[i][j] == 0
        assert lz > 0


def test_host_folding_matches_oracle(allpairs_seqs):
    from locarna_amd import synth
    for seq in allpairs_seqs[:6]:
        P, _, _ = oracle.fold(seq, 100)
        got = synth.base_pair_probs(seq, 100, 1e-6)
        keep = {(i, j) for i in range(len(seq) + 1) for j in range(len(seq) + 1) if P[i][j] > 1e-6}
        assert {(i, j) for i, j, _ in got} == keep
        for i, j, p in got:
            assert abs(p - P[i][j]) <= 1e-12


def test_oracle_fold_pp_text():
    P = np.zeros((13, 13))
    P2 = np.zeros((13, 13))
    P[1][12] = 0.5
    P[2][11] = 0.25
    P2[1][12] = 0.2
    P[3][10] = 0.0005
    t = oracle.fold_pp("x", "gggaaaucccuu", P, P2, 0.0005, stacking=True)
    lines = t.split("\n")
    assert lines[2] == "x" + " " * 17 + " GGGAAAUCCCUU"
    assert "#BPCUT 0.0005" in lines and "#STACK" in lines
    assert "1 12 0.5 0.2" in lines and "2 11 0.25" in lines
    assert not any(l.startswith("3 10") for l in lines)
