"""Alignments with explicit arc-match scores (lcg_pair_prepare_scored,
locarna --read-arcmatch-scores) on the GPU: reading back the written scores
aligns every pair as before (scores, alignments, structures), and with
changed and added arc matches the score and the traced alignment equal the
exhaustive enumeration over the same arc matches and scores
(tests/bruteforce.py) on tiny pairs.  Calls the kernels through the C ABI."""
import random

import pytest

import bruteforce as bf
from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_pair

pytestmark = pytest.mark.gpu


def _run(pairs):
    B = api.Batch(pairs)
    try:
        B.run(mode=0)
        return [(B.score(k), B.alignment(k)) for k in range(len(pairs))]
    finally:
        B.destroy()


def test_round_trip_aligns_the_same():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=48)
    pa = api.locarna_params(max_bp_span=span)
    P = [api.Pair(pps[a], pps[b], pa) for a, b in pairs]
    Q = [api.Pair.from_arcmatch_scores(pps[a], pps[b], p.arcmatch_scores_text(), pa)
         for (a, b), p in zip(pairs, P)]
    assert _run(P) == _run(Q)


@pytest.mark.parametrize("seed", range(10))
def test_changed_scores_match_enumeration(seed):
    a, b = tiny_pair(8100 + seed)
    kw = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0)
    r = oracle.align(a, b, oracle.locarna_params(**kw))
    rng = random.Random(seed)
    lines = []
    for (am, s) in zip(r.arcmatches, r.am_scores):
        lines.append((am, 2 * s - 37 * rng.randrange(4)))
    # arc matches of arcs that are not in the pp files
    for _ in range(3):
        i = rng.randrange(1, r.lenA - 2)
        j = rng.randrange(i + 3, r.lenA + 1)
        k = rng.randrange(1, r.lenB - 2)
        l = rng.randrange(k + 3, r.lenB + 1)
        lines.append(((i, j, k, l), rng.randrange(-300, 900)))
    # one arc match per arc pair (the first line of a repeated pair stays)
    seen, uniq = set(), []
    for am, s in lines:
        if am not in seen:
            seen.add(am)
            uniq.append((am, s))
    text = "".join("%d %d %d %d %d\n" % (am + (s,)) for am, s in uniq)
    Q = api.Pair.from_arcmatch_scores(a, b, text, api.locarna_params(**kw))
    assert Q.num_arcmatches == len(uniq)
    (score, (edges, sA, sB)), = _run([Q])
    prob = bf.Problem(r.lenA, r.lenB, r.sigma, r.gapA, r.gapB, -500, [am for am, _ in uniq],
                      [s for _, s in uniq], r.min_col, r.max_col)
    best, bedges, arcs, nal = bf.best_global(prob)
    print("seed", seed, "score", score, "enumerated", best, "alignments", nal, flush=True)
    assert score == best
    assert edges == bedges
