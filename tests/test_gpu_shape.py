"""Launch shapes and the autotuner (lcg_batch_create_shaped,
lcg_batch_reshape, lcg_batch_autotune): integer results are the same bit for
bit at every shape and equal the oracle's; LocARNA-P results agree across
shapes and with the oracle within the parity tolerances.  Calls the kernels
through the C ABI."""
import math
import os

import pytest

from locarna_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def _results(B):
    return [(B.score(k), B.alignment(k)) for k in range(len(B.pairs))]


def _allpairs(limit, device_compaction):
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=limit)
    pp_pairs = [(pps[a], pps[b]) for a, b in pairs]
    pa = api.locarna_params(max_bp_span=span)
    return pp_pairs, span, [api.Pair(a, b, pa, device_compaction=device_compaction) for a, b in pp_pairs]


@pytest.mark.parametrize("device_compaction", [False, True])
def test_reshape_keeps_results(device_compaction):
    # 40 pairs: fewer pairs than CUs, so several workgroups share a pair by
    # default; shapes with other wave counts, workgroups per pair and top
    # level shapes give the same scores, alignments and structures
    pp_pairs, span, P = _allpairs(40, device_compaction)
    B = api.Batch(P)
    try:
        B.run(mode=2)
        ref = _results(B)
        base = B.shape()
        print("default shape", base, flush=True)
        for shape in [dict(waves=8), dict(wg_per_pair=1), dict(wg_per_pair=2, waves=12),
                      dict(top_waves=16, top_wgs_per_cu=1), dict(top_waves=2, top_wgs_per_cu=4)]:
            print("reshape", shape, flush=True)
            B.reshape(**shape)
            eff = B.shape()
            for key, v in shape.items():
                assert eff[key] == v, (shape, eff)
            for mode in (0, 2):
                B.run(mode=mode)
                assert _results(B) == ref, (shape, mode)
        scores = oracle.align_many(pp_pairs, oracle.locarna_params(max_bp_span=span), nthreads=8)
        assert [r[0] for r in ref] == list(scores)
    finally:
        B.destroy()


def test_create_shaped_matches_default():
    pp_pairs, span, P = _allpairs(24, False)
    A = api.Batch(P)
    S = api.Batch(P, shape=dict(waves=8, wg_per_pair=3, top_waves=8, top_wgs_per_cu=2))
    try:
        assert S.shape()["waves"] == 8 and S.shape()["wg_per_pair"] == 3
        A.run()
        S.run()
        assert _results(A) == _results(S)
    finally:
        A.destroy()
        S.destroy()


def test_autotune_allpairs64():
    # the full all-pairs batch (2016 pairs, device compaction): the tuner's
    # choice keeps every score and alignment, its time is no worse than the
    # default shape's, and the scores equal the oracle's
    pp_pairs, span, P = _allpairs(None, True)
    B = api.Batch(P)
    try:
        B.run(mode=2)
        ref = _results(B)
        shape, best_ms, default_ms = B.autotune(path=0, reps=2)
        print("autotuned", shape, "best ms", best_ms, "default ms", default_ms, flush=True)
        assert best_ms <= default_ms
        assert B.shape() == shape
        B.run(mode=2)
        assert _results(B) == ref
        nthreads = min(16, os.cpu_count() or 1)
        scores = oracle.align_many(pp_pairs, oracle.locarna_params(max_bp_span=span), nthreads=nthreads)
        assert [r[0] for r in ref] == list(scores)
    finally:
        B.destroy()


def test_autotune_long_pairs_local():
    # four 500-nt pairs, sequence-local: many workgroups per pair
    pps, pairs, span = synth.workload("long500", limit_pairs=2)
    pa = api.locarna_params(max_bp_span=span, sequ_local=1)
    P = [api.Pair(pps[a], pps[b], pa) for a, b in pairs]
    B = api.Batch(P)
    try:
        B.run()
        ref = _results(B)
        shape, best_ms, default_ms = B.autotune(path=0, reps=1)
        print("autotuned", shape, best_ms, default_ms, flush=True)
        B.run()
        assert _results(B) == ref
        po = oracle.locarna_params(max_bp_span=span, sequ_local=1)
        for k, (a, b) in enumerate(pairs):
            assert ref[k][0] == oracle.align(pps[a], pps[b], po, trace=False).score
    finally:
        B.destroy()


def test_autotune_pf():
    pps, pairs, span = synth.workload("pf16x150", limit_pairs=24)
    pa = api.locarna_p_params(max_bp_span=span)
    P = [api.Pair(pps[a], pps[b], pa) for a, b in pairs]
    B = api.Batch(P)
    try:
        B.pf_run()
        z0 = [B.pf_result(k)[0] for k in range(len(P))]
        shape, best_ms, default_ms = B.autotune(path=1, reps=1)
        print("autotuned", shape, best_ms, default_ms, flush=True)
        assert best_ms <= default_ms
        B.pf_run()
        po = oracle.locarna_p_params(max_bp_span=span)
        for k in range(len(P)):
            z = B.pf_result(k)[0]
            assert abs(z - z0[k]) <= 1e-12 * abs(z0[k]), k
        for k in range(0, len(P), 8):
            o = oracle.pf(pps[pairs[k][0]], pps[pairs[k][1]], po)
            z = B.pf_result(k)[0]
            assert abs(math.log(z) - math.log(o.pf)) <= 1e-9 * max(1.0, abs(math.log(o.pf))), k
            for x, y in zip(B.am_probs(k), o.am_probs):
                assert abs(x - y) <= 1e-6, k
    finally:
        B.destroy()
