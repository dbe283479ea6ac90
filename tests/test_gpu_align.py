"""GPU parity of the integer alignment path against the oracle: scores,
alignments (edges in order), structure strings and the D table, bit for
bit.  Calls the kernels through the C ABI."""
import os

import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_pair

pytestmark = pytest.mark.gpu


def _compare(batch, k, o, label):
    assert batch.score(k) == o.score, label
    edges, sA, sB = batch.alignment(k)
    assert edges == o.edges, label
    assert sA == o.structA and sB == o.structB, label
    D = batch.D(k)
    for t, (x, y) in enumerate(zip(D, o.D)):
        if y < -(1 << 60):  # oracle minus infinity (InftyInt range)
            assert x < -(1 << 28), (label, t)
        else:
            assert x == y, (label, t)


def _run_set(pp_pairs, kw, modes=(0,)):
    po = oracle.locarna_params(**kw)
    pa = api.locarna_params(**kw)
    ores = [oracle.align(a, b, po) for a, b in pp_pairs]
    pairs = [api.Pair(a, b, pa) for a, b in pp_pairs]
    batch = api.Batch(pairs)
    for mode in modes:
        print("run mode", mode, "pairs", len(pairs), kw, flush=True)
        batch.run(mode=mode, trace=True)
        for k, o in enumerate(ores):
            _compare(batch, k, o, (mode, k))
    batch.destroy()


def test_pair120():
    pps, pairs, _ = synth.workload("pair120")
    _run_set([(pps[0], pps[1])], {}, modes=(0, 1, 2))


@pytest.mark.parametrize("variant", ["global", "local", "linear", "nobanding", "tau"])
def test_tiny_batch(variant):
    kw = dict(min_prob=0.01, max_bp_span=-1)
    if variant == "local":
        kw.update(sequ_local=1, indel=-150, indel_opening=-200, min_trace_prob=0.0)
    elif variant == "linear":
        kw.update(indel_opening=0)
    elif variant == "nobanding":
        kw.update(min_trace_prob=0.0, use_ribosum=0, indel=-50, indel_opening=-100, struct_weight=100)
    elif variant == "tau":
        kw.update(tau=100, min_trace_prob=0.05)
    pp_pairs = [tiny_pair(5000 + s) for s in range(40)]
    _run_set(pp_pairs, kw)


def test_allpairs_subset():
    pps, pairs, _ = synth.workload("allpairs64", limit_pairs=48)
    _run_set([(pps[a], pps[b]) for a, b in pairs], {}, modes=(0, 2))


def test_allpairs_subset_local():
    pps, pairs, _ = synth.workload("allpairs64", limit_pairs=24)
    _run_set([(pps[a], pps[b]) for a, b in pairs], dict(sequ_local=1))


@pytest.mark.parametrize("local", [0, 1])
def test_long500(local):
    """Four pairs of 500-nt RNAs, max-bp-span 200, global and sequence-local."""
    pps, pairs, span = synth.workload("long500")
    _run_set([(pps[a], pps[b]) for a, b in pairs], dict(max_bp_span=span, sequ_local=local), modes=(0, 2))


def test_counters_match_oracle():
    """The kernel's own count of DP cells and arc candidates equals the
    oracle's (C and K of the work model), and the host's cell estimate."""
    pps, pairs, _ = synth.workload("allpairs64", limit_pairs=6)
    pp_pairs = [(pps[a], pps[b]) for a, b in pairs]
    ores = [oracle.align(a, b, oracle.locarna_params(), trace=False) for a, b in pp_pairs]
    P = [api.Pair(a, b, api.locarna_params()) for a, b in pp_pairs]
    B = api.Batch(P)
    B.run(mode=0, trace=True, count=True)
    for k, o in enumerate(ores):
        c, q = B.counters(k)
        assert c == o.counters["C"] == P[k].cells
        assert q == o.counters["K"]
    B.destroy()
