"""The allocation-free run (lcg_batch_run_noalloc, C ABI) against the
ordinary run of the same batch (lcg_batch_run mode 0): scores, alignments,
structures and D tables equal bit for bit on two calls in a row, the second
with a smaller batch, including a pair that several workgroups share (its
arrival counters live in the library's static array too), and a third call
on a caller's stream."""
import pytest
import torch  # before the library: one HIP runtime for torch's streams and the library

from locarna_amd import api, synth

pytestmark = pytest.mark.gpu


def _results(batch, n):
    return [(batch.score(k), batch.alignment(k), batch.D(k)) for k in range(n)]


def _check(pp, params, label):
    pairs = [api.Pair(a, b, params) for a, b in pp]
    fresh = api.Batch(pairs)
    print(label, "noalloc run of", len(pairs), "pairs", flush=True)
    fresh.run_noalloc(trace=True)  # first run of this batch: its results are the noalloc run's
    got = _results(fresh, len(pairs))
    print(label, "ordinary run", flush=True)
    fresh.run(mode=0, trace=True)
    want = _results(fresh, len(pairs))
    assert got == want, label
    fresh.destroy()
    return got


def test_noalloc_equals_run_on_two_calls():
    pps, pairs, span = synth.workload("allpairs64")
    pa = api.locarna_params(max_bp_span=span)
    first = [(pps[a], pps[b]) for a, b in pairs[:200]]
    _check(first, pa, "call 1")
    # the second call: a smaller batch, the array keeps the first call's contents
    second = [(pps[a], pps[b]) for a, b in pairs[500:530]]
    _check(second, pa, "call 2")
    # a one-pair batch: several workgroups share the pair
    one = [(pps[pairs[7][0]], pps[pairs[7][1]])]
    b1 = api.Batch([api.Pair(*one[0], pa)])
    assert b1.workgroups > 1
    b1.destroy()
    _check(one, pa, "call 3 (one pair, several workgroups)")
    # sequence-local and a long pair
    lps, lpairs, lspan = synth.workload("long500", limit_pairs=1)
    _check([(lps[a], lps[b]) for a, b in lpairs], api.locarna_params(max_bp_span=lspan, sequ_local=1), "call 4")


def test_noalloc_on_caller_stream():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=40)
    pa = api.locarna_params(max_bp_span=span)
    P = [api.Pair(pps[a], pps[b], pa) for a, b in pairs]
    ref = api.Batch(P)
    ref.run(mode=0, trace=True)
    want = _results(ref, len(P))
    ref.destroy()
    s = torch.cuda.Stream()
    b = api.Batch(P)
    b.set_stream(s.cuda_stream)
    print("noalloc on a torch stream", flush=True)
    b.run_noalloc(trace=True)
    assert _results(b, len(P)) == want
    b.set_stream(None)
    b.destroy()
