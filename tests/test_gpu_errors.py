"""Error path of the pairs that several workgroups share: a bounded
cross-workgroup wait that gives up marks the run invalid, and every result
request of that run fails with LCG_E_STATE, the first one and all later
ones, on the integer and the LocARNA-P path; the next run that completes
its waits is valid again.  The condition is forced from the host (a spin
limit of -1 makes every wait give up at once, so every kernel still exits
normally); one-pair batches run on several workgroups per pair."""
import pytest

from locarna_amd import api, synth

pytestmark = pytest.mark.gpu


def _pp_pair(seed, n):
    seqs = synth.sequences(seed, 2, n, n, 0.5, 0.5)
    return synth.make_pp("a", seqs[0], 100), synth.make_pp("b", seqs[1], 100)


def test_expired_wait_integer_path():
    ppA, ppB = _pp_pair(11, 120)
    batch = api.Batch([api.Pair(ppA, ppB, api.locarna_params())])
    assert batch.workgroups > 1  # one pair: several workgroups share it
    print("valid run", flush=True)
    batch.run(mode=0, trace=True)
    score, aln, D = batch.score(0), batch.alignment(0), batch.D(0)
    batch.set_spin_limit(-1)
    print("run with every wait given up", flush=True)
    batch.run(mode=0, trace=False)
    for _ in range(3):
        for getter in (lambda: batch.score(0), lambda: batch.D(0), lambda: batch.counters(0)):
            with pytest.raises(api.LcgError, match=r"\(-5\)"):
                getter()
    batch.set_spin_limit(0)
    print("valid run again", flush=True)
    batch.run(mode=0, trace=True)
    assert batch.score(0) == score and batch.alignment(0) == aln and batch.D(0) == D
    batch.destroy()


def test_expired_wait_pf_path():
    ppA, ppB = _pp_pair(12, 100)
    batch = api.Batch([api.Pair(ppA, ppB, api.locarna_p_params())])
    wg, nst, _ = batch.pf_config()
    assert wg > 1 and nst == 1
    print("valid pf run", flush=True)
    batch.pf_run()
    z, grad = batch.pf_result(0), batch.pf_grad(0)
    batch.set_spin_limit(-1)
    print("pf run with every wait given up", flush=True)
    with pytest.raises(api.LcgError, match=r"\(-5\)"):
        batch.pf_run()
    for _ in range(2):
        for getter in (lambda: batch.pf_result(0), lambda: batch.am_probs(0), lambda: batch.pf_D(0),
                       lambda: batch.bm_probs(0), lambda: batch.pf_grad(0), lambda: batch.am_probs_text(0)):
            with pytest.raises(api.LcgError, match=r"\(-5\)"):
                getter()
    batch.set_spin_limit(0)
    print("valid pf run again", flush=True)
    batch.pf_run()
    # several workgroups add into the same accumulators: equal up to the order
    # of floating-point additions
    z2, grad2 = batch.pf_result(0), batch.pf_grad(0)
    assert z2[0] == pytest.approx(z[0], rel=1e-12)
    assert grad2 == pytest.approx(grad, rel=1e-9, abs=1e-12)
    batch.destroy()


def test_penalty_out_of_range_refused():
    """A position penalty that would push a pair's penalized scores past the
    kernels' int32 range is refused before any launch (LCG_E_CAPACITY)."""
    pps, pairs, span = synth.workload("long500", limit_pairs=1)
    a, b = pairs[0]
    batch = api.Batch([api.Pair(pps[a], pps[b], api.locarna_params(max_bp_span=span, sequ_local=1))])
    with pytest.raises(api.LcgError, match=r"\(-4\)"):
        batch.run_penalized(100000)
    with pytest.raises(api.LcgError, match=r"\(-4\)"):
        batch.run_penalized(-100000)
    print("penalized run in range", flush=True)
    batch.run_penalized(50)
    batch.score(0)
    batch.destroy()
