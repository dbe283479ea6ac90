"""GPU parity of the device arc-match compaction (lcg_pair_prepare_device +
lcg_batch_create): the device image of every pair -- arc-match scores in the
reference's index order, groups by common left ends with their LDS
schedule, candidate lists by common right ends, fill list -- equals the
host packing of the same pairs int for int, and both batches align to the
same scores, alignments and D tables.  The host packing itself is checked
against the oracle in test_gpu_align.py; the tiny pairs are also checked
against the oracle here.  Calls the kernels through the C ABI."""
import numpy as np
import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_pair

pytestmark = pytest.mark.gpu


def _same_images(hb, db, label):
    for k in range(len(hb.pairs)):
        assert hb.pair_info(k) == db.pair_info(k), (label, k)
        h, d = hb.segment(k), db.segment(k)
        assert len(h) == len(d), (label, k)
        bad = np.nonzero(h != d)[0]
        if len(bad):
            raise AssertionError("%s pair %d: %d ints differ, first at %d (%d != %d)"
                                 % (label, k, len(bad), bad[0], h[bad[0]], d[bad[0]]))


def _same_results(hb, db, label, modes=(0,)):
    for mode in modes:
        hb.run(mode=mode, trace=True)
        db.run(mode=mode, trace=True)
        for k in range(len(hb.pairs)):
            assert hb.score(k) == db.score(k), (label, mode, k)
            assert hb.alignment(k) == db.alignment(k), (label, mode, k)
            assert hb.D(k) == db.D(k), (label, mode, k)


def _run(pp_pairs, kw, label, modes=(0,), oracle_check=False):
    pa = api.locarna_params(**kw)
    print("compaction", label, "pairs", len(pp_pairs), kw, flush=True)
    hb = api.Batch([api.Pair(a, b, pa) for a, b in pp_pairs])
    db = api.Batch([api.Pair(a, b, pa, device_compaction=True) for a, b in pp_pairs])
    try:
        _same_images(hb, db, label)
        # the rerun reproduces the image
        db.compact_run()
        _same_images(hb, db, label + " rerun")
        _same_results(hb, db, label, modes)
        if oracle_check:
            po = oracle.locarna_params(**kw)
            for k, (a, b) in enumerate(pp_pairs):
                o = oracle.align(a, b, po)
                assert db.score(k) == o.score, (label, k)
                assert db.alignment(k)[0] == o.edges, (label, k)
    finally:
        hb.destroy()
        db.destroy()


def test_compact_pair120():
    pps, pairs, _ = synth.workload("pair120")
    _run([(pps[0], pps[1])], {}, "pair120", modes=(0, 1, 2), oracle_check=True)


@pytest.mark.parametrize("variant", ["global", "local", "tau", "tau_sigma", "nobanding"])
def test_compact_tiny(variant):
    kw = dict(min_prob=0.01, max_bp_span=-1)
    if variant == "local":
        kw.update(sequ_local=1, indel=-150, indel_opening=-200, min_trace_prob=0.0)
    elif variant == "tau":
        kw.update(tau=100, min_trace_prob=0.05)
    elif variant == "tau_sigma":  # arc-match term from sigma sums (no RIBOSUM)
        kw.update(tau=100, use_ribosum=0, min_trace_prob=0.05)
    elif variant == "nobanding":
        kw.update(min_trace_prob=0.0, use_ribosum=0, indel=-50, indel_opening=-100, struct_weight=100)
    _run([tiny_pair(5000 + s) for s in range(40)], kw, "tiny-" + variant, oracle_check=True)


def test_compact_allpairs_subset():
    pps, pairs, span = synth.workload("allpairs64")
    sel = pairs[::16]  # 126 pairs over the whole length range
    _run([(pps[a], pps[b]) for a, b in sel], dict(max_bp_span=span), "allpairs64/16")


def test_compact_grad_tau50():
    # locarna_p scoring (tau 50): the RIBOSUM arc-match term
    pps, pairs, span = synth.workload("grad256")
    sel = pairs[:48]
    _run([(pps[a], pps[b]) for a, b in sel], dict(max_bp_span=span, tau=50, min_trace_prob=1e-5,
                                                  band_extended=0), "grad256/48")


@pytest.mark.parametrize("local", [0, 1])
def test_compact_long500(local):
    # 4 pairs: several workgroups per pair, the schedule deals groups to members
    pps, pairs, span = synth.workload("long500")
    _run([(pps[a], pps[b]) for a, b in pairs], dict(max_bp_span=span, sequ_local=local), "long500-%d" % local)
