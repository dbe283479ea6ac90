"""The trace band on the device (lcg_pairs_prepare with
LCG_PREPARE_DEVICE_BAND): every pair's band (min_col / max_col per row)
equals the host band of lcg_pair_prepare (long double for locarna, double
for locarna_p), and so do the arc matches and the DP work that follow from
it; flagged pairs are rechecked on the host, which a huge band_eps forces
for every pair.  The host band itself is pinned against enumeration in
test_oracle_band.py.  Calls the kernel through the C ABI."""
import pytest

from locarna_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def _same(dev, host, label):
    for k, (x, y) in enumerate(zip(dev, host)):
        assert x.band() == y.band(), (label, k)
        assert (x.num_arcmatches, x.num_groups, x.cells) == (y.num_arcmatches, y.num_groups, y.cells), (label, k)


def _check(workload, params, label, limit=None, device_compaction=False):
    pps, pairs, span = synth.workload(workload, limit_pairs=limit)
    pp = [(pps[a], pps[b]) for a, b in pairs]
    pa = params(max_bp_span=span)
    print("band", label, len(pp), "pairs", flush=True)
    dev, st = api.prepare_pairs(pp, pa, device_band=True, device_compaction=device_compaction)
    host, st_h = api.prepare_pairs(pp, pa, device_band=False, device_compaction=device_compaction)
    print(label, "device", st, flush=True)
    assert st["device_banded"] + st["host_banded"] == len(pp)
    assert st_h["host_banded"] == len(pp) and st_h["device_banded"] == 0
    if not device_compaction:
        _same(dev, host, label)
    else:
        for k, (x, y) in enumerate(zip(dev, host)):
            assert x.band() == y.band(), (label, k)
    return pp, pa, dev, st


def test_band_allpairs64():
    pp, pa, dev, st = _check("allpairs64", api.locarna_params, "allpairs64")
    # the kernel bands all of these pairs itself
    assert st["device_banded"] >= len(pp) * 0.99


def test_band_long_pairs():
    _check("long500", api.locarna_params, "long500 global")
    _check("long500", lambda **kw: api.locarna_params(sequ_local=1, **kw), "long500 local")


def test_band_locarna_p_double():
    pp, pa, dev, st = _check("pf16x150", api.locarna_p_params, "pf16x150")
    assert st["device_banded"] >= len(pp) * 0.99
    _check("grad256", api.locarna_p_params, "grad256", limit=64)


def test_band_recheck_all_on_host():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=40)
    pp = [(pps[a], pps[b]) for a, b in pairs]
    pa = api.locarna_params(max_bp_span=span)
    dev, st = api.prepare_pairs(pp, pa, device_band=True, band_eps=1e30)
    assert st["device_banded"] == 0 and st["host_banded"] == len(pp)
    host = [api.Pair(a, b, pa) for a, b in pp]
    _same(dev, host, "recheck")


def test_band_then_device_compaction_aligns_as_oracle():
    pp, pa, dev, st = _check("allpairs64", api.locarna_params, "allpairs64 compaction", limit=120,
                             device_compaction=True)
    B = api.Batch(dev)
    try:
        B.run(mode=2)
        po = oracle.locarna_params(max_bp_span=pa.max_bp_span)
        scores = oracle.align_many(pp, po, nthreads=8)
        assert [B.score(k) for k in range(len(pp))] == list(scores)
    finally:
        B.destroy()


@pytest.mark.parametrize("spec", ["+-+-", "-+-+", "++++", "---+", "+---"])
def test_band_free_endgaps(spec):
    # the free-end-gap terms of the band kernel (free starts on row 0 /
    # column 0, reversed flags in the reverse pass, the extra Z terms)
    fe = api.free_endgaps(spec)
    pp, pa, dev, st = _check("allpairs64", lambda **kw: api.locarna_params(free_endgaps=fe, **kw),
                             "allpairs64 free " + spec, limit=200)
    assert st["device_banded"] >= len(pp) * 0.95
