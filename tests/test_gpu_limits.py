"""Size limits.  A 1500-nt pair (longer than the band kernel takes, so its
band is computed on the host; groups near the LDS limit of one workgroup)
aligns as the oracle does on every packing path, global and
sequence-local.  A pair whose longer sequence needs an LDS header larger
than a workgroup's LDS (6000 nt: 8 ints per position) is refused at batch
creation with LCG_E_CAPACITY, before any launch, for LocARNA and LocARNA-P
parameters.  Calls the kernels through the C ABI."""
import pytest

from locarna_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

SPAN = 200


def _long_pair():
    seqs = synth.sequences(11, 2, 1500, 1500, 0.5, 0.5)
    return [synth.make_pp("s%d" % k, s, SPAN) for k, s in enumerate(seqs)]


@pytest.mark.parametrize("local", [0, 1])
def test_long_pair_1500(local):
    a, b = _long_pair()
    kw = dict(max_bp_span=SPAN, sequ_local=local)
    if local:
        kw.update(indel=-150, indel_opening=-200)
    o = oracle.align(a, b, oracle.locarna_params(**kw))
    pa = api.locarna_params(**kw)
    dev, st = api.prepare_pairs([(a, b)], pa, device_band=True, device_compaction=True)
    assert st["host_banded"] == 1  # longer than the band kernel takes
    for label, P in [("host", api.Pair(a, b, pa)), ("device", dev[0])]:
        print("long1500", "local" if local else "global", label, flush=True)
        B = api.Batch([P])
        try:
            B.run(mode=0)
            assert B.score(0) == o.score, label
            edges, sA, sB = B.alignment(0)
            assert edges == o.edges, label
            assert (sA, sB) == (o.structA, o.structB), label
        finally:
            B.destroy()


@pytest.mark.parametrize("params", [api.locarna_params, api.locarna_p_params])
def test_header_beyond_lds_is_refused(params):
    long_seq = synth.sequences(12, 1, 6000, 6000, 0.5, 0.5)[0]
    a = synth.pp_text("long", long_seq, [])
    b = synth.pp_text("short", "GGGAAACCCA", [(1, 9, 0.8)])
    P = api.Pair(a, b, params())
    print("capacity", params.__name__, flush=True)
    with pytest.raises(api.LcgError) as e:
        api.Batch([P])
    assert "(%d)" % -4 in str(e.value)  # LCG_E_CAPACITY
