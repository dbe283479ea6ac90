"""--max-diff / --max-diff-am / --max-diff-at-am on the GPU: scores,
alignments and D tables equal the oracle's with host packing, with the
device arc-match compaction (its predicate takes the two arc-match limits)
and with the trace band computed on the device over the max-diff band.
Calls the kernels through the C ABI."""
import pytest

from locarna_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

CASES = [dict(max_diff=12), dict(max_diff=20, max_diff_am=10, max_diff_at_am=15), dict(max_diff_am=3),
         dict(max_diff=8, sequ_local=1)]


@pytest.mark.parametrize("kw", CASES, ids=[str(c) for c in CASES])
def test_max_diff_equal_oracle(kw):
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=24)
    pp = [(pps[a], pps[b]) for a, b in pairs]
    pa = api.locarna_params(max_bp_span=span, **kw)
    po = oracle.locarna_params(max_bp_span=span, **kw)
    ores = [oracle.align(a, b, po) for a, b in pp]
    for label, (P, st) in (("host", (api.prepare_pairs(pp, pa)[0], None)),
                           ("device", api.prepare_pairs(pp, pa, device_compaction=True, device_band=True))):
        B = api.Batch(P)
        try:
            B.run(mode=2)
            for k, o in enumerate(ores):
                edges, sA, sB = B.alignment(k)
                assert B.score(k) == o.score, (kw, label, k)
                assert edges == o.edges and sA == o.structA and sB == o.structB, (kw, label, k)
                D = B.D(k)
                assert len(D) == len(o.D)
                for x, y in zip(D, o.D):
                    assert (x < -(1 << 28)) if y < -(1 << 60) else x == y, (kw, label, k)
        finally:
            B.destroy()
