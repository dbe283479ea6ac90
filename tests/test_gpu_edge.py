"""GPU parity on the edge-case pairs of tests/edge.py (single nucleotides,
no base pairs, base pairs on one side only, N symbols, an identical pair,
the smallest arc span), alone and mixed with ordinary tiny pairs: scores,
alignments, structures and D bit for bit on every packing path (host
packing, device compaction, device band), global, sequence-local and with
free end gaps; LocARNA-P log Z and probabilities within the parity
tolerances.  Calls the kernels through the C ABI."""
import math

import pytest

from edge import CASES
from locarna_amd import api
from oracle import oracle
from tiny import tiny_pair

pytestmark = pytest.mark.gpu

KW = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0)
VARIANTS = {
    "global": {},
    "local": dict(sequ_local=1, indel=-150, indel_opening=-200),
    "free": dict(free_endgaps=api.free_endgaps("++++")),
    "band": dict(min_trace_prob=0.01),
}


def _pairs():
    names = sorted(CASES)
    pp = [CASES[n] for n in names] + [tiny_pair(6100 + s) for s in range(4)]
    return names + ["tiny%d" % s for s in range(4)], pp


def _check(B, ores, labels):
    for k, o in enumerate(ores):
        assert B.score(k) == o.score, labels[k]
        edges, sA, sB = B.alignment(k)
        assert edges == o.edges, labels[k]
        assert (sA, sB) == (o.structA, o.structB), labels[k]
        for t, (x, y) in enumerate(zip(B.D(k), o.D)):
            if y < -(1 << 60):
                assert x < -(1 << 28), (labels[k], t)
            else:
                assert x == y, (labels[k], t)


@pytest.mark.parametrize("variant", sorted(VARIANTS))
@pytest.mark.parametrize("path", ["host", "device_compaction", "device_band"])
def test_edge_batch(variant, path):
    labels, pp = _pairs()
    kw = dict(KW, **VARIANTS[variant])
    po = oracle.locarna_params(**kw)
    pa = api.locarna_params(**kw)
    ores = [oracle.align(a, b, po) for a, b in pp]
    if path == "device_band":
        P, st = api.prepare_pairs(pp, pa, device_band=True, device_compaction=True)
    else:
        P = [api.Pair(a, b, pa, device_compaction=(path == "device_compaction")) for a, b in pp]
    print("edge batch", variant, path, len(P), "pairs", flush=True)
    B = api.Batch(P)
    try:
        for mode in (0, 2):
            B.run(mode=mode)
            _check(B, ores, labels)
    finally:
        B.destroy()


@pytest.mark.parametrize("name", ["len1", "noarcs", "identical"])
def test_edge_single_pair_batch(name):
    # one pair per batch: several workgroups share it (wg_per_pair > 1)
    a, b = CASES[name]
    po = oracle.locarna_params(**KW)
    pa = api.locarna_params(**KW)
    o = oracle.align(a, b, po)
    print("edge single", name, flush=True)
    B = api.Batch([api.Pair(a, b, pa)])
    try:
        B.run(mode=0)
        _check(B, [o], [name])
    finally:
        B.destroy()


def test_edge_partition_function():
    labels, pp = _pairs()
    kw = dict(KW, tau=50)
    po = oracle.locarna_p_params(**kw)
    pa = api.locarna_p_params(**kw)
    ores = [oracle.pf(a, b, po) for a, b in pp]
    print("edge pf", len(pp), "pairs", flush=True)
    B = api.Batch([api.Pair(a, b, pa) for a, b in pp])
    try:
        B.pf_run()
        for k, o in enumerate(ores):
            Z, openings = B.pf_result(k)
            assert abs(math.log(Z) - math.log(o.pf)) <= 1e-9 * max(1.0, abs(math.log(o.pf))), labels[k]
            am = B.am_probs(k)
            assert len(am) == len(o.am_probs), labels[k]
            for x, y in zip(am, o.am_probs):
                assert abs(x - y) <= 1e-6, labels[k]
            for x, y in zip(B.bm_probs(k, 0), o.bm_probs):
                assert abs(x - y) <= 1e-6, labels[k]
            assert abs(openings - o.openings) <= 1e-9 * max(1.0, o.openings), labels[k]
    finally:
        B.destroy()
