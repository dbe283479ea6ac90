"""Every pair of the benchmark workloads against the oracle:

- allpairs64 (64 RNAs of 60-200 nt, 2016 pairs), prepared as the all-pairs
  driver prepares them (trace band and arc-match compaction on the device)
  and run as one captured-graph launch: score, alignment edges and both
  structures of all 2016 pairs bit for bit, global; sequence-local for every
  third pair;
- pf16x150 (16 RNAs of 150 nt, 120 pairs): log Z within 1e-9 relative, D and
  D' within 1e-9 relative, arc-match and base-match probabilities within
  1e-6 absolute, for all 120 pairs;
- grad256 (80-250 nt): the gradient of log Z of its four longest pairs
  against central finite differences (h = 1e-4) and against the oracle's
  marginals through the identities of the adjoint (T = temperature):
  d log Z / d sigma_xy = sum of match-edge probabilities of symbol pair
  (x, y) / T; d / d AM_uv = (tau / 100) sum of P(am) over arc matches of
  base-pair types (u, v) / T; d / d indel = (n + m - 2 E[#matched
  positions]) / T; d / d opening = E[#gap openings] / T; d / d struct
  weight = sum P(am) (f_A + f_B) / T with f = 1 - ln p / ln p_e."""
import concurrent.futures as cf
import math
import os

import pytest

from locarna_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _oracle_many(fn, items):
    with cf.ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(fn, items))


@pytest.mark.parametrize("local", [0, 1])
def test_allpairs64_every_pair(local):
    pps, pairs, span = synth.workload("allpairs64")
    if local:
        pairs = pairs[::3]
    pp = [(pps[a], pps[b]) for a, b in pairs]
    pa = api.locarna_params(max_bp_span=span, sequ_local=local)
    print("prepare", len(pp), "pairs, local", local, flush=True)
    prep, st = api.prepare_pairs(pp, pa, device_compaction=True, device_band=True, threads=THREADS)
    batch = api.Batch(prep)
    print("run (graph)", flush=True)
    batch.run(mode=2, trace=True)
    print("oracle", flush=True)
    po = oracle.locarna_params(max_bp_span=span, sequ_local=local)
    ores = _oracle_many(lambda ab: oracle.align(ab[0], ab[1], po), pp)
    bad = []
    for k, o in enumerate(ores):
        edges, sA, sB = batch.alignment(k)
        if batch.score(k) != o.score or edges != o.edges or sA != o.structA or sB != o.structB:
            bad.append(k)
    assert not bad, ("pairs differing from the oracle", bad[:20], len(bad))
    batch.destroy()


def _pf_compare(B, k, o, label):
    Z, openings = B.pf_result(k)
    assert abs(math.log(Z) - math.log(o.pf)) <= 1e-9 * max(1.0, abs(math.log(o.pf))), label
    D, Dp = B.pf_D(k)
    assert len(D) == len(o.pf_D)
    for t, (x, y) in enumerate(zip(D, o.pf_D)):
        assert abs(x - y) <= 1e-9 * abs(y), (label, "D", t)
    for t, (x, y) in enumerate(zip(Dp, o.pf_Dprime)):
        assert abs(x - y) <= 1e-9 * abs(y), (label, "D'", t)
    for t, (x, y) in enumerate(zip(B.am_probs(k), o.am_probs)):
        assert abs(x - y) <= 1e-6, (label, "am", t)
    for t, (x, y) in enumerate(zip(B.bm_probs(k, 0), o.bm_probs)):
        assert abs(x - y) <= 1e-6, (label, "bm", t)
    for t, (x, y) in enumerate(zip(B.bm_probs(k, 2), o.bm_full)):
        assert abs(x - y) <= 1e-6, (label, "bm_full", t)
    assert abs(openings - o.openings) <= 1e-9 * max(1.0, o.openings), label


def test_pf16x150_every_pair():
    pps, pairs, span = synth.workload("pf16x150")
    pp = [(pps[a], pps[b]) for a, b in pairs]
    pa = api.locarna_p_params(max_bp_span=span)
    P = [api.Pair(a, b, pa) for a, b in pp]
    B = api.Batch(P)
    print("pf run", len(P), flush=True)
    B.pf_run()
    po = oracle.locarna_p_params(max_bp_span=span)
    print("oracle", flush=True)
    ores = _oracle_many(lambda ab: oracle.pf(ab[0], ab[1], po), pp)
    for k, o in enumerate(ores):
        _pf_compare(B, k, o, k)
    B.destroy()


def _seq_and_probs(pp_text):
    lines = pp_text.split("\n")
    seq = None
    probs = {}
    section = False
    for ln in lines:
        f = ln.split()
        if not f:
            continue
        if f[0] == "#SECTION":
            section = True
        elif section and not f[0].startswith("#"):
            probs[(int(f[0]), int(f[1]))] = float(f[2])
        elif not section and not f[0].startswith("#") and seq is None and len(f) == 2:
            seq = f[1]
    return seq, probs


def test_grad256_gradient_identities_and_fd():
    pps, pairs, span = synth.workload("grad256")
    lens = {k: len(_seq_and_probs(pps[k])[0]) for k in pps}
    sel = sorted(pairs, key=lambda ab: -(lens[ab[0]] * lens[ab[1]]))[:4]
    pp = [(pps[a], pps[b]) for a, b in sel]
    pa = api.locarna_p_params(max_bp_span=span)
    P = [api.Pair(a, b, pa) for a, b in pp]
    B = api.Batch(P)
    print("pf run, pairs of lengths", [(lens[a], lens[b]) for a, b in sel], flush=True)
    B.pf_run()
    g = [B.pf_grad(k) for k in range(len(P))]
    T = float(pa.temperature_alipf)
    tau = float(pa.tau)
    po = oracle.locarna_p_params(max_bp_span=span)
    ores = _oracle_many(lambda ab: oracle.pf(ab[0], ab[1], po), pp)
    code = {c: x for x, c in enumerate("ACGU")}
    for k, (o, (ppA, ppB)) in enumerate(zip(ores, pp)):
        sA, probA = _seq_and_probs(ppA)
        sB, probB = _seq_and_probs(ppB)
        n, m = len(sA), len(sB)
        W = m + 1
        # base-match scores: match-edge marginals by symbol pair
        sig = [0.0] * 16
        for i in range(1, n + 1):
            for j in range(1, m + 1):
                x, y = code.get(sA[i - 1]), code.get(sB[j - 1])
                if x is not None and y is not None:
                    sig[4 * x + y] += o.bm_full[i * W + j]
        for c in range(16):
            assert g[k][c] == pytest.approx(sig[c] / T, rel=1e-6, abs=1e-10), (k, "sigma", c)
        # arc-match terms, struct weight, matched positions
        am = [0.0] * 256
        sw = 0.0
        pam = 0.0
        peA, peB = 1.0 / (2 * n), 1.0 / (2 * m)
        for (al, ar, bl, br), p in zip(o.arcmatches, o.am_probs):
            u = 4 * code[sA[al - 1]] + code[sA[ar - 1]]
            v = 4 * code[sB[bl - 1]] + code[sB[br - 1]]
            am[16 * u + v] += p
            fA = 1 - math.log(probA[(al, ar)]) / math.log(peA)
            fB = 1 - math.log(probB[(bl, br)]) / math.log(peB)
            sw += p * (fA + fB)
            pam += p
        for c in range(256):
            assert g[k][16 + c] == pytest.approx(tau / 100.0 * am[c] / T, rel=1e-6, abs=1e-10), (k, "am", c)
        assert g[k][274] == pytest.approx(sw / T, rel=1e-6), (k, "struct weight")
        matched = sum(o.bm_full) + 2 * pam  # match edges and arc-match ends
        assert g[k][272] == pytest.approx((n + m - 2 * matched) / T, rel=1e-6), (k, "indel")
        assert g[k][273] == pytest.approx(o.openings / T, rel=1e-6), (k, "opening")
    # central finite differences of the GPU's own log Z
    h = 1e-4
    for comp in [2 * 4 + 1, 0 * 4 + 3, 16 + (2 * 4 + 1) * 16 + (1 * 4 + 2), 272, 273, 274]:
        d = [0.0] * api.NTHETA
        d[comp] = h
        B.pf_run(d)
        up = [math.log(B.pf_result(k)[0]) for k in range(len(P))]
        d[comp] = -h
        B.pf_run(d)
        dn = [math.log(B.pf_result(k)[0]) for k in range(len(P))]
        for k in range(len(P)):
            fd = (up[k] - dn[k]) / (2 * h)
            assert abs(fd - g[k][comp]) <= 1e-6 * max(1.0, abs(g[k][comp])), (comp, k, fd, g[k][comp])
    B.destroy()
