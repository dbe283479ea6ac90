"""GPU parity of the LocARNA-P path against the oracle: log Z within 1e-9
relative, D and D' per arc match within 1e-9 relative, arc-match and base-match probabilities within 1e-6 absolute; the
gradient of log Z against central finite differences of the GPU's own
log Z (h = 1e-4) and against the oracle's marginals."""
import math

import pytest

from locarna_amd import api, synth
from oracle import oracle
from tiny import tiny_pair

pytestmark = pytest.mark.gpu

TOL_LOGZ = 1e-9
TOL_P = 1e-6
TOL_REL_D = 1e-9  # inside D and outside D' per arc match, relative


def _compare(B, k, o, label, include_am=False):
    Z, openings = B.pf_result(k)
    assert abs(math.log(Z) - math.log(o.pf)) <= TOL_LOGZ * max(1.0, abs(math.log(o.pf))), label
    D, Dp = B.pf_D(k)
    for t, (x, y) in enumerate(zip(D, o.pf_D)):
        assert abs(x - y) <= TOL_REL_D * abs(y), (label, "D", t, x, y)
    for t, (x, y) in enumerate(zip(Dp, o.pf_Dprime)):
        assert abs(x - y) <= TOL_REL_D * abs(y), (label, "D'", t, x, y)
    am = B.am_probs(k)
    assert len(am) == len(o.am_probs)
    for t, (x, y) in enumerate(zip(am, o.am_probs)):
        assert abs(x - y) <= TOL_P, (label, t, x, y)
    bm = B.bm_probs(k, 1 if include_am else 0)
    for t, (x, y) in enumerate(zip(bm, o.bm_probs)):
        assert abs(x - y) <= TOL_P, (label, t, x, y)
    full = B.bm_probs(k, 2)
    for t, (x, y) in enumerate(zip(full, o.bm_full)):
        assert abs(x - y) <= TOL_P, (label, "full", t, x, y)
    assert abs(openings - o.openings) <= 1e-9 * max(1.0, o.openings), label


def _run(pp_pairs, kw, include_am=False):
    po = oracle.locarna_p_params(**kw)
    pa = api.locarna_p_params(**kw)
    ores = [oracle.pf(a, b, po, include_am_in_bm=include_am) for a, b in pp_pairs]
    P = [api.Pair(a, b, pa) for a, b in pp_pairs]
    B = api.Batch(P)
    print("pf run", len(P), kw, flush=True)
    B.pf_run()
    for k, o in enumerate(ores):
        _compare(B, k, o, k, include_am)
    return B, P, ores


@pytest.mark.parametrize("variant", ["default", "band", "nothreshold", "include_am"])
def test_pf_tiny_batch(variant):
    kw = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0)
    inc = False
    if variant == "band":
        kw.update(min_trace_prob=0.01)
    elif variant == "nothreshold":
        kw.update(min_am_prob=0.0)
    elif variant == "include_am":
        inc = True
    _run([tiny_pair(7000 + s) for s in range(30)], kw, include_am=inc)


def test_pf_workload_subset():
    pps, pairs, span = synth.workload("pf16x150", limit_pairs=6)
    B, P, ores = _run([(pps[a], pps[b]) for a, b in pairs], dict(max_bp_span=span))
    B.destroy()


def test_gradient_finite_differences():
    pps, pairs, span = synth.workload("pf16x150", limit_pairs=2)
    pa = api.locarna_p_params(max_bp_span=span)
    P = [api.Pair(pps[a], pps[b], pa) for a, b in pairs]
    B = api.Batch(P)
    B.pf_run()
    g = [B.pf_grad(k) for k in range(len(P))]
    h = 1e-4
    # components: sigma GC, sigma AU, an arc-match term, indel, opening, struct weight
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


def test_pf_split_pairs_match_single(monkeypatch):
    # 160 pairs of 80-250 nt: more pairs than half the CUs, so the largest
    # pairs run on two workgroups each and the rest from the queue; every
    # result matches the same batch with one workgroup per pair, and a
    # sample matches the oracle
    pps, pairs, span = synth.workload("grad256")
    sel = pairs[:160]
    pa = api.locarna_p_params(max_bp_span=span)
    P = [api.Pair(pps[a], pps[b], pa) for a, b in sel]
    mixed = api.Batch(P)
    K, nst, grid = mixed.pf_config()
    print("pf config", K, nst, grid, flush=True)
    assert K == 2 and nst > 0
    monkeypatch.setenv("LCG_PF_WG_PER_PAIR", "1")
    single = api.Batch(P)
    assert single.pf_config()[1] == 0
    mixed.pf_run()
    single.pf_run()
    for k in range(len(P)):
        Zm, om = mixed.pf_result(k)
        Zs, os_ = single.pf_result(k)
        assert abs(Zm - Zs) <= 1e-12 * abs(Zs), k
        assert abs(om - os_) <= 1e-9 * max(1.0, abs(os_)), k
        for x, y in zip(mixed.am_probs(k), single.am_probs(k)):
            assert abs(x - y) <= 1e-10, k
        for x, y in zip(mixed.pf_grad(k), single.pf_grad(k)):
            assert abs(x - y) <= 1e-9 * max(1.0, abs(y)), k
    po = oracle.locarna_p_params(max_bp_span=span)
    for k in range(0, len(P), 40):
        o = oracle.pf(pps[sel[k][0]], pps[sel[k][1]], po)
        _compare(mixed, k, o, ("split", k))
    mixed.destroy()
    single.destroy()


def _check_samples(B, k, T, S, label):
    mean, var = B.pf_sample(k, 1234 + k, S)
    g = B.pf_grad(k)
    worst = 0.0
    for t in range(api.NTHETA):
        expect = T * g[t]
        # sample variance, floored by the Poisson variance of a count with
        # this mean (rare features may not occur in S samples)
        sd = math.sqrt(max(var[t], abs(expect)) / S)
        dev = abs(mean[t] - expect)
        # 5 standard errors (275 components) plus rounding
        assert dev <= 5 * sd + 1e-7 * max(1.0, abs(expect)), (label, k, t, mean[t], expect, sd)
        if sd > 0:
            worst = max(worst, dev / sd)
    return worst


@pytest.mark.parametrize("variant", ["default", "tau100"])
def test_sampled_alignments_match_gradient(variant):
    # the mean features of sampled alignments estimate T dlogZ/dtheta (the
    # adjoint's gradient, itself checked against finite differences)
    kw = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0)
    if variant == "tau100":
        kw.update(tau=100)
    pairs = [tiny_pair(7300 + s) for s in range(8)]
    pa = api.locarna_p_params(**kw)
    P = [api.Pair(a, b, pa) for a, b in pairs]
    B = api.Batch(P)
    B.pf_run()
    for k in range(len(P)):
        _check_samples(B, k, pa.temperature_alipf, 20000, variant)
    B.destroy()


def test_sampled_alignments_workload():
    pps, pairs, span = synth.workload("pf16x150", limit_pairs=3)
    pa = api.locarna_p_params(max_bp_span=span)
    P = [api.Pair(pps[a], pps[b], pa) for a, b in pairs]
    B = api.Batch(P)
    B.pf_run()
    for k in range(len(P)):
        w = _check_samples(B, k, pa.temperature_alipf, 3000, "pf16x150")
        print("pair", k, "largest deviation / standard error", round(w, 2), flush=True)
    B.destroy()


@pytest.mark.parametrize("include_am", [False, True])
def test_probability_files(include_am):
    # locarna_p --write-arcmatch-probs / --write-basematch-probs
    # (aligner_p.icc:1404-1435): lines in the reference's order with p above
    # the cut, numbers formatted as an ostream does (%g); values against the
    # oracle within 1e-6 plus the 6-digit rounding
    kw = dict(min_prob=0.01, max_bp_span=-1, min_trace_prob=0.0, min_am_prob=0.01, min_bm_prob=0.01)
    pairs = [tiny_pair(7600 + s) for s in range(10)]
    B, P, ores = _run(pairs, kw, include_am=include_am)
    pa = api.locarna_p_params(**kw)
    for k, o in enumerate(ores):
        am = B.am_probs(k)
        want = "".join("%d %d %d %d %g\n" % (ends + (p,)) for ends, p in zip(o.arcmatches, am)
                       if p >= pa.min_am_prob)
        text = B.am_probs_text(k)
        assert text == want, k
        got = {tuple(map(int, l.split()[:4])): float(l.split()[4]) for l in text.splitlines()}
        for ends, p in zip(o.arcmatches, o.am_probs):
            if abs(p - pa.min_am_prob) <= 1e-6:
                continue
            assert (tuple(ends) in got) == (p >= pa.min_am_prob), (k, ends)
            if tuple(ends) in got:
                assert abs(got[tuple(ends)] - p) <= 1e-6 + 5e-6 * p, (k, ends)
        bm_text = B.bm_probs_text(k, include_am)
        W = o.lenB + 1
        got_bm = {tuple(map(int, l.split()[:2])): float(l.split()[2]) for l in bm_text.splitlines()}
        dense = B.bm_probs(k, 1 if include_am else 0)
        assert bm_text == "".join("%d %d %g\n" % (i, j, dense[i * W + j]) for i in range(1, o.lenA + 1)
                                  for j in range(1, o.lenB + 1) if dense[i * W + j] >= pa.min_bm_prob), k
        for i in range(1, o.lenA + 1):
            for j in range(1, o.lenB + 1):
                p = o.bm_probs[i * W + j]
                if abs(p - pa.min_bm_prob) <= 1e-6:
                    continue
                assert ((i, j) in got_bm) == (p >= pa.min_bm_prob), (k, i, j)
                if (i, j) in got_bm:
                    assert abs(got_bm[(i, j)] - p) <= 1e-6 + 5e-6 * p, (k, i, j)
    B.destroy()
