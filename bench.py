#!/usr/bin/env python3
"""Benchmark of the MI355X LocARNA path: batched all-pairs pairwise
alignment (mlocarna's all-pairs stage, reference src/Utils/mlocarna:
3547-3640, one `locarna` per pair) on one GPU.

Workload "allpairs64": 64 seeded random RNAs, lengths uniform in [60, 200],
GC uniform in [0.35, 0.65]; base-pair probabilities from the library's
McCaskill-style folding (cut 0.0005, span <= 100); all 2016 pairs aligned
globally with locarna's flags --indel=-350 --indel-opening=-500
--struct-weight=200 --tau=0 --min-prob=0.0005 --maxBPspan=100 (trace band
at the default min-trace-probability 1e-4).  One step = score and
traceback of all 2016 pairs, inputs resident in HBM, as one hipGraph-
replayed persistent launch.

Prints one JSON line (metric pairs/s).  Also reported: GCUPS, the plain
one-pair-per-launch rate of the same kernel, the integer-op roofline of the
alignment kernel (HIP events on its stream), and the CPU oracle timed on the
host cores over the same pairs (scores compared bit for bit).
"""
import argparse
import concurrent.futures as cf
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Peak int32 VALU issue rate: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz.
P_INT32_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="allpairs64")
    ap.add_argument("--mode", type=int, default=2, help="0 persistent launch, 2 hipGraph replay")
    ap.add_argument("--cpu-baseline", choices=["auto", "skip"], default="auto")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=0, help="pairs timed on the CPU (0: all)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--sequ-local", action="store_true", help="locarna --sequ-local (sequence-local top level)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the secondary configurations (long500, LocARNA-P)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes (no lcg_batch_autotune before timing)")
    ap.add_argument("--only", default=None,
                    choices=["long500_global", "long500_local", "pf16x150", "grad256", "progressive64", "fold64", "mlocarna64"],
                    help="run only this secondary configuration and print its record (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; --gpus %d is not supported (the all-pairs "
              "stage would shard the pair list across GPUs by estimated work, one share "
              "per GPU, and gather the scores)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from locarna_amd import api, synth

    if args.only:
        nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(json.dumps(extra_configs(args, api, synth, torch, nthreads, only=args.only)))
        return

    pps, pairs, span = synth.workload(args.workload)
    params = api.locarna_params(max_bp_span=span, sequ_local=1 if args.sequ_local else 0)
    nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
    # end to end: pp texts in host memory -> prepared pairs (lcg_pairs_prepare
    # on host threads) -> device arena -> one launch -> scores and alignments
    # back on the host.  Three times: trace band on the device and arc
    # matches compacted on the device (the default path), band on the host
    # with device compaction, and band, arc matches and packing on the host.
    pp_list = [(pps[a], pps[b]) for a, b in pairs]

    def end_to_end(device_compaction, device_band=False):
        torch.cuda.synchronize()
        e0 = time.perf_counter()
        prep, pst = api.prepare_pairs(pp_list, params, device_compaction=device_compaction,
                                      device_band=device_band, threads=nthreads,
                                      device=torch.cuda.current_device())
        e1 = time.perf_counter()
        b = api.Batch(prep, device=torch.cuda.current_device())
        e2 = time.perf_counter()
        # digest of every pair's device image before any run (untimed)
        image = [hashlib.blake2b(b.segment(k).tobytes(), digest_size=16).digest() for k in range(len(prep))]
        e2x = time.perf_counter()
        e0 += e2x - e2
        e1 += e2x - e2
        e2 = e2x
        b.run(mode=0, trace=True)
        e3 = time.perf_counter()
        for k in range(len(prep)):
            b.score(k)
            b.alignment(k)
        e4 = time.perf_counter()
        rec = {"pairs_per_sec": round(len(prep) / (e4 - e0), 2), "prepare_s": round(e1 - e0, 3),
               "upload_s": round(e2 - e1, 3), "run_s": round(e3 - e2, 3), "fetch_s": round(e4 - e3, 3)}
        if device_band:
            rec["band"] = pst
        return prep, b, rec, image

    prepared_host, batch_host, e2e_host, image_host = end_to_end(False)
    batch_host.destroy()
    prepared_hb, batch_hb, e2e_dev, image_hb = end_to_end(True)
    batch_hb.destroy()
    # the band kernel's one-time load and first allocation outside the timing
    api.prepare_pairs(pp_list[:64], params, device_compaction=True, device_band=True, threads=nthreads,
                      device=torch.cuda.current_device())
    prepared, batch, e2e_band, image_dev = end_to_end(True, device_band=True)
    npairs = len(prepared)
    arcm = sum(batch.pair_info(k)[0] for k in range(npairs))
    groups = sum(batch.pair_info(k)[1] for k in range(npairs))
    same_image = image_host == image_dev == image_hb
    same_band = all(x.band() == y.band() for x, y in zip(prepared, prepared_host))
    del image_host, image_dev, image_hb
    end_to_end = {"device_band_and_compaction": e2e_band, "device_compaction": e2e_dev,
                  "host_compaction": e2e_host, "host_threads": nthreads, "images_identical": same_image,
                  "bands_identical": same_band,
                  "note": "the host and device compaction paths include their kernels' one-time "
                          "loading (the band kernel is loaded before its path); upload_s of the device "
                          "compaction paths includes their two compaction kernels; prepare_s of the "
                          "device band path includes the band kernel"}

    # trace band kernel (lcg_band_pairs) over all pairs, timed by HIP events
    # around it; algorithmic bytes: the forward values written once and read
    # once (48 B per matrix cell); CPU baseline: the host band (long double,
    # as the reference) of the same pairs on the same threads
    band = None
    if args.steps > 0:
        runs = [api.prepare_pairs(pp_list, params, device_compaction=True, device_band=True, threads=nthreads,
                                  device=torch.cuda.current_device())[1] for _ in range(3)]
        b_ms = sum(r["band_ms"] for r in runs) / len(runs)
        mcells = sum((p.lenA + 1) * (p.lenB + 1) for p in prepared)
        t0 = time.perf_counter()
        api.prepare_pairs(pp_list, params, device_compaction=True, device_band=False, threads=nthreads)
        t_host = time.perf_counter() - t0
        t0 = time.perf_counter()
        api.prepare_pairs(pp_list, api.locarna_params(max_bp_span=span, min_trace_prob=0.0,
                                                      sequ_local=1 if args.sequ_local else 0),
                          device_compaction=True, device_band=False, threads=nthreads)
        t_noband = time.perf_counter() - t0
        host_band_s = t_host - t_noband
        band = {"kernel": "lcg_band_pairs", "kernel_ms": round(b_ms, 3),
                "pairs_per_sec": round(npairs / (b_ms / 1e3), 1), "matrix_cells": mcells,
                "device_banded": runs[-1]["device_banded"], "host_rechecked": runs[-1]["host_banded"],
                "roofline": {"bound": "hbm", "achieved": round(48 * mcells / (b_ms / 1e3) / 1e9, 2),
                             "peak": 8000.0, "unit": "GB/s",
                             "frac": round(48 * mcells / (b_ms / 1e3) / 1e9 / 8000.0, 5),
                             "algorithmic_bytes": 48 * mcells},
                "cpu_baseline": {"value": round(npairs / host_band_s, 1) if host_band_s > 0 else None,
                                 "unit": "pairs/s", "cores": nthreads, "kind": "port",
                                 "sample": "all %d pairs: host preparation with the long-double band minus "
                                           "host preparation without a band, same threads" % npairs},
                "bands_identical_to_host": same_band}

    # untimed counting run: DP cells and arc candidates each kernel evaluates
    batch.run(mode=0, trace=True, count=True)
    fill_cells = fill_cands = top_cells = top_cands = 0
    for k in range(npairs):
        fc, fq, tc, tq = batch.counters_split(k)
        fill_cells += fc
        fill_cands += fq
        top_cells += tc
        top_cands += tq
    cells, cands = fill_cells + top_cells, fill_cands + top_cands

    # launch shape: the autotuner times candidate shapes on this batch (each
    # must reproduce the default shape's scores and alignments) and keeps the
    # fastest; untimed
    tune = None
    if not args.no_autotune:
        a0 = time.perf_counter()
        shape, best_ms, default_ms = batch.autotune(path=0, reps=2)
        tune = {"shape": shape, "tuned_ms": round(best_ms, 3), "default_ms": round(default_ms, 3),
                "tuning_s": round(time.perf_counter() - a0, 2)}

    for _ in range(args.warmup):
        batch.run(mode=args.mode, trace=True)
    torch.cuda.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev_ms.append(batch.run(mode=args.mode, trace=True))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # per-kernel device times: the same step as persistent launches (mode 0),
    # HIP events recorded on the launch stream before, between and after the
    # fill-D and top-level kernels
    fill_ms, top_ms = [], []
    for _ in range(max(1, min(args.steps, 3))):
        batch.run(mode=0, trace=True)
        f, t = batch.kernel_times()
        fill_ms.append(f)
        top_ms.append(t)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = npairs / (ms_per_step / 1e3)
    kern_ms = sum(dev_ms) / len(dev_ms)
    gpu_scores = [batch.score(k) for k in range(npairs)]

    # plain baseline: one pair per launch, as mlocarna runs one locarna per
    # pair (src/Utils/mlocarna:2991-3022).  Each pair is its own one-pair
    # batch, at the launch shape a one-pair batch takes (its groups dealt to
    # several workgroups), timed by HIP events around its launch after one
    # warm-up run; the sum over all pairs is the plain step.  Secondary: mode
    # 1 of the batch itself (one launch per pair at the batch's shape, one
    # workgroup per pair for this workload).
    plain = plain_one_cu = plain_wg = None
    if not args.no_plain:
        plain_ms = 0.0
        for p in prepared_host:
            b1 = api.Batch([p], device=torch.cuda.current_device())
            b1.run(mode=0, trace=True)
            plain_ms += b1.run(mode=0, trace=True)
            plain_wg = b1.workgroups
            b1.destroy()
        plain = npairs / (plain_ms / 1e3)
        plain_one_cu = npairs / (batch.run(mode=1, trace=True) / 1e3)

    # integer-op roofline of the dominant kernel, lcg_fill_pairs (fill of D
    # over all arc-match groups): W_int = 11 C + 2 K + A over its own cells
    # C and candidates K, divided by its device time (HIP events around it
    # on its stream)
    w_fill = 11 * fill_cells + 2 * fill_cands + arcm
    fill_avg = sum(fill_ms) / len(fill_ms) if fill_ms else None
    achieved = w_fill / (fill_avg / 1e3) / 1e12 if fill_avg else None
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "lcg_fill_pairs_%s_traffic.json" % args.workload)
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {"bound": "valu", "achieved": round(achieved, 4) if achieved else None,
                "peak": round(P_INT32_TOPS, 2), "unit": "TOPS",
                "frac": round(achieved / P_INT32_TOPS, 5) if achieved else None, "traffic": traffic,
                "kernel": "lcg_fill_pairs", "kernel_ms": round(fill_avg, 3) if fill_avg else None,
                "int_ops_per_launch": w_fill,
                "other_kernel": {"name": "lcg_top_pairs",
                                 "ms": round(sum(top_ms) / len(top_ms), 3) if top_ms else None,
                                 "cells": top_cells, "candidates": top_cands}}

    # device arc-match compaction (count + build kernels) rerun on the same
    # inputs; algorithmic bytes: the image structures it writes (per arc
    # match 52 B: score 4, candidate 16, fill entry 12, group entry 16, fill
    # -> candidate 4; per group 32 B; per band cell 4 B), the validity bitmap
    # (written, read back: 16 B per 64-bit word) and the arc inputs
    comp = None
    if args.steps > 0:
        cms = [batch.compact_run() for _ in range(max(3, min(args.steps, 10)))]
        c_ms = sum(a for a, _ in cms) / len(cms)
        b_ms = sum(b for _, b in cms) / len(cms)
        words = sum(p.compaction_words for p in prepared)
        band_cells = sum(p.band_cells for p in prepared)
        arcs = sum(p.num_arcs for p in prepared)
        alg = 52 * arcm + 32 * groups + 4 * band_cells + 16 * words + 20 * arcs
        t_s = (c_ms + b_ms) / 1e3
        prep_diff = e2e_host["prepare_s"] - e2e_dev["prepare_s"]
        comp = {"kernels": ["lcg_compact_count", "lcg_compact_build"],
                "count_ms": round(c_ms, 3), "build_ms": round(b_ms, 3),
                "pairs_per_sec": round(npairs / t_s, 1), "arc_matches": arcm, "groups": groups,
                "roofline": {"bound": "hbm", "achieved": round(alg / t_s / 1e9, 2), "peak": 8000.0,
                             "unit": "GB/s", "frac": round(alg / t_s / 1e9 / 8000.0, 5),
                             "algorithmic_bytes": alg},
                "cpu_baseline": {"value": round(npairs / prep_diff, 1) if prep_diff > 0 else None,
                                 "unit": "pairs/s", "cores": nthreads, "kind": "port",
                                 "sample": "all %d pairs: host preparation with the host arc-match "
                                           "enumeration and packing minus host preparation without "
                                           "them, same threads" % npairs},
                "images_identical_to_host": same_image}
        batch.run(mode=args.mode, trace=True)

    cpu = None
    if args.cpu_baseline == "auto":
        from oracle import oracle
        n_s = args.cpu_sample or npairs
        idx = list(range(0, npairs, max(1, npairs // n_s)))[:n_s]
        sample = [(pps[pairs[k][0]], pps[pairs[k][1]]) for k in idx]
        t0 = time.perf_counter()
        cpu_scores = oracle.align_many(sample, oracle.locarna_params(max_bp_span=span,
                                                                     sequ_local=1 if args.sequ_local else 0),
                                       nthreads=nthreads)
        dt = time.perf_counter() - t0
        exact = sum(1 for k, s in zip(idx, cpu_scores) if gpu_scores[k] == s)
        cpu = {"value": round(len(sample) / dt, 3), "unit": "pairs/s", "cores": nthreads, "kind": "port",
               "sample": "%d of %d pairs (score + traceback per pair, CPU oracle restating locarna)"
                         % (len(sample), npairs),
               "scores_bit_exact": "%d/%d" % (exact, len(sample))}

    out = {
        "metric": "pairs/sec (batched all-pairs LocARNA alignment, score + traceback)",
        "value": round(value, 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": args.workload, "pairs": npairs, "max_bp_span": span,
                   "mode": "sequence-local" if args.sequ_local else "global",
                   "launch": {0: "persistent", 2: "hipGraph"}.get(args.mode, str(args.mode)),
                   "waves_per_workgroup": batch.waves, "workgroups": batch.workgroups,
                   "lds_bytes": batch.lds_bytes, "autotune": tune},
        "gcups": round(cells / (ms_per_step / 1e3) / 1e9, 3),
        "dp_cells_per_step": cells, "arc_candidates_per_step": cands,
        "plain_pairs_per_sec": round(plain, 2) if plain else None,
        "plain_workgroups_per_pair": plain_wg,
        "batched_vs_plain": round(value / plain, 2) if plain else None,
        "plain_one_workgroup_pairs_per_sec": round(plain_one_cu, 2) if plain_one_cu else None,
        "roofline": roofline,
        "end_to_end": end_to_end,
        "arc_match_compaction": comp,
        "trace_band": band,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["vs_cpu_baseline"] = round(value / cpu["value"], 2)
    batch.destroy()
    if not args.no_extra:
        out["other_configs"] = extra_configs(args, api, synth, torch, nthreads)
    print(json.dumps(out))


# FP64 peak of MI355X (vector and MFMA alike), TFLOP/s
P_FP64_TFLOPS = 78.6


def extra_configs(args, api, synth, torch, nthreads, only=None):
    """Secondary figures: the long-pair integer workload (global and
    sequence-local), mlocarna's progressive stage over the 64 all-pairs
    sequences, and the LocARNA-P workloads (inside + outside + gradient),
    each with its CPU oracle timed on a sample of its work."""
    from oracle import oracle
    res = {}
    if not only or only == "progressive64":
        res["progressive64"] = progressive_stage(api, synth, oracle)
    if not only or only == "fold64":
        res["fold64"] = fold_stage(api, synth, oracle, cpu_threads=nthreads)
    if not only or only == "mlocarna64":
        res["mlocarna64"] = mlocarna_stage(api, synth, oracle, cpu_threads=nthreads)
    for local in (0, 1):
        if only and only != "long500_" + ("local" if local else "global"):
            continue
        pps, pairs, span = synth.workload("long500")
        pa = api.locarna_params(max_bp_span=span, sequ_local=local)
        P = [api.Pair(pps[a], pps[b], pa) for a, b in pairs]
        B = api.Batch(P)
        B.run(mode=0, trace=True, count=True)
        cells = sum(B.counters(k)[0] for k in range(len(P)))
        tune = None
        if not args.no_autotune:
            shape, best_ms, default_ms = B.autotune(path=0, reps=1)
            tune = {"shape": shape, "tuned_ms": round(best_ms, 3), "default_ms": round(default_ms, 3)}
        B.run(mode=2, trace=True)
        ms = [B.run(mode=2, trace=True) for _ in range(3)]
        t = sum(ms) / len(ms)
        scores = [B.score(k) for k in range(len(P))]
        B.destroy()
        t0 = time.perf_counter()
        cs = oracle.align_many([(pps[a], pps[b]) for a, b in pairs],
                               oracle.locarna_params(max_bp_span=span, sequ_local=local), nthreads=nthreads)
        dt = time.perf_counter() - t0
        res["long500_" + ("local" if local else "global")] = {
            "pairs": len(P), "ms_per_step": round(t, 3), "pairs_per_sec": round(len(P) / (t / 1e3), 2),
            "gcups": round(cells / (t / 1e3) / 1e9, 3), "workgroups_per_pair": B.workgroups // len(P) if B.workgroups else None,
            "cpu_baseline_pairs_per_sec": round(len(P) / dt, 3), "cpu_cores": nthreads,
            "scores_bit_exact": "%d/%d" % (sum(1 for x, y in zip(scores, cs) if x == y), len(P)),
            "autotune": tune}
    for wl, nsample in (("pf16x150", 32), ("grad256", 32)):
        if only and only != wl:
            continue
        pps, pairs, span = synth.workload(wl)
        pa = api.locarna_p_params(max_bp_span=span)
        P = [api.Pair(pps[a], pps[b], pa) for a, b in pairs]
        B = api.Batch(P)
        B.run(mode=0, trace=False, count=True)
        C = K = 0
        for k in range(len(P)):
            c, q = B.counters(k)
            C += c
            K += q
        A = sum(p.num_arcmatches for p in P)
        tune = None
        if not args.no_autotune:
            shape, best_ms, default_ms = B.autotune(path=1, reps=1)
            tune = {"shape": {k: shape[k] for k in ("pf_waves", "pf_wg_per_pair")},
                    "tuned_ms": round(best_ms, 3), "default_ms": round(default_ms, 3)}
        B.pf_run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev = []
        steps = 3
        for _ in range(steps):
            dev.append(B.pf_run())
            grads = [B.pf_grad(k) for k in range(len(P))]
        wall = (time.perf_counter() - t0) / steps
        dev_ms = sum(dev) / len(dev)
        w_pf = 3 * (13 * C + 2 * K + A)
        ach = w_pf / (dev_ms / 1e3) / 1e12
        idx = list(range(0, len(P), max(1, len(P) // nsample)))[:nsample]
        po = oracle.locarna_p_params(max_bp_span=span)
        t0 = time.perf_counter()
        import concurrent.futures as cf2
        with cf2.ThreadPoolExecutor(nthreads) as ex:
            ores = list(ex.map(lambda k: oracle.pf(pps[pairs[k][0]], pps[pairs[k][1]], po), idx))
        dt = time.perf_counter() - t0
        import math
        ok = sum(1 for k, o in zip(idx, ores)
                 if abs(math.log(B.pf_result(k)[0]) - math.log(o.pf)) <= 1e-9 * max(1.0, abs(math.log(o.pf))))
        fit = None
        if wl == "grad256":
            # 20 gradient steps through the torch.autograd binding, fitting theta
            # to 64 alignments per pair sampled (stochastic traceback, seed 6)
            # at a seeded generating theta
            from locarna_amd.torch_grad import fit_to_target, generating_theta, sampled_statistics
            theta_gen = generating_theta(seed=6, scale=30.0)
            t0 = time.perf_counter()
            target = sampled_statistics(B, theta_gen, samples=64, seed=6).sum(0)
            t_sample = time.perf_counter() - t0
            t0 = time.perf_counter()
            losses, dist, _ = fit_to_target(B, theta_gen, target, 20, 0.7, None, method="newton", samples=32,
                                            ridge=1.0 / (30.0 * 30.0 * 64))
            fit = {"steps": 20, "s_per_step": round((time.perf_counter() - t0) / 20, 3),
                   "method": "Newton steps, diagonal Hessian from 32 sampled alignments per pair and step; "
                             "Gaussian prior of the generating theta (sd 30) as a ridge term",
                   "target": "64 sampled alignments per pair at the generating theta",
                   "sampling_s": round(t_sample, 3),
                   "loss_first": round(losses[0], 4), "loss_last": round(losses[-1], 4),
                   "theta_dist_first": round(dist[0], 3), "theta_dist_last": round(dist[-1], 3)}
        B.destroy()
        res[wl] = {"pairs": len(P), "pf_ms_per_step": round(dev_ms, 3),
                   "grad_pairs_per_sec": round(len(P) / wall, 2),
                   "pf_pairs_per_sec": round(len(P) / (dev_ms / 1e3), 2), "dtype": "f64",
                   "roofline": {"bound": "valu", "achieved": round(ach, 5), "peak": P_FP64_TFLOPS, "unit": "TFLOP/s",
                                "frac": round(ach / P_FP64_TFLOPS, 6), "kernel": "lcg_pf_pairs",
                                "flops_per_launch": w_pf},
                   "cpu_baseline_pairs_per_sec": round(len(idx) / dt, 3), "cpu_cores": nthreads,
                   "cpu_sample": "%d of %d pairs (inside + outside, CPU oracle)" % (len(idx), len(P)),
                   "logZ_within_1e-9": "%d/%d" % (ok, len(idx)), "autotune": tune}
        if fit:
            res[wl]["fit"] = fit
    return res


def progressive_stage(api, synth, oracle, cpu_budget_s=15.0):
    """mlocarna after the all-pairs stage: UPGMA guide tree from the 2016
    scores, then the 63 merges of the progressive alignment, each a
    profile-profile alignment whose consensus pp feeds the next; merges
    whose children are done run as one batch.  The CPU oracle runs the
    merges in tree order for about cpu_budget_s seconds on one core; each
    of them is checked against the GPU's consensus text."""
    pps, pairs, span = synth.workload("allpairs64")
    n = 64
    texts = [pps[x] for x in range(n)]
    pa = api.locarna_params(max_bp_span=span)
    B = api.Batch([api.Pair(texts[a], texts[b], pa) for a, b in pairs])
    B.run(mode=2)
    S = [[0] * n for _ in range(n)]
    for k, (a, b) in enumerate(pairs):
        S[a][b] = S[b][a] = B.score(k)
    B.destroy()
    t0 = time.perf_counter()
    merges = api.guide_tree_upgma(S)
    t_tree = time.perf_counter() - t0
    api.Progressive(texts, merges, pa).destroy()  # warm-up (kernel load, first allocations)
    t0 = time.perf_counter()
    R = api.Progressive(texts, merges, pa)
    wall = time.perf_counter() - t0
    st = R.stats()
    po = oracle.locarna_params(max_bp_span=span)
    text = list(texts)
    done, ok = 0, 0
    t0 = time.perf_counter()
    for t, (a, b) in enumerate(merges):
        if text[a] is None or text[b] is None:
            text.append(None)
            continue
        r = oracle.align(text[a], text[b], po)
        text.append(oracle.consensus_pp(text[a], text[b], r.edges, po))
        done += 1
        ok += int(R.pp(n + t) == text[-1])
        if time.perf_counter() - t0 > cpu_budget_s:
            break
    dt = time.perf_counter() - t0
    rows = len([l for l in R.pp(2 * n - 2).split("#END")[0].split("\n")[2:] if l.strip()])
    R.destroy()
    return {"sequences": n, "merges": n - 1, "rows_final": rows, "rounds": st["rounds"], "batches": st["batches"],
            "guide_tree_ms": round(t_tree * 1e3, 3), "wall_ms": round(wall * 1e3, 2),
            "device_ms": round(st["device_ms"], 3), "prepare_ms": round(st["prepare_ms"], 2),
            "merges_per_sec": round((n - 1) / wall, 2),
            "cpu_baseline_merges_per_sec": round(done / dt, 3), "cpu_cores": 1,
            "cpu_sample": "%d of %d merges in tree order (CPU oracle, one core)" % (done, n - 1),
            "consensus_bit_exact": "%d/%d" % (ok, done)}


def fold_stage(api, synth, oracle, reps=5, cpu_threads=16):
    """mlocarna's input step before the all-pairs stage: base-pair
    probabilities of the 64 allpairs64 sequences (60-200 nt, pairs of span
    <= 100, pairs with p > 0.0005 written as pp texts), one kernel launch for
    all of them.  The CPU oracle folds the same 64 sequences on
    cpu_threads host threads; every GPU probability is checked against it."""
    seqs = synth.sequences(2, 64, 60, 200, 0.35, 0.65)
    api.Fold(seqs, max_bp_span=100, min_prob=synth.BPCUT).destroy()  # warm-up
    dev, wall = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        F = api.Fold(seqs, max_bp_span=100, min_prob=synth.BPCUT)
        texts = [F.pp(k, "s%03d" % k) for k in range(len(seqs))]
        wall.append(time.perf_counter() - t0)
        dev.append(F.stats()["device_ms"])
        if _ < reps - 1:
            F.destroy()
    t0 = time.perf_counter()
    ref = oracle.fold_many(seqs, 100, cpu_threads)
    dt = time.perf_counter() - t0
    err = 0.0
    for k in range(len(seqs)):
        P, P2 = F.probs(k)
        err = max(err, float(abs(P - ref[k][0]).max()), float(abs(P2 - ref[k][1]).max()))
    F.destroy()
    dms = sorted(dev)[len(dev) // 2]
    return {"sequences": len(seqs), "nt": sum(len(x) for x in seqs), "max_bp_span": 100,
            "kernel_ms": round(dms, 3), "wall_ms_with_pp_texts": round(sorted(wall)[len(wall) // 2] * 1e3, 2),
            "sequences_per_sec": round(len(seqs) / (dms / 1e3), 1),
            "cpu_baseline_sequences_per_sec": round(len(seqs) / dt, 2), "cpu_cores": cpu_threads,
            "cpu_sample": "all 64 sequences (CPU oracle, %d threads)" % cpu_threads,
            "max_abs_diff_vs_oracle": err, "pp_text_bytes": sum(len(t) for t in texts)}


def mlocarna_stage(api, synth, oracle, cpu_threads=16):
    """mlocarna's whole default run from the 64 allpairs64 sequences to the
    final multiple alignment on one GPU (lcg_mlocarna_run: folding, the
    2016 pairs, UPGMA, 63 progressive merges).  The CPU oracle runs the same
    pipeline (folding and all pairs on cpu_threads threads, the merges on
    one core, as mlocarna runs them in tree order) and its final consensus
    text is compared with the GPU's."""
    seqs = synth.sequences(2, 64, 60, 200, 0.35, 0.65)
    names = ["s%03d" % k for k in range(len(seqs))]
    n = len(seqs)
    api.Mlocarna(names, seqs, max_bp_span=100).destroy()  # warm-up
    t0 = time.perf_counter()
    M = api.Mlocarna(names, seqs, max_bp_span=100)
    wall = time.perf_counter() - t0
    st = M.stats()
    final = M.pp(2 * n - 2)
    merges = M.tree()
    M.destroy()
    po = oracle.locarna_params(max_bp_span=100)
    t0 = time.perf_counter()
    folded = oracle.fold_many(seqs, 100, cpu_threads)
    pps = [oracle.fold_pp(names[k], seqs[k], folded[k][0], folded[k][1], 0.0005) for k in range(n)]
    pairs = [(a, b) for a in range(n) for b in range(a)]
    sc = oracle.align_many([(pps[a], pps[b]) for a, b in pairs], po, cpu_threads)
    S = [[0] * n for _ in range(n)]
    for (a, b), v in zip(pairs, sc):
        S[a][b] = S[b][a] = v
    ref_merges = oracle.upgma(S)
    ref_final, _ = oracle.progressive(pps, ref_merges, po)
    dt = time.perf_counter() - t0
    return {"sequences": n, "pairs": len(pairs), "merges": n - 1, "wall_ms": round(wall * 1e3, 2),
            "fold_ms": round(st["fold_ms"], 2), "pairs_ms": round(st["pairs_ms"], 2),
            "progressive_ms": round(st["progressive_ms"], 2), "runs_per_sec": round(1.0 / wall, 3),
            "cpu_baseline_s": round(dt, 2), "cpu_cores": cpu_threads,
            "cpu_sample": "the whole pipeline (CPU oracle: folding and pairs on %d threads, merges on one core)"
                          % cpu_threads,
            "speedup_vs_cpu": round(dt / wall, 1),
            "same_guide_tree": [tuple(m) for m in ref_merges] == merges,
            "final_alignment_equal": ref_final == final}


if __name__ == "__main__":
    main()
