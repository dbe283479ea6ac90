"""Seeded synthetic workloads: random RNAs and their ``#PP 2.0`` files.

Sequences come from a splitmix64 stream; base-pair probabilities from the
library's McCaskill-style folding (``lcg_synth_bpp``), written with
``#BPCUT`` and filtered by span, in the format read by LocARNA's pp reader
(reference: src/LocARNA/rna_data.cc:984-1124).
"""
import ctypes
import random

from . import _lib

BPCUT = 0.0005


def sequences(seed, count, len_min, len_max, gc_min, gc_max):
    lib = _lib.load()
    stride = len_max + 1
    buf = ctypes.create_string_buffer(stride * count)
    lens = (ctypes.c_int * count)()
    rc = lib.lcg_synth_sequences(ctypes.c_uint64(seed), count, len_min, len_max,
                                 ctypes.c_double(gc_min), ctypes.c_double(gc_max),
                                 buf, stride, lens)
    if rc != 0:
        raise ValueError("lcg_synth_sequences failed")
    raw = buf.raw
    return [raw[k * stride:k * stride + lens[k]].decode() for k in range(count)]


def base_pair_probs(seq, max_span, cutoff=BPCUT):
    lib = _lib.load()
    cap = max(1024, len(seq) * 64)
    I = (ctypes.c_int * cap)()
    J = (ctypes.c_int * cap)()
    P = (ctypes.c_double * cap)()
    n = lib.lcg_synth_bpp(seq.encode(), len(seq), max_span, ctypes.c_double(cutoff), I, J, P, cap)
    if n < 0:
        raise ValueError("lcg_synth_bpp failed (%d)" % n)
    return [(I[k], J[k], P[k]) for k in range(n)]


def pp_text(name, seq, pairs, bpcut=BPCUT):
    """#PP 2.0 text; pairs (i, j, p), or (i, j, p, p2) with the joint
    probability p2 of (i, j) and (i+1, j-1) (then the section carries the
    #STACK keyword, rna_data.cc:1055-1057)."""
    stack = any(len(t) > 3 for t in pairs)
    lines = ["#PP 2.0", "", "%s %s" % (name, seq), "", "#END", "",
             "#SECTION BASEPAIRS", "", "#BPCUT %g" % bpcut] + (["#STACK"] if stack else []) + [""]
    lines += [("%d %d %.8g %.8g" % t) if len(t) > 3 else ("%d %d %.8g" % t[:3]) for t in pairs]
    lines += ["", "#END", ""]
    return "\n".join(lines)


def pp_text_rows(rows, pairs, bpcut=BPCUT):
    """#PP 2.0 text of a multiple alignment (a profile): rows (name,
    aligned sequence) of one length, base pairs (i, j, p) on its columns."""
    width = max([18] + [len(n) for n, _ in rows])
    lines = ["#PP 2.0", ""] + ["%-*s %s" % (width, n, r) for n, r in rows] + ["", "#END", "",
                                                                             "#SECTION BASEPAIRS", "",
                                                                             "#BPCUT %g" % bpcut, ""]
    lines += ["%d %d %.8g" % t[:3] for t in pairs]
    lines += ["", "#END", ""]
    return "\n".join(lines)


def with_stacking(pairs, seed=0):
    """The pairs (i, j, p) with a joint probability of (i, j) and its inner
    pair (i+1, j-1) where that pair is listed: a seeded fraction in [0.2, 1]
    of min(p, p_inner), so that it never exceeds either."""
    rng = random.Random(seed)
    prob = {(i, j): p for (i, j, p) in pairs}
    out = []
    for (i, j, p) in pairs:
        q = prob.get((i + 1, j - 1))
        if q is None:
            out.append((i, j, p))
        else:
            out.append((i, j, p, min(p, q) * rng.uniform(0.2, 1.0)))
    return out


def make_pp(name, seq, max_span, bpcut=BPCUT):
    return pp_text(name, seq, base_pair_probs(seq, max_span, bpcut), bpcut)


# Workloads.  Pair order for all-pairs sets follows mlocarna's job order
# (reference: src/Utils/mlocarna:3574-3577): a = 0..n-1, b = 0..a-1.
WORKLOADS = {
    # name: (seed, count, len_min, len_max, gc_min, gc_max, max_span, pairing)
    "pair120": (1, 2, 120, 120, 0.5, 0.5, 100, "adjacent"),
    "allpairs64": (2, 64, 60, 200, 0.35, 0.65, 100, "all"),
    "long500": (3, 8, 500, 500, 0.5, 0.5, 200, "adjacent"),
    "pf16x150": (4, 16, 150, 150, 0.5, 0.5, 100, "all"),
    "grad256": (5, 512, 80, 250, 0.5, 0.5, 100, "adjacent"),
}


def pair_list(count, pairing):
    if pairing == "all":
        return [(a, b) for a in range(count) for b in range(a)]
    return [(2 * k, 2 * k + 1) for k in range(count // 2)]


def workload(name, limit_pairs=None):
    """Returns (pp_texts, pairs, max_span) for a named workload."""
    seed, count, lmin, lmax, g0, g1, span, pairing = WORKLOADS[name]
    seqs = sequences(seed, count, lmin, lmax, g0, g1)
    pairs = pair_list(count, pairing)
    if limit_pairs is not None:
        pairs = pairs[:limit_pairs]
    used = sorted({x for p in pairs for x in p})
    pps = {k: make_pp("s%03d" % k, seqs[k], span) for k in used}
    return pps, pairs, span
