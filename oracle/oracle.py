"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
RIBOSUM = os.path.join(HERE, "data", "RIBOSUM85_60")


class Params(ctypes.Structure):
    _fields_ = [
        ("indel", ctypes.c_int), ("indel_opening", ctypes.c_int),
        ("struct_weight", ctypes.c_int), ("tau", ctypes.c_int),
        ("match", ctypes.c_int), ("mismatch", ctypes.c_int),
        ("temperature_alipf", ctypes.c_int), ("sequ_local", ctypes.c_int),
        ("max_bp_span", ctypes.c_int), ("band_extended", ctypes.c_int),
        ("use_ribosum", ctypes.c_int), ("free_endgaps", ctypes.c_int),
        ("max_diff", ctypes.c_int), ("max_diff_am", ctypes.c_int), ("max_diff_at_am", ctypes.c_int),
        ("min_prob", ctypes.c_double), ("min_trace_prob", ctypes.c_double),
        ("exp_prob", ctypes.c_double), ("pf_scale", ctypes.c_double),
        ("min_am_prob", ctypes.c_double), ("min_bm_prob", ctypes.c_double),
        ("no_lonely_pairs", ctypes.c_int), ("stacking", ctypes.c_int),
        ("struct_local", ctypes.c_int), ("exclusion", ctypes.c_int),
    ]


def locarna_params(**kw):
    """Defaults of the benchmark configurations (locarna CLI flags
    --indel=-350 --indel-opening=-500 --struct-weight=200 --tau=0
    --min-prob=0.0005 --maxBPspan=100, min-trace-probability 1e-4)."""
    d = dict(indel=-350, indel_opening=-500, struct_weight=200, tau=0, match=50,
             mismatch=0, temperature_alipf=300, sequ_local=0, max_bp_span=100,
             band_extended=1, use_ribosum=1, free_endgaps=0, max_diff=-1, max_diff_am=-1, max_diff_at_am=-1,
             min_prob=0.0005, min_trace_prob=1e-4,
             exp_prob=-1.0, pf_scale=1.0, min_am_prob=0.001, min_bm_prob=0.001, no_lonely_pairs=0)
    d.update(kw)
    return Params(**d)


def locarna_p_params(**kw):
    """locarna_p defaults used by the probabilistic configurations: tau 50,
    double band PF, min-trace-probability 1e-5."""
    d = dict(tau=50, band_extended=0, min_trace_prob=1e-5)
    d.update(kw)
    return locarna_params(**d)


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.oracle_align.restype = ctypes.c_void_p
        L.oracle_align.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Params),
                                   ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.oracle_pf.restype = ctypes.c_void_p
        L.oracle_pf.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Params),
                                ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.oracle_align_variant.restype = ctypes.c_void_p
        L.oracle_align_variant.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Params),
                                           ctypes.c_char_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_char_p,
                                           ctypes.c_int]
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_score.restype = ctypes.c_longlong
        L.oracle_score.argtypes = [ctypes.c_void_p]
        L.oracle_partition_function.restype = ctypes.c_double
        L.oracle_partition_function.argtypes = [ctypes.c_void_p]
        for f in ("oracle_lengths", "oracle_edges", "oracle_structures", "oracle_band",
                  "oracle_arcmatches", "oracle_counters", "oracle_am_probs",
                  "oracle_bm_probs", "oracle_pf_D", "oracle_tables", "oracle_pf_extra"):
            getattr(L, f).argtypes = None
        _lib = L
    return _lib


class Result:
    """Integer (or PF) result of one pair."""

    def __init__(self, handle):
        L = load()
        self._h = ctypes.c_void_p(handle)
        a, b = ctypes.c_int(), ctypes.c_int()
        L.oracle_lengths(self._h, ctypes.byref(a), ctypes.byref(b))
        self.lenA, self.lenB = a.value, b.value
        self.score = L.oracle_score(self._h)
        self.pf = L.oracle_partition_function(self._h)
        cap = self.lenA + self.lenB + 2
        ea, eb = (ctypes.c_int * cap)(), (ctypes.c_int * cap)()
        n = L.oracle_edges(self._h, ea, eb, cap)
        self.edges = [(ea[k], eb[k]) for k in range(n)]
        sa = ctypes.create_string_buffer(cap + 2)
        sb = ctypes.create_string_buffer(cap + 2)
        L.oracle_structures(self._h, sa, sb, cap + 2)
        self.structA, self.structB = sa.value.decode(), sb.value.decode()
        mn, mx = (ctypes.c_int * (self.lenA + 1))(), (ctypes.c_int * (self.lenA + 1))()
        L.oracle_band(self._h, mn, mx, self.lenA + 1)
        self.min_col, self.max_col = list(mn), list(mx)
        c6 = (ctypes.c_longlong * 6)()
        L.oracle_counters(self._h, c6)
        self.counters = dict(zip(["N_A", "N_B", "A", "G", "C", "K"], list(c6)))
        na = self.counters["A"]
        arrs = [(ctypes.c_int * max(na, 1))() for _ in range(4)]
        sc, D = (ctypes.c_longlong * max(na, 1))(), (ctypes.c_longlong * max(na, 1))()
        L.oracle_arcmatches(self._h, *arrs, sc, D, na)
        self.arcmatches = [(arrs[0][k], arrs[1][k], arrs[2][k], arrs[3][k]) for k in range(na)]
        self.am_scores = list(sc)[:na]
        self.D = list(D)[:na]
        st, stk = (ctypes.c_longlong * max(na, 1))(), (ctypes.c_int * max(na, 1))()
        ns = L.oracle_stacked_scores(self._h, st, stk, max(na, 1))
        self.am_scores_stacked = [st[k] if stk[k] else None for k in range(max(ns, 0))]
        nb = (self.lenA + 1) * (self.lenB + 1)
        sig = (ctypes.c_longlong * nb)()
        gA = (ctypes.c_longlong * (self.lenA + 1))()
        gB = (ctypes.c_longlong * (self.lenB + 1))()
        L.oracle_tables(self._h, sig, gA, gB)
        self.sigma = {(i, j): sig[i * (self.lenB + 1) + j]
                      for i in range(1, self.lenA + 1) for j in range(1, self.lenB + 1)}
        self.gapA, self.gapB = list(gA), list(gB)
        pr = (ctypes.c_double * max(na, 1))()
        npr = L.oracle_am_probs(self._h, pr, na)
        self.am_probs = list(pr)[:npr]
        nb = (self.lenA + 1) * (self.lenB + 1)
        bm = (ctypes.c_double * nb)()
        nbm = L.oracle_bm_probs(self._h, bm, nb)
        self.bm_probs = list(bm)[:nbm]
        pd, pdp = (ctypes.c_double * max(na, 1))(), (ctypes.c_double * max(na, 1))()
        npd = L.oracle_pf_D(self._h, pd, pdp, na)
        self.pf_D, self.pf_Dprime = list(pd)[:npd], list(pdp)[:npd]
        bf = (ctypes.c_double * nb)()
        op = ctypes.c_double()
        nbf = L.oracle_pf_extra(self._h, bf, nb, ctypes.byref(op))
        self.bm_full = list(bf)[:nbf]
        self.openings = op.value
        L.oracle_free(self._h)
        self._h = None


def align(ppA, ppB, params=None, trace=True):
    L = load()
    params = params or locarna_params()
    err = ctypes.create_string_buffer(512)
    h = L.oracle_align(ppA.encode(), ppB.encode(), ctypes.byref(params), RIBOSUM.encode(),
                       1 if trace else 0, err, 512)
    if not h:
        raise RuntimeError("oracle_align: " + err.value.decode())
    return Result(h)


def align_variant(ppA, ppB, params=None, penalized=None, normalized=None):
    """Penalized (position penalty) or normalized (L, sequence-local)
    alignment with traceback (locarna --penalized / --normalized)."""
    L = load()
    params = params or locarna_params()
    mode, arg = (1, penalized) if penalized is not None else (2, normalized)
    err = ctypes.create_string_buffer(512)
    h = L.oracle_align_variant(ppA.encode(), ppB.encode(), ctypes.byref(params), RIBOSUM.encode(), mode, arg,
                               err, 512)
    if not h:
        raise RuntimeError("oracle_align_variant: " + err.value.decode())
    return Result(h)


def pf(ppA, ppB, params=None, include_am_in_bm=False):
    L = load()
    params = params or locarna_p_params()
    err = ctypes.create_string_buffer(512)
    h = L.oracle_pf(ppA.encode(), ppB.encode(), ctypes.byref(params), RIBOSUM.encode(),
                    1 if include_am_in_bm else 0, err, 512)
    if not h:
        raise RuntimeError("oracle_pf: " + err.value.decode())
    return Result(h)


def suboptimal(ppA, ppB, k, threshold=-1000000, params=None, cap=64):
    """k-best suboptimal alignments (locarna --kbest k --better threshold):
    list of (score, edges, (startA, endA, startB, endB))."""
    L = load()
    params = params or locarna_params()
    ce = len(ppA) + len(ppB) + 2  # above n + m
    scores = (ctypes.c_longlong * cap)()
    ea, eb = (ctypes.c_int * (cap * ce))(), (ctypes.c_int * (cap * ce))()
    cnt = (ctypes.c_int * cap)()
    restr = (ctypes.c_int * (4 * cap))()
    err = ctypes.create_string_buffer(512)
    L.oracle_suboptimal.restype = ctypes.c_int
    L.oracle_suboptimal.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Params), ctypes.c_char_p,
                                    ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_char_p, ctypes.c_int]
    n = L.oracle_suboptimal(ppA.encode(), ppB.encode(), ctypes.byref(params), RIBOSUM.encode(), k, threshold,
                            scores, restr, ea, eb, cnt, cap, ce, err, 512)
    if n < 0:
        raise RuntimeError("oracle_suboptimal: " + err.value.decode())
    return [(scores[t], [(ea[t * ce + e], eb[t * ce + e]) for e in range(cnt[t])],
             tuple(restr[4 * t + x] for x in range(4))) for t in range(n)]


def pf_fragments(ppA, ppB, frags, params=None):
    """Fragment match probabilities (locarna_p --fragment-match-probs) of
    the fragments (i, j, k, l): A[i..j] matched with B[k..l]."""
    L = load()
    params = params or locarna_p_params()
    err = ctypes.create_string_buffer(512)
    h = L.oracle_pf(ppA.encode(), ppB.encode(), ctypes.byref(params), RIBOSUM.encode(), 0, err, 512)
    if not h:
        raise RuntimeError("oracle_pf: " + err.value.decode())
    try:
        nq = len(frags)
        q = (ctypes.c_int * max(1, 4 * nq))(*[x for f in frags for x in f])
        prob = (ctypes.c_double * max(1, nq))()
        L.oracle_pf_fragments.restype = ctypes.c_int
        L.oracle_pf_fragments.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_char_p, ctypes.c_int]
        rc = L.oracle_pf_fragments(h, nq, q, prob, err, 512)
        if rc < 0:
            raise RuntimeError("oracle_pf_fragments: " + err.value.decode())
        return list(prob)[:nq]
    finally:
        L.oracle_free(ctypes.c_void_p(h))


def align_many(pp_pairs, params=None, nthreads=1):
    L = load()
    params = params or locarna_params()
    n = len(pp_pairs)
    A = (ctypes.c_char_p * n)(*[a.encode() for a, _ in pp_pairs])
    B = (ctypes.c_char_p * n)(*[b.encode() for _, b in pp_pairs])
    scores = (ctypes.c_longlong * n)()
    rc = L.oracle_align_many(A, B, n, ctypes.byref(params), RIBOSUM.encode(), nthreads, scores)
    if rc < 0:
        raise RuntimeError("oracle_align_many failed")
    return list(scores)


def trace_probs(ppA, ppB, params=None):
    """Band step of the oracle over the full matrix: (probs, sim, (w_open,
    w_ext, T)) with probs[i][j] the trace probability of cell (i, j) and
    sim[i][j] the similarity sigma(i, j) it used (i, j >= 1)."""
    L = load()
    params = params or locarna_params()
    cap = (len(ppA) + 2) * (len(ppB) + 2)  # the lengths are below the text sizes
    probs = (ctypes.c_double * cap)()
    sim = (ctypes.c_double * cap)()
    w = (ctypes.c_double * 5)()
    L.oracle_trace_probs.restype = ctypes.c_int
    L.oracle_trace_probs.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Params), ctypes.c_char_p,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    cnt = L.oracle_trace_probs(ppA.encode(), ppB.encode(), ctypes.byref(params), RIBOSUM.encode(), probs, sim, w,
                               cap)
    if cnt < 0:
        raise RuntimeError("oracle_trace_probs failed (%d)" % cnt)
    n, m = int(w[3]), int(w[4])
    P = [[probs[i * (m + 1) + j] for j in range(m + 1)] for i in range(n + 1)]
    S = [[sim[i * (m + 1) + j] for j in range(m + 1)] for i in range(n + 1)]
    return P, S, (w[0], w[1], w[2])


def consensus_pp(ppA, ppB, edges, params=None):
    """pp text of the pair aligned by `edges` ((a, b), -1 for a gap), as
    locarna --pp writes it (consensus base-pair probabilities)."""
    L = load()
    params = params or locarna_params()
    n = len(edges)
    ea = (ctypes.c_int * max(n, 1))(*[e[0] for e in edges])
    eb = (ctypes.c_int * max(n, 1))(*[e[1] for e in edges])
    err = ctypes.create_string_buffer(512)
    L.oracle_consensus_pp.restype = ctypes.c_longlong
    L.oracle_consensus_pp.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Params), ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_longlong,
                                      ctypes.c_char_p, ctypes.c_int]
    need = L.oracle_consensus_pp(ppA.encode(), ppB.encode(), ctypes.byref(params), ea, eb, n, None, 0, err, 512)
    if need < 0:
        raise RuntimeError("oracle_consensus_pp: " + err.value.decode())
    buf = ctypes.create_string_buffer(need + 1)
    L.oracle_consensus_pp(ppA.encode(), ppB.encode(), ctypes.byref(params), ea, eb, n, buf, need + 1, err, 512)
    return buf.value.decode()


def format_prob(p):
    L = load()
    buf = ctypes.create_string_buffer(64)
    L.oracle_format_prob.restype = ctypes.c_int
    L.oracle_format_prob.argtypes = [ctypes.c_double, ctypes.c_char_p, ctypes.c_int]
    L.oracle_format_prob(p, buf, 64)
    return buf.value.decode()


def upgma(scores):
    """Guide tree of mlocarna from a symmetric similarity matrix (lists of
    lists; the diagonal is ignored and read as 0): MLocarna::Tree::_upgma
    (lib/perl/MLocarna/Tree.pm:164-315).  Distances are max - score, max
    over the upper triangle and the (zeroed) [0][0] entry; the closest pair
    is the first strict minimum in the scan over the current cluster list,
    whose order follows the Perl list updates.  Returns the merges as a
    list of (child0, child1): leaves 0..n-1, merge t is node n + t."""
    n = len(scores)
    if n < 2:
        return []
    mx = 0
    for i in range(n):
        for j in range(i + 1, n):
            mx = max(scores[i][j], mx)
    dist = [[mx - (0 if i == j else scores[i][j]) for j in range(n)] for i in range(n)]
    clusters = list(range(n))
    size = [1] * n
    node = list(range(n))  # cluster id -> tree node
    merges = []
    while len(clusters) > 1:
        min_i = min_j = None
        min_d = 1e10
        for i in range(len(clusters)):
            for j in range(i + 1, len(clusters)):
                d = dist[clusters[i]][clusters[j]]
                if d < min_d:
                    min_i, min_j, min_d = i, j, d
        ci, cj = clusters[min_i], clusters[min_j]
        last = len(clusters) - 1
        clusters[min_j] = clusters[last]
        clusters[min_i] = clusters[0]
        clusters[0] = ci
        for i in range(1, last):
            dist[clusters[0]][clusters[i]] = ((size[ci] * dist[ci][clusters[i]] + size[cj] * dist[cj][clusters[i]])
                                              / (size[ci] + size[cj]))
            dist[clusters[i]][clusters[0]] = dist[clusters[0]][clusters[i]]
        merges.append((node[ci], node[cj]))
        node[clusters[0]] = n + len(merges) - 1
        size[clusters[0]] = size[ci] + size[cj]
        clusters.pop()
    return merges


def progressive(pps, merges, params=None):
    """mlocarna's progressive stage over the guide tree `merges`
    (src/Utils/mlocarna:3660-3707, call_locarna): every merge aligns the
    pp files of its two children and writes the consensus pp of the result
    (locarna --pp); returns (final pp text, per-merge (score, edges))."""
    params = params or locarna_params()
    n = len(pps)
    text = list(pps) + [None] * len(merges)
    out = []
    for t, (a, b) in enumerate(merges):
        r = align(text[a], text[b], params)
        out.append((r.score, r.edges))
        text[n + t] = consensus_pp(text[a], text[b], r.edges, params)
    return (text[-1] if merges else pps[0]), out


def fold(seq, max_span=-1):
    """Base-pair probabilities of one sequence (McCaskill over the library's
    energy model, oracle/fold.cc): (P, P2, log Z) with P and P2 dense
    (n+1) x (n+1) numpy arrays (P2: joint probability of (i,j) and
    (i+1,j-1))."""
    import numpy as np
    L = load()
    L.oracle_fold.restype = ctypes.c_int
    L.oracle_fold.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.POINTER(ctypes.c_double)]
    n = len(seq)
    P = np.zeros((n + 1, n + 1))
    P2 = np.zeros((n + 1, n + 1))
    lz = ctypes.c_double()
    r = L.oracle_fold(seq.encode(), max_span, P.ctypes.data, P2.ctypes.data, ctypes.byref(lz))
    if r != n:
        raise ValueError("oracle_fold failed")
    return P, P2, lz.value


def fold_many(seqs, max_span=-1, nthreads=1):
    """fold() of every sequence on nthreads host threads (the C function
    runs without the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=max(1, nthreads)) as ex:
        return list(ex.map(lambda s: fold(s, max_span), seqs))


def fold_pp(name, seq, P, P2, min_prob, stacking=False):
    """The pp text locarna_rnafold_pp writes for these probabilities
    (RnaData::write_pp, rna_data.cc:1242-1350; pairs with p > min_prob,
    i and j ascending, the joint probability where stacking and above the
    cut)."""
    seq = seq.upper().replace("T", "U")
    lines = ["#PP 2.0", "", "%-18s %s" % (name, seq), "", "#END", "", "#SECTION BASEPAIRS", "",
             "#BPCUT " + format_prob(min_prob)] + (["#STACK"] if stacking else []) + [""]
    n = len(seq)
    for i in range(1, n + 1):
        for j in range(i + 1, n + 1):
            p = P[i][j]
            if p > min_prob:
                s = "%d %d %s" % (i, j, format_prob(p))
                if stacking and P2[i][j] > min_prob:
                    s += " " + format_prob(P2[i][j])
                lines.append(s)
    return "\n".join(lines + ["", "#END", ""])
