"""Thin Python binding of the C ABI (include/locarna_amd.h), used by the
tests and bench.py.  All compute happens in liblocarna_amd.so; there is no
Python or CPU fallback for the alignment kernels."""
import ctypes

from . import _lib


class Params(ctypes.Structure):
    """lcg_params (include/locarna_amd.h)."""
    _fields_ = [
        ("indel", ctypes.c_int32), ("indel_opening", ctypes.c_int32),
        ("struct_weight", ctypes.c_int32), ("tau", ctypes.c_int32),
        ("match", ctypes.c_int32), ("mismatch", ctypes.c_int32),
        ("temperature_alipf", ctypes.c_int32), ("sequ_local", ctypes.c_int32),
        ("max_bp_span", ctypes.c_int32), ("band_extended", ctypes.c_int32),
        ("use_ribosum", ctypes.c_int32), ("free_endgaps", ctypes.c_int32),
        ("max_diff", ctypes.c_int32), ("max_diff_am", ctypes.c_int32), ("max_diff_at_am", ctypes.c_int32),
        ("min_prob", ctypes.c_double), ("min_trace_prob", ctypes.c_double),
        ("exp_prob", ctypes.c_double), ("pf_scale", ctypes.c_double),
        ("min_am_prob", ctypes.c_double), ("min_bm_prob", ctypes.c_double),
        ("no_lonely_pairs", ctypes.c_int32), ("stacking", ctypes.c_int32),
        ("struct_local", ctypes.c_int32), ("exclusion", ctypes.c_int32),
    ]


def locarna_params(**kw):
    """locarna flags of the integer workloads: --indel=-350
    --indel-opening=-500 --struct-weight=200 --tau=0 --min-prob=0.0005
    --maxBPspan=100; other options at their locarna defaults
    (min-trace-probability 1e-4, trace band in long double)."""
    d = dict(indel=-350, indel_opening=-500, struct_weight=200, tau=0, match=50,
             mismatch=0, temperature_alipf=300, sequ_local=0, max_bp_span=100,
             band_extended=1, use_ribosum=1, free_endgaps=0, max_diff=-1, max_diff_am=-1, max_diff_at_am=-1,
             min_prob=0.0005, min_trace_prob=1e-4,
             exp_prob=-1.0, pf_scale=1.0, min_am_prob=0.001, min_bm_prob=0.001)
    d.update(kw)
    return Params(**d)


def free_endgaps(spec):
    """Bits of a locarna --free-endgaps string ("++--": left and right end
    of the first sequence, left and right end of the second)."""
    if len(spec) != 4 or any(c not in "+-" for c in spec):
        raise ValueError("free end gaps: 4 characters of + and -")
    return sum(1 << k for k, c in enumerate(spec) if c == "+")


def locarna_p_params(**kw):
    """locarna_p defaults of the probabilistic workloads: tau 50, trace band
    in double, min-trace-probability 1e-5."""
    d = dict(tau=50, band_extended=0, min_trace_prob=1e-5)
    d.update(kw)
    return locarna_params(**d)


class PrepareOpts(ctypes.Structure):
    """lcg_prepare_opts (include/locarna_amd.h)."""
    _fields_ = [("flags", ctypes.c_int32), ("device", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("band_eps", ctypes.c_double)]


class PrepareStats(ctypes.Structure):
    """lcg_prepare_stats (include/locarna_amd.h)."""
    _fields_ = [("band_ms", ctypes.c_float), ("device_banded", ctypes.c_int32),
                ("host_banded", ctypes.c_int32), ("parse_ms", ctypes.c_float), ("band_wall_ms", ctypes.c_float),
                ("finish_ms", ctypes.c_float)]


class Shape(ctypes.Structure):
    """lcg_shape (include/locarna_amd.h): launch shape, 0 = default."""
    _fields_ = [(n, ctypes.c_int32) for n in ("waves", "wgs_per_cu", "wg_per_pair", "top_waves",
                                              "top_wgs_per_cu", "pf_waves", "pf_wg_per_pair", "reserved")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


class LcgError(RuntimeError):
    pass


class FoldParams(ctypes.Structure):
    """lcg_fold_params (include/locarna_amd.h): locarna_rnafold_pp's
    --maxBPspan, --stacking and -p/--min-prob."""
    _fields_ = [("max_bp_span", ctypes.c_int32), ("stacking", ctypes.c_int32), ("min_prob", ctypes.c_double),
                ("scale_kcal_per_nt", ctypes.c_double)]


def _sig(L):
    c = ctypes
    L.lcg_pair_prepare.argtypes = [c.c_char_p, c.c_char_p, c.POINTER(Params), c.POINTER(c.c_void_p),
                                   c.c_char_p, c.c_int]
    L.lcg_pair_prepare_device.argtypes = L.lcg_pair_prepare.argtypes
    L.lcg_pair_free.argtypes = [c.c_void_p]
    L.lcg_pair_info.argtypes = [c.c_void_p, c.POINTER(c.c_int32), c.POINTER(c.c_int32),
                                c.POINTER(c.c_int32), c.POINTER(c.c_int32), c.POINTER(c.c_int64)]
    L.lcg_pair_num_arcs.argtypes = [c.c_void_p, c.POINTER(c.c_int32), c.POINTER(c.c_int32)]
    L.lcg_pair_band.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p, c.c_int32]
    L.lcg_pair_arcmatches.argtypes = [c.c_void_p] + [c.c_void_p] * 5 + [c.c_int32]
    L.lcg_pair_score_tables.argtypes = [c.c_void_p] + [c.c_void_p] * 5
    L.lcg_pair_schedule.argtypes = [c.c_void_p, c.c_int32, c.c_int32, c.c_int32, c.c_void_p, c.c_void_p, c.c_int32]
    L.lcg_batch_create.argtypes = [c.POINTER(c.c_void_p), c.c_int32, c.c_int32, c.POINTER(c.c_void_p),
                                   c.c_char_p, c.c_int]
    L.lcg_batch_destroy.argtypes = [c.c_void_p]
    L.lcg_batch_run.argtypes = [c.c_void_p, c.c_int32, c.c_int32, c.POINTER(c.c_float)]
    L.lcg_batch_set_spin_limit.argtypes = [c.c_void_p, c.c_int32]
    L.lcg_batch_run_noalloc.argtypes = [c.c_void_p, c.c_int32]
    L.lcg_batch_set_stream.argtypes = [c.c_void_p, c.c_void_p]
    L.lcg_batch_config.argtypes = [c.c_void_p, c.POINTER(c.c_int32), c.POINTER(c.c_int32),
                                   c.POINTER(c.c_int32)]
    L.lcg_batch_score.argtypes = [c.c_void_p, c.c_int32, c.POINTER(c.c_int64)]
    L.lcg_batch_alignment.argtypes = [c.c_void_p, c.c_int32, c.c_void_p, c.c_void_p, c.c_int32,
                                      c.c_char_p, c.c_char_p, c.c_int32]
    L.lcg_batch_D.argtypes = [c.c_void_p, c.c_int32, c.c_void_p, c.c_int32]
    L.lcg_batch_counters.argtypes = [c.c_void_p, c.c_int32, c.POINTER(c.c_int64), c.POINTER(c.c_int64)]
    L.lcg_batch_counters_split.argtypes = [c.c_void_p, c.c_int32] + [c.POINTER(c.c_int64)] * 4
    L.lcg_batch_kernel_times.argtypes = [c.c_void_p, c.POINTER(c.c_float), c.POINTER(c.c_float)]
    L.lcg_batch_pf_run.argtypes = [c.c_void_p, c.c_void_p, c.POINTER(c.c_float)]
    L.lcg_batch_pf_result.argtypes = [c.c_void_p, c.c_int32, c.POINTER(c.c_double), c.POINTER(c.c_double)]
    L.lcg_batch_am_probs.argtypes = [c.c_void_p, c.c_int32, c.c_void_p, c.c_int32]
    L.lcg_batch_pf_D.argtypes = [c.c_void_p, c.c_int32, c.c_void_p, c.c_void_p, c.c_int32]
    L.lcg_batch_bm_probs.argtypes = [c.c_void_p, c.c_int32, c.c_int32, c.c_void_p, c.c_int32]
    L.lcg_batch_pf_grad.argtypes = [c.c_void_p, c.c_int32, c.c_void_p, c.c_int32]
    L.lcg_batch_pf_sample.argtypes = [c.c_void_p, c.c_int32, c.c_uint64, c.c_int32, c.c_void_p, c.c_void_p,
                                      c.c_int32]
    L.lcg_batch_pf_config.argtypes = [c.c_void_p] + [c.POINTER(c.c_int32)] * 3
    L.lcg_batch_suboptimal.argtypes = [c.c_void_p, c.c_int32, c.c_int32, c.c_int64, c.c_void_p, c.c_int32]
    L.lcg_batch_suboptimal_alignment.argtypes = [c.c_void_p, c.c_int32, c.c_void_p, c.c_void_p, c.c_int32,
                                                 c.c_char_p, c.c_char_p, c.c_int32, c.c_void_p]
    L.lcg_batch_pf_set_fragments.argtypes = [c.c_void_p, c.c_int32, c.c_int32, c.c_void_p]
    L.lcg_batch_pf_fragment_probs.argtypes = [c.c_void_p, c.c_int32, c.c_void_p, c.c_int32]
    L.lcg_batch_pair_info.argtypes = [c.c_void_p, c.c_int32, c.POINTER(c.c_int32), c.POINTER(c.c_int32),
                                      c.POINTER(c.c_int64)]
    L.lcg_batch_segment.argtypes = [c.c_void_p, c.c_int32, c.c_void_p, c.c_int64]
    L.lcg_batch_compact_run.argtypes = [c.c_void_p, c.POINTER(c.c_float), c.POINTER(c.c_float)]
    L.lcg_batch_run_penalized.argtypes = [c.c_void_p, c.c_int64, c.POINTER(c.c_float)]
    L.lcg_batch_run_normalized.argtypes = [c.c_void_p, c.c_int64, c.c_int32, c.POINTER(c.c_int32),
                                           c.POINTER(c.c_float)]
    L.lcg_batch_write_clustal.argtypes = [c.c_void_p, c.c_int32, c.c_int32, c.c_int32, c.c_char_p, c.c_int64]
    L.lcg_batch_write_clustal.restype = c.c_int64
    L.lcg_batch_write_consensus_pp.argtypes = [c.c_void_p, c.c_int32, c.c_char_p, c.c_int64]
    L.lcg_batch_write_consensus_pp.restype = c.c_int64
    L.lcg_guide_tree_upgma.argtypes = [c.c_void_p, c.c_int32, c.c_void_p]
    L.lcg_progressive_run.argtypes = [c.POINTER(c.c_char_p), c.c_int32, c.c_void_p, c.POINTER(Params), c.c_int32,
                                      c.POINTER(c.c_void_p), c.c_char_p, c.c_int]
    L.lcg_progressive_pp.argtypes = [c.c_void_p, c.c_int32, c.c_char_p, c.c_int64]
    L.lcg_progressive_pp.restype = c.c_int64
    L.lcg_progressive_clustal.argtypes = [c.c_void_p, c.c_int32, c.c_int32, c.c_char_p, c.c_int64]
    L.lcg_progressive_clustal.restype = c.c_int64
    L.lcg_progressive_scores.argtypes = [c.c_void_p, c.c_void_p, c.c_int32]
    L.lcg_progressive_stats.argtypes = [c.c_void_p] + [c.c_void_p] * 5
    L.lcg_progressive_free.argtypes = [c.c_void_p]
    L.lcg_progressive_free.restype = None
    L.lcg_fold_run.argtypes = [c.POINTER(c.c_char_p), c.c_int32, c.POINTER(FoldParams), c.c_int32,
                               c.POINTER(c.c_void_p), c.c_char_p, c.c_int]
    L.lcg_fold_write_pp.argtypes = [c.c_void_p, c.c_int32, c.c_char_p, c.c_char_p, c.c_int64]
    L.lcg_fold_write_pp.restype = c.c_int64
    L.lcg_fold_probs.argtypes = [c.c_void_p, c.c_int32, c.c_void_p, c.c_void_p, c.c_int64]
    L.lcg_fold_log_z.argtypes = [c.c_void_p, c.c_int32, c.POINTER(c.c_double)]
    L.lcg_fold_stats.argtypes = [c.c_void_p, c.POINTER(c.c_float), c.POINTER(c.c_float)]
    L.lcg_fold_free.argtypes = [c.c_void_p]
    L.lcg_mlocarna_run.argtypes = [c.POINTER(c.c_char_p), c.POINTER(c.c_char_p), c.c_int32, c.POINTER(FoldParams),
                                   c.POINTER(Params), c.c_int32, c.POINTER(c.c_void_p), c.c_char_p, c.c_int]
    L.lcg_mlocarna_clustal.argtypes = [c.c_void_p, c.c_int32, c.c_char_p, c.c_int64]
    L.lcg_mlocarna_clustal.restype = c.c_int64
    L.lcg_mlocarna_pp.argtypes = [c.c_void_p, c.c_int32, c.c_char_p, c.c_int64]
    L.lcg_mlocarna_pp.restype = c.c_int64
    L.lcg_mlocarna_scores.argtypes = [c.c_void_p, c.c_void_p, c.c_int64]
    L.lcg_mlocarna_tree.argtypes = [c.c_void_p, c.c_void_p, c.c_int32]
    L.lcg_mlocarna_stats.argtypes = [c.c_void_p] + [c.POINTER(c.c_float)] * 4
    L.lcg_mlocarna_free.argtypes = [c.c_void_p]
    L.lcg_mlocarna_free.restype = None
    L.lcg_fold_free.restype = None
    L.lcg_batch_write_am_probs.argtypes = [c.c_void_p, c.c_int32, c.c_char_p, c.c_int64]
    L.lcg_batch_write_am_probs.restype = c.c_int64
    L.lcg_batch_write_bm_probs.argtypes = [c.c_void_p, c.c_int32, c.c_int32, c.c_char_p, c.c_int64]
    L.lcg_batch_write_bm_probs.restype = c.c_int64
    L.lcg_pair_prepare_scored.argtypes = [c.c_char_p, c.c_char_p, c.POINTER(Params), c.c_char_p, c.c_int32,
                                          c.POINTER(c.c_void_p), c.c_char_p, c.c_int]
    L.lcg_pair_write_arcmatch_scores.argtypes = [c.c_void_p, c.c_char_p, c.c_int64]
    L.lcg_pair_write_arcmatch_scores.restype = c.c_int64
    L.lcg_pairs_prepare.argtypes = [c.POINTER(c.c_char_p), c.POINTER(c.c_char_p), c.c_int32, c.POINTER(Params),
                                    c.POINTER(PrepareOpts), c.POINTER(c.c_void_p), c.POINTER(PrepareStats),
                                    c.c_char_p, c.c_int]
    L.lcg_batch_create_shaped.argtypes = [c.POINTER(c.c_void_p), c.c_int32, c.c_int32, c.POINTER(Shape),
                                          c.POINTER(c.c_void_p), c.c_char_p, c.c_int]
    L.lcg_batch_get_shape.argtypes = [c.c_void_p, c.POINTER(Shape)]
    L.lcg_batch_reshape.argtypes = [c.c_void_p, c.POINTER(Shape), c.c_char_p, c.c_int]
    L.lcg_batch_autotune.argtypes = [c.c_void_p, c.c_int32, c.c_int32, c.POINTER(Shape), c.POINTER(c.c_float),
                                     c.POINTER(c.c_float), c.c_char_p, c.c_int]
    return L


NTHETA = 16 + 256 + 3  # sigma 4x4, arc-match term 16x16, indel, indel opening, struct weight


_L = None


def lib():
    global _L
    if _L is None:
        _L = _sig(_lib.load())
    return _L


class Pair:
    """A prepared pair (lcg_pair_prepare).  With device_compaction=True the
    arc matches are not enumerated on the host (lcg_pair_prepare_device):
    lcg_batch_create compacts them on the device."""

    def __init__(self, ppA, ppB, params=None, device_compaction=False, _handle=None):
        L = lib()
        self.params = params or locarna_params()
        self.device_compaction = device_compaction
        if _handle is None:
            h = ctypes.c_void_p()
            err = ctypes.create_string_buffer(512)
            fn = L.lcg_pair_prepare_device if device_compaction else L.lcg_pair_prepare
            rc = fn(ppA.encode(), ppB.encode(), ctypes.byref(self.params), ctypes.byref(h), err, 512)
            if rc != 0:
                raise LcgError("lcg_pair_prepare (%d): %s" % (rc, err.value.decode()))
        else:
            h = _handle
        self.h = h
        a, b, na, ng = (ctypes.c_int32() for _ in range(4))
        cells = ctypes.c_int64()
        L.lcg_pair_info(h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(na), ctypes.byref(ng),
                        ctypes.byref(cells))
        self.lenA, self.lenB, self.num_arcmatches, self.num_groups = a.value, b.value, na.value, ng.value
        self.cells = cells.value

    @classmethod
    def from_arcmatch_scores(cls, ppA, ppB, scores_text, params=None, probability_scale=-1):
        """Pair whose arc matches and scores come from an arc-match scores
        text (lcg_pair_prepare_scored)."""
        params = params or locarna_params()
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = lib().lcg_pair_prepare_scored(ppA.encode(), ppB.encode(), ctypes.byref(params), scores_text.encode(),
                                           probability_scale, ctypes.byref(h), err, 512)
        if rc != 0:
            raise LcgError("lcg_pair_prepare_scored (%d): %s" % (rc, err.value.decode()))
        return cls(None, None, params, False, _handle=h)

    def arcmatch_scores_text(self):
        """Arc-match scores text (lcg_pair_write_arcmatch_scores)."""
        n = lib().lcg_pair_write_arcmatch_scores(self.h, None, 0)
        if n < 0:
            raise LcgError("lcg_pair_write_arcmatch_scores (%d)" % n)
        buf = ctypes.create_string_buffer(n + 1)
        lib().lcg_pair_write_arcmatch_scores(self.h, buf, n + 1)
        return buf.value.decode()

    @property
    def num_arcs(self):
        """arcs of A plus arcs of B (BasePairs::num_bps)"""
        a, b = ctypes.c_int32(), ctypes.c_int32()
        lib().lcg_pair_num_arcs(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value + b.value

    @property
    def compaction_words(self):
        """64-bit words of the arc-pair validity bitmap: N_A * ceil(N_B / 64)"""
        a, b = ctypes.c_int32(), ctypes.c_int32()
        lib().lcg_pair_num_arcs(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value * ((b.value + 63) // 64)

    @property
    def band_cells(self):
        mn, mx = self.band()
        return sum(max(0, y - x + 1) for x, y in zip(mn, mx))

    def band(self):
        n = self.lenA + 1
        mn, mx = (ctypes.c_int32 * n)(), (ctypes.c_int32 * n)()
        lib().lcg_pair_band(self.h, mn, mx, n)
        return list(mn), list(mx)

    def score_tables(self):
        """(sigma dict (i, j) -> score, gapA list, gapB list, rowsA, rowsB)."""
        n, m = self.lenA, self.lenB
        sig = (ctypes.c_int64 * ((n + 1) * (m + 1)))()
        gA, gB = (ctypes.c_int64 * (n + 1))(), (ctypes.c_int64 * (m + 1))()
        rA, rB = ctypes.c_int32(), ctypes.c_int32()
        rc = lib().lcg_pair_score_tables(self.h, sig, gA, gB, ctypes.byref(rA), ctypes.byref(rB))
        if rc != 0:
            raise LcgError("lcg_pair_score_tables (%d)" % rc)
        return ({(i, j): sig[i * (m + 1) + j] for i in range(1, n + 1) for j in range(1, m + 1)},
                list(gA), list(gB), rA.value, rB.value)

    def arcmatches(self):
        n = max(self.num_arcmatches, 1)
        arrs = [(ctypes.c_int32 * n)() for _ in range(4)]
        sc = (ctypes.c_int64 * n)()
        lib().lcg_pair_arcmatches(self.h, *arrs, sc, n)
        k = self.num_arcmatches
        return [(arrs[0][t], arrs[1][t], arrs[2][t], arrs[3][t]) for t in range(k)], list(sc)[:k]

    def schedule(self, waves, pool_ints, members=1):
        """LDS schedule of the fill of D: list of (round, wave, member, lds
        offset, lds ints, al, bl) per group, and the rounds per row."""
        cap = max(4 * self.num_groups, self.lenA + 2, 1)
        words, rounds = (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)()
        n = lib().lcg_pair_schedule(self.h, waves, pool_ints, members, words, rounds, cap)
        if n < 0:
            raise LcgError("lcg_pair_schedule (%d)" % n)
        out = []
        for g in range(n):
            w, off, size, lr = words[4 * g], words[4 * g + 1], words[4 * g + 2], words[4 * g + 3]
            out.append((w & 0xfff, (w >> 12) & 0xf, w >> 16, off, size, lr & 0xffff, lr >> 16))
        return out, list(rounds)[:self.lenA + 2]

    def free(self):
        if self.h:
            lib().lcg_pair_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def prepare_pairs(pp_pairs, params=None, device_compaction=False, device_band=False, threads=0, device=0,
                  band_eps=0.0):
    """Prepares many pairs with the same parameters (lcg_pairs_prepare):
    host threads, and with device_band=True the trace bands of all pairs in
    one kernel launch (flagged pairs rechecked on the host).  Returns the
    list of Pair and a dict of the preparation statistics."""
    L = lib()
    params = params or locarna_params()
    n = len(pp_pairs)
    # one buffer per distinct text, so that the library parses it once
    enc = {}
    A = (ctypes.c_char_p * n)(*[enc.setdefault(a, a.encode()) for a, _ in pp_pairs])
    B = (ctypes.c_char_p * n)(*[enc.setdefault(b, b.encode()) for _, b in pp_pairs])
    out = (ctypes.c_void_p * max(n, 1))()
    flags = (1 if device_compaction else 0) | (2 if device_band else 0)
    opts = PrepareOpts(flags=flags, device=device, threads=threads, reserved=0, band_eps=band_eps)
    st = PrepareStats()
    err = ctypes.create_string_buffer(512)
    rc = L.lcg_pairs_prepare(A, B, n, ctypes.byref(params), ctypes.byref(opts), out, ctypes.byref(st), err, 512)
    if rc != 0:
        raise LcgError("lcg_pairs_prepare (%d): %s" % (rc, err.value.decode()))
    pairs = [Pair(None, None, params, device_compaction, _handle=ctypes.c_void_p(out[k])) for k in range(n)]
    return pairs, {"band_ms": st.band_ms, "device_banded": st.device_banded, "host_banded": st.host_banded,
                   "parse_ms": st.parse_ms, "band_wall_ms": st.band_wall_ms, "finish_ms": st.finish_ms}


class Batch:
    """Device batch (lcg_batch_create / lcg_batch_run)."""

    def __init__(self, pairs, device=0, shape=None):
        """shape: dict of lcg_shape fields (missing or 0: default)."""
        L = lib()
        self.pairs = pairs
        arr = (ctypes.c_void_p * len(pairs))(*[p.h.value for p in pairs])
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        if shape is None:
            rc = L.lcg_batch_create(arr, len(pairs), device, ctypes.byref(h), err, 512)
        else:
            rc = L.lcg_batch_create_shaped(arr, len(pairs), device, ctypes.byref(Shape(**shape)), ctypes.byref(h),
                                           err, 512)
        if rc != 0:
            raise LcgError("lcg_batch_create (%d): %s" % (rc, err.value.decode()))
        self.h = h
        self._config()

    def _config(self):
        w, g, lds = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        lib().lcg_batch_config(self.h, ctypes.byref(w), ctypes.byref(g), ctypes.byref(lds))
        self.waves, self.workgroups, self.lds_bytes = w.value, g.value, lds.value

    def shape(self):
        """Effective launch shape (dict)."""
        s = Shape()
        rc = lib().lcg_batch_get_shape(self.h, ctypes.byref(s))
        if rc != 0:
            raise LcgError("lcg_batch_get_shape (%d)" % rc)
        return s.as_dict()

    def reshape(self, **shape):
        err = ctypes.create_string_buffer(512)
        rc = lib().lcg_batch_reshape(self.h, ctypes.byref(Shape(**shape)), err, 512)
        if rc != 0:
            raise LcgError("lcg_batch_reshape (%d): %s" % (rc, err.value.decode()))
        self._config()

    def autotune(self, path=0, reps=3):
        """Times candidate launch shapes (path 0: integer, 1: LocARNA-P) and
        keeps the fastest; returns (shape, best_ms, default_ms)."""
        s, b, d = Shape(), ctypes.c_float(), ctypes.c_float()
        err = ctypes.create_string_buffer(512)
        rc = lib().lcg_batch_autotune(self.h, path, reps, ctypes.byref(s), ctypes.byref(b), ctypes.byref(d), err, 512)
        if rc != 0:
            raise LcgError("lcg_batch_autotune (%d): %s" % (rc, err.value.decode()))
        self._config()
        return s.as_dict(), b.value, d.value

    def pair_info(self, k):
        """(arc matches, groups, image ints) of pair k."""
        nam, ng, sz = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        rc = lib().lcg_batch_pair_info(self.h, k, ctypes.byref(nam), ctypes.byref(ng), ctypes.byref(sz))
        if rc != 0:
            raise LcgError("lcg_batch_pair_info (%d)" % rc)
        return nam.value, ng.value, sz.value

    def segment(self, k):
        """Device image of pair k (numpy int32 array)."""
        import numpy as np
        n = self.pair_info(k)[2]
        buf = np.zeros(max(n, 1), dtype=np.int32)
        rc = lib().lcg_batch_segment(self.h, k, buf.ctypes.data_as(ctypes.c_void_p), n)
        if rc < 0:
            raise LcgError("lcg_batch_segment (%d)" % rc)
        return buf[:rc]

    def compact_run(self):
        """Reruns the device compaction; returns (count_ms, build_ms)."""
        a, b = ctypes.c_float(), ctypes.c_float()
        rc = lib().lcg_batch_compact_run(self.h, ctypes.byref(a), ctypes.byref(b))
        if rc != 0:
            raise LcgError("lcg_batch_compact_run failed (%d)" % rc)
        return a.value, b.value

    def set_spin_limit(self, polls):
        """Bound of the cross-workgroup waits (lcg_batch_set_spin_limit):
        0 default, -1 every wait gives up at once (error-path tests)."""
        rc = lib().lcg_batch_set_spin_limit(self.h, polls)
        if rc != 0:
            raise LcgError("lcg_batch_set_spin_limit (%d)" % rc)

    def run_noalloc(self, trace=True):
        """lcg_batch_run_noalloc: enqueues the run with the library's static
        scratch; the getters synchronize."""
        rc = lib().lcg_batch_run_noalloc(self.h, 1 if trace else 0)
        if rc != 0:
            raise LcgError("lcg_batch_run_noalloc (%d)" % rc)

    def set_stream(self, stream):
        """Runs on a caller's HIP stream (an integer hipStream_t handle of
        the process's HIP runtime, e.g. torch.cuda.Stream().cuda_stream when
        torch was imported before the library; 0 / None: the batch's own)."""
        rc = lib().lcg_batch_set_stream(self.h, ctypes.c_void_p(stream or None))
        if rc != 0:
            raise LcgError("lcg_batch_set_stream (%d)" % rc)

    def run(self, mode=0, trace=True, count=False):
        ms = ctypes.c_float()
        flags = (1 if trace else 0) | (2 if count else 0)
        rc = lib().lcg_batch_run(self.h, mode, flags, ctypes.byref(ms))
        if rc != 0:
            raise LcgError("lcg_batch_run failed (%d)" % rc)
        return ms.value

    def run_penalized(self, position_penalty):
        """locarna --penalized (lcg_batch_run_penalized); device ms."""
        ms = ctypes.c_float()
        rc = lib().lcg_batch_run_penalized(self.h, position_penalty, ctypes.byref(ms))
        if rc != 0:
            raise LcgError("lcg_batch_run_penalized (%d)" % rc)
        return ms.value

    def run_normalized(self, L, max_iterations=100):
        """locarna --normalized L (lcg_batch_run_normalized); returns the
        Dinkelbach iterations and the device ms."""
        it, ms = ctypes.c_int32(), ctypes.c_float()
        rc = lib().lcg_batch_run_normalized(self.h, L, max_iterations, ctypes.byref(it), ctypes.byref(ms))
        if rc != 0:
            raise LcgError("lcg_batch_run_normalized (%d)" % rc)
        return it.value, ms.value

    def suboptimal(self, k, kbest, threshold=-1000000):
        """k-best suboptimal alignments of pair k (locarna --kbest kbest
        --better threshold): list of (score, edges, (startA, endA, startB,
        endB), structA, structB)."""
        cap = 4096
        scores = (ctypes.c_int64 * cap)()
        n = lib().lcg_batch_suboptimal(self.h, k, kbest, threshold, scores, cap)
        if n < 0:
            raise LcgError("lcg_batch_suboptimal (%d)" % n)
        p = self.pairs[k]
        ec = p.lenA + p.lenB + 2
        out = []
        for t in range(n):
            a, b = (ctypes.c_int32 * ec)(), (ctypes.c_int32 * ec)()
            sa = ctypes.create_string_buffer(max(p.lenA, p.lenB) + 2)
            sb = ctypes.create_string_buffer(max(p.lenA, p.lenB) + 2)
            r = (ctypes.c_int32 * 4)()
            ne = lib().lcg_batch_suboptimal_alignment(self.h, t, a, b, ec, sa, sb, len(sa), r)
            if ne < 0:
                raise LcgError("lcg_batch_suboptimal_alignment (%d)" % ne)
            out.append((scores[t] if t < cap else None, [(a[e], b[e]) for e in range(ne)], tuple(r),
                        sa.value.decode(), sb.value.decode()))
        return out

    def score(self, k):
        s = ctypes.c_int64()
        rc = lib().lcg_batch_score(self.h, k, ctypes.byref(s))
        if rc != 0:
            raise LcgError("lcg_batch_score (%d)" % rc)
        return s.value

    def alignment(self, k):
        p = self.pairs[k]
        cap = p.lenA + p.lenB + 2
        a, b = (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)()
        sa = ctypes.create_string_buffer(cap + 2)
        sb = ctypes.create_string_buffer(cap + 2)
        n = lib().lcg_batch_alignment(self.h, k, a, b, cap, sa, sb, cap + 2)
        if n < 0:
            raise LcgError("lcg_batch_alignment (%d)" % n)
        return [(a[t], b[t]) for t in range(n)], sa.value.decode(), sb.value.decode()

    def counters(self, k):
        c, q = ctypes.c_int64(), ctypes.c_int64()
        rc = lib().lcg_batch_counters(self.h, k, ctypes.byref(c), ctypes.byref(q))
        if rc != 0:
            raise LcgError("lcg_batch_counters (%d)" % rc)
        return c.value, q.value

    def counters_split(self, k):
        """(fill cells, fill candidates, top cells, top candidates)"""
        v = [ctypes.c_int64() for _ in range(4)]
        rc = lib().lcg_batch_counters_split(self.h, k, *[ctypes.byref(x) for x in v])
        if rc != 0:
            raise LcgError("lcg_batch_counters_split (%d)" % rc)
        return tuple(x.value for x in v)

    def kernel_times(self):
        """(fill-D ms, top-level/traceback ms) of the last mode-0 run"""
        f, t = ctypes.c_float(), ctypes.c_float()
        rc = lib().lcg_batch_kernel_times(self.h, ctypes.byref(f), ctypes.byref(t))
        if rc != 0:
            raise LcgError("lcg_batch_kernel_times (%d)" % rc)
        return f.value, t.value

    def D(self, k):
        n = max(self.pair_info(k)[0], 1)
        out = (ctypes.c_int64 * n)()
        cnt = lib().lcg_batch_D(self.h, k, out, n)
        if cnt < 0:
            raise LcgError("lcg_batch_D (%d)" % cnt)
        return list(out)[:cnt]

    # ---- LocARNA-P
    def pf_sample(self, k, seed, nsamples):
        """(mean, variance) of the feature vectors of `nsamples` alignments of
        pair k drawn from the Boltzmann distribution of the last pf_run."""
        mean, var = (ctypes.c_double * NTHETA)(), (ctypes.c_double * NTHETA)()
        rc = lib().lcg_batch_pf_sample(self.h, k, seed, nsamples, mean, var, NTHETA)
        if rc < 0:
            raise LcgError("lcg_batch_pf_sample (%d)" % rc)
        return list(mean), list(var)

    def _text(self, fn, *args):
        n = fn(self.h, *args, None, 0)
        if n < 0:
            raise LcgError("probability writer (%d)" % n)
        buf = ctypes.create_string_buffer(n + 1)
        fn(self.h, *args, buf, n + 1)
        return buf.value.decode()

    def clustal(self, k, width=120, local_only=False, structures=False):
        """locarna --clustal text of pair k (lcg_batch_write_clustal)."""
        return self._text(lib().lcg_batch_write_clustal, k, width, (1 if local_only else 0) | (2 if structures else 0))

    def consensus_pp(self, k):
        """locarna --pp text of pair k (lcg_batch_write_consensus_pp)."""
        return self._text(lib().lcg_batch_write_consensus_pp, k)

    def am_probs_text(self, k):
        """locarna_p --write-arcmatch-probs text of pair k."""
        return self._text(lib().lcg_batch_write_am_probs, k)

    def bm_probs_text(self, k, include_am_in_bm=False):
        """locarna_p --write-basematch-probs text of pair k."""
        return self._text(lib().lcg_batch_write_bm_probs, k, 1 if include_am_in_bm else 0)

    def pf_config(self):
        """(workgroups per split pair, split pairs, workgroups) of LocARNA-P"""
        a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        rc = lib().lcg_batch_pf_config(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        if rc != 0:
            raise LcgError("lcg_batch_pf_config (%d)" % rc)
        return a.value, b.value, c.value

    def pf_run(self, dtheta=None):
        """Inside + adjoint for all pairs; dtheta: offsets of the NTHETA score
        parameters from the reference scores (None: reference)."""
        ms = ctypes.c_float()
        arr = None
        if dtheta is not None:
            arr = (ctypes.c_double * NTHETA)(*[float(x) for x in dtheta])
        rc = lib().lcg_batch_pf_run(self.h, arr, ctypes.byref(ms))
        if rc != 0:
            raise LcgError("lcg_batch_pf_run failed (%d)" % rc)
        return ms.value

    def pf_result(self, k):
        z, op = ctypes.c_double(), ctypes.c_double()
        rc = lib().lcg_batch_pf_result(self.h, k, ctypes.byref(z), ctypes.byref(op))
        if rc != 0:
            raise LcgError("lcg_batch_pf_result (%d)" % rc)
        return z.value, op.value

    def am_probs(self, k):
        n = max(self.pairs[k].num_arcmatches, 1)
        out = (ctypes.c_double * n)()
        cnt = lib().lcg_batch_am_probs(self.h, k, out, n)
        if cnt < 0:
            raise LcgError("lcg_batch_am_probs (%d)" % cnt)
        return list(out)[:cnt]

    def pf_D(self, k):
        """(inside D, outside D') per arc match, reference arc-match order"""
        n = max(self.pairs[k].num_arcmatches, 1)
        D, Dp = (ctypes.c_double * n)(), (ctypes.c_double * n)()
        cnt = lib().lcg_batch_pf_D(self.h, k, D, Dp, n)
        if cnt < 0:
            raise LcgError("lcg_batch_pf_D (%d)" % cnt)
        return list(D)[:cnt], list(Dp)[:cnt]

    def bm_probs(self, k, mode=0):
        """dense (lenA+1)*(lenB+1): mode 0 reference base-match probabilities,
        1 with arc-match ends (--include-am-in-bm), 2 all match edges"""
        p = self.pairs[k]
        n = (p.lenA + 1) * (p.lenB + 1)
        out = (ctypes.c_double * n)()
        cnt = lib().lcg_batch_bm_probs(self.h, k, mode, out, n)
        if cnt < 0:
            raise LcgError("lcg_batch_bm_probs (%d)" % cnt)
        return list(out)

    def pf_set_fragments(self, k, frags):
        """Fragments (i, j, k, l) of pair k, A[i..j] matched with B[k..l],
        whose probabilities the next pf_run computes (locarna_p
        --fragment-match-probs)."""
        arr = (ctypes.c_int32 * max(1, 4 * len(frags)))(*[int(x) for f in frags for x in f])
        rc = lib().lcg_batch_pf_set_fragments(self.h, k, len(frags), arr)
        if rc != 0:
            raise LcgError("lcg_batch_pf_set_fragments (%d)" % rc)

    def pf_fragment_probs(self, k):
        """Fragment match probabilities of pair k from the last pf_run."""
        cap = 4096
        out = (ctypes.c_double * cap)()
        cnt = lib().lcg_batch_pf_fragment_probs(self.h, k, out, cap)
        if cnt < 0:
            raise LcgError("lcg_batch_pf_fragment_probs (%d)" % cnt)
        return list(out)[:cnt]

    def pf_grad(self, k):
        out = (ctypes.c_double * NTHETA)()
        rc = lib().lcg_batch_pf_grad(self.h, k, out, NTHETA)
        if rc < 0:
            raise LcgError("lcg_batch_pf_grad (%d)" % rc)
        return list(out)

    def destroy(self):
        if self.h:
            lib().lcg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def guide_tree_upgma(scores):
    """mlocarna's UPGMA guide tree of a symmetric score matrix (list of
    lists): list of merges (child0, child1), leaves 0..n-1, merge t = node
    n + t (lcg_guide_tree_upgma)."""
    n = len(scores)
    flat = (ctypes.c_int64 * (n * n))(*[int(scores[i][j]) if i != j else 0 for i in range(n) for j in range(n)])
    m = (ctypes.c_int32 * max(2 * (n - 1), 1))()
    rc = lib().lcg_guide_tree_upgma(flat, n, m)
    if rc != 0:
        raise LcgError("lcg_guide_tree_upgma (%d)" % rc)
    return [(m[2 * t], m[2 * t + 1]) for t in range(n - 1)]


class Progressive:
    """mlocarna's progressive alignment along a guide tree on one GPU
    (lcg_progressive_run): pp(node), clustal(t), scores, stats."""

    def __init__(self, pps, merges, params=None, device=0):
        params = params or locarna_params()
        n = len(pps)
        arr = (ctypes.c_char_p * n)(*[p.encode() for p in pps])
        flat = (ctypes.c_int32 * max(2 * len(merges), 1))(*[x for m in merges for x in m])
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        self.n = n
        rc = lib().lcg_progressive_run(arr, n, flat, ctypes.byref(params), device, ctypes.byref(h), err, 1024)
        if rc != 0:
            raise LcgError("lcg_progressive_run (%d): %s" % (rc, err.value.decode()))
        self.h = h

    def _text(self, fn, *args):
        need = fn(self.h, *args, None, 0)
        if need < 0:
            raise LcgError("%s (%d)" % (fn.__name__, need))
        buf = ctypes.create_string_buffer(need + 1)
        fn(self.h, *args, buf, need + 1)
        return buf.value.decode()

    def pp(self, node):
        return self._text(lib().lcg_progressive_pp, node)

    def clustal(self, t=None, width=120):
        return self._text(lib().lcg_progressive_clustal, self.n - 2 if t is None else t, width)

    def scores(self):
        out = (ctypes.c_int64 * max(self.n - 1, 1))()
        lib().lcg_progressive_scores(self.h, out, max(self.n - 1, 1))
        return list(out)[:self.n - 1]

    def stats(self):
        r, b = ctypes.c_int32(), ctypes.c_int32()
        d, p, w = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        lib().lcg_progressive_stats(self.h, ctypes.byref(r), ctypes.byref(b), ctypes.byref(d), ctypes.byref(p),
                                    ctypes.byref(w))
        return dict(rounds=r.value, batches=b.value, device_ms=d.value, prepare_ms=p.value, wall_ms=w.value)

    def destroy(self):
        if self.h:
            lib().lcg_progressive_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Fold:
    """Base-pair probabilities of single sequences on one GPU
    (lcg_fold_run; the step locarna_rnafold_pp performs for mlocarna,
    src/locarna_rnafold_pp.cc:209-271): pp(k, name), probs(k), log_z(k),
    stats()."""

    def __init__(self, seqs, max_bp_span=-1, min_prob=0.001, stacking=False, device=0, scale_kcal_per_nt=0.0):
        n = len(seqs)
        arr = (ctypes.c_char_p * n)(*[s.encode() for s in seqs])
        fp = FoldParams(max_bp_span, 1 if stacking else 0, min_prob, scale_kcal_per_nt)
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        self.h = None
        self.lens = [len(s) for s in seqs]
        rc = lib().lcg_fold_run(arr, n, ctypes.byref(fp), device, ctypes.byref(h), err, 1024)
        if rc != 0:
            raise LcgError("lcg_fold_run (%d): %s" % (rc, err.value.decode()))
        self.h = h

    def pp(self, k, name):
        need = lib().lcg_fold_write_pp(self.h, k, name.encode(), None, 0)
        if need < 0:
            raise LcgError("lcg_fold_write_pp (%d)" % need)
        buf = ctypes.create_string_buffer(need + 1)
        lib().lcg_fold_write_pp(self.h, k, name.encode(), buf, need + 1)
        return buf.value.decode()

    def probs(self, k):
        """(P, P2) as dense (n+1) x (n+1) numpy arrays."""
        import numpy as np
        n = self.lens[k]
        P = np.zeros((n + 1, n + 1))
        P2 = np.zeros((n + 1, n + 1))
        rc = lib().lcg_fold_probs(self.h, k, P.ctypes.data, P2.ctypes.data, (n + 1) * (n + 1))
        if rc != n:
            raise LcgError("lcg_fold_probs (%d)" % rc)
        return P, P2

    def log_z(self, k):
        v = ctypes.c_double()
        rc = lib().lcg_fold_log_z(self.h, k, ctypes.byref(v))
        if rc != 0:
            raise LcgError("lcg_fold_log_z (%d)" % rc)
        return v.value

    def stats(self):
        d, w = ctypes.c_float(), ctypes.c_float()
        lib().lcg_fold_stats(self.h, ctypes.byref(d), ctypes.byref(w))
        return dict(device_ms=d.value, wall_ms=w.value)

    def destroy(self):
        if self.h:
            lib().lcg_fold_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Mlocarna:
    """mlocarna's default pipeline from sequences on one GPU
    (lcg_mlocarna_run): base-pair probabilities, all pairs, UPGMA guide
    tree, progressive alignment.  clustal(), pp(node), scores(), tree(),
    stats()."""

    def __init__(self, names, seqs, params=None, max_bp_span=-1, min_prob=0.0005, device=0):
        params = params or locarna_params(max_bp_span=max_bp_span, min_prob=min_prob)
        n = len(seqs)
        self.n = n
        nm = (ctypes.c_char_p * n)(*[x.encode() for x in names])
        sq = (ctypes.c_char_p * n)(*[x.encode() for x in seqs])
        fp = FoldParams(max_bp_span, 0, min_prob, 0.0)
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        self.h = None
        rc = lib().lcg_mlocarna_run(nm, sq, n, ctypes.byref(fp), ctypes.byref(params), device, ctypes.byref(h),
                                    err, 1024)
        if rc != 0:
            raise LcgError("lcg_mlocarna_run (%d): %s" % (rc, err.value.decode()))
        self.h = h

    def _text(self, fn, *args):
        need = fn(self.h, *args, None, 0)
        if need < 0:
            raise LcgError("%s (%d)" % (fn.__name__, need))
        buf = ctypes.create_string_buffer(need + 1)
        fn(self.h, *args, buf, need + 1)
        return buf.value.decode()

    def clustal(self, width=120):
        return self._text(lib().lcg_mlocarna_clustal, width)

    def pp(self, node):
        return self._text(lib().lcg_mlocarna_pp, node)

    def scores(self):
        out = (ctypes.c_int64 * (self.n * self.n))()
        lib().lcg_mlocarna_scores(self.h, out, self.n * self.n)
        return [list(out[a * self.n:(a + 1) * self.n]) for a in range(self.n)]

    def tree(self):
        out = (ctypes.c_int32 * (2 * (self.n - 1)))()
        lib().lcg_mlocarna_tree(self.h, out, 2 * (self.n - 1))
        return [(out[2 * t], out[2 * t + 1]) for t in range(self.n - 1)]

    def stats(self):
        v = [ctypes.c_float() for _ in range(4)]
        lib().lcg_mlocarna_stats(self.h, *[ctypes.byref(x) for x in v])
        return dict(zip(("fold_ms", "pairs_ms", "progressive_ms", "wall_ms"), [x.value for x in v]))

    def destroy(self):
        if self.h:
            lib().lcg_mlocarna_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
