/* C ABI of the MI355X-native LocARNA pairwise alignment path.
 *
 * Plain C types only; every call returns an int status (0 = ok, negative =
 * error class, see LCG_E_*).  A caller that replaces LocARNA's in-process
 * alignment (src/locarna.cc:727-785: Aligner(AlignerParams).align() /
 * trace() / get_alignment()) or mlocarna's per-pair subprocess loop
 * (src/Utils/mlocarna:3547-3640 -> :2991-3022) binds these entry points.
 *
 * Flow: lcg_pair_prepare() reads two #PP 2.0 texts and builds, on the host,
 * what the reference builds before Aligner::align(): RnaData
 * (rna_data.cc:984-1124), the trace band (TraceController +
 * restrict_trace_by_probabilities, main_helper.icc:408-426), BasePairs /
 * ArcMatches (basepairs.cc:156-220, arc_matches.cc:130-188) and Scoring
 * (scoring.cc:28-554).  lcg_batch_create() uploads any number of prepared
 * pairs into one device arena; lcg_batch_run() runs the gfx950 kernels
 * (fill of the arc-match table D, top level, traceback) for all pairs;
 * results are then read per pair.
 */
#ifndef LOCARNA_AMD_H
#define LOCARNA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LCG_OK 0
#define LCG_E_ARG (-1)      /* invalid argument                              */
#define LCG_E_INPUT (-2)    /* malformed input (LocARNA::failure equivalent)  */
#define LCG_E_HIP (-3)      /* HIP runtime error                             */
#define LCG_E_CAPACITY (-4) /* problem exceeds the kernel's capacity          */
#define LCG_E_STATE (-5)    /* result requested before it was computed        */

/* Parameters of a run; field meaning and defaults follow the locarna /
 * locarna_p command lines (src/locarna.cc:83-272, src/locarna_p.cc:95-200). */
typedef struct {
    int32_t indel;             /* --indel                                   */
    int32_t indel_opening;     /* --indel-opening                           */
    int32_t struct_weight;     /* --struct-weight                           */
    int32_t tau;               /* --tau                                     */
    int32_t match, mismatch;   /* --match / --mismatch (non-RIBOSUM symbols) */
    int32_t temperature_alipf; /* --temperature-alipf                       */
    int32_t sequ_local;        /* --sequ-local                              */
    int32_t max_bp_span;       /* --maxBPspan (-1: unlimited)               */
    int32_t band_extended;     /* trace band PF in long double (locarna) or
                                  double (locarna_p default)                 */
    int32_t use_ribosum;       /* --use-ribosum (built-in RIBOSUM85_60)      */
    int32_t free_endgaps;      /* --free-endgaps "++--" as bits LCG_FREE_*;
                                  not with sequ_local (src/locarna.cc:375),
                                  integer path only                          */
    int32_t max_diff;          /* --max-diff: band around the diagonal
                                  (TraceController::constrain_wo_ref,
                                  trace_controller.cc:361-390); -1: none    */
    int32_t max_diff_am;       /* --max-diff-am (-1: the longer length)     */
    int32_t max_diff_at_am;    /* --max-diff-at-am (-1: the longer length)  */
    double min_prob;           /* --min-prob                                */
    double min_trace_prob;     /* --min-trace-probability (<= 0: no band)    */
    double exp_prob;           /* --exp-prob (<= 0: 1/(2 length))            */
    double pf_scale;           /* --pf-scale                                */
    double min_am_prob;        /* --min-am-prob                             */
    double min_bm_prob;        /* --min-bm-prob                             */
    int32_t no_lonely_pairs;   /* --noLP: arc matches only in stacks of two
                                  or more (aligner.cc:612-657, 1030-1078);
                                  integer path with host arc matches
                                  (lcg_pair_prepare, lcg_pairs_prepare without
                                  LCG_PREPARE_DEVICE_COMPACTION); LocARNA-P
                                  refuses it                                */
    int32_t stacking;          /* bit 0 --stacking, bit 1 --new-stacking:
                                  the #STACK joint probabilities of the pp
                                  input give an arc match whose inner arc
                                  match exists a stacked alternative of D
                                  (aligner.cc:574-610; weights
                                  scoring.cc:200-236); integer path with host
                                  arc matches and pp input (not arc-match
                                  score files); LocARNA-P refuses it        */
    int32_t struct_local;      /* --struct-local: inside every arc match one
                                  block of A and one of B positions may be
                                  excluded (aligner.cc:373-565, eight
                                  states; trace :1233-1343); integer path
                                  with host arc matches, single sequences,
                                  not with --noLP or stacking             */
    int32_t exclusion;         /* --exclusion: score of one excluded block
                                  (ScoringParams::exclusion, scoring.hh:107) */
} lcg_params;

/* free end gaps (FreeEndgaps, free_endgaps.hh: the characters of the
 * --free-endgaps string in order) */
#define LCG_FREE_LEFT_1 1  /* left end of the first sequence  */
#define LCG_FREE_RIGHT_1 2 /* right end of the first sequence */
#define LCG_FREE_LEFT_2 4  /* left end of the second sequence */
#define LCG_FREE_RIGHT_2 8 /* right end of the second sequence */

typedef struct lcg_pair lcg_pair;
typedef struct lcg_batch lcg_batch;

/* Host preparation of one pair (replaces the construction of RnaData,
 * TraceController, ArcMatches and Scoring in src/locarna.cc:455-703). */
int lcg_pair_prepare(const char *ppA, const char *ppB, const lcg_params *params, lcg_pair **out, char *err,
                     int errlen);
void lcg_pair_free(lcg_pair *pair);
/* lengths, number of arc matches, number of arc-match groups and an
 * estimate of the DP cells of the pair */
int lcg_pair_info(const lcg_pair *pair, int32_t *lenA, int32_t *lenB, int32_t *num_arcmatches, int32_t *num_groups,
                  int64_t *cells);
/* arcs of the two sequences after the probability cut
 * (BasePairs::num_bps, basepairs.hh) */
int lcg_pair_num_arcs(const lcg_pair *pair, int32_t *num_arcsA, int32_t *num_arcsB);
/* trace band: min_col[0..lenA], max_col[0..lenA]
 * (TraceController::min_col / max_col, trace_controller.hh) */
int lcg_pair_band(const lcg_pair *pair, int32_t *min_col, int32_t *max_col, int32_t cap);
/* score tables of a prepared pair: sigma[(i)(lenB+1) + j] = base-match
 * score of positions (columns) i, j (Scoring::basematch, scoring.cc:140-198:
 * the sum-of-pairs average for alignment profiles), gapA[0..lenA] and
 * gapB[0..lenB] the position-specific gap costs (Scoring::gapA / gapB,
 * scoring.cc:271-311); rowsA / rowsB the rows of each side (1 for a single
 * sequence).  Any output may be NULL. */
int lcg_pair_score_tables(const lcg_pair *pair, int64_t *sigma, int64_t *gapA, int64_t *gapB, int32_t *rowsA,
                          int32_t *rowsB);
/* arc matches in the reference's index order (ArcMatches::arcmatch(idx)),
 * with their score (Scoring::arcmatch, scoring.cc:440-554) */
int lcg_pair_arcmatches(const lcg_pair *pair, int32_t *al, int32_t *ar, int32_t *bl, int32_t *br, int64_t *score,
                        int32_t cap);

/* Host preparation of one pair whose arc matches and their scores come from
 * an arc-match scores text (locarna --read-arcmatch-scores /
 * --read-arcmatch-probs, src/locarna.cc:599-612;
 * ArcMatches::read_arcmatch_scores, arc_matches.cc:190-273): one line
 * "i j k l score" per arc match of arc (i, j) of A with arc (k, l) of B; with
 * probability_scale >= 0 the fifth field is a probability p and the score
 * (long)(p * probability_scale).  The arcs of each sequence are the arcs of
 * the text (in sorted order), the arc matches its lines in order with the
 * invalid ones (trace band) left out, and the arc-match score is the given
 * one (Scoring::arcmatch, scoring.cc:541-546).  The pp texts give the
 * sequences and the trace band.  Malformed lines and positions fail with
 * LCG_E_INPUT as in the reference, and so do arcs with j - i < 3, which the
 * reference's BasePairs never registers.  LocARNA-P runs of such pairs treat
 * the arc-match scores as constants in the gradient. */
int lcg_pair_prepare_scored(const char *ppA, const char *ppB, const lcg_params *params, const char *arcmatch_scores,
                            int32_t probability_scale, lcg_pair **out, char *err, int errlen);
/* Arc-match scores text of a pair (ArcMatches::write_arcmatch_scores,
 * arc_matches.cc:275-302, locarna --write-arcmatch-scores): "al ar bl br
 * score" per arc match in index order.  Returns the text length; writes
 * the text and its terminating 0 into buf when cap exceeds the length.
 * LCG_E_STATE for pairs of lcg_pair_prepare_device. */
int64_t lcg_pair_write_arcmatch_scores(const lcg_pair *pair, char *buf, int64_t cap);

/* Host preparation of one pair without the arc-match enumeration: RnaData,
 * trace band, BasePairs and the score tables as lcg_pair_prepare, while the
 * arc matches (ArcMatches constructor, arc_matches.cc:130-188: predicate
 * :19-48 over all arc pairs, index order, lists by common left and right
 * ends) are compacted on the device by lcg_batch_create, which also builds
 * every arc-match structure of the pair's device image there.  Such pairs
 * report -1 arc matches and groups from lcg_pair_info (see
 * lcg_batch_pair_info), and lcg_pair_arcmatches / lcg_pair_schedule return
 * LCG_E_STATE.  A batch takes either kind of pair, not both; LocARNA-P
 * (lcg_batch_pf_run) needs pairs from lcg_pair_prepare. */
int lcg_pair_prepare_device(const char *ppA, const char *ppB, const lcg_params *params, lcg_pair **out, char *err,
                            int errlen);

/* Preparation of many pairs with the same parameters (mlocarna's all-pairs
 * stage, src/Utils/mlocarna:3547-3640, which runs the per-pair preparation
 * of src/locarna.cc:455-703 once per pair): pp parsing, score tables and arc
 * matches on host threads, and with LCG_PREPARE_DEVICE_BAND the trace band
 * (PFGotoh + PFTraceProbs, edge_probs.icc:76-268, and the cut of
 * trace_controller.cc:565-597) of all pairs in one kernel launch.  The
 * kernel computes in double with a power-of-two scale per anti-diagonal; a
 * pair with a trace probability within band_eps (relative) of the cut, or a
 * value the host precision (long double for locarna, double for
 * locarna_p) could not represent or that the kernel did not keep exactly,
 * is banded on the host instead, so every band equals lcg_pair_prepare's.
 * Pairs longer than 1024 are banded on the host.  out[k] is the pair of
 * ppA[k], ppB[k], as lcg_pair_prepare (or lcg_pair_prepare_device with
 * LCG_PREPARE_DEVICE_COMPACTION) returns it. */
#define LCG_PREPARE_DEVICE_COMPACTION 1
#define LCG_PREPARE_DEVICE_BAND 2
typedef struct {
    int32_t flags;   /* LCG_PREPARE_* */
    int32_t device;  /* HIP device of the band kernel */
    int32_t threads; /* host threads; <= 0: the hardware's, at most 16 */
    int32_t reserved;
    double band_eps; /* <= 0: 1e-9 */
} lcg_prepare_opts;
typedef struct {
    float band_ms;         /* device time of the band kernel(s) */
    int32_t device_banded; /* pairs whose band the kernel computed */
    int32_t host_banded;   /* pairs banded on the host (too long, flagged, or no device band asked) */
    float parse_ms;        /* wall time: pp parsing and score inputs (host threads) */
    float band_wall_ms;    /* wall time: band stage (device: inputs, copies, kernel; host bands) */
    float finish_ms;       /* wall time: base pairs, arc matches or compaction inputs, images */
} lcg_prepare_stats;
int lcg_pairs_prepare(const char *const *ppA, const char *const *ppB, int32_t npairs, const lcg_params *params,
                      const lcg_prepare_opts *opts, lcg_pair **out, lcg_prepare_stats *stats, char *err,
                      int errlen);

/* Device batch of prepared pairs (device >= 0: HIP device ordinal).  The
 * pairs may be freed after the call. */
int lcg_batch_create(lcg_pair *const *pairs, int32_t npairs, int32_t device, lcg_batch **out, char *err,
                     int errlen);

/* Launch shape of a batch's kernels (no counterpart in the reference, which
 * runs on the CPU).  A field of 0 takes the library's default; the
 * environment variables LCG_WAVES, LCG_WGS_PER_CU, LCG_WG_PER_PAIR,
 * LCG_TOP_WAVES, LCG_TOP_WGS_PER_CU, LCG_PF_WAVES and LCG_PF_WG_PER_PAIR
 * override the defaults, and nonzero fields override both.  Results do not
 * depend on the shape (integer results bit for bit). */
typedef struct {
    int32_t waves;          /* fill of D: waves per workgroup, 1..16 (default 16) */
    int32_t wgs_per_cu;     /* fill: workgroups per CU, 1..4, with at least as many pairs as CUs (1) */
    int32_t wg_per_pair;    /* fill: workgroups per pair, 1..64, with fewer pairs than CUs (CUs / pairs) */
    int32_t top_waves;      /* top level + traceback: waves per workgroup, 1..16 (4) */
    int32_t top_wgs_per_cu; /* top level + traceback: workgroups per CU, 1..4 (4) */
    int32_t pf_waves;       /* LocARNA-P: waves per workgroup, 1..8 (8) */
    int32_t pf_wg_per_pair; /* LocARNA-P: workgroups per pair, 1..64; 0: two for the largest pairs where
                               that shortens the estimated makespan, else CUs / pairs */
    int32_t reserved;
} lcg_shape;
/* lcg_batch_create with a requested launch shape (NULL: defaults). */
int lcg_batch_create_shaped(lcg_pair *const *pairs, int32_t npairs, int32_t device, const lcg_shape *shape,
                            lcg_batch **out, char *err, int errlen);
/* Effective launch shape (shapes whose regions do not fit the LDS are
 * reduced to fewer workgroups per CU). */
int lcg_batch_get_shape(const lcg_batch *batch, lcg_shape *shape);
/* Applies another launch shape to a batch: recomputes the LDS schedule of
 * the fill when its shape changes (on the host, or by the device
 * compaction) and invalidates the results of earlier runs.  On an error the
 * batch keeps its previous shape. */
int lcg_batch_reshape(lcg_batch *batch, const lcg_shape *shape, char *err, int errlen);
/* Autotuner: times candidate launch shapes on this batch (minimum of `reps`
 * runs after one warm-up run each) and leaves the batch at the fastest.
 * path 0: integer path (score + traceback); fill shapes (waves, workgroups
 * per CU or per pair) timed by the fill kernel, then top-level shapes with
 * the best fill, the winner kept only if its graph replay beats the
 * starting shape's; every candidate's scores, alignments and structures
 * must equal the starting shape's, else LCG_E_STATE.  path 1: LocARNA-P
 * (waves, workgroups per pair), every candidate's partition functions within
 * 1e-12 relative of the starting shape's.  best: the chosen shape;
 * best_ms / default_ms: device time of one run at the chosen and at the
 * starting shape. */
int lcg_batch_autotune(lcg_batch *batch, int32_t path, int32_t reps, lcg_shape *best, float *best_ms,
                       float *default_ms, char *err, int errlen);
/* Arc matches and groups of pair k of a batch (ArcMatches::num_arc_matches;
 * the groups of AlignerImpl::align_D, aligner.cc:659-732) and the size of
 * its device image in ints. */
int lcg_batch_pair_info(const lcg_batch *batch, int32_t k, int32_t *num_arcmatches, int32_t *num_groups,
                        int64_t *image_ints);
/* Copies the device image of pair k (device_layout.h) to dst; returns its
 * size in ints (no counterpart in the reference; for the comparison of the
 * device compaction with the host packing). */
int lcg_batch_segment(const lcg_batch *batch, int32_t k, int32_t *dst, int64_t cap);
/* Reruns the device compaction of a batch of lcg_pair_prepare_device pairs
 * (both passes, same inputs, same result) and reports the device time of
 * each pass (no counterpart in the reference; for timing).  Results of an
 * earlier lcg_batch_run are invalidated. */
int lcg_batch_compact_run(lcg_batch *batch, float *count_ms, float *build_ms);
void lcg_batch_destroy(lcg_batch *batch);
/* Runs the kernels for all pairs of the batch.
 *   mode 0: batched, one persistent launch over a work queue (all pairs)
 *   mode 1: plain, one launch per pair (same kernel), pairs in order
 *   mode 2: batched, replay of a captured hipGraph of mode 0
 * trace: bit 0: also run the traceback; bit 1: count the DP cells and arc
 * candidates the fill evaluates (lcg_batch_counters; not with mode 2).
 * device_ms (may be NULL): device time of the run from HIP events. */
int lcg_batch_run(lcg_batch *batch, int32_t mode, int32_t trace, float *device_ms);
/* Penalized alignment (Aligner::penalized_align, aligner.cc:1603-1622;
 * locarna --penalized P): fill of D, then the top level with every
 * position charged P (sigma - 2P, gap costs - P, arc matches - P times both
 * arc lengths; Scoring::modify_by_parameter, scoring.cc:77-90,
 * ModifiedScoringView, aligner_impl.hh:197-262), sequence-local or global
 * as the pairs were prepared, and the traceback; lcg_batch_score then
 * returns the penalized score. */
int lcg_batch_run_penalized(lcg_batch *batch, int64_t position_penalty, float *device_ms);
/* Normalized alignment (Aligner::normalized_align, aligner.cc:1526-1593;
 * locarna --normalized L; sequence-local pairs only, else LCG_E_ARG): fill
 * of D once, then Dinkelbach iterations of the top level per pair from
 * lambda = 0: lambda <- (score + length lambda) / (length + L) with length
 * = (max_i - min_i + 1) + (max_j - min_j + 1) of the traced alignment, until
 * lambda repeats; lcg_batch_score returns the normalized score lambda and
 * the alignment is the last traced one.  LCG_E_STATE if a pair needs more
 * than max_iterations iterations. */
int lcg_batch_run_normalized(lcg_batch *batch, int64_t L, int32_t max_iterations, int32_t *iterations,
                             float *device_ms);
/* k-best suboptimal alignments by interval splitting (Aligner::suboptimal,
 * aligner.cc:1382-1514; locarna --kbest k / --better t, locarna.cc:203-206,
 * 740-746; not normalized): the best alignment of pair k, then the best of
 * each half of its restriction, split at the middle of the alignment's
 * extent in the longer sequence (both halves keep that position), best
 * first, until kbest alignments (kbest < 0: no limit) or until the best
 * queued score is at most threshold; empty alignments are skipped.  Runs
 * the batch's fill of D once (lcg_batch_run) and then one top-level launch
 * of pair k per restriction (AlignerRestriction, aligner_restriction.hh:
 * 26-120); afterwards the batch's results of pair k are those of its last
 * restricted launch.  Writes the first cap scores; returns the number of
 * alignments. */
int lcg_batch_suboptimal(lcg_batch *batch, int32_t k, int32_t kbest, int64_t threshold, int64_t *scores,
                         int32_t cap);
/* Alignment t of the last lcg_batch_suboptimal: edges a[e], b[e] (-1 for a
 * gap) and structures as lcg_batch_alignment, and its restriction (startA,
 * endA, startB, endB; may be NULL).  Returns the number of edges. */
int lcg_batch_suboptimal_alignment(const lcg_batch *batch, int32_t t, int32_t *a, int32_t *b, int32_t cap,
                                   char *structA, char *structB, int32_t scap, int32_t *restriction);
/* Allocation-free run (no counterpart in the reference; for callers that
 * cannot allocate device memory: code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc): the fill of D, the top level and,
 * with trace bit 0, the traceback of every pair of the batch, as
 * lcg_batch_run mode 0, with the work queues, the cross-workgroup arrival
 * counters and the top-level per-wave regions in a static 64 MiB device
 * array of the library instead of the batch's buffers.  Calls no hipMalloc
 * or hipFree, records no event and does not wait: it enqueues kernels on
 * the batch's stream (lcg_batch_set_stream) and returns; the result getters
 * synchronize.  LCG_E_CAPACITY when the batch needs more than the array
 * holds (one top-level workgroup's regions plus one counter per pair).
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (different batches, streams or threads) race. */
int lcg_batch_run_noalloc(lcg_batch *batch, int32_t trace);
/* Stream of a batch (a hipStream_t of the batch's device, or NULL for the
 * batch's own stream; no counterpart in the reference, which runs on the
 * CPU): every later launch and copy of the batch is ordered on it, after
 * the work the caller queued there before the call.  Calls stay blocking:
 * each run returns when its results are complete.  The caller keeps the
 * stream alive while the batch uses it.  The stream belongs to the HIP
 * runtime the library runs on: the libamdhip64.so.7 the process loaded
 * first (a PyTorch that is imported before the library brings its own). */
int lcg_batch_set_stream(lcg_batch *batch, void *stream);
/* Bound of the cross-workgroup waits of pairs that several workgroups share
 * (no counterpart in the reference; for tests of the error path): polls of
 * the arrival counter before a wait gives up (0: the default, 2^22; -1:
 * every wait gives up at once without waiting).  A wait that gives up marks
 * the pair's results invalid; every result request of that run (integer
 * and LocARNA-P getters alike) then fails with LCG_E_STATE, the first one
 * and all later ones, until a run completes without such a wait. */
int lcg_batch_set_spin_limit(lcg_batch *batch, int32_t polls);
/* launch configuration chosen for the batch: waves per workgroup,
 * workgroups, LDS bytes per workgroup */
int lcg_batch_config(const lcg_batch *batch, int32_t *waves, int32_t *workgroups, int32_t *lds_bytes);
/* Score of pair k (Aligner::align(), aligner.cc:924-962).  Minus infinity
 * (the reference's normalized InftyInt -inf, infty_int.cc:7-23: no
 * alignment reaches the end cell within the band, possible with free end
 * gaps) is reported as -2^29; every value below -2^28 means minus infinity,
 * here and in lcg_batch_D. */
int lcg_batch_score(const lcg_batch *batch, int32_t k, int64_t *score);
/* Alignment of pair k (Aligner::trace() + get_alignment(),
 * aligner.cc:1345-1363): edges a[t], b[t] (1-based positions, -1 = gap) in
 * alignment order, and the two structure strings (dot bracket).  Returns the
 * number of edges (>= 0) or an error. */
int lcg_batch_alignment(const lcg_batch *batch, int32_t k, int32_t *a, int32_t *b, int32_t cap, char *structA,
                        char *structB, int32_t scap);
/* Pairwise alignment of pair k as locarna writes it with --clustal
 * (src/LocARNA/main_helper.icc:557-580, MultipleAlignment::write,
 * multiple_alignment.cc:994-1085), after a run with traceback: the header
 * "CLUSTAL W --- LocARNA 2.0.1 --- Score: <score or -inf>" that mlocarna
 * reads the pair's score from (src/Utils/mlocarna:3516-3527; without the
 * score when a side is an alignment profile), then the rows of A and of B
 * ("A." / "B." name prefixes on a name clash, multiple_alignment.cc:
 * 186-247) in blocks of `width` columns (locarna --width, 120), names
 * padded to 18.  Positions
 * that a local or free-end-gap alignment leaves out are gap columns
 * (Alignment::alignment_edges, alignment.cc:121-167) unless
 * LCG_CLUSTAL_LOCAL_ONLY (--local-file-output); LCG_CLUSTAL_STRUCTURES adds
 * the structure rows (--write-structure).  Returns the text length; writes
 * the text and its terminating 0 into buf when cap exceeds the length. */
#define LCG_CLUSTAL_LOCAL_ONLY 1
#define LCG_CLUSTAL_STRUCTURES 2
int64_t lcg_batch_write_clustal(const lcg_batch *batch, int32_t k, int32_t width, int32_t flags, char *buf,
                                int64_t cap);
/* D table of pair k in the reference's arc-match order
 * (AlignerImpl::D(am), aligner.cc:574-610).  Returns the count. */
int lcg_batch_D(const lcg_batch *batch, int32_t k, int64_t *D, int32_t cap);

/* LDS schedule of the fill of D for a prepared pair, as lcg_batch_create
 * computes it (no counterpart in the reference; exported for inspection):
 * per group 4 ints (round | wave<<12 | member<<16, LDS offset, LDS ints,
 * al | bl<<16),
 * per row al the number of rounds (rounds[0..n+1]).  `members` workgroups
 * share the pair.  Returns the number of groups. */
int lcg_pair_schedule(const lcg_pair *pair, int32_t waves, int32_t pool_ints, int32_t members,
                      int32_t *group_words, int32_t *rounds, int32_t cap);

/* Work counted by the last run with trace bit 1 set: DP cells and arc-match
 * candidates evaluated by the fill of D and the top level of pair k. */
int lcg_batch_counters(const lcg_batch *batch, int32_t k, int64_t *cells, int64_t *candidates);
/* The same split by kernel: the fill of D (all arc-match groups) and the top
 * level (no counterpart in the reference; for the work figures of a
 * roofline). */
int lcg_batch_counters_split(const lcg_batch *batch, int32_t k, int64_t *fill_cells, int64_t *fill_candidates,
                             int64_t *top_cells, int64_t *top_candidates);
/* Device times of the fill-D kernel and of the top-level/traceback kernel in
 * the last mode-0 run (no counterpart in the reference). */
int lcg_batch_kernel_times(const lcg_batch *batch, float *fill_ms, float *top_ms);

/* ---- LocARNA-P (FP64): inside partition function and its adjoint.
 * Score parameters theta (LCG_NTHETA = 275 doubles): base-match scores of
 * the 4x4 ACGU pairs (16, row-major A,C,G,U), the RIBOSUM arc-match term of
 * the 16x16 base-pair pairs (256, index (uA*16 + uB), pair index = left*4 +
 * right), indel, indel opening, struct weight.  Every score of a pair is
 * affine in theta around the reference scores (theta offsets `dtheta` = 0
 * give the reference's rounded integer scores).
 * lcg_batch_pf_run: runs inside + adjoint for all pairs (AlignerP::
 * align_inside / align_outside, aligner_p.icc:413-438, 1113-1139). */
#define LCG_NTHETA 275
int lcg_batch_pf_run(lcg_batch *batch, const double *dtheta, float *device_ms);
/* Text of the probability files of locarna_p (AlignerP::
 * write_arcmatch_probabilities / write_basematch_probabilities,
 * aligner_p.icc:1404-1435; --write-arcmatch-probs / --write-basematch-probs,
 * src/locarna_p.cc:499-526) for pair k after lcg_batch_pf_run: "al ar bl br
 * p" per arc match in index order with p >= min_am_prob, "i j p" per base
 * match in row-major order with p >= min_bm_prob (with or without the arc
 * ends, --include-am-in-bm); numbers as an ostream writes them (%g).
 * Returns the text length; writes the text and its terminating 0 into buf
 * when cap exceeds the length. */
int64_t lcg_batch_write_am_probs(const lcg_batch *batch, int32_t k, char *buf, int64_t cap);
int64_t lcg_batch_write_bm_probs(const lcg_batch *batch, int32_t k, int32_t include_am_in_bm, char *buf,
                                 int64_t cap);
/* Stochastic traceback (no counterpart in the reference's AlignerP; for
 * fitting score parameters to sampled alignments): draws `nsamples`
 * alignments of pair k from the Boltzmann distribution of the last
 * lcg_batch_pf_run (global top level), seeded, and returns the mean and
 * variance of their feature vectors over the LCG_NTHETA parameters (score of
 * an alignment = reference score + theta . features: sequence match edges by
 * ACGU pair, tau/100 per arc match by base-pair types, gap positions, gap
 * openings, struct-weight terms).  The mean estimates
 * T d log Z / d theta (lcg_batch_pf_grad). */
int lcg_batch_pf_sample(const lcg_batch *batch, int32_t k, uint64_t seed, int32_t nsamples, double *mean,
                        double *var, int32_t cap);
/* LocARNA-P launch configuration of the batch (no counterpart in the
 * reference): the first `nstatic` pairs of the work order (largest first) run
 * on `wg_per_pair` workgroups each, the others one workgroup each from a
 * queue; `workgroups` in all. */
int lcg_batch_pf_config(lcg_batch *batch, int32_t *wg_per_pair, int32_t *nstatic, int32_t *workgroups);
/* Z (AlignerP::align_inside) and the expected number of gap openings */
int lcg_batch_pf_result(const lcg_batch *batch, int32_t k, double *Z, double *openings);
/* arc-match probabilities in the reference's arc-match order
 * (compute_arcmatch_probabilities, aligner_p.icc:1150-1195) */
int lcg_batch_am_probs(const lcg_batch *batch, int32_t k, double *prob, int32_t cap);
/* inside D(am) and outside D'(am) of every arc match in the reference's
 * arc-match order (AlignerP::D / Dprime over Dmat and Dmatprime,
 * aligner_p.icc:117-128, 450-1139; D'(am) = dZ/dD(am) e^{s(am)/T} / pf_scale).
 * Either pointer may be NULL.  Returns the count. */
int lcg_batch_pf_D(const lcg_batch *batch, int32_t k, double *D, double *Dprime, int32_t cap);
/* base-match probabilities, dense (lenA+1)*(lenB+1): mode 0 as
 * compute_basematch_probabilities(false) (aligner_p.icc:1201-1399), mode 1
 * with --include-am-in-bm, mode 2 marginals of all match edges */
int lcg_batch_bm_probs(const lcg_batch *batch, int32_t k, int32_t mode, double *dense, int32_t cap);
/* d log Z / d theta of pair k (no counterpart in the reference) */
int lcg_batch_pf_grad(const lcg_batch *batch, int32_t k, double *grad, int32_t cap);
/* Fragment match probabilities (AlignerP::compute_fragment_match_prob,
 * aligner_p.hh:605-609, aligner_p.icc:1440-1479; locarna_p
 * --fragment-match-probs, locarna_p.cc:537-561): the fragments of pair k
 * whose probabilities the next lcg_batch_pf_run computes, A[i..j] matched
 * with B[k..l] as one block for each ijkl[4q .. 4q+3] (1-based, i <= j <= n,
 * k <= l <= m; LCG_E_ARG otherwise); nq = 0 clears them.  A list longer
 * than the one the LocARNA-P buffers were laid out for discards the results
 * of the last run (the next run lays them out again).  A block with a
 * corner cell, (i-1, k-1) or (j, l), outside the trace band has
 * probability 0. */
int lcg_batch_pf_set_fragments(lcg_batch *batch, int32_t k, int32_t nq, const int32_t *ijkl);
/* The fragment match probabilities of pair k from the last lcg_batch_pf_run,
 * in the order they were set; returns their count. */
int lcg_batch_pf_fragment_probs(const lcg_batch *batch, int32_t k, double *prob, int32_t cap);

/* Consensus pp text of pair k's alignment, after a run with traceback
 * (locarna --pp, main_helper.icc:617-632; RnaData consensus constructor,
 * rna_data.cc:75-122, 1475-1570; write_pp, :1242-1339): the rows of both
 * inputs with the alignment's gaps and the consensus base-pair
 * probabilities, the input of the next step of mlocarna's progressive
 * alignment.  Base-pair lines in (i, j) order (the reference writes them
 * in hash order).  Returns the text length; writes the text and its
 * terminating 0 into buf when cap exceeds the length. */
int64_t lcg_batch_write_consensus_pp(const lcg_batch *batch, int32_t k, char *buf, int64_t cap);

/* mlocarna's guide tree (MLocarna::Tree UPGMA on the similarity matrix,
 * lib/perl/MLocarna/Tree.pm:164-315; mlocarna:2352-2390): scores[i n + j]
 * the pairwise alignment scores (diagonal ignored), distances max - score,
 * average linkage, ties to the first pair in the reference's cluster-list
 * order.  merges[2t], merges[2t+1] receive the two children of merge t
 * (leaves 0..n-1, merge t is node n + t), in the order the reference's
 * tree lists them. */
int lcg_guide_tree_upgma(const int64_t *scores, int32_t n, int32_t *merges);

/* mlocarna's progressive alignment (src/Utils/mlocarna:3660-3707,
 * call_locarna :3462-3500) along a guide tree: each merge aligns the pp
 * texts of its two children (leaves: pp[0..n-1]; merge t: the consensus pp
 * of its result, lcg_batch_write_consensus_pp) with `params`.  Merges
 * whose children are done run in one round: their pairs are prepared
 * together (lcg_pairs_prepare with the device band for pairs of single
 * sequences) and aligned in one batch per sigma kind.  The reference runs
 * one locarna process per merge; the results are the same. */
typedef struct lcg_progressive lcg_progressive;
int lcg_progressive_run(const char *const *pp, int32_t n, const int32_t *merges, const lcg_params *params,
                        int32_t device, lcg_progressive **out, char *err, int errlen);
/* pp text of a node (leaf i or merge n + t); returns its length, writes it
 * with its terminating 0 when cap exceeds the length */
int64_t lcg_progressive_pp(const lcg_progressive *r, int32_t node, char *buf, int64_t cap);
/* CLUSTAL text of merge t (locarna --clustal of that step; the last merge is
 * mlocarna's result alignment) */
int64_t lcg_progressive_clustal(const lcg_progressive *r, int32_t t, int32_t width, char *buf, int64_t cap);
/* alignment score of every merge; returns n - 1 */
int lcg_progressive_scores(const lcg_progressive *r, int64_t *scores, int32_t cap);
/* rounds, batches, summed device time of the alignment runs, host time of
 * the pair preparation, wall time of lcg_progressive_run */
int lcg_progressive_stats(const lcg_progressive *r, int32_t *rounds, int32_t *batches, float *device_ms,
                          float *prepare_ms, float *wall_ms);
void lcg_progressive_free(lcg_progressive *r);

/* Base-pair probabilities of single sequences on the device: the input
 * step mlocarna runs per sequence (compute_dotplot_rnafold_pp,
 * mlocarna:3148-3188 -> locarna_rnafold_pp, locarna_rnafold_pp.cc:209-271:
 * RnaEnsemble's partition function folding, rna_ensemble.cc:141-193, the
 * p > min_prob filter of rna_data.cc:509-547 and RnaData::write_pp,
 * rna_data.cc:1242-1350).  The partition function is McCaskill's, over this
 * library's reduced nearest-neighbour energy model (canonical pairs,
 * Turner 2004 Watson-Crick stacks, size-dependent hairpin, bulge and
 * interior loops up to 30 unpaired bases, linear multiloop, terminal AU/GU
 * penalties; no dangles), not ViennaRNA's full Turner 2004 set: the
 * probabilities are McCaskill probabilities of that model.  Sequences are
 * upper-cased and T read as U (normalize_rna_sequence, aux.cc:65-70);
 * other symbols do not pair.  LCG_E_CAPACITY for sequences longer than
 * 4096, or when no scale keeps a sequence's values in the double range. */
typedef struct lcg_fold lcg_fold;
typedef struct {
    int32_t max_bp_span; /* --maxBPspan: pairs with j - i + 1 <= span; -1: no limit */
    int32_t stacking;    /* --stacking: joint probabilities of (i,j) and (i+1,j-1) (#STACK) */
    double min_prob;     /* -p / --min-prob: pairs with p > min_prob are written
                            (locarna_rnafold_pp 0.001; mlocarna passes its --min-prob) */
    double scale_kcal_per_nt; /* free energy per base (kcal/mol) the partition
                            function is scaled by (ViennaRNA's pf_scale: a
                            segment of L bases is stored times
                            exp(L g / kT)); 0: chosen per sequence, retried
                            by bisection when a value leaves the double
                            range.  The probabilities do not depend on it. */
} lcg_fold_params;
int lcg_fold_run(const char *const *seqs, int32_t count, const lcg_fold_params *params, int32_t device,
                 lcg_fold **out, char *err, int errlen);
/* pp text of sequence k under `name` (RnaData::write_pp): pairs with
 * p > min_prob in order i, j ascending; returns the length, writes the text
 * with its terminating 0 when cap exceeds it.  The handle keeps the last
 * text per sequence (a length query followed by the copy builds it once),
 * so calls on one handle must not overlap. */
int64_t lcg_fold_write_pp(const lcg_fold *f, int32_t k, const char *name, char *buf, int64_t cap);
/* pair probabilities P (RnaEnsemble::arc_prob) and joint probabilities P2
 * of (i,j) and (i+1,j-1) (RnaEnsemble::arc_2_prob, rna_ensemble.cc:415-436)
 * as dense (n+1) x (n+1) row-major matrices (either may be NULL); returns n */
int lcg_fold_probs(const lcg_fold *f, int32_t k, double *P, double *P2, int64_t cap);
/* natural logarithm of the partition function of sequence k */
int lcg_fold_log_z(const lcg_fold *f, int32_t k, double *log_z);
/* device time of the folding kernel, wall time of lcg_fold_run */
int lcg_fold_stats(const lcg_fold *f, float *device_ms, float *wall_ms);
void lcg_fold_free(lcg_fold *f);

/* mlocarna's default pipeline from sequences to the multiple alignment
 * (src/Utils/mlocarna: compute_all_dotplots :3099-3110, the all-pairs
 * stage :3547-3640, the UPGMA guide tree :2352-2420 and the progressive
 * steps :3660-3707) on one GPU: lcg_fold_run on all inputs (pp texts under
 * names[k]), every pair a > b prepared together (device band; device
 * arc-match compaction unless --noLP, stacking or --struct-local) and
 * aligned in one batch, lcg_guide_tree_upgma on the scores, then
 * lcg_progressive_run along the tree.  n >= 2. */
typedef struct lcg_mlocarna lcg_mlocarna;
int lcg_mlocarna_run(const char *const *names, const char *const *seqs, int32_t n, const lcg_fold_params *fold,
                     const lcg_params *params, int32_t device, lcg_mlocarna **out, char *err, int errlen);
/* the result alignment (CLUSTAL text of the last merge) */
int64_t lcg_mlocarna_clustal(const lcg_mlocarna *r, int32_t width, char *buf, int64_t cap);
/* pp text of node k: an input (k < n, as lcg_fold_write_pp wrote it) or a
 * merge (n + t) */
int64_t lcg_mlocarna_pp(const lcg_mlocarna *r, int32_t node, char *buf, int64_t cap);
/* all-pairs scores, n x n row major (0 on the diagonal); returns n */
int lcg_mlocarna_scores(const lcg_mlocarna *r, int64_t *scores, int64_t cap);
/* the guide tree as lcg_guide_tree_upgma returns it; returns n - 1 */
int lcg_mlocarna_tree(const lcg_mlocarna *r, int32_t *merges, int32_t cap);
/* wall times of the three stages and of the whole run */
int lcg_mlocarna_stats(const lcg_mlocarna *r, float *fold_ms, float *pairs_ms, float *progressive_ms,
                       float *wall_ms);
void lcg_mlocarna_free(lcg_mlocarna *r);

/* Synthetic inputs (no counterpart in the reference; they replace
 * RNAfold -p, which the reference calls through ViennaRNA,
 * rna_data.cc:52-61). */
int lcg_synth_sequences(uint64_t seed, int count, int len_min, int len_max, double gc_min, double gc_max,
                        char *out_seqs, int stride, int *out_lens);
int lcg_synth_bpp(const char *seq, int len, int max_span, double cutoff, int *out_i, int *out_j, double *out_p,
                  int cap);

#ifdef __cplusplus
}
#endif

#endif
