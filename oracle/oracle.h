/* CPU oracle: a plain restatement of LocARNA 2.0.1's pairwise alignment
 * path, used ONLY as the checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product links or calls it.
 *
 * Each function in oracle.cc cites the reference file:line it follows
 * (paths relative to the LocARNA source tree).
 *
 * Pinning (how the oracle itself is checked, see tests/test_oracle_*.py):
 *   - pp reader filter (p == cut dropped, span == maxBPspan kept, #BPCUT
 *     only raising the cut), BasePairs arc order (left end descending, right
 *     end ascending), arc weights (0 at p = pe, struct_weight at p = 1,
 *     half of it at sqrt(pe), monotone, --exp-prob), the truncating tau
 *     division, the band's StralScore similarity (left ends with left ends)
 *     and the trace cut (the smallest monotone band holding every cell with
 *     probability >= the cut, ties included): closed forms on hand-built
 *     pp inputs (tests/test_oracle_tables.py), each failing when a
 *     plausible mistake is planted in this file;
 *   - RIBOSUM reader: the data file's printed base-match and arc-match
 *     matrices are reproduced from its frequency tables, and the band's
 *     similarity equals the printed base-match matrix at unpaired cells;
 *   - parity unpinned: the corrected RIBOSUM base-match score sigma
 *     (ribosum.cc:325-331) and the riboX arc-match term
 *     (scoring.cc:368-438) follow the reference's formulas, which no check
 *     available here fixes independently;
 *   - trace band (PFGotoh, PFTraceProbs): enumeration of all alignments
 *     (global) or of all prefixes and suffixes (sequence-local) on tiny
 *     inputs, trace probabilities to 1e-12 in long double and double
 *     (tests/test_oracle_band.py, through oracle_trace_probs); with free
 *     end gaps, enumeration of the prefix paths of the forward and reversed
 *     problems (free starts included) and of Z;
 *   - free end gaps of the integer DP: exhaustive enumeration with free
 *     leading runs and free ends (score and traceback, 9 end
 *     specifications);
 *   - penalized alignment: exhaustive enumeration over the penalized score
 *     tables (local and global); normalized alignment: Dinkelbach
 *     iterations over pinned penalized alignments;
 *   - --max-diff band (Band::constrain_wo_ref, trace_controller.cc:361-390):
 *     cell membership against the distance to the scaled diagonal in exact
 *     integers, |i - j| <= max-diff for equal lengths, non-empty connected
 *     rows (tests/test_host_max_diff.py; the reference's own TraceController
 *     test covers only the band of a reference alignment);
 *     --max-diff-am / --max-diff-at-am: the predicate
 *     of arc_matches.cc:19-48, checked through the arc-match lists;
 *   - integer DP: exhaustive enumeration of all alignments x nested
 *     arc-match sets on tiny inputs (max score and the traceback's tie rule),
 *     reduction to textbook Gotoh / Smith-Waterman-Gotoh without arcs,
 *     re-scoring of every traced alignment;
 *   - partition function: exhaustive enumeration of Z and of match
 *     probabilities on tiny inputs, reduction to the Gotoh forward algorithm
 *     without arcs.
 */
#ifndef LOCARNA_ORACLE_H
#define LOCARNA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int indel;             /* --indel            (score, e.g. -350)  */
    int indel_opening;     /* --indel-opening    (score, e.g. -500)  */
    int struct_weight;     /* --struct-weight    (200)               */
    int tau;               /* --tau              (0 locarna cfgs, 50 locarna_p) */
    int match, mismatch;   /* fallback for non-ACGU symbols (50, 0)  */
    int temperature_alipf; /* --temperature-alipf (300)              */
    int sequ_local;        /* --sequ-local                           */
    int max_bp_span;       /* --maxBPspan, -1 = unlimited            */
    int band_extended;     /* 1: band PF in long double (locarna), 0: double (locarna_p) */
    int use_ribosum;       /* 1: RIBOSUM85_60, 0: match/mismatch      */
    int free_endgaps;      /* --free-endgaps: bit 0 left end of the first
                              sequence, 1 its right end, 2 left end of the
                              second, 3 its right end (free_endgaps.hh) */
    int max_diff;          /* --max-diff (-1: no band around the diagonal) */
    int max_diff_am;       /* --max-diff-am (-1: the longer length)       */
    int max_diff_at_am;    /* --max-diff-at-am (-1: the longer length)    */
    double min_prob;       /* --min-prob                             */
    double min_trace_prob; /* --min-trace-probability (<=0: no band)  */
    double exp_prob;       /* --exp-prob; <=0: 1/(2 len)              */
    double pf_scale;       /* --pf-scale                             */
    double min_am_prob;    /* --min-am-prob                          */
    double min_bm_prob;    /* --min-bm-prob                          */
    int no_lonely_pairs;   /* --noLP (integer path): arc matches only in
                              stacks of two or more                   */
    int stacking;          /* integer path: bit 0 --stacking, bit 1
                              --new-stacking (stacked arc-match scores from
                              the #STACK joint probabilities)         */
    int struct_local;      /* --struct-local (integer path): at most one
                              excluded block per sequence inside every
                              arc match                               */
    int exclusion;         /* --exclusion: score of one exclusion      */
} oracle_params;

typedef struct oracle_result oracle_result;

/* Integer pairwise alignment (locarna).  ribosum_path: RIBOSUM file.  On
 * error returns NULL and writes a message into err. */
oracle_result *oracle_align(const char *ppA, const char *ppB,
                            const oracle_params *p, const char *ribosum_path,
                            int do_trace, char *err, int errlen);

/* LocARNA-P: inside, outside, arc-match and base-match probabilities. */
oracle_result *oracle_pf(const char *ppA, const char *ppB,
                         const oracle_params *p, const char *ribosum_path,
                         int include_am_in_bm, char *err, int errlen);

/* Fragment match probabilities (AlignerP::compute_fragment_match_prob,
 * locarna_p --fragment-match-probs) of the problem of an oracle_pf result:
 * prob[q] for fragments A[i..j], B[k..l] given as ijkl[4q..4q+3].  Returns
 * nq or -1 (message in err). */
int oracle_pf_fragments(oracle_result *r, int nq, const int *ijkl, double *prob, char *err, int errlen);

/* k-best suboptimal alignments by interval splitting (Aligner::suboptimal,
 * locarna --kbest / --better): at most cap_alns alignments; scores[t], the
 * restriction restr[4t..4t+3] = startA, endA, startB, endB (may be NULL) and
 * edge_count[t] edges of alignment t at edge_a / edge_b + t * cap_edges
 * (-1 for a gap).  Returns the number of alignments or -1. */
int oracle_suboptimal(const char *ppA, const char *ppB, const oracle_params *p, const char *ribosum_path, int k,
                      long long threshold, long long *scores, int *restr, int *edge_a, int *edge_b,
                      int *edge_count, int cap_alns, int cap_edges, char *err, int errlen);

/* Penalized (mode 1, arg = position penalty; locarna --penalized) and
 * normalized (mode 2, arg = L, sequence-local; --normalized) alignment with
 * traceback; the score of a normalized alignment is the normalized score. */
oracle_result *oracle_align_variant(const char *ppA, const char *ppB, const oracle_params *p,
                                    const char *ribosum_path, int mode, long long arg, char *err, int errlen);

void oracle_free(oracle_result *r);
/* band step: trace probabilities over the full matrix, the similarity of
 * every cell and w = (exp(opening/T), exp(extension/T), T, lenA, lenB) */
int oracle_trace_probs(const char *ppA, const char *ppB, const oracle_params *p, const char *ribosum_path,
                       double *probs, double *sim, double *w, int cap);

long long oracle_score(const oracle_result *r);
double oracle_partition_function(const oracle_result *r);
int oracle_lengths(const oracle_result *r, int *lenA, int *lenB);
/* alignment edges: a[k], b[k] with -1 for a gap; returns count */
int oracle_edges(const oracle_result *r, int *a, int *b, int cap);
/* structure strings of the traced alignment (dot bracket per sequence) */
int oracle_structures(const oracle_result *r, char *sa, char *sb, int cap);
/* band: min_col[0..lenA], max_col[0..lenA] */
int oracle_band(const oracle_result *r, int *min_col, int *max_col, int cap);
/* arc matches in reference index order: ends, score s(am), D(am) (integer
 * path) */
int oracle_arcmatches(const oracle_result *r, int *al, int *ar, int *bl,
                      int *br, long long *score, long long *D, int cap);
/* stacked arc-match scores (--stacking) in arc-match order, stackable[k] = 0
 * where the arc match has no stacked score; returns the count */
int oracle_stacked_scores(const oracle_result *r, long long *score, int *stackable, int cap);
/* score tables: sigma (lenA+1)*(lenB+1) row major, gapA[lenA+1], gapB[lenB+1] */
int oracle_tables(const oracle_result *r, long long *sigma, long long *gapA, long long *gapB);
/* work counters: [N_A, N_B, A, G, C, K] */
int oracle_counters(const oracle_result *r, long long *out6);
/* probabilities (pf): arc matches as in oracle_arcmatches order */
int oracle_am_probs(const oracle_result *r, double *prob, int cap);
/* base match probabilities as dense (lenA+1)*(lenB+1) row major */
int oracle_bm_probs(const oracle_result *r, double *prob, int cap);
/* arc match inside D (pf) and outside D' in arcmatch order */
int oracle_pf_D(const oracle_result *r, double *D, double *Dprime, int cap);

/* pf extras: unrestricted base-match (match edge, not arc end) marginals,
 * dense (lenA+1)*(lenB+1), and the expected number of gap openings */
int oracle_pf_extra(const oracle_result *r, double *bm_full, int cap, double *openings);
/* pp text of the aligned pair given by its edges (ea[k], eb[k]: 1-based
 * positions, -1 for a gap): rows of both inputs with the alignment's gaps,
 * consensus base-pair probabilities (locarna --pp; rna_data.cc:1475-1570,
 * :1242-1339).  Returns the text length (written with its terminating zero
 * when it fits cap), -1 on an error (message in err). */
long long oracle_consensus_pp(const char *ppA, const char *ppB, const oracle_params *p, const int *ea,
                              const int *eb, int nedges, char *buf, long long cap, char *err, int errlen);
/* probability text of pp files (format_prob, rna_data.cc:1274-1303) */
int oracle_format_prob(double prob, char *buf, int cap);

/* Base-pair probabilities of one sequence (oracle/fold.cc: McCaskill over
 * the library's energy model): P and P2 dense (n+1)*(n+1) row major (P2:
 * joint probability of (i,j) and (i+1,j-1); either may be NULL), log Z.
 * max_span < 0: no span limit.  Returns n or -1. */
int oracle_fold(const char *seq, int max_span, double *P, double *P2, double *log_z);

/* CPU baseline: align many pairs with nthreads worker threads (scores
 * only, with traceback).  Returns number of pairs done or <0. */
int oracle_align_many(const char **ppA, const char **ppB, int npairs,
                      const oracle_params *p, const char *ribosum_path,
                      int nthreads, long long *scores);

#ifdef __cplusplus
}
#endif

#endif
