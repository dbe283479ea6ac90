// Device-resident layout of a batch of pairwise alignment problems.
//
// Everything of a batch lives in one int32 arena in HBM; a PairDesc holds
// int32 element offsets into it.  Positions are 1-based as in the reference
// (index 0 of per-position arrays is unused or the DP origin).
//
// Per pair:
//   band:     min_col[n+1], max_col[n+1], cell_off[n+2]   (band cells of row i
//             are cell_off[i] + (j - min_col[i]); rows may be empty)
//   codes:    codeA[n+1], codeB[m+1]    symbol codes (0..3 = ACGU)
//   gaps:     gapA[n+1], gapB[m+1]      position-specific gap scores
//   sigma:    8 x 8 base-match scores by symbol code; for a pair of
//             alignment profiles (n+1) x (m+1) scores by column pair
//   rows:     row_begin[n+2]            groups of left end al are
//                                       [row_begin[al+1], row_begin[al])
//                                       (groups are stored al descending)
//   rounds:   n+2                       LDS packing rounds per row
//   groups:   LCG_GROUP_INTS per group  (al | bl<<16, ar | br<<16, W,
//                                        am_begin, am_end,
//                                        round | wave<<12 | member<<16,
//                                        lds offset, lds ints) where ar, br are
//                                       the maximal right ends and W the
//                                       stored row width of the group region
//   gams:     4 per arc match, group order: (ar | br<<16, candidate entry,
//             arc match, fill entry)
//   amscore:  s(am) per arc match (reference index order)
//   cptr:     ncells + 1                CSR over band cells of the arc matches
//   cand:     4 per entry               with those right ends, sorted by cell
//                                       then (al desc, bl desc):
//                                       (al | bl<<16, j, D(am),
//                                        min_col(al - 1))
//   fptr:     n + 2                     first fill block of row i
//   fblk:     2 per block + sentinel    one block per arc of A (left end p,
//                                       right end i), rows i ascending, p
//                                       descending within a row: (p, first
//                                       fill entry); sentinel (0, nam)
//   fcand:    3 per entry               the arc matches again, by block,
//                                       j ascending: (p | q<<16,
//                                       j | min_col(p-1)<<16, D(am)); a group
//                                       with left end al reads the prefix
//                                       p > al of each row
//   f2c:      per fill entry            its index in cand (cell order)
//   D:        per arc match (output of the fill, reference index order)
//   edges:    n + m + 1 slots (output): slot i+j holds the edge that reaches
//             cell (i, j): (a << 16) | b with 0 for a gap side, -1 if empty
//   structA/B: n + m ints (output): 0 '.', 1 '(', 2 ')'
//   tstack:   2 * (n + 1) ints scratch (traceback work list)
//   out:      [score, max_i, max_j, status (0 ok, 2 a bounded wait of the
//             multi-workgroup fill expired), fill-D cells, candidates (u64),
//             top-level cells, candidates (u64)]
//   inner:    --noLP only: per arc match the index of the arc match of the
//             arcs directly inside both of its arcs, -1 if none
//   c2a:      --noLP only: per candidate entry its arc match
// With --noLP a group (al, bl) computes the D of the arc matches with left
// ends (al-1, bl-1) that have an inner arc match (their gams entries hold
// the inner arc match's right ends); the others keep D = minus infinity.
//
// A group region in LDS (or scratch) is [E: br-bl][S: W][M: (ar-al) * W]:
// E the deletion vector over columns bl..br-1, S the arc-candidate maxima of
// the current row, M the matrix rows al..ar-1 (row i stored from column
// max(bl, min_col(i))).
#pragma once
#include <stdint.h>

#define LCG_NEG (-(1 << 29))        /* normalized minus infinity */
#define LCG_NEG_LIMIT (-(1 << 28))  /* values below are minus infinity */
#define LCG_GROUP_INTS 8
#define LCG_NCODES 8

typedef struct {
    int32_t n, m;
    int32_t nam, ngroups;
    int32_t band, codes, gaps, sigma;
    int32_t rows, groups, gams, amscore;
    int32_t cptr, cand, D, edges;
    int32_t structs, tstack, out, rounds;
    int32_t wtop;      // stored width of the top-level rows
    int32_t gmax;      // max LDS ints of a group region
    int32_t ncells;
    int32_t top_ints;  // LDS ints of the top-level region
    int32_t fptr, fblk, fcand;
    int32_t nfblk;     // fill blocks including the sentinel
    int32_t f2c;       // fill entry -> candidate entry (cell order)
    int32_t inner;     // --noLP / --stacking only: per arc match the arc match directly inside it, -1 if none
    int32_t c2a;       // --noLP / --stacking only: candidate entry -> arc match
    int32_t amst;      // --stacking only: per arc match its stacked score, LCG_NEG if not stackable
    int32_t slmax;     // --struct-local: ints of the largest group region (eight states)
} PairDesc;  // 33 ints

typedef struct {
    int32_t alpha;           // indel opening
    int32_t local;           // sequence-local top level
    int32_t trace;           // run the traceback
    int32_t npairs;
    int32_t hdr_ints;        // per-pair LDS header (band, codes, gaps, sigma)
    int32_t pool_ints;       // LDS pool for group regions
    int32_t max_len;         // max sequence length over the batch
    int32_t count;           // count DP cells and arc candidates (out[4..11])
    int32_t scratch_stride;  // ints of global scratch per (workgroup, wave)
    int32_t fixed_cols;      // fill narrow groups with the fixed-column kernel
    int32_t fptr_lds;        // LDS offsets (ints) of the pair's fptr and fblk
    int32_t fblk_lds;
    int32_t wg_per_pair;     // workgroups sharing one pair in the fill of D
    int32_t free_endgaps;    // top level: LCG_FREE_* bits (left_1 1, right_1 2, left_2 4, right_2 8)
    int32_t spin_limit;      // polls of a bounded cross-workgroup wait (0: default,
                             // -1: every wait gives up at once, for error-path tests)
    int32_t nolp;            // --noLP: D of stacks (groups keyed by the inner left ends)
    int32_t stacking;        // --stacking / --new-stacking: stacked alternative of D
    int32_t sigma_table;     // pairs of alignment profiles: sigma is a per-pair
                             // (n+1) x (m+1) table at PairDesc.sigma (row i at
                             // sigma + i * (m+1)), read from global memory
    int32_t struct_local;    // --struct-local: eight states inside arc matches,
                             // group regions in global scratch
    int32_t exclusion;       // --exclusion: score of one excluded block
    int32_t sl_stride;       // ints of global scratch per (workgroup, wave) of the
                             // fill with --struct-local
} BatchParams;
