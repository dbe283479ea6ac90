// Device-resident layout of a batch of pairwise alignment problems.
//
// Everything of a batch lives in one int32 arena in HBM; a PairDesc holds
// int32 element offsets into it.  Positions are 1-based as in the reference
// (index 0 of per-position arrays is unused or the DP origin).
//
// Per pair:
//   band:     min_col[n+1], max_col[n+1], cell_off[n+2]   (band cells of row i
//             are cell_off[i] + (j - min_col[i]))
//   codes:    codeA[n+1], codeB[m+1]    symbol codes (0..3 = ACGU)
//   gaps:     gapA[n+1], gapB[m+1]      position-specific gap scores
//   sigma:    8 x 8 base-match scores by symbol code
//   rows:     row_begin[n+2]            groups of left end al are
//                                       [row_begin[al+1], row_begin[al])
//                                       (groups are stored al descending)
//   groups:   GROUP_INTS per group      (al | bl<<16, ar | br<<16, W, am_begin,
//                                        am_end, 0, 0, 0) where ar, br are the
//                                       maximal right ends and W the stored
//                                       row width of the group region
//   gams:     2 per arc match, group order: (ar | br<<16, am)
//   amscore:  s(am) per arc match (reference index order)
//   cptr:     ncells + 1                CSR over band cells of arc matches by
//   cand:     2 per entry               right ends, sorted (al desc, bl desc):
//                                       (al | bl<<16, am)
//   D:        per arc match (output of the fill, kernel-written)
//   edges:    n + m + 1 slots (output): slot i+j holds the edge that reaches
//             cell (i, j): (a << 16) | b with 0 for a gap side, -1 if empty
//   structA/B: n + m ints (output): 0 '.', 1 '(', 2 ')'
//   tstack:   2 * (n + 1) ints scratch (traceback work list)
//   topM:     scratch for the top-level matrix when it does not fit in LDS
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
    int32_t structs, tstack, topM, out;   // out: [score, max_i, max_j, status, cells(u64), candidates(u64)]
    int32_t wtop;                         // stored width of the top-level rows
    int32_t gbuf;                         // max over groups of rows*W (ints)
    int32_t ncells;
    int32_t pad[9];
} PairDesc;  // 32 ints

typedef struct {
    int32_t alpha;        // indel opening
    int32_t local;        // sequence-local top level
    int32_t trace;        // run the traceback
    int32_t npairs;
    int32_t lds_wave_ints;  // per-wave LDS area (E + M), ints
    int32_t lds_hdr_ints;   // per-pair header area (band arrays), ints
    int32_t max_len;        // max n over the batch (header sizing)
    int32_t ecap;           // per-wave E capacity (ints); M follows it
    int32_t count;          // count DP cells and arc candidates (out[4..7])
    int32_t pad[3];
} BatchParams;
