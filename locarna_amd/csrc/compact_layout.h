// Inputs and workspace of the device-side arc-match compaction and packing.
//
// The host keeps what is O(n + m) per pair: the trace band, the two arc
// lists (BasePairs), per-arc weights and base-pair types, and the adjacency
// of arcs by left and by right end.  The device evaluates the arc-match
// predicate over all N_A x N_B arc pairs (reference: ArcMatches,
// src/LocARNA/arc_matches.cc:19-48, 130-188) and builds every arc-match
// structure of the pair's arena image (device_layout.h): the arc-match
// scores in the reference's index order, the groups (common left ends) with
// their LDS schedule, the candidate lists by right-end cell and the fill
// list.
//
// All offsets are int32 element offsets: CompactDesc fields into the input
// arena (`cin`), ws_* fields into the two workspaces (ints; ws_bits is 2-int
// aligned and addressed as uint64).
#pragma once
#include <stdint.h>

typedef struct {
    int32_t n, m, NA, NB;
    int32_t W64;           // 64-bit words per bitmap row: ceil(NB / 64)
    int32_t arcsA, arcsB;  // left | right<<16 per arc, index order
    int32_t wA, wB;        // arc weights (Scoring::probToWeight)
    int32_t ptA, ptB;      // base-pair type 4*code(left)+code(right), -1 if
                           // an end is not A, C, G or U
    int32_t lA, lB;        // 2 per position i: [begin, end) of the arcs with
                           // left end i (contiguous in index order)
    int32_t rpA, rpB;      // n+2 / m+2: CSR start of right end i in radj
    int32_t radjA, radjB;  // arcs by right end ascending, left end descending
    int32_t band;          // min_col[0..n], max_col[0..n]
    int32_t tab;           // 256 ints: riboX term per pair of base-pair types
    int32_t tau, ribosum;  // --tau; RIBOSUM scoring (else sigma sums)
    int32_t dlen, dend;    // max-diff-am, max-diff-at-am (>= lengths: off)
    int32_t ws_bits;       // NA * W64 uint64: valid arc pairs, row a
    int32_t ws_wpre;       // NA * W64: index of the first arc match of word
    int32_t ws_rowcnt;     // NA: arc matches of arc a of A
    int32_t ws_sblk;       // NA + 1: fill blocks before radj position u
    int32_t ws_blka;       // NA: arc of A of fill block b
    int32_t ws_blks;       // NA: first fill entry of block b
    // second workspace (sized after the counting pass)
    int32_t ws_eof;        // nam: candidate entry of arc match k
    int32_t ws_fof;        // nam: fill entry of arc match k
    int32_t ws_sched;      // 3 * ngroups + K * (n + 2): schedule scratch
    int32_t hdr, hdr_len;  // head of the pair's image (band, codes, gap
                           // scores, sigma), copied to the arena by pass 2
    int32_t pad[2];
} CompactDesc;  // 36 ints

// Per-pair results of the counting pass (and status of the build pass).
#define LCG_CS_NAM 0
#define LCG_CS_NGROUPS 1
#define LCG_CS_NBLK 2      // arcs of A with at least one arc match
#define LCG_CS_GMAX 3      // max LDS ints of a group region
#define LCG_CS_ECAP 4      // max columns of a group region
#define LCG_CS_CELLS 5     // group DP cells, uint64 in [5], [6]
#define LCG_CS_STATUS 7    // 0 ok; 1 a row needs 4096 LDS rounds or more
#define LCG_CS_INTS 8

typedef struct {
    int32_t npairs;
    int32_t max_len;   // LDS: band of the longest sequence
    int32_t waves;     // LDS schedule: regions per round
    int32_t pool;      // LDS schedule: pool ints
    int32_t members;   // LDS schedule: workgroups per pair
    int32_t pad[3];
} CompactParams;
