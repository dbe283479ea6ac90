// Device layout of the LocARNA-P (FP64) data of a batch: a double arena
// next to the int32 arena of device_layout.h (whose band, groups and
// candidate lists the PF kernels reuse).
//
// Per pair (offsets in doubles):
//   egA[n+1], egB[m+1]  exp(gap score / T)
//   es[64]              exp(sigma / T) by symbol codes
//   eam[nam]            exp(s(am) / T) per candidate entry
//   Dc[nam], Dbar[nam]  inside D and its adjoint dZ/dD per candidate entry
//   edge_full[ncells]   weight of match edges into each band cell (all groups)
//   edge_ref[ncells]    the same restricted to the reference's base-match
//                       probability cases (top level and arc matches with
//                       P(am) > sqrt(min_am_prob))
//   scal[4]             exp(indel_opening / T), pf_scale, sqrt(min_am_prob)
//   out[4]              Z, expected-openings weight
//   fD[nam]             inside D per fill entry (the fill list of the int
//                       arena, for the candidate gathers of the group kernels)
//   frag[8 cap]         fragment match queries, 8 doubles each: i, j, k, l,
//                       W (the stored row width of the block's region; 0:
//                       a corner outside the band, probability 0), in (the
//                       block's inside), out (the weight of its outside),
//                       unused
//
// Per pair in the int schedule buffer (offsets in ints): per group 4 ints
// (inside round | wave<<12, inside LDS offset in doubles, outside round |
// wave<<12, outside LDS offset; offset -1: the group runs in global
// scratch), the number of rounds of every row for both passes, and per
// group the offset of its forward values in the store (units of 4 doubles;
// -1: none, the outside pass recomputes them): E and s = M - F of every
// computed cell and the region's M rows, (ar - al) W doubles each, written
// by the inside pass and read by the outside pass instead of a second
// forward fill.
#pragma once
#include <stdint.h>

typedef struct {
    int32_t egA, egB, es, eam;
    int32_t Dc, Dbar, edge_full, edge_ref;
    int32_t scal, out;
    int32_t fD;
    int32_t sched, rin, rout;  // int schedule buffer
    int32_t frag, nfrag;       // fragment match queries of the run
    int32_t fso;               // int schedule buffer: forward-store offset per group
} PfDesc;  // 17 ints

typedef struct {
    int32_t npairs;
    int32_t max_len;
    int32_t scratch_stride;  // doubles of global scratch per (workgroup, wave)
    int32_t pool;            // doubles of LDS pool for group regions
    int32_t nfblk_max;       // fill blocks of the largest pair (LDS header)
    int32_t wg_per_pair;     // workgroups sharing one of the first nstatic pairs
    int32_t nstatic;         // pairs (in queue order) with wg_per_pair
                             // workgroups each, assigned by block index; the
                             // other workgroups take the rest from the queue
    int32_t spin_limit;      // polls of a bounded cross-workgroup wait (0: default,
                             // -1: every wait gives up at once, for error-path tests)
} PfBatchParams;

// score parameters theta: sigma 4x4, arc-match term 16x16, indel,
// indel opening, struct weight
#define LCG_NTHETA (16 + 256 + 3)
