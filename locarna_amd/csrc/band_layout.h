// Device layout of the trace-band kernel (band_kernels.hip), shared with the
// host code that prepares it (batch.cpp).
//
// Per pair: an int arena block with the symbol codes, the initial band and
// the output rows, and a double arena block with the structure profiles of
// the similarity (StralScore, stral_score.cc:29-60).  The forward values of
// all pairs live in one double scratch buffer.
#ifndef LOCARNA_AMD_BAND_LAYOUT_H
#define LOCARNA_AMD_BAND_LAYOUT_H

#include <stdint.h>

// output flags of a pair (BandDesc::rows + 2 (n + 1))
#define LCG_BAND_AMBIGUOUS 1  // a trace probability within eps (relative) of the cut, or not finite
#define LCG_BAND_RANGE 2      // a value outside the range of the host precision

// Largest sequence length the kernel takes (LDS: about 136 bytes per
// position); longer pairs are banded on the host.
#define LCG_BAND_MAX_LEN 1024

struct BandDesc {
    int32_t n, m;
    int32_t codes;  // ints: codeA[0..n], codeB[0..m]; -1 = not one of ACGU
    int32_t band;   // ints: min_col[0..n], max_col[0..n] of the initial band
    int32_t rows;   // ints, output: lo[0..n], hi[0..n], flags
    int32_t prof;   // doubles: upA[0..n], dnA[0..n], upB[0..m], dnB[0..m]
    int64_t fwd;    // doubles of the scratch: 3 (n + 1)(m + 1) forward values (M, A, B per cell)
};

struct BandParams {
    double temp;     // temperature_alipf / 100
    double g_open;   // exp(indel_opening / 100 / temp)
    double g_ext;    // exp(indel / 100 / temp)
    double sw;       // struct_weight / 100
    double thr;      // min_trace_prob
    double eps;      // relative distance to thr below which a cell is ambiguous
    double sim[16];  // base similarity of ACGU x ACGU
    int32_t npairs;
    int32_t local;     // sequence-local partition function
    int32_t extended;  // host precision long double (1) or double (0)
    int32_t lmax;      // largest length in the batch (LDS arrays)
    int32_t free_endgaps;  // LCG_FREE_* bits (forward orientation)
    int32_t pad;
};

#endif
