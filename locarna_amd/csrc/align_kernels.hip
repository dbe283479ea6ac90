// gfx950 kernels of the integer LocARNA alignment (max-plus DP, int32).
//
// A workgroup of NW wave64s aligns one pair at a time, pulling pairs from a
// work queue (persistent grid, longest pairs first).  Per pair:
//   1. fill D: rows of left ends al = n .. 1.  The arc-match groups (al, bl)
//      of one row are independent; the host packs them into the workgroup's
//      LDS pool (rounds of at most NW groups) and each wave computes its
//      group's matrix M row by row, lanes over columns (reference:
//      AlignerImpl::align_D / align_in_arcmatch / align_noex /
//      fill_D_entries, src/LocARNA/aligner.cc:153-233, 264-437, 574-732);
//   2. the top level (global or sequence-local, aligner.cc:736-880);
//   3. the traceback (aligner.cc:967-1358), on the device: one wave traces
//      the top level, every traced arc match is re-aligned and traced by a
//      free wave (breadth first).  Each edge is written to slot i+j of the
//      cell it reaches, which orders the edges without a sort.
//
// Row recurrence of a group (origin (al, bl)), lane l owns column j:
//   s(i,j) = max(M(i-1,j-1) + sigma(i,j), E(i,j), S(i,j))
//   S(i,j) = max over arc matches (p..i, q..j), p > al, q > bl, of
//            M(p-1,q-1) + D(p..i, q..j)
//   E(i,j) = max(E(i-1,j) + gA(i), M(i-1,j) + gA(i) + alpha)
//   F(i,j) = max(F(i,j-1) + gB(j) + max(0,alpha), s(i,j-1) + gB(j) + alpha)
//   M(i,j) = max(s(i,j), F(i,j))
// S is evaluated candidate-parallel: the arc matches ending in row i within
// the group's columns are one contiguous run of the candidate list; lanes
// stride over it and combine with LDS atomic max.  The F chain is a max-plus
// linear recurrence along the row, solved by a prefix sum and a prefix max
// across the wave (DPP row_shr / row_bcast):
//   F(j) = P(j) + max_{k<=j} (c(k) - P(k)),  P = inclusive sum of a.
// Scores are int32; minus infinity is any value below LCG_NEG_LIMIT = -2^28.
// The general fill normalizes such values to LCG_NEG = -2^29, the
// fixed-column fill bounds them below at LCG_NEG (one max instead of a
// compare and select); neither changes a finite result.
#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>

#include "device_layout.h"

namespace {

__device__ __forceinline__ int clampneg(int x) { return x < LCG_NEG_LIMIT ? LCG_NEG : x; }

template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWMASK, 0xf, false);
}

// inclusive prefix sum over the 64 lanes
__device__ __forceinline__ int wave_scan_add(int v) {
    v += dpp<0x111, 0xf>(0, v);  // row_shr:1
    v += dpp<0x112, 0xf>(0, v);  // row_shr:2
    v += dpp<0x114, 0xf>(0, v);  // row_shr:4
    v += dpp<0x118, 0xf>(0, v);  // row_shr:8
    v += dpp<0x142, 0xa>(0, v);  // row_bcast:15 -> rows 1, 3
    v += dpp<0x143, 0xc>(0, v);  // row_bcast:31 -> rows 2, 3
    return v;
}

// inclusive prefix max over the 64 lanes.  With INT_MIN (the identity of
// max) as the value of lanes without a source, each step compiles to one
// v_max_i32_dpp instead of a constant move, a DPP move and a max.
__device__ __forceinline__ int wave_scan_max(int v) {
    const int id = INT_MIN;
    v = max(v, dpp<0x111, 0xf>(id, v));
    v = max(v, dpp<0x112, 0xf>(id, v));
    v = max(v, dpp<0x114, 0xf>(id, v));
    v = max(v, dpp<0x118, 0xf>(id, v));
    v = max(v, dpp<0x142, 0xa>(id, v));
    v = max(v, dpp<0x143, 0xc>(id, v));
    return v;
}

// lane l receives lane l-1's value; lane 0 receives `first`
__device__ __forceinline__ int wave_shift_up(int v, int first) {
    return __builtin_amdgcn_update_dpp(first, v, 0x138, 0xf, 0xf, false);  // wave_shr:1
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Orders the memory accesses of one wave to a region it owns: LDS
// instructions of a wave complete in order, so only the compiler needs a
// barrier; a region in global scratch belongs to one wave of the workgroup,
// whose vector L1 (write-through, one per CU) serves all its accesses, so a
// workgroup-scope fence (wait for the outstanding accesses) suffices.
template <bool GLOBAL>
__device__ __forceinline__ void region_sync() {
    if (GLOBAL)
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    else
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
}

struct Pair {
    const int *__restrict__ mn;     // LDS: min_col[0..n]
    const int *__restrict__ mx;     // LDS: max_col[0..n]
    const int *__restrict__ coff;   // LDS: cell_off[0..n+1]
    const int *__restrict__ codeA;  // LDS
    const int *__restrict__ codeB;  // LDS
    const int *__restrict__ gapA;   // LDS
    const int *__restrict__ gapB;   // LDS
    const int *__restrict__ sig;    // LDS: 8x8 sigma
    const int *__restrict__ gsig;   // global: (n+1) x (m+1) sigma of a pair of profiles
    const int *__restrict__ cptr;   // global
    int4 *cand;                     // global: (al|bl<<16, j, D, min_col(al-1))
    const int *fbp, *fbk;           // LDS: fill blocks of row i, (p, first entry)
    int *fcand;                     // global: fill entries, 3 ints each
    const int *__restrict__ amscore;
    int *D;
    int *edges;
    int *structs;
    const int *inner;  // --noLP / --stacking: arc match -> arc match directly inside it (-1: none)
    const int *c2a;    // --noLP / --stacking: candidate entry -> arc match
    const int *amst;   // --stacking: stacked score per arc match, LCG_NEG if not stackable
    int n, m, nam;
    int alpha, alpha_pos;
    int nolp, stacking;
    int struct_local, exclusion;  // --struct-local, --exclusion
};

// Base-match score sigma(i, j): by symbol codes from the LDS table, or (TAB,
// pairs of alignment profiles) from the pair's column-pair table in global
// memory, row i at gsig + i * (m + 1).  sig_row gives the row, sig_col the
// index of column j in it.
template <bool TAB>
__device__ __forceinline__ const int *sig_row(const Pair &P, int i) {
    return TAB ? P.gsig + i * (P.m + 1) : P.sig + P.codeA[i] * LCG_NCODES;
}
template <bool TAB>
__device__ __forceinline__ int sig_col(const Pair &P, int j) {
    return TAB ? j : P.codeB[j];
}

// Row i's stored column range in the group with origin (al, bl) and
// column bound br (exclusive): [lo, hi].
__device__ __forceinline__ int row_lo(const Pair &P, int i, int bl) { return max(bl, P.mn[i]); }
__device__ __forceinline__ int row_hi(const Pair &P, int i, int br) { return min(br - 1, P.mx[i]); }

// fill-list entry: (p | q<<16, j | min_col(p-1)<<16, D)
struct FillEntry {
    int pq, jm, D;
};

struct TopBest {
    int score, i, j;
};

// Fills the matrix of the group with origin (al, bl), rows al..ar-1, columns
// bl..br-1 (within the band), into the region at `base` (layout E | S | M,
// M row r = i - al at M + r*W, column j at j - row_lo(i)).
// GLOBAL: the region is global scratch; TOP_LOCAL: the sequence-local top
// level (values clamped at 0, first maximum tracked).
template <bool TAB, bool GLOBAL, bool TOP_LOCAL>
__device__ __forceinline__ void group_fill(const Pair &P, int al, int bl, int ar, int br, int W, int *base, bool global_col,
                           bool global_row, TopBest *best, unsigned long long *counts, int lam = 0) {
    const int lane = lane_id();
    const int alpha = P.alpha;
    int *Est = base;
    int *Srow = base + (br - bl);
    int *Mst = Srow + W;
    unsigned cnt_cells = 0, cnt_cand = 0;
    for (int c = lane; c < br - bl; c += 64) Est[c] = LCG_NEG;
    // row al: M(al, bl) = 0, M(al, j) = alpha + sum gapB (global_row) or 0;
    // column bl likewise with global_col (init_state, aligner.cc:264-368)
    const int hi0 = row_hi(P, al, br);
    {
        int carry = global_row ? alpha : 0;
        for (int j0 = bl + 1; j0 <= hi0; j0 += 64) {
            const int j = j0 + lane;
            const bool act = j <= hi0;
            const int g = (act && global_row) ? P.gapB[j] - lam : 0;
            const int ps = wave_scan_add(g);
            if (act) Mst[j - bl] = carry + ps;
            carry += readlane(ps, 63);
        }
        if (lane == 0) Mst[0] = 0;
    }
    region_sync<GLOBAL>();
    int colv = global_col ? alpha : 0;
    bool colopen = true;
    int lo_p = bl, hi_p = hi0;
    // M(i-1, bl) of the init column, also where row i-1 stores no column bl
    // (a restricted top level whose origin lies right of the row's band;
    // the reference's init column runs while min_col <= bl,
    // aligner.cc:264-368); row al: the origin
    int cprev = 0;
    bool cprev_ok = true;
    // Candidate entries of row i are fetched one row ahead: the loads only
    // need the band, the M values they index are read after row i-1 is done.
    auto row_range = [&](int r, int &t0, int &t1) {
        const int jl = max(bl + 1, P.mn[r]), jh = min(br - 1, P.mx[r]);
        if (jl > jh) {
            t0 = t1 = 0;
            return;
        }
        const int cb = P.coff[r] - P.mn[r];
        t0 = P.cptr[cb + jl];
        t1 = P.cptr[cb + jh + 1];
    };
    int nt0 = 0, nt1 = 0;
    int4 ne = make_int4(0, 0, 0, 0);
    if (al + 1 < ar) {
        row_range(al + 1, nt0, nt1);
        if (nt0 + lane < nt1) ne = P.cand[nt0 + lane];
    }
    for (int i = al + 1; i < ar; ++i) {
        const int r = i - al;
        int *row = Mst + r * W;
        const int *prow = row - W;
        const int mni = P.mn[i];
        if (colopen && mni <= bl) {
            if (global_col) colv += P.gapA[i] - lam;
        } else {
            colopen = false;
        }
        const int lo = max(bl, mni), hi = min(br - 1, P.mx[i]);
        if (colopen && lane == 0) row[0] = colv;
        const int jlo = max(bl + 1, mni);
        const int t0 = nt0, t1 = nt1;
        const int4 e0 = ne;
        if (i + 1 < ar) {
            row_range(i + 1, nt0, nt1);
            ne = (nt0 + lane < nt1) ? P.cand[nt0 + lane] : make_int4(0, 0, 0, 0);
        }
        if (jlo <= hi) {
            // arc candidates of the row, candidate-parallel
            for (int c = lane; c <= hi - jlo; c += 64) Srow[c] = LCG_NEG;
            region_sync<GLOBAL>();
            for (int t = t0 + lane; t < t1; t += 64) {
                const int4 e = (t - t0 < 64) ? e0 : P.cand[t];
                const int p = e.x & 0xffff, q = e.x >> 16;
                // first stored column of row p-1: the region's first row is
                // stored from bl, which lies left of the band for a
                // restricted top level whose origin the band excludes
                const int rl = (p - 1 == al) ? bl : max(bl, e.w);
                if (p > al && q > bl && q - 1 >= rl) {
                    // top-level position penalty: D - lam (arc lengths), aligner_impl.hh:249-253
                    const int v = Mst[(p - 1 - al) * W + (q - 1 - rl)] + e.z - lam * ((i - p + 1) + (e.y - q + 1));
                    atomicMax(&Srow[e.y - jlo], v);
                    ++cnt_cand;
                }
            }
            region_sync<GLOBAL>();
            int Fcarry = LCG_NEG;
            int sprev = colopen ? colv : LCG_NEG;  // M(i, jlo - 1)
            const int gA = P.gapA[i] - lam;
            const int *sigrow = sig_row<TAB>(P, i);
            for (int j0 = jlo; j0 <= hi; j0 += 64) {
                const int j = j0 + lane;
                const bool act = j <= hi;
                int s = LCG_NEG, gB = 0;
                if (act) {
                    const int up = (j >= lo_p && j <= hi_p) ? prow[j - lo_p] : LCG_NEG;
                    const int dg = (j - 1 >= lo_p && j - 1 <= hi_p) ? prow[j - 1 - lo_p]
                                   : (j - 1 == bl && cprev_ok)      ? cprev
                                                                    : LCG_NEG;
                    const int e = clampneg(max(Est[j - bl] + gA, up + gA + alpha));
                    Est[j - bl] = e;
                    s = clampneg(max(max(dg + sigrow[sig_col<TAB>(P, j)] - 2 * lam, e), Srow[j - jlo]));
                    if (TOP_LOCAL) s = max(s, 0);
                    gB = P.gapB[j] - lam;
                    ++cnt_cells;
                }
                // F chain
                const int sl = wave_shift_up(s, sprev);
                const int a = act ? gB + P.alpha_pos : 0;
                const int c = act ? sl + gB + alpha : LCG_NEG;
                const int Pa = wave_scan_add(a);
                const int px = wave_scan_max(act ? c - Pa : LCG_NEG);
                const int F = Pa + max(Fcarry, px);
                const int Mv = clampneg(max(s, F));
                if (act) row[j - lo] = Mv;
                if (TOP_LOCAL) {
                    // first strict maximum in row-major order (aligner.cc:870-875)
                    int wmax = act ? Mv : -1;
                    for (int o = 32; o >= 1; o >>= 1) wmax = max(wmax, __shfl_xor(wmax, o));
                    if (wmax > best->score) {
                        const unsigned long long bal = __ballot(act && Mv == wmax);
                        best->score = wmax;
                        best->i = i;
                        best->j = j0 + (int)__ffsll((long long)bal) - 1;
                    }
                }
                Fcarry = clampneg(readlane(F, 63));
                sprev = readlane(s, 63);
            }
        }
        lo_p = lo;
        hi_p = hi;
        cprev = colv;
        cprev_ok = colopen;
        region_sync<GLOBAL>();
    }
    if (counts) {
        atomicAdd(&counts[0], (unsigned long long)cnt_cells);
        atomicAdd(&counts[1], (unsigned long long)cnt_cand);
    }
}

// Fill of an arc-match group whose columns bl+1 .. br-1 fit NC chunks of 64
// lanes and whose rows al+1 .. ar-1 fit 64*RT entries of a lane-indexed row
// table.  Same recurrence and region layout as group_fill (M rows at
// base + (br-bl) + W), but lane l of chunk c owns column bl+1+64c+l for all
// rows: the previous row's M, the column's E, gap score and symbol code stay
// in registers, so a row costs one LDS read (the arc term S) besides the
// stores.  S is kept at base + (j - bl - 1) and reset to minus infinity as it
// is read.  The arc candidates come from the fill list (arc matches by right
// end in A, left end p descending): a row's candidates with p > al are one
// prefix, located from the LDS block table once per group; the first 256
// entries of the next row are fetched while the current row is computed.
template <bool TAB, bool GLOBAL, int NC, int RT, bool COUNT>
__device__ void group_fill_fixed(const Pair &P, int al, int bl, int ar, int br, int W, int *base,
                                 unsigned long long *counts) {
    const int lane = lane_id();
    const int alpha = P.alpha;
    int *Srow = base;
    int *Mst = base + (br - bl) + W;
    const int nrows = ar - al - 1;
    unsigned cnt_cells = 0, cnt_cand = 0;
    // fill-list range [rt0, rt1) of row al+1+r (the blocks with left end
    // p > al), lane r % 64 of table r / 64
    int rt0[RT], rt1[RT];
#pragma unroll
    for (int h = 0; h < RT; ++h) {
        const int r = lane + 64 * h;
        rt0[h] = rt1[h] = 0;
        if (r < nrows) {
            const int i = al + 1 + r;
            if (max(bl + 1, P.mn[i]) <= min(br - 1, P.mx[i])) {
                int b = P.fbp[i];
                const int be = P.fbp[i + 1];
                rt0[h] = P.fbk[2 * b + 1];
                while (b < be && P.fbk[2 * b] > al) ++b;
                rt1[h] = P.fbk[2 * b + 1];
            }
        }
    }
    auto row_cands = [&](int r, int &t0, int &t1) {
        t0 = t1 = 0;
#pragma unroll
        for (int h = 0; h < RT; ++h)
            if ((r >> 6) == h) {
                t0 = readlane(rt0[h], r & 63);
                t1 = readlane(rt1[h], r & 63);
            }
    };
    const FillEntry none{0, 0, 0};
    // one unpredicated 12-byte load per lane: lanes past the row's range
    // read its last entry (or entry 0 of the pair; a group implies arc
    // matches), and arc_slot ignores them through its lane flag
    auto fetch = [&](int t, int t1) {
        const int tt = t < t1 ? t : max(t1 - 1, 0);
        return *reinterpret_cast<const FillEntry *>(P.fcand + 3 * tt);
    };
    for (int c = lane; c < br - bl; c += 64) Srow[c] = LCG_NEG;
    // row al (init row): M(al, bl) = 0, M(al, j) = alpha + sum gapB
    // Pre: prefix over the chunk's columns of gB(j) + max(0, alpha), the
    // slope of the F chain (fixed per column; a row's active columns are one
    // range, whose prefix is Pre minus Pre before the range)
    int Mp[NC], Ec[NC], gBc[NC], cBc[NC], Pre[NC];
    const int hi0 = row_hi(P, al, br);
    {
        int carry = alpha;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int j = bl + 1 + 64 * c + lane;
            const bool in = j <= hi0;
            const bool col = j <= br - 1;
            const int ps = wave_scan_add(in ? P.gapB[j] : 0);
            Mp[c] = in ? carry + ps : LCG_NEG;
            if (in) Mst[j - bl] = carry + ps;
            carry += readlane(ps, 63);
            Ec[c] = LCG_NEG;
            gBc[c] = col ? P.gapB[j] : 0;
            cBc[c] = col ? sig_col<TAB>(P, j) : 0;
            Pre[c] = wave_scan_add(col ? gBc[c] + P.alpha_pos : 0);
        }
        if (lane == 0) Mst[0] = 0;
    }
    region_sync<GLOBAL>();
    int colv = alpha;
    bool colopen = true;
    int Mcol_prev = 0;  // M(i-1, bl)
    // entries of the current row, 4 chunks of 64 (fetched during the
    // previous row); entries beyond 256 are loaded when used
    FillEntry cur[4];
    {
        int u0 = 0, u1 = 0;
        if (nrows > 0) row_cands(0, u0, u1);
#pragma unroll
        for (int k = 0; k < 4; ++k) cur[k] = fetch(u0 + 64 * k + lane, u1);
    }
    for (int i = al + 1; i < ar; ++i) {
        const int r = i - al - 1;
        const int mni = P.mn[i];
        int Mcol = LCG_NEG;  // M(i, bl)
        if (colopen && mni <= bl) {
            colv += P.gapA[i];
            Mcol = colv;
        } else {
            colopen = false;
        }
        const int lo = max(bl, mni), hi = min(br - 1, P.mx[i]);
        int *row = Mst + (r + 1) * W;
        if (colopen && lane == 0) row[0] = colv;
        const int jlo = max(bl + 1, mni);
        int t0, t1, u0 = 0, u1 = 0;
        row_cands(r, t0, t1);
        if (r + 1 < nrows) row_cands(r + 1, u0, u1);
        // the next row's first chunks (the second only if it has more than 64)
        const FillEntry nx0 = fetch(u0 + lane, u1);
        const FillEntry nx1 = (u1 - u0 > 64) ? fetch(u0 + 64 + lane, u1) : none;
        // arc terms, candidate-parallel over the row's prefix p > al: the
        // M gathers of all chunks are issued before the LDS atomics (which
        // the compiler may not reorder them across)
        auto arc_slot = [&](const FillEntry &e, bool in, int &v) {
            const int p = e.pq & 0xffff, q = e.pq >> 16;
            const int j = e.jm & 0xffff;
            const int rl = max(bl, e.jm >> 16);  // first stored column of row p-1
            const bool ok = in && q > bl && j >= jlo && j <= hi && q - 1 >= rl;
            // branch-free gather (cell 0 of the region when unused), so that
            // the gathers of several chunks are in flight together
            v = Mst[ok ? __mul24(p - 1 - al, W) + (q - 1 - rl) : 0] + e.D;
            if (COUNT) cnt_cand += ok;
            return ok ? j - bl - 1 : -1;
        };
        {
            int v[4], slot[4];
            // rows of at most 64 candidates (most) use chunk 0 only, rows of
            // at most 128 skip chunks 2 and 3
            if (t1 - t0 <= 64) {
                slot[0] = arc_slot(cur[0], t0 + lane < t1, v[0]);
                if (slot[0] >= 0) atomicMax(&Srow[slot[0]], v[0]);
            } else if (t1 - t0 <= 128) {
#pragma unroll
                for (int k = 0; k < 2; ++k) slot[k] = arc_slot(cur[k], t0 + 64 * k + lane < t1, v[k]);
#pragma unroll
                for (int k = 0; k < 2; ++k)
                    if (slot[k] >= 0) atomicMax(&Srow[slot[k]], v[k]);
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) slot[k] = arc_slot(cur[k], t0 + 64 * k + lane < t1, v[k]);
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (slot[k] >= 0) atomicMax(&Srow[slot[k]], v[k]);
            }
        }
        for (int t = t0 + 256 + lane; t < t1; t += 64) {
            int v;
            const int sl = arc_slot(fetch(t, t1), true, v);
            if (sl >= 0) atomicMax(&Srow[sl], v);
        }
        if (u1 - u0 > 128) {
            cur[2] = fetch(u0 + 128 + lane, u1);
            cur[3] = fetch(u0 + 192 + lane, u1);
        }
        region_sync<GLOBAL>();
        int Fcarry = LCG_NEG;
        int sprev = Mcol;      // s(i, bl) entering the F chain: M(i, bl)
        int dgin = Mcol_prev;  // M(i-1, bl)
        const int gA = P.gapA[i];
        const int *sigrow = sig_row<TAB>(P, i);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int j = bl + 1 + 64 * c + lane;
            const bool act = j >= jlo && j <= hi;
            const int up = Mp[c];
            const int dg = wave_shift_up(Mp[c], dgin);
            dgin = readlane(Mp[c], 63);
            // both LDS reads unconditional (slot 0 when inactive), issued together
            const int sv = Srow[act ? j - bl - 1 : 0];
            const int sg = sigrow[cBc[c]];
            // minus infinity: E is bounded below by LCG_NEG, so s and M
            // (maxima including E) need no clamp; a minus-infinity value
            // stays below LCG_NEG_LIMIT (it gains at most the positive
            // scores of one path, far less than 2^28)
            int s = LCG_NEG;
            if (act) {
                const int e = max(max(Ec[c] + gA, up + gA + alpha), LCG_NEG);
                Ec[c] = e;
                Srow[j - bl - 1] = LCG_NEG;
                s = max(max(dg + sg, e), sv);
                if (COUNT) ++cnt_cells;
            }
            const int sl = wave_shift_up(s, sprev);
            const int cc = act ? sl + gBc[c] + alpha : LCG_NEG;
            // F(j) = Pa(j) + max(Fcarry, max_k (cc(k) - Pa(k))) with Pa the
            // prefix from the first active column js; Pa = Pre - Pre(js - 1)
            const int x = max(jlo, bl + 1 + 64 * c) - (bl + 1 + 64 * c) - 1;  // lane before js
            const int C = x < 0 ? 0 : readlane(Pre[c], min(x, 63));
            const int px = wave_scan_max(act ? cc - Pre[c] : LCG_NEG);
            const int F = Pre[c] + max(Fcarry - C, px);
            const int Mv = max(s, F);
            if (act) row[j - lo] = Mv;
            Mp[c] = act ? Mv : LCG_NEG;
            Fcarry = max(readlane(F, 63), LCG_NEG);
            sprev = readlane(s, 63);
        }
        cur[0] = nx0;
        cur[1] = nx1;
        Mcol_prev = Mcol;
        region_sync<GLOBAL>();
    }
    if (COUNT && counts) {
        atomicAdd(&counts[0], (unsigned long long)cnt_cells);
        atomicAdd(&counts[1], (unsigned long long)cnt_cand);
    }
}

// Fill of an arc-match group (fill-D and trace re-fills), by the fixed-column
// kernel when the group fits it, else by the general one.  WIDE enables the
// 4-chunk variant (columns up to 256), left out of the traceback kernel for
// its register budget.  COUNT keeps the per-lane work counters of the
// fixed-column kernel (the counting run); without it they cost two vector
// adds per row chunk and candidate slot.
template <bool TAB, bool GLOBAL, bool WIDE = true, bool COUNT = false>
__device__ void fill_group(const Pair &P, int al, int bl, int ar, int br, int W, int *base,
                           unsigned long long *counts, bool fixed_cols) {
    const int cols = br - bl - 1, rows = ar - al - 1;
    if (!fixed_cols)
        group_fill<TAB, GLOBAL, false>(P, al, bl, ar, br, W, base, true, true, nullptr, counts);
    else if (cols <= 64 && rows <= 128)
        group_fill_fixed<TAB, GLOBAL, 1, 2, COUNT>(P, al, bl, ar, br, W, base, counts);
    else if (cols <= 128 && rows <= 128)
        group_fill_fixed<TAB, GLOBAL, 2, 2, COUNT>(P, al, bl, ar, br, W, base, counts);
    else if (WIDE && cols <= 256 && rows <= 256)
        group_fill_fixed<TAB, GLOBAL, 4, 4, COUNT>(P, al, bl, ar, br, W, base, counts);
    else
        group_fill<TAB, GLOBAL, false>(P, al, bl, ar, br, W, base, true, true, nullptr, counts);
}

// stored value of cell (i, j) of the group (origin al, bl; bound br)
__device__ __forceinline__ int Mget(const Pair &P, const int *Mst, int W, int al, int bl, int br, int i, int j) {
    const int lo = (i == al) ? bl : row_lo(P, i, bl);
    const int hi = row_hi(P, i, br);
    if (j < lo || j > hi) {
        // the init column where the row's band ends left of it (a restricted
        // top level): group_fill keeps its value in the row's first slot
        if (j == bl && i > al && hi < bl && P.mn[i] <= bl) return Mst[(i - al) * W];
        return LCG_NEG;
    }
    return Mst[(i - al) * W + (j - lo)];
}

__device__ __forceinline__ bool valid(const Pair &P, int i, int j) { return P.mn[i] <= j && j <= P.mx[i]; }

__device__ __forceinline__ void emit(const Pair &P, int a, int b, int slot) { P.edges[slot] = (a << 16) | b; }

// Traces one group back from (i, j) to its origin (al, bl) (aligner.cc:
// trace_in_arcmatch / trace_noex, state E_NO_NO).  Mst is the group's
// matrix.  Arc matches taken are appended to the work list (tstack, counter
// *tcount in LDS).
template <bool TAB, bool GLOBAL>
__device__ __forceinline__ void group_trace(const Pair &P, const int *Mst, int W, int al, int bl, int br, int i, int j, bool top,
                            bool local, int free_endgaps, int *tstack, int *tcount, int lam = 0) {
    const int lane = lane_id();
    region_sync<GLOBAL>();
    for (;;) {
        // a top-level cell outside the band was never written and holds the
        // matrix's initial 0 (InftyInt(), infty_int.hh:80, 348): the trace
        // can start there when free end gaps left (n, m) out of the band
        const int Mij = (top && i > al && j > bl && !valid(P, i, j)) ? 0 : Mget(P, Mst, W, al, bl, br, i, j);
        if (top && local && Mij == 0) break;
        // no padding of free leading gaps at the top level (aligner.cc:1260-1289)
        if (i <= al) {
            if (!(top && (local || (free_endgaps & 1))))
                for (int k = bl + 1 + lane; k <= j; k += 64) emit(P, 0, k, al + k);
            break;
        }
        if (j <= bl) {
            if (!(top && (local || (free_endgaps & 4))))
                for (int k = al + 1 + lane; k <= i; k += 64) emit(P, k, 0, k + bl);
            break;
        }
        // match
        if (valid(P, i - 1, j - 1) &&
            Mij == Mget(P, Mst, W, al, bl, br, i - 1, j - 1) + sig_row<TAB>(P, i)[sig_col<TAB>(P, j)] - 2 * lam) {
            if (lane == 0) emit(P, i, j, i + j);
            --i;
            --j;
            continue;
        }
        bool done = false;
        if (P.alpha != 0) {
            // deletion of length k = 1, 2, ...: lanes test k = base + lane;
            // the reference stops at the first k outside the band
            int carry = P.alpha;
            for (int base = 1; !done && i - base >= al; base += 64) {
                const int k = base + lane;
                const bool inr = (i - k >= al) && valid(P, i - k, j);
                const int g = inr ? P.gapA[i - k + 1] - lam : 0;
                const int cost = carry + wave_scan_add(g);
                const bool hit = inr && Mij == Mget(P, Mst, W, al, bl, br, i - k, j) + cost;
                const unsigned long long bh = __ballot(hit);
                const unsigned long long nv = ~__ballot(inr);
                const unsigned long long valid_prefix = nv ? (nv & (0ull - nv)) - 1 : ~0ull;
                const unsigned long long h = bh & valid_prefix;
                if (h) {
                    const int kk = base + (int)__ffsll((long long)h) - 1;
                    for (int t = i - kk + 1 + lane; t <= i; t += 64) emit(P, t, 0, t + j);
                    i -= kk;
                    done = true;
                } else if (valid_prefix != ~0ull) {
                    break;
                }
                carry = readlane(cost, 63);
            }
            if (done) continue;
            carry = P.alpha;
            for (int base = 1; !done && j - base >= bl; base += 64) {
                const int k = base + lane;
                const bool inr = (j - k >= bl) && valid(P, i, j - k);
                const int g = inr ? P.gapB[j - k + 1] - lam : 0;
                const int cost = carry + wave_scan_add(g);
                const bool hit = inr && Mij == Mget(P, Mst, W, al, bl, br, i, j - k) + cost;
                const unsigned long long bh = __ballot(hit);
                const unsigned long long nv = ~__ballot(inr);
                const unsigned long long valid_prefix = nv ? (nv & (0ull - nv)) - 1 : ~0ull;
                const unsigned long long h = bh & valid_prefix;
                if (h) {
                    const int kk = base + (int)__ffsll((long long)h) - 1;
                    for (int t = j - kk + 1 + lane; t <= j; t += 64) emit(P, 0, t, i + t);
                    j -= kk;
                    done = true;
                } else if (valid_prefix != ~0ull) {
                    break;
                }
                carry = readlane(cost, 63);
            }
            if (done) continue;
        } else {
            if (valid(P, i - 1, j) && Mij == Mget(P, Mst, W, al, bl, br, i - 1, j) + P.gapA[i] - lam) {
                if (lane == 0) emit(P, i, 0, i + j);
                --i;
                continue;
            }
            if (valid(P, i, j - 1) && Mij == Mget(P, Mst, W, al, bl, br, i, j - 1) + P.gapB[j] - lam) {
                if (lane == 0) emit(P, 0, j, i + j);
                --j;
                continue;
            }
        }
        // no arc match ends outside the band (arc_matches.cc:19-48), where
        // only a top-level trace can start (free end gaps)
        if (!valid(P, i - 1, j - 1) || !valid(P, i, j)) break;
        // arc matches with right ends (i, j) in list order (al desc, bl desc)
        const int cell = P.coff[i] + (j - P.mn[i]);
        const int t0 = P.cptr[cell], t1 = P.cptr[cell + 1];
        bool found = false;
        int fp = 0, fq = 0, ft = 0;
        for (int base = t0; base < t1 && !found; base += 64) {
            const int t = base + lane;
            bool hit = false;
            int p = 0, q = 0;
            if (t < t1) {
                const int4 e = P.cand[t];
                p = e.x & 0xffff;
                q = e.x >> 16;
                hit = p > al && q > bl &&
                      Mij == Mget(P, Mst, W, al, bl, br, p - 1, q - 1) + e.z - lam * ((i - p + 1) + (j - q + 1));
            }
            const unsigned long long bh = __ballot(hit);
            if (bh) {
                const int src = (int)__ffsll((long long)bh) - 1;
                found = true;
                fp = readlane(p, src);
                fq = readlane(q, src);
                ft = readlane(t, src);
            }
        }
        if (!found) break;
        if (lane == 0) {
            emit(P, fp, fq, fp + fq);
            emit(P, i, j, i + j);
            P.structs[fp - 1] = 1;
            P.structs[i - 1] = 2;
            P.structs[P.n + fq - 1] = 1;
            P.structs[P.n + j - 1] = 2;
            int a = fp, b = fq, c = i, e = j;
            if (P.nolp) {
                // aligner.cc:1030-1078: the stack downwards; each inner arc
                // match continues it while D(x) == D(inner) + s(x), the last
                // one's region is traced as a task
                int x = P.c2a[ft];
                for (int depth = 0; depth < P.n && (unsigned)x < (unsigned)P.nam; ++depth) {
                    // each step nests deeper (at most n / 2 steps); a traced
                    // arc match always has an inner one, else its D is minus
                    // infinity
                    const int y = P.inner[x];
                    if ((unsigned)y >= (unsigned)P.nam) break;
                    ++a;
                    ++b;
                    --c;
                    --e;
                    emit(P, a, b, a + b);
                    emit(P, c, e, c + e);
                    P.structs[a - 1] = 1;
                    P.structs[c - 1] = 2;
                    P.structs[P.n + b - 1] = 1;
                    P.structs[P.n + e - 1] = 2;
                    if (P.D[x] != P.D[y] + (P.stacking ? P.amst[x] : P.amscore[x])) break;
                    x = y;
                }
            } else if (P.stacking) {
                // aligner.cc:967-1028: while the arc match's D is its inner
                // one's plus its stacked score, the inner arc match is traced
                // directly (only stackable arc matches have a stacked score)
                int x = P.c2a[ft];
                for (int depth = 0; depth < P.n && (unsigned)x < (unsigned)P.nam; ++depth) {
                    const int y = P.inner[x];
                    if ((unsigned)y >= (unsigned)P.nam) break;
                    const int st = P.amst[x];
                    if (st == LCG_NEG || P.D[x] != P.D[y] + st) break;
                    ++a;
                    ++b;
                    --c;
                    --e;
                    emit(P, a, b, a + b);
                    emit(P, c, e, c + e);
                    P.structs[a - 1] = 1;
                    P.structs[c - 1] = 2;
                    P.structs[P.n + b - 1] = 1;
                    P.structs[P.n + e - 1] = 2;
                    x = y;
                }
            }
            const int slot = atomicAdd(tcount, 1);
            tstack[2 * slot] = a | (b << 16);
            tstack[2 * slot + 1] = c | (e << 16);
        }
        i = fp - 1;
        j = fq - 1;
    }
}


// ---- structure-local alignment (--struct-local; aligner.cc:118-128,
// 264-565, 586-592, 1012-1025, 1233-1343).  Inside an arc match the
// alignment may exclude one block of A and one of B positions, each at the
// exclusion score: eight matrices, E = exclusion, NO = none, X = one
// (closed), OP = one open; the first part refers to A, the second to B.
// D takes the best non-open state at (ar-1, br-1); the top level is the
// plain recursion.  A group's region lives in global scratch (one region
// per wave): [E: br-bl] [S: W] [M of state 0 .. 7: (ar-al) W each], rows
// stored as in group_fill.
enum { SL_NO_NO, SL_X_NO, SL_NO_X, SL_X_X, SL_OP_NO, SL_NO_OP, SL_OP_X, SL_X_OP };

__device__ __forceinline__ int *sl_M(int *base, int al, int bl, int ar, int br, int W, int k) {
    return base + (br - bl) + W + k * (ar - al) * W;
}

// init_state (aligner.cc:264-368) of one state's matrix: origin 0; the
// first row and column by mode: 0 zero (an exclusion open from the origin,
// no global flag), 1 indel opening plus gap costs, 2 the exclusion score
// (the prefix excluded)
__device__ void sl_init(const Pair &P, int *Mst, int al, int bl, int ar, int br, int W, int colmode, int rowmode) {
    const int lane = lane_id();
    const int hi0 = row_hi(P, al, br);
    int carry = rowmode == 1 ? P.alpha : rowmode == 2 ? P.exclusion : 0;
    for (int j0 = bl + 1; j0 <= hi0; j0 += 64) {
        const int j = j0 + lane;
        const bool act = j <= hi0;
        const int ps = wave_scan_add((act && rowmode == 1) ? P.gapB[j] : 0);
        if (act) Mst[j - bl] = carry + ps;
        carry += readlane(ps, 63);
    }
    if (lane == 0) {
        Mst[0] = 0;
        int colv = colmode == 1 ? P.alpha : colmode == 2 ? P.exclusion : 0;
        for (int i = al + 1; i < ar && P.mn[i] <= bl; ++i) {
            if (colmode == 1) colv += P.gapA[i];
            Mst[(i - al) * W] = colv;  // rows reaching column bl are stored from bl
        }
    }
}

// value of the init column at row i (its mode), valid where mn(i) <= bl
__device__ __forceinline__ int sl_colv(const Pair &P, int al, int i, int colmode) {
    if (colmode == 0) return 0;
    if (colmode == 2) return P.exclusion;
    int v = P.alpha;
    for (int x = al + 1; x <= i; ++x) v += P.gapA[x];
    return v;
}

// One state k < 4: the recursion of group_fill over M_k (align_noex,
// aligner.cc:153-233), raised to X1 + exclusion and X2 + exclusion where
// given (the closing of an open exclusion, aligner.cc:476-563)
template <bool TAB>
__device__ void sl_fill_state(const Pair &P, int al, int bl, int ar, int br, int W, int *base, int k, int colmode,
                              int rowmode, const int *X1, const int *X2) {
    const int lane = lane_id();
    const int alpha = P.alpha, ex = P.exclusion;
    int *Est = base;
    int *Srow = base + (br - bl);
    int *Mst = sl_M(base, al, bl, ar, br, W, k);
    for (int c = lane; c < br - bl; c += 64) Est[c] = LCG_NEG;
    sl_init(P, Mst, al, bl, ar, br, W, colmode, rowmode);
    region_sync<true>();
    const int hi0 = row_hi(P, al, br);
    int lo_p = bl, hi_p = hi0;
    int colv = sl_colv(P, al, al, colmode);
    bool colopen = true;
    int cprev = 0;  // M(i-1, bl) where row i-1 reaches column bl; row al: the origin
    bool cprev_ok = true;
    for (int i = al + 1; i < ar; ++i) {
        const int r = i - al;
        int *row = Mst + r * W;
        const int *prow = row - W;
        const int mni = P.mn[i];
        if (colopen && mni <= bl) {
            if (colmode == 1) colv += P.gapA[i];
        } else {
            colopen = false;
        }
        const int lo = max(bl, mni), hi = min(br - 1, P.mx[i]);
        const int jlo = max(bl + 1, mni);
        if (jlo <= hi) {
            for (int c = lane; c <= hi - jlo; c += 64) Srow[c] = LCG_NEG;
            region_sync<true>();
            const int cb0 = P.coff[i] - P.mn[i];
            const int t0 = P.cptr[cb0 + jlo], t1 = P.cptr[cb0 + hi + 1];
            for (int t = t0 + lane; t < t1; t += 64) {
                const int4 e = P.cand[t];
                const int p = e.x & 0xffff, q = e.x >> 16;
                const int rl = (p - 1 == al) ? bl : max(bl, e.w);
                if (p > al && q > bl && q - 1 >= rl) atomicMax(&Srow[e.y - jlo], Mst[(p - 1 - al) * W + (q - 1 - rl)] + e.z);
            }
            region_sync<true>();
            int Fcarry = LCG_NEG;
            int sprev = colopen ? colv : LCG_NEG;  // M(i, jlo - 1)
            const int gA = P.gapA[i];
            const int *sigrow = sig_row<TAB>(P, i);
            for (int j0 = jlo; j0 <= hi; j0 += 64) {
                const int j = j0 + lane;
                const bool act = j <= hi;
                int s = LCG_NEG, gB = 0;
                if (act) {
                    const int up = (j >= lo_p && j <= hi_p) ? prow[j - lo_p] : LCG_NEG;
                    const int dg = (j - 1 >= lo_p && j - 1 <= hi_p) ? prow[j - 1 - lo_p]
                                   : (j - 1 == bl && cprev_ok)      ? cprev
                                                                    : LCG_NEG;
                    const int e = clampneg(max(Est[j - bl] + gA, up + gA + alpha));
                    Est[j - bl] = e;
                    s = clampneg(max(max(dg + sigrow[sig_col<TAB>(P, j)], e), Srow[j - jlo]));
                    if (X1) s = clampneg(max(s, X1[r * W + (j - lo)] + ex));
                    if (X2) s = clampneg(max(s, X2[r * W + (j - lo)] + ex));
                    gB = P.gapB[j];
                }
                const int sl = wave_shift_up(s, sprev);
                const int a = act ? gB + P.alpha_pos : 0;
                const int c = act ? sl + gB + alpha : LCG_NEG;
                const int Pa = wave_scan_add(a);
                const int px = wave_scan_max(act ? c - Pa : LCG_NEG);
                const int F = Pa + max(Fcarry, px);
                if (act) row[j - lo] = clampneg(max(s, F));
                Fcarry = clampneg(readlane(F, 63));
                sprev = readlane(s, 63);
            }
        }
        lo_p = lo;
        hi_p = hi;
        cprev = colv;
        cprev_ok = colopen;
        region_sync<true>();
    }
}

// An open state: M(i, j) = max(M(i-1, j), Src(i, j)) (vertical: an open
// exclusion of A, aligner.cc:443-460) or max(M(i, j-1), Src(i, j))
// (horizontal: of B, :462-479)
__device__ void sl_chain(const Pair &P, int al, int bl, int ar, int br, int W, int *Mst, const int *Src, bool vertical,
                         int colmode, int rowmode) {
    const int lane = lane_id();
    sl_init(P, Mst, al, bl, ar, br, W, colmode, rowmode);
    region_sync<true>();
    const int hi0 = row_hi(P, al, br);
    int lo_p = bl, hi_p = hi0;
    int colv = sl_colv(P, al, al, colmode);
    for (int i = al + 1; i < ar; ++i) {
        const int r = i - al;
        int *row = Mst + r * W;
        const int *prow = row - W;
        const int *srow = Src + r * W;
        const int mni = P.mn[i];
        if (mni <= bl && colmode == 1) colv += P.gapA[i];
        const int lo = max(bl, mni), hi = min(br - 1, P.mx[i]);
        const int jlo = max(bl + 1, mni);
        int carry = mni <= bl ? colv : LCG_NEG;  // M(i, jlo - 1)
        for (int j0 = jlo; j0 <= hi; j0 += 64) {
            const int j = j0 + lane;
            const bool act = j <= hi;
            const int sv = act ? srow[j - lo] : LCG_NEG;
            if (vertical) {
                const int up = (act && j >= lo_p && j <= hi_p) ? prow[j - lo_p] : LCG_NEG;
                if (act) row[j - lo] = max(up, sv);
            } else {
                const int v = max(carry, wave_scan_max(sv));
                if (act) row[j - lo] = v;
                carry = readlane(v, 63);
            }
        }
        lo_p = lo;
        hi_p = hi;
        region_sync<true>();
    }
}

// all eight states of a group, in the reference's order (aligner.cc:409-563)
template <bool TAB>
__device__ void group_fill_sl(const Pair &P, int al, int bl, int ar, int br, int W, int *base) {
    int *M0 = sl_M(base, al, bl, ar, br, W, SL_NO_NO);
    int *Mxn = sl_M(base, al, bl, ar, br, W, SL_X_NO), *Mnx = sl_M(base, al, bl, ar, br, W, SL_NO_X);
    int *Mon = sl_M(base, al, bl, ar, br, W, SL_OP_NO), *Mno = sl_M(base, al, bl, ar, br, W, SL_NO_OP);
    int *Mox = sl_M(base, al, bl, ar, br, W, SL_OP_X), *Mxo = sl_M(base, al, bl, ar, br, W, SL_X_OP);
    sl_fill_state<TAB>(P, al, bl, ar, br, W, base, SL_NO_NO, 1, 1, nullptr, nullptr);
    sl_chain(P, al, bl, ar, br, W, Mon, M0, true, 0, 1);
    sl_chain(P, al, bl, ar, br, W, Mno, M0, false, 1, 0);
    sl_fill_state<TAB>(P, al, bl, ar, br, W, base, SL_NO_X, 1, 2, Mno, nullptr);
    sl_chain(P, al, bl, ar, br, W, Mox, Mnx, true, 0, 2);
    sl_fill_state<TAB>(P, al, bl, ar, br, W, base, SL_X_NO, 2, 1, Mon, nullptr);
    sl_chain(P, al, bl, ar, br, W, Mxo, Mxn, false, 2, 0);
    sl_fill_state<TAB>(P, al, bl, ar, br, W, base, SL_X_X, 2, 2, Mox, Mxo);
}

// best non-open state value at (i, j) and the first state holding it
// (fill_D_entries, aligner.cc:586-592; trace_arcmatch, :1015-1024)
__device__ __forceinline__ int sl_best(const Pair &P, int *base, int al, int bl, int ar, int br, int W, int i, int j,
                                       int *state) {
    int best = INT_MIN, k0 = 0;
    for (int k = 0; k < 4; ++k) {
        const int v = Mget(P, sl_M(base, al, bl, ar, br, W, k), W, al, bl, br, i, j);
        if (v > best) {
            best = v;
            k0 = k;
        }
    }
    if (state) *state = k0;
    return best;
}

// Traceback of a structure-local arc match's region from (ar-1, br-1)
// (trace_in_arcmatch / trace_noex in states, aligner.cc:1082-1343):
// excluded positions emit no edge; arc matches found go to the work list
template <bool TAB>
__device__ void group_trace_sl(const Pair &P, int *base, int W, int al, int bl, int ar, int br, int *tstack,
                               int *tcount) {
    const int lane = lane_id();
    const int ex = P.exclusion;
    int i = ar - 1, j = br - 1, k = 0;
    sl_best(P, base, al, bl, ar, br, W, i, j, &k);
    region_sync<true>();
    auto Mk = [&](int s, int x, int y) { return Mget(P, sl_M(base, al, bl, ar, br, W, s), W, al, bl, br, x, y); };
    for (int step = 0; step < 4 * (ar - al + br - bl) + 16; ++step) {
        if (i <= al) {
            if (k == SL_NO_NO || k == SL_OP_NO || k == SL_X_NO)
                for (int t = bl + 1 + lane; t <= j; t += 64) emit(P, 0, t, al + t);
            break;
        }
        if (j <= bl) {
            if (k == SL_NO_NO || k == SL_NO_OP || k == SL_NO_X)
                for (int t = al + 1 + lane; t <= i; t += 64) emit(P, t, 0, t + bl);
            break;
        }
        const int Mij = Mk(k, i, j);
        if (k == SL_OP_NO || k == SL_OP_X) {  // open exclusion of A
            if (Mij == Mk(k, i - 1, j)) {
                --i;
            } else {
                k = (k == SL_OP_NO) ? SL_NO_NO : SL_NO_X;
            }
            continue;
        }
        if (k == SL_NO_OP || k == SL_X_OP) {  // open exclusion of B
            if (Mij == Mk(k, i, j - 1)) {
                --j;
            } else {
                k = (k == SL_NO_OP) ? SL_NO_NO : SL_X_NO;
            }
            continue;
        }
        if (k == SL_NO_X && Mij == Mk(SL_NO_OP, i, j) + ex) {
            k = SL_NO_OP;
            continue;
        }
        if (k == SL_X_NO && Mij == Mk(SL_OP_NO, i, j) + ex) {
            k = SL_OP_NO;
            continue;
        }
        if (k == SL_X_X) {
            if (Mij == Mk(SL_OP_X, i, j) + ex) {
                k = SL_OP_X;
                continue;
            }
            if (Mij == Mk(SL_X_OP, i, j) + ex) {
                k = SL_X_OP;
                continue;
            }
        }
        // trace_noex in state k
        if (valid(P, i - 1, j - 1) && Mij == Mk(k, i - 1, j - 1) + sig_row<TAB>(P, i)[sig_col<TAB>(P, j)]) {
            if (lane == 0) emit(P, i, j, i + j);
            --i;
            --j;
            continue;
        }
        bool done = false;
        if (P.alpha != 0) {
            for (int dir = 0; dir < 2 && !done; ++dir) {
                int carry = P.alpha;
                const int lim = dir == 0 ? i - al : j - bl;
                for (int b0 = 1; !done && b0 <= lim; b0 += 64) {
                    const int t = b0 + lane;
                    const bool inr = t <= lim && (dir == 0 ? valid(P, i - t, j) : valid(P, i, j - t));
                    const int g = inr ? (dir == 0 ? P.gapA[i - t + 1] : P.gapB[j - t + 1]) : 0;
                    const int cost = carry + wave_scan_add(g);
                    const bool hit = inr && Mij == (dir == 0 ? Mk(k, i - t, j) : Mk(k, i, j - t)) + cost;
                    const unsigned long long bh = __ballot(hit);
                    const unsigned long long nv = ~__ballot(inr);
                    const unsigned long long valid_prefix = nv ? (nv & (0ull - nv)) - 1 : ~0ull;
                    const unsigned long long h = bh & valid_prefix;
                    if (h) {
                        const int kk = b0 + (int)__ffsll((long long)h) - 1;
                        if (dir == 0) {
                            for (int u = i - kk + 1 + lane; u <= i; u += 64) emit(P, u, 0, u + j);
                            i -= kk;
                        } else {
                            for (int u = j - kk + 1 + lane; u <= j; u += 64) emit(P, 0, u, i + u);
                            j -= kk;
                        }
                        done = true;
                    } else if (valid_prefix != ~0ull) {
                        break;
                    }
                    carry = readlane(cost, 63);
                }
            }
        } else {
            if (valid(P, i - 1, j) && Mij == Mk(k, i - 1, j) + P.gapA[i]) {
                if (lane == 0) emit(P, i, 0, i + j);
                --i;
                done = true;
            } else if (valid(P, i, j - 1) && Mij == Mk(k, i, j - 1) + P.gapB[j]) {
                if (lane == 0) emit(P, 0, j, i + j);
                --j;
                done = true;
            }
        }
        if (done) continue;
        if (!valid(P, i - 1, j - 1) || !valid(P, i, j)) break;
        const int cell = P.coff[i] + (j - P.mn[i]);
        const int t0 = P.cptr[cell], t1 = P.cptr[cell + 1];
        bool found = false;
        int fp = 0, fq = 0;
        for (int b0 = t0; b0 < t1 && !found; b0 += 64) {
            const int t = b0 + lane;
            bool hit = false;
            int p = 0, q = 0;
            if (t < t1) {
                const int4 e = P.cand[t];
                p = e.x & 0xffff;
                q = e.x >> 16;
                hit = p > al && q > bl && Mij == Mk(k, p - 1, q - 1) + e.z;
            }
            const unsigned long long bh = __ballot(hit);
            if (bh) {
                const int src = (int)__ffsll((long long)bh) - 1;
                found = true;
                fp = readlane(p, src);
                fq = readlane(q, src);
            }
        }
        if (!found) break;
        if (lane == 0) {
            emit(P, fp, fq, fp + fq);
            emit(P, i, j, i + j);
            P.structs[fp - 1] = 1;
            P.structs[i - 1] = 2;
            P.structs[P.n + fq - 1] = 1;
            P.structs[P.n + j - 1] = 2;
            const int slot = atomicAdd(tcount, 1);
            tstack[2 * slot] = fp | (fq << 16);
            tstack[2 * slot + 1] = i | (j << 16);
        }
        i = fp - 1;
        j = fq - 1;
    }
}

// stored width of the rows al..ar-1 of a region (origin al, bl; bound br)
__device__ __forceinline__ int region_width(const Pair &P, int al, int bl, int ar, int br) {
    int w = 0;
    for (int i = al; i < ar; ++i) {
        const int lo = (i == al) ? bl : row_lo(P, i, bl);
        w = max(w, row_hi(P, i, br) - lo + 1);
    }
    return w;
}

template <bool TAB, bool GLOBAL>
__device__ __forceinline__ void fill_and_trace(const Pair &P, int al, int bl, int ar, int br, int W, int *base, int *tstack,
                               int *tcount) {
    fill_group<TAB, GLOBAL, false>(P, al, bl, ar, br, W, base, nullptr, true);
    group_trace<TAB, GLOBAL>(P, base + (br - bl) + W, W, al, bl, br, ar - 1, br - 1, false, false, 0, tstack, tcount);
}

}  // namespace

// Loads the per-pair header into LDS (band, cell offsets, codes, gaps, sigma,
// fill-list blocks) and sets up the pair's view.  Ends with a barrier.
__device__ __forceinline__ Pair load_pair(const PairDesc &d, int *arena, int *lds, const BatchParams &bp) {
    const int L = bp.max_len;
    const int n = d.n, m = d.m;
    int *mn = lds, *mx = lds + (L + 1), *coff = lds + 2 * (L + 1);
    int *codeA = lds + 3 * L + 4, *codeB = lds + 4 * L + 5;
    int *gapA = lds + 5 * L + 6, *gapB = lds + 6 * L + 7;
    int *sig = lds + 7 * L + 8;
    int *fbp = lds + bp.fptr_lds, *fbk = lds + bp.fblk_lds;
    for (int t = threadIdx.x; t <= n; t += blockDim.x) {
        mn[t] = arena[d.band + t];
        mx[t] = arena[d.band + n + 1 + t];
        codeA[t] = arena[d.codes + t];
        gapA[t] = arena[d.gaps + t];
    }
    for (int t = threadIdx.x; t <= m; t += blockDim.x) {
        codeB[t] = arena[d.codes + n + 1 + t];
        gapB[t] = arena[d.gaps + n + 1 + t];
    }
    for (int t = threadIdx.x; t <= n + 1; t += blockDim.x) {
        coff[t] = arena[d.band + 2 * (n + 1) + t];
        fbp[t] = arena[d.fptr + t];
    }
    if (!bp.sigma_table)
        for (int t = threadIdx.x; t < LCG_NCODES * LCG_NCODES; t += blockDim.x) sig[t] = arena[d.sigma + t];
    for (int t = threadIdx.x; t < 2 * d.nfblk; t += blockDim.x) fbk[t] = arena[d.fblk + t];
    Pair P;
    P.mn = mn;
    P.mx = mx;
    P.coff = coff;
    P.codeA = codeA;
    P.codeB = codeB;
    P.gapA = gapA;
    P.gapB = gapB;
    P.sig = sig;
    P.gsig = arena + d.sigma;
    P.cptr = arena + d.cptr;
    P.cand = reinterpret_cast<int4 *>(arena + d.cand);
    P.fbp = fbp;
    P.fbk = fbk;
    P.fcand = arena + d.fcand;
    P.amscore = arena + d.amscore;
    P.D = arena + d.D;
    P.edges = arena + d.edges;
    P.structs = arena + d.structs;
    P.inner = arena + d.inner;
    P.c2a = arena + d.c2a;
    P.amst = arena + d.amst;
    P.n = n;
    P.m = m;
    P.nam = d.nam;
    P.alpha = bp.alpha;
    P.alpha_pos = bp.alpha > 0 ? bp.alpha : 0;
    P.nolp = bp.nolp;
    P.stacking = bp.stacking;
    P.struct_local = bp.struct_local;
    P.exclusion = bp.exclusion;
    return P;
}

// Kernel 1: fill D of every pair (rows al = n .. 1, the groups of a row
// packed into LDS rounds by the host).  With K = bp.wg_per_pair > 1 (batches
// of fewer pairs than CUs), K workgroups share each pair: workgroup
// blockIdx.x serves pair blockIdx.x / K as member blockIdx.x % K, takes the
// groups the host assigned to that member, and the K members meet at a
// bounded barrier (a global arrival counter) after every row with groups.
// Group schedule word G[5] = round | wave << 12 | member << 16.
#define LCG_SPIN_LIMIT (1 << 22)

template <bool TAB, bool SL>
__device__ __forceinline__ void fill_pairs(const PairDesc *__restrict__ descs, const int *__restrict__ order,
                                           int *queue, int *arena, int *barriers, int *scratch, BatchParams bp) {
    extern __shared__ int lds[];
    __shared__ int s_pair;
    const int wave = threadIdx.x >> 6;
    const int lane = lane_id();
    int *pool = lds + bp.hdr_ints;
    const int K = bp.wg_per_pair;
    const int member = K > 1 ? (int)blockIdx.x % K : 0;
    const int spin_lim = bp.spin_limit ? bp.spin_limit : LCG_SPIN_LIMIT;
    for (int iter = 0;; ++iter) {
        if (K > 1) {
            if (iter > 0) break;
            if (threadIdx.x == 0) s_pair = (int)blockIdx.x / K;
        } else if (threadIdx.x == 0) {
            s_pair = atomicAdd(queue, 1);
        }
        __syncthreads();
        const int k = s_pair;
        if (k >= bp.npairs) break;
        const PairDesc d = descs[order[k]];
        if (member == 0) {
            if (bp.count && threadIdx.x < 4) arena[d.out + 4 + threadIdx.x] = 0;
            if (threadIdx.x == 0) arena[d.out + 3] = 0;  // status
        }
        const Pair P = load_pair(d, arena, lds, bp);
        unsigned long long *counts = bp.count ? (unsigned long long *)(arena + d.out + 4) : nullptr;
        if (K > 1 && counts) {
            // members count into the pair's words after member 0 cleared them
            if (threadIdx.x == 0) {
                __threadfence();
                atomicAdd(&barriers[k], 1);
                for (int spin = 0; spin_lim < 0 || atomicAdd(&barriers[k], 0) < K; ++spin) {
                    if (spin > spin_lim) {
                        atomicExch(&arena[d.out + 3], 2);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
        }
        int passed = (K > 1 && counts) ? 1 : 0;  // barrier generations passed
        __threadfence_block();
        __syncthreads();
        const int *rows = arena + d.rows;
        const int *rounds = arena + d.rounds;
        const int *groups = arena + d.groups;
        const int4 *gams = reinterpret_cast<const int4 *>(arena + d.gams);
        const int mine = member << 16;
        for (int al = d.n; al >= 1; --al) {
            const int g0 = rows[al + 1], g1 = rows[al];
            if (g0 == g1) continue;
            const int nr = rounds[al];
            for (int rnd = 0; rnd < nr; ++rnd) {
                // the wave's group of this round (at most one): the lanes
                // compare the schedule words of 64 groups at a time
                const int want = rnd | (wave << 12) | mine;
                int mg = -1;
                for (int gb = g0; gb < g1 && mg < 0; gb += 64) {
                    const int g = gb + lane;
                    const unsigned long long hit = __ballot(g < g1 && groups[g * LCG_GROUP_INTS + 5] == want);
                    if (hit) mg = gb + __ffsll((long long)hit) - 1;
                }
                if (SL && mg >= 0) {
                    // structure-local: eight states in the wave's global scratch
                    const int *G = groups + mg * LCG_GROUP_INTS;
                    const int gal = G[0] & 0xffff, gbl = G[0] >> 16;
                    const int gar = G[1] & 0xffff, gbr = G[1] >> 16;
                    const int W = G[2];
                    int *base = scratch + (size_t)(blockIdx.x * (blockDim.x >> 6) + wave) * bp.sl_stride;
                    group_fill_sl<TAB>(P, gal, gbl, gar, gbr, W, base);
                    for (int t = G[3] + lane; t < G[4]; t += 64) {
                        const int4 ga = gams[t];
                        const int ar = ga.x & 0xffff, br = ga.x >> 16;
                        const int v = sl_best(P, base, gal, gbl, gar, gbr, W, ar - 1, br - 1, nullptr) + P.amscore[ga.z];
                        P.cand[ga.y].z = v;
                        P.fcand[3 * ga.w + 2] = v;
                        P.D[ga.z] = v;
                    }
                    region_sync<true>();
                } else if (!SL && mg >= 0) {
                    const int *G = groups + mg * LCG_GROUP_INTS;
                    const int gal = G[0] & 0xffff, gbl = G[0] >> 16;
                    const int gar = G[1] & 0xffff, gbr = G[1] >> 16;
                    const int W = G[2];
                    int *base = pool + G[6];
                    if (counts)
                        fill_group<TAB, false, true, true>(P, gal, gbl, gar, gbr, W, base, counts, bp.fixed_cols != 0);
                    else
                        fill_group<TAB, false, true, false>(P, gal, gbl, gar, gbr, W, base, nullptr, bp.fixed_cols != 0);
                    const int *Mst = base + (gbr - gbl) + W;
                    for (int t = G[3] + lane; t < G[4]; t += 64) {
                        const int4 ga = gams[t];
                        const int ar = ga.x & 0xffff, br = ga.x >> 16;
                        const int mv = Mst[(ar - 1 - gal) * W + (br - 1 - row_lo(P, ar - 1, gbl))];
                        int v = mv + P.amscore[ga.z];
                        if (P.nolp) {
                            // aligner.cc:612-657: the inner arc match (its right
                            // ends are ar, br) ends the stack or continues it;
                            // its D comes from a later row, written before.  With
                            // stacking only a stackable arc match has a D, with
                            // its stacked score
                            const int y = P.inner[ga.z];  // always one (the host lists no other)
                            const int sx = P.stacking ? P.amst[ga.z] : P.amscore[ga.z];
                            v = (unsigned)y < (unsigned)P.nam && sx != LCG_NEG
                                    ? max(mv + P.amscore[y], P.D[y]) + sx
                                    : LCG_NEG;
                        } else if (P.stacking) {
                            // aligner.cc:574-610: a stackable arc match with an
                            // inner one may take the inner one's D (a later row,
                            // written before) plus its stacked score
                            const int y = P.inner[ga.z];
                            const int st = P.amst[ga.z];
                            if ((unsigned)y < (unsigned)P.nam && st != LCG_NEG) v = max(v, P.D[y] + st);
                        }
                        P.cand[ga.y].z = v;  // D, not normalized (the reference's InftyInt sum)
                        P.fcand[3 * ga.w + 2] = v;
                        P.D[ga.z] = v;
                    }
                }
                __threadfence_block();
                __syncthreads();
            }
            if (K > 1) {
                // every member's D of row al is written before any member
                // reads it in a later row: each wave completes its stores at
                // device scope, then thread 0 arrives and waits for all
                __threadfence();
                __syncthreads();
                ++passed;
                if (threadIdx.x == 0) {
                    atomicAdd(&barriers[k], 1);
                    for (int spin = 0; spin_lim < 0 || atomicAdd(&barriers[k], 0) < K * passed; ++spin) {
                        if (spin > spin_lim) {
                            atomicExch(&arena[d.out + 3], 2);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                }
                __syncthreads();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
        }
        __syncthreads();
    }
}

// One instance per sigma kind: symbol codes, and the column-pair tables of
// pairs of alignment profiles.
extern "C" __global__ void __launch_bounds__(1024) lcg_fill_pairs(const PairDesc *__restrict__ descs,
                                                                  const int *__restrict__ order, int *queue,
                                                                  int *arena, int *barriers, int *scratch,
                                                                  BatchParams bp) {
    fill_pairs<false, false>(descs, order, queue, arena, barriers, scratch, bp);
}
extern "C" __global__ void __launch_bounds__(1024) lcg_fill_pairs_tab(const PairDesc *__restrict__ descs,
                                                                      const int *__restrict__ order, int *queue,
                                                                      int *arena, int *barriers, int *scratch,
                                                                      BatchParams bp) {
    fill_pairs<true, false>(descs, order, queue, arena, barriers, scratch, bp);
}
// structure-local (symbol codes only)
extern "C" __global__ void __launch_bounds__(1024) lcg_fill_pairs_sl(const PairDesc *__restrict__ descs,
                                                                     const int *__restrict__ order, int *queue,
                                                                     int *arena, int *barriers, int *scratch,
                                                                     BatchParams bp) {
    fill_pairs<false, true>(descs, order, queue, arena, barriers, scratch, bp);
}

extern "C" __global__ void lcg_reset_barriers(int *barriers, int n) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) barriers[t] = 0;
}

// Kernel 2: top level (global or sequence-local) and traceback of every
// pair, after kernel 1 has filled D.
template <bool TAB, bool SL>
__device__ __forceinline__ void top_pairs(const PairDesc *__restrict__ descs, const int *__restrict__ order,
                                          int *queue, int *arena, int *scratch, BatchParams bp,
                                          const int *__restrict__ lambda, const int4 *__restrict__ restr) {
    extern __shared__ int lds[];
    __shared__ int s_pair, s_tcount;
    const int nwaves = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = lane_id();
    int *pool = lds + bp.hdr_ints;
    int *wave_scratch = scratch + (size_t)(blockIdx.x * nwaves + wave) * bp.scratch_stride;
    for (;;) {
        if (threadIdx.x == 0) s_pair = atomicAdd(queue, 1);
        __syncthreads();
        const int k = s_pair;
        if (k >= bp.npairs) break;
        const PairDesc d = descs[order[k]];
        const int n = d.n, m = d.m;
        // position penalty of the top level (penalized / normalized
        // alignment, aligner.cc:1526-1622); 0 otherwise
        const int lam = lambda ? lambda[order[k]] : 0;
        if (bp.trace) {
            for (int t = threadIdx.x; t <= n + m; t += blockDim.x) arena[d.edges + t] = -1;
            for (int t = threadIdx.x; t < n + m; t += blockDim.x) arena[d.structs + t] = 0;
        }
        if (threadIdx.x == 0) s_tcount = 0;
        if (bp.count && threadIdx.x < 4) arena[d.out + 8 + threadIdx.x] = 0;
        const Pair P = load_pair(d, arena, lds, bp);
        unsigned long long *counts = bp.count ? (unsigned long long *)(arena + d.out + 8) : nullptr;  // top level
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // the pair is this workgroup's alone
        __syncthreads();
        // ---- top level (wave 0)
        // the top level aligns A[sA..eA] with B[sB..eB]: the whole sequences,
        // or a restriction (AlignerRestriction, aligner_restriction.hh:26-120;
        // the k-best enumeration, aligner.cc:1382-1514), whose region (origin
        // (sA-1, sB-1), bounds (eA+1, eB+1)) the host sized the scratch for
        int sA = 1, eA = n, sB = 1, eB = m;
        if (restr) {
            const int4 rr = restr[order[k]];
            sA = rr.x;
            eA = rr.y;
            sB = rr.z;
            eB = rr.w;
        }
        const int tal = sA - 1, tbl = sB - 1, tar = eA + 1, tbr = eB + 1;
        const int tW = restr ? region_width(P, tal, tbl, tar, tbr) : d.wtop;
        const int t_ints = restr ? (((tbr - tbl) + tW + (tar - tal) * tW + 3) & ~3) : d.top_ints;
        const bool top_in_lds = t_ints <= bp.pool_ints;
        int *topbase = top_in_lds ? pool : wave_scratch;
        int *topM = topbase + (tbr - tbl) + tW;
        if (wave == 0) {
            TopBest best{0, tal, tbl};
            const int fe = bp.free_endgaps;
            if (bp.local) {
                if (top_in_lds)
                    group_fill<TAB, false, true>(P, tal, tbl, tar, tbr, tW, topbase, false, false, &best, counts, lam);
                else
                    group_fill<TAB, true, true>(P, tal, tbl, tar, tbr, tW, topbase, false, false, &best, counts, lam);
            } else {
                // free end gaps: column 0 free with the second sequence's
                // left end, row 0 with the first's (aligner.cc:736-745)
                const bool gcol = !(fe & 4), grow = !(fe & 1);
                if (top_in_lds)
                    group_fill<TAB, false, false>(P, tal, tbl, tar, tbr, tW, topbase, gcol, grow, nullptr, counts, lam);
                else
                    group_fill<TAB, true, false>(P, tal, tbl, tar, tbr, tW, topbase, gcol, grow, nullptr, counts, lam);
                if (top_in_lds)
                    region_sync<false>();
                else
                    region_sync<true>();
                // M(eA, eB) of the reference's matrix; outside the band (free
                // end gaps can cut it out) its never-written initial 0, for a
                // proper restriction minus infinity (the reference reads what
                // an earlier computation left there)
                const bool whole = sA == 1 && sB == 1 && eA == n && eB == m;
                best.score = (P.mn[eA] <= eB && eB <= P.mx[eA]) ? Mget(P, topM, tW, tal, tbl, tbr, eA, eB)
                             : whole                             ? 0
                                                                 : LCG_NEG;
                best.i = eA;
                best.j = eB;
                if (fe & (2 | 8)) {
                    // first strict maximum of the last column (second
                    // sequence's right end; rows with max_col >= eB,
                    // ascending), then of the last row (first sequence's
                    // right end, band columns ascending) (aligner.cc:775-816)
                    best.score = LCG_NEG;
                    best.i = tal;
                    best.j = tbl;
                    auto scan = [&](int cnt, bool col, int base) {
                        for (int t0 = 0; t0 < cnt; t0 += 64) {
                            const int t = t0 + lane;
                            const int ci = col ? base + t : eA, cj = col ? eB : base + t;
                            const bool ok = t < cnt && (!col || P.mx[ci] >= eB);
                            const int v = ok ? Mget(P, topM, tW, tal, tbl, tbr, ci, cj) : INT_MIN;
                            int wmax = v;
                            for (int o = 32; o >= 1; o >>= 1) wmax = max(wmax, __shfl_xor(wmax, o));
                            if (wmax > best.score) {
                                const unsigned long long bal = __ballot(ok && v == wmax);
                                const int l = (int)__ffsll((long long)bal) - 1;
                                best.score = wmax;
                                best.i = col ? base + t0 + l : eA;
                                best.j = col ? eB : base + t0 + l;
                            }
                        }
                    };
                    if (fe & 8) scan(eA - sA + 1, true, sA);
                    if (fe & 2) {
                        const int j0 = max(sB, P.mn[eA]), j1 = min(eB, P.mx[eA]);
                        if (j1 >= j0) scan(j1 - j0 + 1, false, j0);
                    }
                }
            }
            if (lane == 0) {
                arena[d.out + 0] = best.score;
                arena[d.out + 1] = best.i;
                arena[d.out + 2] = best.j;
            }
            // ---- trace the top level
            if (bp.trace) {
                if (top_in_lds)
                    group_trace<TAB, false>(P, topM, tW, tal, tbl, tbr, best.i, best.j, true, bp.local != 0,
                                       bp.free_endgaps, arena + d.tstack, &s_tcount, lam);
                else
                    group_trace<TAB, true>(P, topM, tW, tal, tbl, tbr, best.i, best.j, true, bp.local != 0,
                                      bp.free_endgaps, arena + d.tstack, &s_tcount, lam);
            }
        }
        __threadfence_block();
        __syncthreads();
        // ---- trace inside the traced arc matches, breadth first.  A level's
        // tasks share the LDS pool in equal slots; a task whose region does
        // not fit its slot waits for a pass in which wave 0 has the whole
        // pool (every group region fits the pool, checked at batch creation).
        if (SL && bp.trace) {
            // structure-local regions in each wave's global scratch
            int lo = 0, hi = s_tcount;
            __syncthreads();
            while (lo < hi) {
                for (int e = lo + wave; e < hi; e += nwaves) {
                    const int *ts = arena + d.tstack + 2 * e;
                    const int al = ts[0] & 0xffff, bl = ts[0] >> 16;
                    const int ar = ts[1] & 0xffff, br = ts[1] >> 16;
                    const int W = region_width(P, al, bl, ar, br);
                    group_fill_sl<TAB>(P, al, bl, ar, br, W, wave_scratch);
                    group_trace_sl<TAB>(P, wave_scratch, W, al, bl, ar, br, arena + d.tstack, &s_tcount);
                }
                __threadfence_block();
                __syncthreads();
                lo = hi;
                hi = s_tcount;
                __syncthreads();
            }
        } else if (!SL && bp.trace) {
            int lo = 0, hi = s_tcount;
            __syncthreads();
            while (lo < hi) {
                const int nt = hi - lo;
                const int nw = nt < nwaves ? nt : nwaves;
                const int slot = (bp.pool_ints / nw) & ~3;
                for (int e = lo + wave; wave < nw && e < hi; e += nw) {
                    const int *ts = arena + d.tstack + 2 * e;
                    const int al = ts[0] & 0xffff, bl = ts[0] >> 16;
                    const int ar = ts[1] & 0xffff, br = ts[1] >> 16;
                    const int W = region_width(P, al, bl, ar, br);
                    if ((br - bl) + W + (ar - al) * W <= slot)
                        fill_and_trace<TAB, false>(P, al, bl, ar, br, W, pool + wave * slot, arena + d.tstack, &s_tcount);
                }
                __threadfence_block();
                __syncthreads();
                if (wave == 0) {
                    for (int e = lo; e < hi; ++e) {
                        const int *ts = arena + d.tstack + 2 * e;
                        const int al = ts[0] & 0xffff, bl = ts[0] >> 16;
                        const int ar = ts[1] & 0xffff, br = ts[1] >> 16;
                        const int W = region_width(P, al, bl, ar, br);
                        if ((br - bl) + W + (ar - al) * W > slot)
                            fill_and_trace<TAB, false>(P, al, bl, ar, br, W, pool, arena + d.tstack, &s_tcount);
                    }
                }
                __threadfence_block();
                __syncthreads();
                lo = hi;
                hi = s_tcount;
                __syncthreads();
            }
        }
        __syncthreads();
    }
}

extern "C" __global__ void __launch_bounds__(1024) lcg_top_pairs(const PairDesc *__restrict__ descs,
                                                                 const int *__restrict__ order, int *queue,
                                                                 int *arena, int *scratch, BatchParams bp,
                                                                 const int *__restrict__ lambda,
                                                                 const int4 *__restrict__ restr) {
    top_pairs<false, false>(descs, order, queue, arena, scratch, bp, lambda, restr);
}
extern "C" __global__ void __launch_bounds__(1024) lcg_top_pairs_tab(const PairDesc *__restrict__ descs,
                                                                     const int *__restrict__ order, int *queue,
                                                                     int *arena, int *scratch, BatchParams bp,
                                                                     const int *__restrict__ lambda,
                                                                     const int4 *__restrict__ restr) {
    top_pairs<true, false>(descs, order, queue, arena, scratch, bp, lambda, restr);
}
extern "C" __global__ void __launch_bounds__(1024) lcg_top_pairs_sl(const PairDesc *__restrict__ descs,
                                                                    const int *__restrict__ order, int *queue,
                                                                    int *arena, int *scratch, BatchParams bp,
                                                                    const int *__restrict__ lambda,
                                                                    const int4 *__restrict__ restr) {
    top_pairs<false, true>(descs, order, queue, arena, scratch, bp, lambda, restr);
}

// Copies the outputs of every pair (edges, structure strings, traceback
// stack, out words: one contiguous range of its image) into a packed
// buffer, so that the host reads the results of a run with one copy.
extern "C" __global__ void lcg_gather_outputs(const PairDesc *descs, const int *res_off, const int *arena, int *dst) {
    const PairDesc d = descs[blockIdx.x];
    const int len = d.out + 12 - d.edges;
    int *o = dst + res_off[blockIdx.x];
    for (int t = threadIdx.x; t < len; t += blockDim.x) o[t] = arena[d.edges + t];
}

extern "C" int lcg_launch_gather(const PairDesc *descs, const int *res_off, const int *arena, int *dst, int npairs,
                                 hipStream_t stream) {
    hipLaunchKernelGGL(lcg_gather_outputs, dim3(npairs), dim3(256), 0, stream, descs, res_off, arena, dst);
    return (int)hipGetLastError();
}

extern "C" __global__ void lcg_reset_queue(int *queue) {
    if (threadIdx.x == 0) *queue = 0;
}

// Scratch of the allocation-free run (lcg_batch_run_noalloc in batch.cpp):
// the two work-queue counters, the arrival counters of pairs that several
// workgroups share and the per-wave regions of the top-level kernel.  Every
// GPU the library runs on holds its own copy until the process ends; host
// code reaches it only through the per-device address that
// lcg_noalloc_scratch looks up (the array's name is not a device address
// in host code).
#define LCG_NOALLOC_SCRATCH_INTS (16 << 20) /* 64 MiB */
static __device__ __align__(256) int scratch[LCG_NOALLOC_SCRATCH_INTS];

// Device address and size (ints) of `scratch` on `device`, looked up once
// per device with that device current.
extern "C" int lcg_noalloc_scratch(int device, int **ptr, long long *ints) {
    static std::mutex mu;
    static int *addr[64] = {nullptr};
    if (device < 0 || device >= 64 || !ptr || !ints) return (int)hipErrorInvalidValue;
    std::lock_guard<std::mutex> guard(mu);
    if (!addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return (int)e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return (int)e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        const hipError_t e2 = (cur != device) ? hipSetDevice(cur) : hipSuccess;
        if (e != hipSuccess) return (int)e;
        if (e2 != hipSuccess) return (int)e2;
        if (!p) return (int)hipErrorInvalidSymbol;
        addr[device] = static_cast<int *>(p);
    }
    *ptr = addr[device];
    *ints = LCG_NOALLOC_SCRATCH_INTS;
    return (int)hipSuccess;
}

extern "C" int lcg_launch_align(const PairDesc *descs, const int *order, int *queue, int *arena, int *scratch,
                                int *barriers, BatchParams bp, int grid, int waves, size_t lds_bytes, BatchParams bp_top,
                                int top_grid, int top_waves, size_t top_lds_bytes, const int *lambda, int top_only,
                                hipEvent_t mid, hipStream_t stream, const int4 *restr) {
    static bool attr_set = false;
    if (!attr_set) {
        const void *kernels[] = {(const void *)lcg_fill_pairs,    (const void *)lcg_top_pairs,
                                 (const void *)lcg_fill_pairs_tab, (const void *)lcg_top_pairs_tab,
                                 (const void *)lcg_fill_pairs_sl,  (const void *)lcg_top_pairs_sl};
        for (const void *k : kernels) {
            hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256);
            if (e != hipSuccess) return (int)e;
        }
        attr_set = true;
    }
    // two work counters, 64 bytes apart; the per-pair barrier counters of
    // the multi-workgroup fill
    hipLaunchKernelGGL(lcg_reset_queue, dim3(1), dim3(64), 0, stream, queue + 16);
    if (!top_only) {  // the fill of D (skipped when only the top level is redone)
        hipLaunchKernelGGL(lcg_reset_queue, dim3(1), dim3(64), 0, stream, queue);
        if (bp.wg_per_pair > 1)
            hipLaunchKernelGGL(lcg_reset_barriers, dim3(1), dim3(256), 0, stream, barriers, bp.npairs);
        if (bp.struct_local)
            hipLaunchKernelGGL(lcg_fill_pairs_sl, dim3(grid), dim3(waves * 64), lds_bytes, stream, descs, order,
                               queue, arena, barriers, scratch, bp);
        else if (bp.sigma_table)
            hipLaunchKernelGGL(lcg_fill_pairs_tab, dim3(grid), dim3(waves * 64), lds_bytes, stream, descs, order,
                               queue, arena, barriers, scratch, bp);
        else
            hipLaunchKernelGGL(lcg_fill_pairs, dim3(grid), dim3(waves * 64), lds_bytes, stream, descs, order, queue,
                               arena, barriers, scratch, bp);
    }
    if (mid) hipEventRecord(mid, stream);
    if (bp_top.struct_local)
        hipLaunchKernelGGL(lcg_top_pairs_sl, dim3(top_grid), dim3(top_waves * 64), top_lds_bytes, stream, descs,
                           order, queue + 16, arena, scratch, bp_top, lambda, restr);
    else if (bp_top.sigma_table)
        hipLaunchKernelGGL(lcg_top_pairs_tab, dim3(top_grid), dim3(top_waves * 64), top_lds_bytes, stream, descs,
                           order, queue + 16, arena, scratch, bp_top, lambda, restr);
    else
        hipLaunchKernelGGL(lcg_top_pairs, dim3(top_grid), dim3(top_waves * 64), top_lds_bytes, stream, descs, order,
                           queue + 16, arena, scratch, bp_top, lambda, restr);
    return (int)hipGetLastError();
}
