// gfx950 kernels of the device-side arc-match compaction: the arc-match
// predicate over all arc pairs of a pair of sequences, its ordered
// compaction, and every arc-match structure of the pair's arena image that
// the alignment kernels read (device_layout.h).
//
// Reference: ArcMatches, src/LocARNA/arc_matches.cc:19-48 (is_valid_arcmatch),
// :130-188 (constructor: index order arc of A outer, arc of B inner; lists by
// common left ends in push order, by common right ends sorted (al desc, bl
// desc)), Scoring::arcmatch (scoring.cc:440-554) for the arc-match scores,
// and AlignerImpl::align_D (aligner.cc:659-732) for the group order (al
// descending, bl descending) and the region of a group (max right ends).
//
// One workgroup of 4 wave64s per pair; the grid covers the batch.  Pass 1
// (lcg_compact_count) writes the validity bitmap of the N_A x N_B arc pairs
// (one 64-bit word per 64 arcs of B, a wave ballot per word) and counts arc
// matches, groups and fill blocks, from which the host lays out the arena.
// Pass 2 (lcg_compact_build) turns the bitmap into the structures.  Every
// ordered list is produced by a workgroup prefix sum over its enumeration
// order (band cells, arcs in right-adjacency order, bitmap words), so the
// output does not depend on scheduling and equals the host packing bit for
// bit.  The index of arc match (a, b) in the reference order is the number
// of set bits before it in the row-major bitmap: wpre[word] (the prefix of
// the words before) plus a popcount within the word.
#include <hip/hip_runtime.h>

#include "compact_layout.h"
#include "device_layout.h"

namespace {

constexpr int CT = 256;  // threads per workgroup
constexpr int CW = CT / 64;

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

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ uint64_t lanes_below() {
    const int l = lane_id();
    return l ? (~0ull >> (64 - l)) : 0ull;
}

// Exclusive prefix sums of two values over the workgroup (every thread
// calls it); tmp: 2 * CW ints of LDS.  Returns the totals as well.
__device__ __forceinline__ void block_scan2(int a, int b, int &ea, int &eb, int &ta, int &tb, int *tmp) {
    const int ia = wave_scan_add(a), ib = wave_scan_add(b);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 63) {
        tmp[2 * w] = ia;
        tmp[2 * w + 1] = ib;
    }
    __syncthreads();
    int oa = 0, ob = 0;
    ta = 0;
    tb = 0;
    for (int x = 0; x < CW; ++x) {
        const int va = tmp[2 * x], vb = tmp[2 * x + 1];
        if (x < w) {
            oa += va;
            ob += vb;
        }
        ta += va;
        tb += vb;
    }
    __syncthreads();
    ea = oa + ia - a;
    eb = ob + ib - b;
}

struct CPair {
    int n, m, NA, NB, W64;
    const int *arcsA, *arcsB, *wA, *wB, *ptA, *ptB, *lA, *lB, *rpA, *rpB, *radjA, *radjB, *tab;
    int tau, ribosum, dlen, dend;
    const int *mn, *mx, *coff;  // LDS: band of A's rows, band cell offsets
    uint64_t *bits;
    int *wpre, *rowcnt, *sblk, *blka, *blks, *eof, *fof, *sched;
};

__device__ CPair make_cpair(const CompactDesc &c, const int *cin, int *ws, int *ws2) {
    CPair P;
    P.n = c.n;
    P.m = c.m;
    P.NA = c.NA;
    P.NB = c.NB;
    P.W64 = c.W64;
    P.arcsA = cin + c.arcsA;
    P.arcsB = cin + c.arcsB;
    P.wA = cin + c.wA;
    P.wB = cin + c.wB;
    P.ptA = cin + c.ptA;
    P.ptB = cin + c.ptB;
    P.lA = cin + c.lA;
    P.lB = cin + c.lB;
    P.rpA = cin + c.rpA;
    P.rpB = cin + c.rpB;
    P.radjA = cin + c.radjA;
    P.radjB = cin + c.radjB;
    P.tab = cin + c.tab;
    P.tau = c.tau;
    P.ribosum = c.ribosum;
    P.dlen = c.dlen;
    P.dend = c.dend;
    P.bits = reinterpret_cast<uint64_t *>(ws + c.ws_bits);
    P.wpre = ws + c.ws_wpre;
    P.rowcnt = ws + c.ws_rowcnt;
    P.sblk = ws + c.ws_sblk;
    P.blka = ws + c.ws_blka;
    P.blks = ws + c.ws_blks;
    P.eof = ws2 ? ws2 + c.ws_eof : nullptr;
    P.fof = ws2 ? ws2 + c.ws_fof : nullptr;
    P.sched = ws2 ? ws2 + c.ws_sched : nullptr;
    return P;
}

__device__ __forceinline__ int left_of(int arc) { return arc & 0xffff; }
__device__ __forceinline__ int right_of(int arc) { return arc >> 16; }

// TraceController::is_valid_match (trace_controller.hh): (i, j) and
// (i-1, j-1) both in the band; i >= 1
__device__ __forceinline__ bool valid_match(const CPair &P, int i, int j) {
    return P.mn[i] <= j && j <= P.mx[i] && P.mn[i - 1] <= j - 1 && j - 1 <= P.mx[i - 1];
}

__device__ __forceinline__ bool has_bit(const CPair &P, int a, int b) {
    return (P.bits[(size_t)a * P.W64 + (b >> 6)] >> (b & 63)) & 1ull;
}

// reference index of the valid arc match (a, b)
__device__ __forceinline__ int am_index(const CPair &P, int a, int b) {
    const size_t w = (size_t)a * P.W64 + (b >> 6);
    const int s = b & 63;
    const uint64_t below = s ? (P.bits[w] & (~0ull >> (64 - s))) : 0ull;
    return P.wpre[w] + __popcll(below);
}

// Loads the band of the pair (min_col, max_col of rows 0..n) into LDS and
// the band cell offsets coff[0..n+1] (a prefix sum of the row widths).
__device__ void load_band(int *lds, const int *mn_src, const int *mx_src, int n, int *tmp) {
    int *mn = lds, *mx = lds + (n + 1), *coff = lds + 2 * (n + 1);
    int carry = 0;
    for (int base = 0; base <= n; base += CT) {
        const int i = base + (int)threadIdx.x;
        int w = 0;
        if (i <= n) {
            mn[i] = mn_src[i];
            mx[i] = mx_src[i];
            w = max(0, mx_src[i] - mn_src[i] + 1);
        }
        int e, e2, t, t2;
        block_scan2(w, 0, e, e2, t, t2, tmp);
        if (i <= n) coff[i] = carry + e;
        carry += t;
    }
    if (threadIdx.x == 0) coff[n + 1] = carry;
    __syncthreads();
}

// band cell c -> row i (the last row whose first cell is <= c; rows may be
// empty) and column j
__device__ __forceinline__ void cell_rc(const CPair &P, int c, int &i, int &j) {
    int lo = 0, hi = P.n;  // coff[lo] <= c
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (P.coff[mid] <= c)
            lo = mid;
        else
            hi = mid - 1;
    }
    i = lo;
    j = P.mn[lo] + (c - P.coff[lo]);
}

struct GroupCell {
    int cnt, ar, br;
};

// arc matches with left ends (al, bl) (common_left_end_list) and their
// maximal right ends (arc_matches.cc:333-345)
__device__ GroupCell group_cell(const CPair &P, int al, int bl) {
    GroupCell g{0, al, bl};
    if (al < 1 || bl < 1) return g;
    const int a0 = P.lA[2 * al], a1 = P.lA[2 * al + 1];
    const int b0 = P.lB[2 * bl], b1 = P.lB[2 * bl + 1];
    if (a0 >= a1 || b0 >= b1) return g;
    for (int a = a0; a < a1; ++a) {
        const int ra = right_of(P.arcsA[a]);
        for (int b = b0; b < b1; ++b) {
            if (has_bit(P, a, b)) {
                g.cnt += 1;
                g.ar = max(g.ar, ra);
                g.br = max(g.br, right_of(P.arcsB[b]));
            }
        }
    }
    return g;
}

// stored row width W, LDS ints of the region [E | S | M] and the DP cells of
// the group (al, bl) with maximal right ends (ar, br)
__device__ void group_geom(const CPair &P, int al, int bl, int ar, int br, int &W, int &lds, long long &cells) {
    W = 0;
    cells = 0;
    for (int i = al; i < ar; ++i) {
        const int lo = (i == al) ? bl : max(bl, P.mn[i]);
        const int hi = min(br - 1, P.mx[i]);
        W = max(W, max(0, hi - lo + 1));
        if (i > al) cells += max(0, min(br - 1, P.mx[i]) - max(bl + 1, P.mn[i]) + 1);
    }
    lds = ((br - bl) + W + (ar - al) * W + 3) & ~3;
}

}  // namespace

// Pass 1: validity bitmap and counts.  grid = npairs, CT threads; dynamic
// LDS 3 * (max_len + 1) + 1 ints.
__global__ __launch_bounds__(CT) void lcg_compact_count(const CompactDesc *__restrict__ cds,
                                                        const int *__restrict__ cin, int *ws,
                                                        int *__restrict__ stats) {
    extern __shared__ int lds[];
    __shared__ int tmp[2 * CW];
    __shared__ int s_nam, s_groups, s_blk, s_gmax, s_ecap;
    __shared__ unsigned long long s_cells;
    const CompactDesc cd = cds[blockIdx.x];
    CPair P = make_cpair(cd, cin, ws, nullptr);
    if (threadIdx.x == 0) {
        s_nam = s_groups = s_blk = s_gmax = s_ecap = 0;
        s_cells = 0;
    }
    load_band(lds, cin + cd.band, cin + cd.band + cd.n + 1, cd.n, tmp);
    P.mn = lds;
    P.mx = lds + (P.n + 1);
    P.coff = lds + 2 * (P.n + 1);
    const int lane = lane_id();
    // arc-match predicate (arc_matches.cc:19-48), one bitmap word per ballot
    for (int a = threadIdx.x >> 6; a < P.NA; a += CW) {
        const int arcA = P.arcsA[a];
        const int al = left_of(arcA), ar = right_of(arcA);
        int run = 0;
        for (int w = 0; w < P.W64; ++w) {
            const int b = 64 * w + lane;
            bool ok = false;
            if (b < P.NB) {
                const int arcB = P.arcsB[b];
                const int bl = left_of(arcB), br = right_of(arcB);
                ok = valid_match(P, al, bl) && valid_match(P, ar, br) &&
                     abs((ar - al) - (br - bl)) <= P.dlen && abs(al - bl) <= P.dend && abs(ar - br) <= P.dend;
            }
            const uint64_t mask = __ballot(ok);
            if (lane == 0) {
                P.bits[(size_t)a * P.W64 + w] = mask;
                P.wpre[(size_t)a * P.W64 + w] = run;
            }
            run += __popcll(mask);
        }
        if (lane == 0) {
            P.rowcnt[a] = run;
            atomicAdd(&s_nam, run);
            if (run > 0) atomicAdd(&s_blk, 1);
        }
    }
    __syncthreads();
    // groups: band cells (al, bl) with at least one arc match
    const int ncells = P.coff[P.n + 1];
    for (int c = threadIdx.x; c < ncells; c += CT) {
        int al, bl;
        cell_rc(P, c, al, bl);
        const GroupCell g = group_cell(P, al, bl);
        if (g.cnt > 0) {
            int W, L;
            long long cells;
            group_geom(P, al, bl, g.ar, g.br, W, L, cells);
            atomicAdd(&s_groups, 1);
            atomicMax(&s_gmax, L);
            atomicMax(&s_ecap, g.br - bl);
            atomicAdd(&s_cells, (unsigned long long)cells);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int *S = stats + (size_t)blockIdx.x * LCG_CS_INTS;
        S[LCG_CS_NAM] = s_nam;
        S[LCG_CS_NGROUPS] = s_groups;
        S[LCG_CS_NBLK] = s_blk;
        S[LCG_CS_GMAX] = s_gmax;
        S[LCG_CS_ECAP] = s_ecap;
        S[LCG_CS_CELLS] = (int)(s_cells & 0xffffffffull);
        S[LCG_CS_CELLS + 1] = (int)(s_cells >> 32);
        S[LCG_CS_STATUS] = 0;
    }
}

// Pass 2: the arena structures of each pair (offsets in descs, sizes from
// pass 1 in stats).  grid = npairs, CT threads; dynamic LDS as pass 1.
__global__ __launch_bounds__(CT) void lcg_compact_build(const CompactDesc *__restrict__ cds,
                                                        const PairDesc *__restrict__ descs,
                                                        const int *__restrict__ cin, int *arena, int *ws,
                                                        int *ws2, int *__restrict__ stats, CompactParams cp) {
    extern __shared__ int lds[];
    __shared__ int tmp[2 * CW];
    __shared__ int s_status;
    const int k = blockIdx.x;
    const CompactDesc cd = cds[k];
    const PairDesc d = descs[k];
    CPair P = make_cpair(cd, cin, ws, ws2);
    int *S = stats + (size_t)k * LCG_CS_INTS;
    const int nam = S[LCG_CS_NAM], ngroups = S[LCG_CS_NGROUPS], nblk = S[LCG_CS_NBLK];
    if (threadIdx.x == 0) s_status = 0;
    // head of the image from the host (band, codes, gaps, sigma); the rest
    // of the arena is zero, edges start empty (-1)
    for (int t = threadIdx.x; t < cd.hdr_len; t += CT) arena[d.band + t] = cin[cd.hdr + t];
    for (int t = threadIdx.x; t <= d.n + d.m; t += CT) arena[d.edges + t] = -1;
    __syncthreads();
    load_band(lds, arena + d.band, arena + d.band + d.n + 1, d.n, tmp);
    P.mn = lds;
    P.mx = lds + (P.n + 1);
    P.coff = lds + 2 * (P.n + 1);
    const int lane = lane_id();
    const int ncells = P.coff[P.n + 1];
    const int *codeA = arena + d.codes, *codeB = arena + d.codes + d.n + 1;
    const int *sig = arena + d.sigma;

    // 1. first arc-match index of every bitmap word: prefix of the row
    // counts over the arcs of A (index order) added to the in-row prefix
    {
        int carry = 0;
        for (int base = 0; base < P.NA; base += CT) {
            const int a = base + (int)threadIdx.x;
            const int v = a < P.NA ? P.rowcnt[a] : 0;
            int e, e2, t, t2;
            block_scan2(v, 0, e, e2, t, t2, tmp);
            if (a < P.NA)
                for (int w = 0; w < P.W64; ++w) P.wpre[(size_t)a * P.W64 + w] += carry + e;
            carry += t;
        }
        __syncthreads();
    }
    // 2. arc-match scores in the reference order (Scoring::arcmatch,
    // scoring.cc:440-554: tau * riboX / 100, C division, + weights)
    for (int a = threadIdx.x >> 6; a < P.NA; a += CW) {
        const int arcA = P.arcsA[a];
        for (int w = 0; w < P.W64; ++w) {
            const size_t wi = (size_t)a * P.W64 + w;
            const uint64_t word = P.bits[wi];
            const int b = 64 * w + lane;
            if ((word >> lane) & 1ull) {
                const int t = P.wpre[wi] + __popcll(word & lanes_below());
                int term = 0;
                if (P.tau != 0) {
                    if (P.ribosum) {
                        const int pa = P.ptA[a], pb = P.ptB[b];
                        if (pa >= 0 && pb >= 0) term = P.tab[pa * 16 + pb];
                    } else {
                        const int arcB = P.arcsB[b];
                        term = sig[codeA[left_of(arcA)] * LCG_NCODES + codeB[left_of(arcB)]] +
                               sig[codeA[right_of(arcA)] * LCG_NCODES + codeB[right_of(arcB)]];
                    }
                }
                if (t < nam)
                    arena[d.amscore + t] = (P.tau * term) / 100 + P.wA[a] + P.wB[b];
                else
                    s_status = 2;
            }
        }
    }
    // 3. groups, al descending then bl descending = band cells in reverse
    // order: group index and first arc match from prefix sums over the cells
    {
        int cf = 0, cc = 0;
        for (int base = 0; base < ncells; base += CT) {
            const int c = base + (int)threadIdx.x;
            GroupCell g{0, 0, 0};
            int al = 0, bl = 0;
            if (c < ncells) {
                cell_rc(P, c, al, bl);
                g = group_cell(P, al, bl);
            }
            const int flag = g.cnt > 0;
            int ef, ec, tf, tc;
            block_scan2(flag, g.cnt, ef, ec, tf, tc, tmp);
            if (flag) {
                const int gi = ngroups - 1 - (cf + ef);
                const int amb = nam - (cc + ec) - g.cnt;
                int W, L;
                long long cells;
                group_geom(P, al, bl, g.ar, g.br, W, L, cells);
                if (gi >= 0 && gi < ngroups && amb >= 0) {
                    int *G = arena + d.groups + (size_t)gi * LCG_GROUP_INTS;
                    G[0] = al | (bl << 16);
                    G[1] = g.ar | (g.br << 16);
                    G[2] = W;
                    G[3] = amb;
                    G[4] = amb + g.cnt;
                    G[5] = 0;
                    G[6] = 0;
                    G[7] = L;
                } else {
                    s_status = 2;
                }
            }
            cf += tf;
            cc += tc;
        }
        if (threadIdx.x == 0 && (cf != ngroups || cc != nam)) s_status = 2;
    }
    // 4. candidate lists by right-end cell (common_right_end_list, sorted
    // al desc then bl desc = right adjacency order of A, then of B)
    {
        int carry = 0;
        for (int base = 0; base < ncells; base += CT) {
            const int c = base + (int)threadIdx.x;
            int i = 0, j = 0, cnt = 0;
            if (c < ncells) {
                cell_rc(P, c, i, j);
                if (i >= 1 && j >= 1)
                    for (int u = P.rpA[i]; u < P.rpA[i + 1]; ++u)
                        for (int v = P.rpB[j]; v < P.rpB[j + 1]; ++v) cnt += has_bit(P, P.radjA[u], P.radjB[v]);
            }
            int e, e2, t, t2;
            block_scan2(cnt, 0, e, e2, t, t2, tmp);
            if (c < ncells) {
                int te = carry + e;
                arena[d.cptr + c] = te;
                if (cnt > 0) {
                    for (int u = P.rpA[i]; u < P.rpA[i + 1]; ++u) {
                        const int a = P.radjA[u];
                        const int al = left_of(P.arcsA[a]);
                        for (int v = P.rpB[j]; v < P.rpB[j + 1]; ++v) {
                            const int b = P.radjB[v];
                            if (!has_bit(P, a, b)) continue;
                            const int x = am_index(P, a, b);
                            if (te < nam && x < nam) {
                                int4 *E = reinterpret_cast<int4 *>(arena + d.cand) + te;
                                *E = make_int4(al | (left_of(P.arcsB[b]) << 16), j, LCG_NEG, P.mn[al - 1]);
                                P.eof[x] = te;
                            } else {
                                s_status = 2;
                            }
                            ++te;
                        }
                    }
                }
            }
            carry += t;
        }
        if (threadIdx.x == 0) {
            arena[d.cptr + ncells] = carry;
            if (carry != nam) s_status = 2;
        }
    }
    // 5. fill blocks: arcs of A in right adjacency order (right end
    // ascending, left end descending) with at least one arc match
    {
        int cf = 0, cc = 0;
        for (int base = 0; base < P.NA; base += CT) {
            const int u = base + (int)threadIdx.x;
            int a = 0, cnt = 0;
            if (u < P.NA) {
                a = P.radjA[u];
                cnt = P.rowcnt[a];
            }
            const int flag = cnt > 0;
            int ef, ec, tf, tc;
            block_scan2(flag, cnt, ef, ec, tf, tc, tmp);
            if (u < P.NA) {
                const int bi = cf + ef;
                P.sblk[u] = bi;
                if (flag && bi < nblk) {
                    arena[d.fblk + 2 * bi] = left_of(P.arcsA[a]);
                    arena[d.fblk + 2 * bi + 1] = cc + ec;
                    P.blka[bi] = a;
                    P.blks[bi] = cc + ec;
                } else if (flag) {
                    s_status = 2;
                }
            }
            cf += tf;
            cc += tc;
        }
        if (threadIdx.x == 0) {
            P.sblk[P.NA] = cf;
            arena[d.fblk + 2 * nblk] = 0;
            arena[d.fblk + 2 * nblk + 1] = nam;
            if (cf != nblk || cc != nam) s_status = 2;
        }
        __syncthreads();
    }
    // fill-list row pointers: first block of right end i
    for (int i = threadIdx.x; i <= P.n + 1; i += CT) arena[d.fptr + i] = P.sblk[P.rpA[i]];
    // fill entries of a block: arcs of B in right adjacency order
    for (int bi = threadIdx.x >> 6; bi < nblk; bi += CW) {
        const int a = P.blka[bi];
        const int p = left_of(P.arcsA[a]);
        const int mnp = P.mn[p - 1];
        int tf = P.blks[bi];
        for (int base = 0; base < P.NB; base += 64) {
            const int v = base + lane;
            const int b = v < P.NB ? P.radjB[v] : 0;
            const bool ok = v < P.NB && has_bit(P, a, b);
            const uint64_t mask = __ballot(ok);
            if (ok) {
                const int t = tf + __popcll(mask & lanes_below());
                const int x = am_index(P, a, b);
                const int arcB = P.arcsB[b];
                if (t < nam && x < nam) {
                    int *F = arena + d.fcand + 3 * (size_t)t;
                    F[0] = p | (left_of(arcB) << 16);
                    F[1] = right_of(arcB) | (mnp << 16);
                    F[2] = LCG_NEG;
                    P.fof[x] = t;
                } else {
                    s_status = 2;
                }
            }
            tf += __popcll(mask);
        }
    }
    // group pointers of the rows: rows[al] = groups with left end >= al
    for (int al = threadIdx.x; al <= P.n + 1; al += CT) {
        int lo = 0, hi = ngroups;  // first group with left end < al
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((arena[d.groups + (size_t)mid * LCG_GROUP_INTS] & 0xffff) >= al)
                lo = mid + 1;
            else
                hi = mid;
        }
        arena[d.rows + al] = lo;
    }
    __syncthreads();
    // 6. fill entry -> candidate entry; per group its arc matches in the
    // reference order (ar | br<<16, candidate entry, arc match, fill entry)
    for (int x = threadIdx.x; x < nam; x += CT) {
        const int t = P.fof[x];
        if (t >= 0 && t < nam)
            arena[d.f2c + t] = P.eof[x];
        else
            s_status = 2;
    }
    for (int gi = threadIdx.x; gi < ngroups; gi += CT) {
        const int *G = arena + d.groups + (size_t)gi * LCG_GROUP_INTS;
        const int al = G[0] & 0xffff, bl = G[0] >> 16;
        int t = G[3];
        const int te = G[4];
        for (int a = P.lA[2 * al]; a < P.lA[2 * al + 1]; ++a) {
            const int ra = right_of(P.arcsA[a]);
            for (int b = P.lB[2 * bl]; b < P.lB[2 * bl + 1]; ++b) {
                if (!has_bit(P, a, b)) continue;
                const int x = am_index(P, a, b);
                if (t < te && x < nam) {
                    int4 *Q = reinterpret_cast<int4 *>(arena + d.gams) + t;
                    *Q = make_int4(ra | (right_of(P.arcsB[b]) << 16), P.eof[x], x, P.fof[x]);
                } else {
                    s_status = 2;
                }
                ++t;
            }
        }
    }
    // 7. LDS schedule of each row (one thread per row): groups dealt to the
    // members longest first to the least loaded one; per member, largest
    // region first into the first round with a free wave and room in the
    // pool.  The same rule as the host schedule, so the same result.
    const int K = cp.members;
    for (int al = 1 + (int)threadIdx.x; al <= P.n; al += CT) {
        const int g0 = arena[d.rows + al + 1], g1 = arena[d.rows + al];
        const int cnt = g1 - g0;
        if (cnt <= 0) {
            arena[d.rounds + al] = 0;
            continue;
        }
        int *idx = P.sched + g0, *used = P.sched + ngroups + g0, *cnr = P.sched + 2 * ngroups + g0;
        int *load = P.sched + 3 * ngroups + al * K;
        int *Gt = arena + d.groups;
        auto rows_of = [&](int g) {
            return (Gt[(size_t)g * LCG_GROUP_INTS + 1] & 0xffff) - (Gt[(size_t)g * LCG_GROUP_INTS] & 0xffff);
        };
        auto mem_of = [&](int g) { return Gt[(size_t)g * LCG_GROUP_INTS + 5] >> 16; };
        auto lds_of = [&](int g) { return Gt[(size_t)g * LCG_GROUP_INTS + 7]; };
        for (int t = 0; t < cnt; ++t) idx[t] = g0 + t;
        for (int t = 1; t < cnt; ++t) {  // stable: rows descending
            const int x = idx[t], kx = rows_of(x);
            int s = t;
            while (s > 0 && rows_of(idx[s - 1]) < kx) {
                idx[s] = idx[s - 1];
                --s;
            }
            idx[s] = x;
        }
        for (int mm = 0; mm < K; ++mm) load[mm] = 0;
        for (int t = 0; t < cnt; ++t) {
            const int g = idx[t];
            int best = 0;
            for (int mm = 1; mm < K; ++mm)
                if (load[mm] < load[best]) best = mm;
            Gt[(size_t)g * LCG_GROUP_INTS + 5] = best << 16;
            load[best] += rows_of(g);
        }
        for (int t = 1; t < cnt; ++t) {  // stable: member ascending, LDS descending
            const int x = idx[t], mx_ = mem_of(x), lx = lds_of(x);
            int s = t;
            while (s > 0) {
                const int y = idx[s - 1], my = mem_of(y);
                if (my > mx_ || (my == mx_ && lds_of(y) < lx)) {
                    idx[s] = y;
                    --s;
                } else {
                    break;
                }
            }
            idx[s] = x;
        }
        int nr = 0;
        for (int t = 0; t < cnt;) {
            const int mm = mem_of(idx[t]);
            int nrounds = 0;
            for (; t < cnt && mem_of(idx[t]) == mm; ++t) {
                const int g = idx[t], L = lds_of(g);
                int r = 0;
                while (r < nrounds && (cnr[r] >= cp.waves || used[r] + L > cp.pool)) ++r;
                if (r == nrounds) {
                    used[r] = 0;
                    cnr[r] = 0;
                    ++nrounds;
                }
                Gt[(size_t)g * LCG_GROUP_INTS + 5] = r | (cnr[r] << 12) | (mm << 16);
                Gt[(size_t)g * LCG_GROUP_INTS + 6] = used[r];
                used[r] += L;
                cnr[r] += 1;
            }
            nr = max(nr, nrounds);
        }
        arena[d.rounds + al] = nr;
        if (nr >= 4096) s_status = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) S[LCG_CS_STATUS] = s_status;
}

// Runs both passes over a batch on `stream`.  The build pass needs the
// arena laid out from the counts, so the host synchronizes in between:
// lcg_launch_compact_count, then (host) layout, then lcg_launch_compact_build.
namespace {
int set_lds_limit() {
    static bool done = false;
    if (done) return 0;
    hipError_t e = hipFuncSetAttribute((const void *)lcg_compact_count, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       160 * 1024 - 256);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)lcg_compact_build, hipFuncAttributeMaxDynamicSharedMemorySize,
                                160 * 1024 - 256);
    if (e != hipSuccess) return (int)e;
    done = true;
    return 0;
}
}  // namespace

extern "C" int lcg_launch_compact_count(const CompactDesc *cds, const int *cin, int *ws, int *stats, int npairs,
                                        int lds_bytes, hipStream_t stream) {
    if (int e = set_lds_limit()) return e;
    hipLaunchKernelGGL(lcg_compact_count, dim3(npairs), dim3(CT), lds_bytes, stream, cds, cin, ws, stats);
    return (int)hipGetLastError();
}

extern "C" int lcg_launch_compact_build(const CompactDesc *cds, const PairDesc *descs, const int *cin, int *arena,
                                        int *ws, int *ws2, int *stats, CompactParams cp, int lds_bytes,
                                        hipStream_t stream) {
    if (int e = set_lds_limit()) return e;
    hipLaunchKernelGGL(lcg_compact_build, dim3(cp.npairs), dim3(CT), lds_bytes, stream, cds, descs, cin, arena, ws,
                       ws2, stats, cp);
    return (int)hipGetLastError();
}
