// Packing of prepared pairs into the device arena, launch configuration and
// the C ABI (include/locarna_amd.h).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <queue>
#include <unordered_map>
#include <string>
#include <thread>
#include <vector>

#include "band_layout.h"
#include "compact_layout.h"
#include "device_layout.h"
#include "pf_layout.h"
#include "host.hpp"
#include "locarna_amd.h"

using namespace locarna_amd;

extern "C" int lcg_launch_gather(const PairDesc *descs, const int *res_off, const int *arena, int *dst, int npairs,
                                 hipStream_t stream);
extern "C" int lcg_launch_compact_count(const CompactDesc *cds, const int *cin, int *ws, int *stats, int npairs,
                                        int lds_bytes, hipStream_t stream);
extern "C" int lcg_launch_compact_build(const CompactDesc *cds, const PairDesc *descs, const int *cin, int *arena,
                                        int *ws, int *ws2, int *stats, CompactParams cp, int lds_bytes,
                                        hipStream_t stream);

extern "C" int lcg_launch_align(const PairDesc *descs, const int *order, int *queue, int *arena, int *scratch,
                                int *barriers, BatchParams bp, int grid, int waves, size_t lds_bytes, BatchParams bp_top,
                                int top_grid,
                                int top_waves, size_t top_lds_bytes, const int *lambda, int top_only, hipEvent_t mid,
                                hipStream_t stream, const int32_t *restr = nullptr);  // int4 per pair

struct lcg_pair {
    lcg_params params;
    ScoringParams sp;
    std::shared_ptr<const RnaData> A, B;  // shared with the batches made of the pair (output writers)
    std::shared_ptr<const BasePairs> bA, bB;  // shared by the pairs of one text (lcg_pairs_prepare)
    std::unique_ptr<TraceController> tc;
    std::unique_ptr<ArcMatches> am;
    std::unique_ptr<Scoring> sc;
    // packed device image, offsets relative to the blob start
    std::vector<int32_t> blob;
    PairDesc desc;
    int64_t cells = 0;  // DP cells (groups + top level) as a work estimate
    int ecap = 0;       // max columns of a group region (E vector)
    std::vector<int> entry_of;  // arc match -> candidate entry
    // prepared for the device compaction (lcg_pair_prepare_device): no arc
    // matches on the host, no blob; the compaction inputs instead
    bool deferred = false;
    std::vector<int32_t> cimg;  // compaction inputs, offsets of cdesc relative to it
    CompactDesc cdesc;
};

namespace {

void set_err(char *err, int errlen, const std::string &msg) {
    if (err && errlen > 0) {
        std::strncpy(err, msg.c_str(), errlen - 1);
        err[errlen - 1] = 0;
    }
}

// Appends an array of `count` ints to the blob; returns its offset.
int32_t reserve(std::vector<int32_t> &blob, size_t count, int32_t fill = 0) {
    size_t off = blob.size();
    blob.resize(off + count, fill);
    return (int32_t)off;
}

// Offsets of a pair's arena image (device_layout.h) from its sizes, in the
// order the host packing reserves them; the device compaction lays out its
// pairs with it, so both produce the same image.  Returns the image size.
size_t layout_pair(PairDesc &d, int n, int m, int ncells, int ngroups, int nam, int nfblk, bool nolp = false,
                   bool stacking = false, bool table = false) {
    size_t off = 0;
    auto take = [&](size_t count) {
        const int32_t o = (int32_t)off;
        off += count;
        return o;
    };
    auto align = [&](size_t a) { off = (off + a - 1) / a * a; };
    d.band = take(3 * (size_t)(n + 1) + 1);
    d.codes = take((size_t)n + m + 2);
    d.gaps = take((size_t)n + m + 2);
    d.sigma = take(table ? (size_t)(n + 1) * (m + 1) : LCG_NCODES * LCG_NCODES);
    d.rows = take((size_t)n + 2);
    d.rounds = take((size_t)n + 2);
    d.groups = take((size_t)LCG_GROUP_INTS * ngroups);
    align(4);
    d.gams = take(4 * (size_t)nam);
    d.amscore = take(nam);
    d.cptr = take((size_t)ncells + 1);
    align(4);
    d.cand = take(4 * (size_t)nam);
    d.fptr = take((size_t)n + 2);
    d.fcand = take(3 * (size_t)nam);
    d.fblk = take(2 * (size_t)nfblk);
    d.f2c = take(nam);
    d.D = take(nam);
    d.edges = take((size_t)n + m + 1);
    d.structs = take((size_t)n + m);
    d.tstack = take(2 * (size_t)(n + 1));
    align(2);
    d.out = take(12);
    d.inner = take(nolp || stacking ? nam : 0);
    d.c2a = take(nolp || stacking ? nam : 0);
    d.amst = take(stacking ? nam : 0);
    align(4);
    return off;
}

// Moves every array offset of a pair's image (PairDesc fields that
// layout_pair sets) by `base`, the image's place in the batch arena.
void rebase(PairDesc &d, int32_t base) {
    int32_t *offs[] = {&d.band, &d.codes, &d.gaps, &d.sigma, &d.rows, &d.groups, &d.gams, &d.amscore,
                       &d.cptr, &d.cand, &d.D, &d.edges, &d.structs, &d.tstack, &d.out, &d.rounds,
                       &d.fptr, &d.fblk, &d.fcand, &d.f2c, &d.inner, &d.c2a, &d.amst};
    for (int32_t *o : offs) *o += base;
}

// A rebased image lies in the arena at [base, base + size): every array
// offset in that range and, with --noLP, its arc-match indices in range
// (checked on the host before any launch).
bool image_in_place(const PairDesc &d, int32_t base, int32_t size, const std::vector<int32_t> &arena,
                    bool stacking = false) {
    const int32_t offs[] = {d.band, d.codes, d.gaps, d.sigma, d.rows, d.groups, d.gams, d.amscore, d.cptr, d.cand,
                            d.D, d.edges, d.structs, d.tstack, d.out, d.rounds, d.fptr, d.fblk, d.fcand, d.f2c,
                            d.inner, d.c2a, d.amst};
    for (int32_t o : offs)
        if (o < base || o > base + size) return false;
    if (stacking && d.amst + d.nam > base + size) return false;  // --stacking scores present
    if (d.c2a > d.inner) {  // --noLP / --stacking arrays present
        if (d.inner + d.nam > base + size || d.c2a + d.nam > base + size) return false;
        for (int k = 0; k < d.nam; ++k) {
            if (arena[d.inner + k] < -1 || arena[d.inner + k] >= d.nam) return false;
            if (arena[d.c2a + k] < 0 || arena[d.c2a + k] >= d.nam) return false;
        }
    }
    return true;
}

// Top-level region of a pair: stored row width and LDS ints; adds the
// top-level DP cells to `cells`.
void top_region(const std::vector<int> &mn, const std::vector<int> &mx, int n, int m, int &wtop, int &top_ints,
                int64_t &cells) {
    wtop = 0;
    for (int i = 0; i <= n; ++i) {
        const int lo = (i == 0) ? 0 : std::max(0, mn[i]);
        const int hi = std::min(m, mx[i]);
        wtop = std::max(wtop, std::max(0, hi - lo + 1));
        if (i > 0) cells += std::max(0, std::min(m, mx[i]) - std::max(1, mn[i]) + 1);
    }
    top_ints = ((m + 1) + wtop + (n + 1) * wtop + 3) & ~3;
}

void pack(lcg_pair &P) {
    const TraceController &tc = *P.tc;
    const Scoring &sc = *P.sc;
    const BasePairs &bA = *P.bA, &bB = *P.bB;
    const ArcMatches &ams = *P.am;
    const int n = P.A->length(), m = P.B->length();
    const int nam = ams.num_arc_matches();
    auto &blob = P.blob;
    PairDesc d;
    std::memset(&d, 0, sizeof d);
    d.n = n;
    d.m = m;
    d.nam = nam;
    // band + band cell offsets
    d.band = reserve(blob, 3 * (size_t)(n + 1) + 1);
    int ncells = 0;
    for (int i = 0; i <= n; ++i) {
        blob[d.band + i] = tc.min_col(i);
        blob[d.band + n + 1 + i] = tc.max_col(i);
        blob[d.band + 2 * (n + 1) + i] = ncells;
        // rows of a band can be empty (min_col > max_col), e.g. at the ends
        // of a sequence-local band (trace_controller.cc:565-597)
        ncells += std::max(0, tc.max_col(i) - tc.min_col(i) + 1);
    }
    blob[d.band + 3 * (n + 1)] = ncells;
    d.ncells = ncells;
    d.codes = reserve(blob, (size_t)n + m + 2);
    for (int i = 1; i <= n; ++i) blob[d.codes + i] = sc.codeA(i);
    for (int j = 1; j <= m; ++j) blob[d.codes + n + 1 + j] = sc.codeB(j);
    d.gaps = reserve(blob, (size_t)n + m + 2);
    for (int i = 1; i <= n; ++i) blob[d.gaps + i] = (int32_t)sc.gapA(i);
    for (int j = 1; j <= m; ++j) blob[d.gaps + n + 1 + j] = (int32_t)sc.gapB(j);
    // sigma: by symbol codes, or for profiles by column pair (row i at
    // sigma + i * (m + 1))
    const bool table = sc.profile();
    if (table) {
        d.sigma = reserve(blob, (size_t)(n + 1) * (m + 1));
        const std::vector<score_t> &st = sc.sigma_table();
        for (size_t x = 0; x < st.size(); ++x) blob[d.sigma + x] = (int32_t)st[x];
    } else {
        d.sigma = reserve(blob, LCG_NCODES * LCG_NCODES);
        for (int a = 0; a < LCG_NCODES; ++a)
            for (int b = 0; b < LCG_NCODES; ++b) blob[d.sigma + a * LCG_NCODES + b] = (int32_t)sc.sigma_code(a, b);
    }

    // arc matches by left ends (groups) and by right ends (candidates)
    std::vector<std::vector<int>> by_left((size_t)(n + 1) * (m + 1));
    std::vector<std::vector<int>> by_right((size_t)(n + 1) * (m + 1));
    for (int k = 0; k < nam; ++k) {
        const ArcMatch &x = ams.arcmatch(k);
        const Arc &a = bA.arc(x.arcA), &b = bB.arc(x.arcB);
        by_left[(size_t)a.left * (m + 1) + b.left].push_back(k);
        by_right[(size_t)a.right * (m + 1) + b.right].push_back(k);
    }
    // --noLP: the arc match of the arcs directly inside both arcs, the first
    // of its left-end list (arc_matches.cc:50-74); the group (al, bl) then
    // computes the D of the arc matches with left ends (al-1, bl-1) that have
    // one, over the inner arc match's region (aligner.cc:612-657,
    // arc_matches.cc:313-355)
    const bool nolp = P.params.no_lonely_pairs != 0;
    // --stacking: the inner arc match of every arc match too, for the
    // stacked alternative of its D (aligner.cc:574-610)
    const bool stacking = P.params.stacking != 0;
    std::vector<int> inner(nolp || stacking ? nam : 0, -1);
    std::vector<std::vector<int>> by_group;
    if (nolp || stacking) {
        for (int k = 0; k < nam; ++k) {
            const ArcMatch &x = ams.arcmatch(k);
            const Arc &a = bA.arc(x.arcA), &b = bB.arc(x.arcB);
            if (a.left + 1 > n || b.left + 1 > m) continue;
            for (int y : by_left[(size_t)(a.left + 1) * (m + 1) + (b.left + 1)]) {
                const ArcMatch &z = ams.arcmatch(y);
                if (bA.arc(z.arcA).right == a.right - 1 && bB.arc(z.arcB).right == b.right - 1) {
                    inner[k] = y;
                    break;
                }
            }
        }
    }
    if (nolp) {
        by_group.assign((size_t)(n + 1) * (m + 1), {});
        for (int k = 0; k < nam; ++k) {
            if (inner[k] < 0) continue;
            const ArcMatch &x = ams.arcmatch(k);
            by_group[(size_t)(bA.arc(x.arcA).left + 1) * (m + 1) + (bB.arc(x.arcB).left + 1)].push_back(k);
        }
    }
    const auto &group_lists = nolp ? by_group : by_left;
    // groups, al descending, bl descending (aligner.cc:659-732)
    struct G {
        int al, bl, ar, br, W, b, e, lds;
    };
    std::vector<G> groups;
    std::vector<int32_t> gams;
    std::vector<int32_t> rows(n + 2, 0);
    int64_t cells = 0;
    int ecap = 0, gbuf = 0, slmax = 0;
    const bool sl = P.params.struct_local != 0;
    auto row_width = [&](int i, int al, int bl, int br) {
        int lo = (i == al) ? bl : std::max(bl, tc.min_col(i));
        int hi = std::min(br - 1, tc.max_col(i));
        return std::max(0, hi - lo + 1);
    };
    rows[n + 1] = 0;
    for (int al = n; al >= 1; --al) {
        const int max_bl = std::min(m, tc.max_col(al)), min_bl = std::max(1, tc.min_col(al));
        for (int bl = max_bl; bl >= min_bl; --bl) {
            if (!tc.is_valid_match(al, bl)) continue;
            const auto &lst = group_lists[(size_t)al * (m + 1) + bl];
            if (lst.empty()) continue;
            // right ends of the region's arc match: the arc match itself, or
            // with --noLP its inner one
            const int dr = nolp ? 1 : 0;
            int ar = al, br = bl;
            for (int k : lst) {
                const ArcMatch &x = ams.arcmatch(k);
                ar = std::max(ar, bA.arc(x.arcA).right - dr);
                br = std::max(br, bB.arc(x.arcB).right - dr);
            }
            int W = 0;
            for (int i = al; i < ar; ++i) {
                int w = row_width(i, al, bl, br);
                W = std::max(W, w);
                if (i > al) cells += std::max(0, std::min(br - 1, tc.max_col(i)) - std::max(bl + 1, tc.min_col(i)) + 1);
            }
            // LDS region E | S | M, rounded to 4 ints; --struct-local: no LDS,
            // eight matrices in global scratch
            const int lds = sl ? 0 : ((br - bl) + W + (ar - al) * W + 3) & ~3;
            if (sl) slmax = std::max(slmax, ((br - bl) + W + 8 * (ar - al) * W + 3) & ~3);
            G g{al, bl, ar, br, W, (int)gams.size() / 4, 0, lds};
            for (int k : lst) {
                const ArcMatch &x = ams.arcmatch(k);
                gams.push_back((bA.arc(x.arcA).right - dr) | ((bB.arc(x.arcB).right - dr) << 16));
                gams.push_back(0);  // candidate entry, set below
                gams.push_back(k);
                gams.push_back(0);
            }
            g.e = (int)gams.size() / 4;
            groups.push_back(g);
            ecap = std::max(ecap, br - bl);
            gbuf = std::max(gbuf, lds);
        }
        rows[al] = (int)groups.size();
    }
    rows[0] = (int)groups.size();
    d.ngroups = (int)groups.size();
    d.rows = reserve(blob, rows.size());
    std::copy(rows.begin(), rows.end(), blob.begin() + d.rows);
    d.rounds = reserve(blob, rows.size());  // filled by the batch's LDS schedule
    d.groups = reserve(blob, (size_t)LCG_GROUP_INTS * groups.size());
    for (size_t g = 0; g < groups.size(); ++g) {
        int32_t *G8 = &blob[d.groups + LCG_GROUP_INTS * g];
        G8[0] = groups[g].al | (groups[g].bl << 16);
        G8[1] = groups[g].ar | (groups[g].br << 16);
        G8[2] = groups[g].W;
        G8[3] = groups[g].b;
        G8[4] = groups[g].e;
        G8[7] = groups[g].lds;
    }
    while (blob.size() & 3) blob.push_back(0);
    d.gams = reserve(blob, 4 * (size_t)nam);  // (--noLP: only the arc matches with an inner one have an entry)
    d.amscore = reserve(blob, nam);
    for (int k = 0; k < nam; ++k) blob[d.amscore + k] = (int32_t)sc.arcmatch(k);
    // candidate CSR over band cells, each list sorted (al desc, bl desc);
    // entries (al | bl<<16, j, D, min_col(al - 1)) are 16-byte aligned; the
    // last field locates row al - 1 of a group region without a lookup
    d.cptr = reserve(blob, (size_t)ncells + 1);
    while (blob.size() & 3) blob.push_back(0);
    d.cand = reserve(blob, 4 * (size_t)nam);
    std::vector<int> entry_of(nam, -1);
    int t = 0, next_cell = 0;
    for (int i = 0; i <= n; ++i) {
        for (int j = tc.min_col(i); j <= tc.max_col(i); ++j) {
            const int cell = blob[d.band + 2 * (n + 1) + i] + (j - tc.min_col(i));
            if (cell != next_cell++) throw failure("internal: band cell numbering");
            blob[d.cptr + cell] = t;
            if (i == 0 || j == 0) continue;
            auto lst = by_right[(size_t)i * (m + 1) + j];
            std::sort(lst.begin(), lst.end(), [&](int x, int y) {
                const Arc &ax = bA.arc(ams.arcmatch(x).arcA), &ay = bA.arc(ams.arcmatch(y).arcA);
                const Arc &bx = bB.arc(ams.arcmatch(x).arcB), &by = bB.arc(ams.arcmatch(y).arcB);
                return ax.left > ay.left || (ax.left == ay.left && bx.left > by.left);
            });
            for (int k : lst) {
                const ArcMatch &x = ams.arcmatch(k);
                blob[d.cand + 4 * t] = bA.arc(x.arcA).left | (bB.arc(x.arcB).left << 16);
                blob[d.cand + 4 * t + 1] = j;
                blob[d.cand + 4 * t + 2] = LCG_NEG;
                blob[d.cand + 4 * t + 3] = tc.min_col(bA.arc(x.arcA).left - 1);
                entry_of[k] = t;
                ++t;
            }
        }
    }
    if (next_cell != ncells) throw failure("internal: band cell count");
    blob[d.cptr + ncells] = t;
    if (t != nam) throw failure("internal: arc match right-end lists incomplete");
    // fill-D candidate list: the arc matches of row i (right end in A),
    // blocks of one arc of A by left end p descending, j ascending in a block
    std::vector<std::vector<int>> by_rightA(n + 1);
    for (int k = 0; k < nam; ++k) by_rightA[bA.arc(ams.arcmatch(k).arcA).right].push_back(k);
    d.fptr = reserve(blob, (size_t)n + 2);
    d.fcand = reserve(blob, 3 * (size_t)nam);
    std::vector<int32_t> blk;
    std::vector<int> fentry_of(nam, -1);
    int tf = 0;
    for (int i = 0; i <= n; ++i) {
        blob[d.fptr + i] = (int32_t)(blk.size() / 2);
        auto lst = by_rightA[i];
        std::sort(lst.begin(), lst.end(), [&](int x, int y) {
            const ArcMatch &ax = ams.arcmatch(x), &ay = ams.arcmatch(y);
            const int px = bA.arc(ax.arcA).left, py = bA.arc(ay.arcA).left;
            if (px != py) return px > py;
            const int jx = bB.arc(ax.arcB).right, jy = bB.arc(ay.arcB).right;
            if (jx != jy) return jx < jy;
            return bB.arc(ax.arcB).left > bB.arc(ay.arcB).left;
        });
        int last_p = -1;
        for (int k : lst) {
            const ArcMatch &x = ams.arcmatch(k);
            const int p = bA.arc(x.arcA).left, q = bB.arc(x.arcB).left, j = bB.arc(x.arcB).right;
            if (p != last_p) {
                blk.push_back(p);
                blk.push_back(tf);
                last_p = p;
            }
            blob[d.fcand + 3 * tf] = p | (q << 16);
            blob[d.fcand + 3 * tf + 1] = j | (tc.min_col(p - 1) << 16);
            blob[d.fcand + 3 * tf + 2] = LCG_NEG;
            fentry_of[k] = tf++;
        }
    }
    blob[d.fptr + n + 1] = (int32_t)(blk.size() / 2);
    blk.push_back(0);
    blk.push_back(tf);
    d.nfblk = (int32_t)(blk.size() / 2);
    d.fblk = reserve(blob, blk.size());
    std::copy(blk.begin(), blk.end(), blob.begin() + d.fblk);
    for (size_t q = 0; q < gams.size(); q += 4) {
        gams[q + 1] = entry_of[gams[q + 2]];
        gams[q + 3] = fentry_of[gams[q + 2]];
    }
    d.f2c = reserve(blob, nam);
    for (int k = 0; k < nam; ++k) blob[d.f2c + fentry_of[k]] = entry_of[k];
    P.entry_of = entry_of;
    std::copy(gams.begin(), gams.end(), blob.begin() + d.gams);
    // every D is written by the fill, except with --noLP those of the arc
    // matches without an inner one, which stay minus infinity
    d.D = reserve(blob, nam, nolp ? LCG_NEG : 0);
    d.edges = reserve(blob, (size_t)n + m + 1, -1);
    d.structs = reserve(blob, (size_t)n + m);
    d.tstack = reserve(blob, 2 * (size_t)(n + 1));
    int wtop = 0;
    for (int i = 0; i <= n; ++i) {
        wtop = std::max(wtop, row_width(i, 0, 0, m + 1));
        if (i > 0) cells += std::max(0, std::min(m, tc.max_col(i)) - std::max(1, tc.min_col(i)) + 1);
    }
    d.wtop = wtop;
    d.top_ints = ((m + 1) + wtop + (n + 1) * wtop + 3) & ~3;
    // out[4..11] holds four uint64 counters (cells and candidates of the
    // fill-D kernel, then of the top-level kernel): keep them 8-byte aligned (every
    // blob is a multiple of 4 ints, so blob offsets survive concatenation)
    if (blob.size() & 1) blob.push_back(0);
    d.out = reserve(blob, 12);
    d.inner = reserve(blob, nolp || stacking ? nam : 0);
    d.c2a = reserve(blob, nolp || stacking ? nam : 0);
    if (nolp || stacking) {
        std::copy(inner.begin(), inner.end(), blob.begin() + d.inner);
        for (int k = 0; k < nam; ++k) blob[d.c2a + entry_of[k]] = k;
    }
    d.amst = reserve(blob, stacking ? nam : 0);
    if (stacking)
        for (int k = 0; k < nam; ++k) blob[d.amst + k] = sc.stackable(k) ? (int32_t)sc.arcmatch_stacked(k) : LCG_NEG;
    while (blob.size() & 3) blob.push_back(0);
    d.gmax = gbuf;
    d.slmax = slmax;
    {
        PairDesc L = d;
        const size_t size = layout_pair(L, n, m, ncells, d.ngroups, nam, d.nfblk, nolp, stacking, table);
        if (size != blob.size() || std::memcmp(&L, &d, sizeof d) != 0) throw failure("internal: arena layout");
    }
    P.desc = d;
    P.cells = cells;
    P.ecap = ecap;
}

// Head of the arena image of a pair (offsets of d relative to the image):
// band (min_col, max_col, band cell offsets), codes, gap scores, sigma.
void write_header(const lcg_pair &P, const PairDesc &d, int32_t *img) {
    const TraceController &tc = *P.tc;
    const Scoring &sc = *P.sc;
    const int n = d.n, m = d.m;
    int ncells = 0;
    for (int i = 0; i <= n; ++i) {
        img[d.band + i] = tc.min_col(i);
        img[d.band + n + 1 + i] = tc.max_col(i);
        img[d.band + 2 * (n + 1) + i] = ncells;
        ncells += std::max(0, tc.max_col(i) - tc.min_col(i) + 1);
    }
    img[d.band + 3 * (n + 1)] = ncells;
    for (int i = 1; i <= n; ++i) img[d.codes + i] = sc.codeA(i);
    for (int j = 1; j <= m; ++j) img[d.codes + n + 1 + j] = sc.codeB(j);
    for (int i = 1; i <= n; ++i) img[d.gaps + i] = (int32_t)sc.gapA(i);
    for (int j = 1; j <= m; ++j) img[d.gaps + n + 1 + j] = (int32_t)sc.gapB(j);
    for (int a = 0; a < LCG_NCODES; ++a)
        for (int b = 0; b < LCG_NCODES; ++b) img[d.sigma + a * LCG_NCODES + b] = (int32_t)sc.sigma_code(a, b);
}

// Inputs of the device compaction of one pair: arcs in index order with
// their weights and base-pair types, the arcs of each left end (a range of
// the index order) and of each right end (right adjacency order), the band
// and the riboX table.
void prepare_compaction_inputs(lcg_pair &P) {
    const TraceController &tc = *P.tc;
    const Scoring &sc = *P.sc;
    const BasePairs &bA = *P.bA, &bB = *P.bB;
    const int n = P.A->length(), m = P.B->length();
    const int NA = bA.num_bps(), NB = bB.num_bps();
    auto &img = P.cimg;
    img.clear();
    CompactDesc c;
    std::memset(&c, 0, sizeof c);
    c.n = n;
    c.m = m;
    c.NA = NA;
    c.NB = NB;
    c.W64 = (NB + 63) / 64;
    auto side = [&](const BasePairs &bp, const RnaData &rna, int len, bool isA) {
        const int N = bp.num_bps();
        const int32_t arcs = reserve(img, N), w = reserve(img, N), pt = reserve(img, N);
        const int32_t la = reserve(img, 2 * (size_t)(len + 2)), rp = reserve(img, (size_t)len + 2);
        const int32_t radj = reserve(img, N);
        for (int a = 0; a < N; ++a) {
            const Arc &x = bp.arc(a);
            img[arcs + a] = x.left | (x.right << 16);
            img[w + a] = (int32_t)(isA ? sc.weightA(a) : sc.weightB(a));
            const int u = nt_code(rna.nt(x.left)), v = nt_code(rna.nt(x.right));
            img[pt + a] = (u >= 0 && v >= 0) ? u * 4 + v : -1;
        }
        for (int i = 0; i <= len + 1; ++i) {
            const auto &lst = bp.left_adj(i);
            const int b0 = lst.empty() ? 0 : lst.front();
            for (size_t t = 0; t < lst.size(); ++t)
                if (lst[t] != b0 + (int)t) throw failure("internal: arcs of a left end not contiguous");
            img[la + 2 * i] = b0;
            img[la + 2 * i + 1] = b0 + (int)lst.size();
        }
        int u = 0;
        for (int i = 0; i <= len + 1; ++i) {
            img[rp + i] = u;
            if (i <= len)
                for (int a : bp.right_adj(i)) img[radj + u++] = a;
        }
        if (u != N) throw failure("internal: right adjacency incomplete");
        if (isA) {
            c.arcsA = arcs, c.wA = w, c.ptA = pt, c.lA = la, c.rpA = rp, c.radjA = radj;
        } else {
            c.arcsB = arcs, c.wB = w, c.ptB = pt, c.lB = la, c.rpB = rp, c.radjB = radj;
        }
    };
    side(bA, *P.A, n, true);
    side(bB, *P.B, m, false);
    c.band = reserve(img, 2 * (size_t)(n + 1));
    for (int i = 0; i <= n; ++i) {
        img[c.band + i] = tc.min_col(i);
        img[c.band + n + 1 + i] = tc.max_col(i);
    }
    {
        PairDesc L;
        std::memset(&L, 0, sizeof L);
        layout_pair(L, n, m, 0, 0, 0, 1);
        c.hdr_len = L.rows;  // band, codes, gaps, sigma are the head of the image
        c.hdr = reserve(img, c.hdr_len);
        L.n = n;
        L.m = m;
        write_header(P, L, img.data() + c.hdr);
    }
    c.tab = reserve(img, 256);
    c.tau = P.params.tau;
    c.ribosum = P.params.use_ribosum != 0;
    if (c.tau != 0 && c.ribosum)
        for (int x = 0; x < 256; ++x)
            img[c.tab + x] = (int32_t)Scoring::ribosum_arc_term((x >> 6) & 3, (x >> 4) & 3, (x >> 2) & 3, x & 3);
    // ArcMatches without max-diff-am / max-diff-at-am (arc_matches.cc:19-34)
    // is_valid_arcmatch limits (locarna.cc:606-611: -1 means the longer length)
    c.dlen = P.params.max_diff_am < 0 ? std::max(n, m) : P.params.max_diff_am;
    c.dend = P.params.max_diff_at_am < 0 ? std::max(n, m) : P.params.max_diff_at_am;
    while (img.size() & 3) img.push_back(0);
    P.cdesc = c;
}

}  // namespace

namespace {
// Stage 1 of the preparation of a pair: both RnaData (pp reader), the
// score parameters and the unrestricted band.
// A and B, when given, are the already parsed inputs (shared between the
// pairs of one lcg_pairs_prepare call that name the same text)
void prepare_parse(const char *ppA, const char *ppB, const lcg_params *p, lcg_pair &P,
                   std::shared_ptr<const RnaData> A = nullptr, std::shared_ptr<const RnaData> B = nullptr) {
    P.params = *p;
    if (p->sequ_local && p->free_endgaps)  // src/locarna.cc:375-378
        throw failure("Free endgaps cannot be combined with local alignment.");
    if (p->free_endgaps & ~15) throw failure("free end gaps: bits 1, 2, 4, 8 only");
    if (p->stacking & ~3) throw failure("stacking: bits 1 (--stacking) and 2 (--new-stacking) only");
    if (p->struct_local && (p->no_lonely_pairs || p->stacking))
        throw failure("--struct-local with --noLP or stacking is not supported");
    P.A = A ? A : std::make_shared<const RnaData>(ppA, p->min_prob, p->max_bp_span, p->stacking != 0);
    P.B = B ? B : std::make_shared<const RnaData>(ppB, p->min_prob, p->max_bp_span, p->stacking != 0);
    if (P.A->length() >= 0xffff || P.B->length() >= 0xffff) throw failure("sequence too long");
    ScoringParams &sp = P.sp;
    sp.indel = p->indel;
    sp.indel_opening = p->indel_opening;
    sp.struct_weight = p->struct_weight;
    sp.tau = p->tau;
    sp.match = p->match;
    sp.mismatch = p->mismatch;
    sp.temperature_alipf = p->temperature_alipf;
    sp.use_ribosum = p->use_ribosum != 0;
    sp.exp_prob = p->exp_prob;
    sp.stacking = p->stacking;
    P.tc.reset(new TraceController(P.A->length(), P.B->length()));
    if (p->max_diff >= 0) P.tc->constrain_wo_ref(p->max_diff);  // locarna.cc:568-569
}

// Stage 2 on the host: the trace-probability band in the precision of the
// reference (long double for locarna, double for locarna_p).
void prepare_band_host(lcg_pair &P) {
    const lcg_params *p = &P.params;
    restrict_trace_by_probabilities(*P.tc, *P.A, *P.B, P.sp, p->min_trace_prob, p->band_extended != 0,
                                    p->sequ_local != 0, p->free_endgaps);
}

// Largest magnitude of any alignment score of pair P or partial sum of one
// (M, E, F, D values).  An alignment of lengths n and m has at most n + m
// edges, each adding one base-match or gap score and at most one gap
// opening, and at most min(n, m) / 2 arc matches, each adding its score.
// The kernels compute in int32 with minus infinity at -2^29 and read values
// below -2^28 as minus infinity, where the reference's InftyInt is 64-bit
// (infty_int.hh:56-591), so pairs whose bound exceeds 2^27 are refused
// rather than narrowed.  For pairs compacted on the device the arc-match
// scores are not formed on the host; their bound is that of the weights and
// the tau term (scoring.cc:440-554).
int64_t score_bound(const lcg_pair &P) {
    const Scoring &sc = *P.sc;
    const int64_t n = P.A->length(), m = P.B->length();
    int64_t smax = 0, gmax = 0, amax = 0;
    smax = sc.max_abs_sigma();
    for (int i = 1; i <= n; ++i) gmax = std::max<int64_t>(gmax, std::llabs(sc.gapA(i)));
    for (int j = 1; j <= m; ++j) gmax = std::max<int64_t>(gmax, std::llabs(sc.gapB(j)));
    if (P.deferred) {
        int64_t wa = 0, wb = 0, seq = 2 * smax;
        for (int k = 0; k < P.bA->num_bps(); ++k) wa = std::max<int64_t>(wa, std::llabs(sc.weightA(k)));
        for (int k = 0; k < P.bB->num_bps(); ++k) wb = std::max<int64_t>(wb, std::llabs(sc.weightB(k)));
        for (int x = 0; x < 256; ++x)
            seq = std::max<int64_t>(seq, std::llabs(Scoring::ribosum_arc_term((x >> 6) & 3, (x >> 4) & 3,
                                                                                (x >> 2) & 3, x & 3)));
        amax = wa + wb + (std::llabs((int64_t)P.sp.tau) * seq) / 100 + 1;
    } else {
        for (int k = 0; k < P.am->num_arc_matches(); ++k) {
            amax = std::max<int64_t>(amax, std::llabs(sc.arcmatch(k)));
            if (sc.stackable(k)) amax = std::max<int64_t>(amax, std::llabs(sc.arcmatch_stacked(k)));
        }
    }
    const int64_t alpha = std::llabs((int64_t)P.sp.indel_opening);
    // --struct-local: two excluded blocks inside every arc match and the
    // exclusion score in the first row and column of the regions
    if (P.params.struct_local) amax += 3 * std::llabs((int64_t)P.params.exclusion);
    return (n + m) * (smax + gmax + alpha) + (std::min(n, m) / 2) * amax;
}

const int64_t LCG_SCORE_BOUND = (int64_t)1 << 27;

void check_score_range(const lcg_pair &P) {
    const int64_t b = score_bound(P);
    if (b > LCG_SCORE_BOUND)
        throw capacity_failure("scores out of range: alignment scores of this pair may reach " + std::to_string(b) +
                               " in magnitude; the kernels' int32 scores hold at most 2^27");
}

// Stage 3: base pairs, arc matches (host) or compaction inputs (device),
// score tables and the device image.
// bA and bB, when given, are the base pairs of the inputs (shared between
// the pairs of one lcg_pairs_prepare call that name the same text)
void prepare_finish(lcg_pair &P, bool deferred, std::shared_ptr<const BasePairs> bA = nullptr,
                    std::shared_ptr<const BasePairs> bB = nullptr) {
    const lcg_params *p = &P.params;
    P.bA = bA ? bA : std::make_shared<const BasePairs>(P.A.get(), p->min_prob);
    P.bB = bB ? bB : std::make_shared<const BasePairs>(P.B.get(), p->min_prob);
    if (deferred) {
        if (P.A->num_rows() > 1 || P.B->num_rows() > 1)  // the compaction reads symbol codes
            throw failure("alignment profiles need the arc matches enumerated on the host");
        P.am.reset(new ArcMatches(*P.bA, *P.bB));
        P.sc.reset(new Scoring(*P.A, *P.B, *P.am, P.sp));
        P.deferred = true;
        std::memset(&P.desc, 0, sizeof P.desc);
        P.desc.n = P.A->length();
        P.desc.m = P.B->length();
        P.desc.nam = P.desc.ngroups = -1;
        check_score_range(P);
        prepare_compaction_inputs(P);
    } else {
        P.am.reset(new ArcMatches(*P.bA, *P.bB, *P.tc, p->max_diff_am, p->max_diff_at_am));
        P.sc.reset(new Scoring(*P.A, *P.B, *P.am, P.sp));
        check_score_range(P);
        pack(P);
    }
}

int prepare(const char *ppA, const char *ppB, const lcg_params *p, lcg_pair **out, char *err, int errlen,
            bool deferred) {
    if (!ppA || !ppB || !p || !out) return LCG_E_ARG;
    if (deferred && p->no_lonely_pairs) {  // the device compaction builds no --noLP groups
        set_err(err, errlen, "--noLP needs the arc matches enumerated on the host");
        return LCG_E_ARG;
    }
    if (deferred && p->stacking) {  // nor the inner arc matches and stacked scores
        set_err(err, errlen, "--stacking needs the arc matches enumerated on the host");
        return LCG_E_ARG;
    }
    if (deferred && p->struct_local) {  // nor the eight-state regions
        set_err(err, errlen, "--struct-local needs the arc matches enumerated on the host");
        return LCG_E_ARG;
    }
    *out = nullptr;
    try {
        auto P = std::make_unique<lcg_pair>();
        prepare_parse(ppA, ppB, p, *P);
        prepare_band_host(*P);
        prepare_finish(*P, deferred);
        *out = P.release();
        return LCG_OK;
    } catch (capacity_failure &f) {
        set_err(err, errlen, f.what());
        return LCG_E_CAPACITY;
    } catch (failure &f) {
        set_err(err, errlen, f.what());
        return LCG_E_INPUT;
    } catch (std::exception &e) {
        set_err(err, errlen, e.what());
        return LCG_E_INPUT;
    }
}
}  // namespace

extern "C" int lcg_pair_prepare(const char *ppA, const char *ppB, const lcg_params *p, lcg_pair **out, char *err,
                                int errlen) {
    return prepare(ppA, ppB, p, out, err, errlen, false);
}

extern "C" int lcg_pair_prepare_device(const char *ppA, const char *ppB, const lcg_params *p, lcg_pair **out,
                                       char *err, int errlen) {
    return prepare(ppA, ppB, p, out, err, errlen, true);
}

extern "C" int lcg_pair_prepare_scored(const char *ppA, const char *ppB, const lcg_params *p,
                                       const char *arcmatch_scores, int32_t probability_scale, lcg_pair **out,
                                       char *err, int errlen) {
    if (!ppA || !ppB || !p || !arcmatch_scores || !out) return LCG_E_ARG;
    *out = nullptr;
    // explicit arc-match scores take no stacked scores (the reference's
    // Scoring::arcmatch ignores stacking for them, scoring.cc:541-546)
    if (p->stacking) {
        set_err(err, errlen, "--stacking needs arc matches from the pp input, not an arc-match scores file");
        return LCG_E_ARG;
    }
    try {
        auto P = std::make_unique<lcg_pair>();
        prepare_parse(ppA, ppB, p, *P);
        prepare_band_host(*P);
        const int n = P->A->length(), m = P->B->length();
        const std::vector<ExplicitArcMatch> lines = parse_arcmatch_scores(arcmatch_scores, n, m, probability_scale);
        // arcs of the file, as sets (sorted, unique)
        std::vector<std::pair<int, int>> arcsA, arcsB;
        for (const ExplicitArcMatch &x : lines) {
            arcsA.push_back({x.i, x.j});
            arcsB.push_back({x.k, x.l});
        }
        for (auto *v : {&arcsA, &arcsB}) {
            std::sort(v->begin(), v->end());
            v->erase(std::unique(v->begin(), v->end()), v->end());
        }
        P->bA.reset(new BasePairs(n, arcsA));
        P->bB.reset(new BasePairs(m, arcsB));
        P->am.reset(new ArcMatches(*P->bA, *P->bB, *P->tc, lines, p->max_diff_am, p->max_diff_at_am));
        P->sc.reset(new Scoring(*P->A, *P->B, *P->am, P->sp));
        check_score_range(*P);
        pack(*P);
        *out = P.release();
        return LCG_OK;
    } catch (capacity_failure &f) {
        set_err(err, errlen, f.what());
        return LCG_E_CAPACITY;
    } catch (failure &f) {
        set_err(err, errlen, f.what());
        return LCG_E_INPUT;
    } catch (std::exception &e) {
        set_err(err, errlen, e.what());
        return LCG_E_INPUT;
    }
}

extern "C" int64_t lcg_pair_write_arcmatch_scores(const lcg_pair *p, char *buf, int64_t cap) {
    if (!p) return LCG_E_ARG;
    if (p->deferred) return LCG_E_STATE;
    std::string text;
    char line[96];
    const BasePairs &bA = *p->bA, &bB = *p->bB;
    for (int k = 0; k < p->am->num_arc_matches(); ++k) {
        const ArcMatch &x = p->am->arcmatch(k);
        const Arc &a = bA.arc(x.arcA), &b = bB.arc(x.arcB);
        std::snprintf(line, sizeof line, "%d %d %d %d %ld\n", a.left, a.right, b.left, b.right,
                      (long)p->sc->arcmatch(k));
        text += line;
    }
    if (buf && cap > (int64_t)text.size()) std::memcpy(buf, text.c_str(), text.size() + 1);
    return (int64_t)text.size();
}

extern "C" void lcg_pair_free(lcg_pair *p) { delete p; }

extern "C" int lcg_pair_info(const lcg_pair *p, int32_t *lenA, int32_t *lenB, int32_t *nam, int32_t *ngroups,
                             int64_t *cells) {
    if (!p) return LCG_E_ARG;
    if (lenA) *lenA = p->desc.n;
    if (lenB) *lenB = p->desc.m;
    if (nam) *nam = p->desc.nam;
    if (ngroups) *ngroups = p->desc.ngroups;
    if (cells) *cells = p->cells;
    return LCG_OK;
}

extern "C" int lcg_pair_num_arcs(const lcg_pair *p, int32_t *num_arcsA, int32_t *num_arcsB) {
    if (!p) return LCG_E_ARG;
    if (num_arcsA) *num_arcsA = p->bA->num_bps();
    if (num_arcsB) *num_arcsB = p->bB->num_bps();
    return LCG_OK;
}

extern "C" int lcg_pair_band(const lcg_pair *p, int32_t *mn, int32_t *mx, int32_t cap) {
    if (!p || cap < p->desc.n + 1) return LCG_E_ARG;
    for (int i = 0; i <= p->desc.n; ++i) {
        mn[i] = p->tc->min_col(i);
        mx[i] = p->tc->max_col(i);
    }
    return p->desc.n + 1;
}

extern "C" int lcg_pair_score_tables(const lcg_pair *p, int64_t *sigma, int64_t *gapA, int64_t *gapB,
                                     int32_t *rowsA, int32_t *rowsB) {
    if (!p || !p->sc) return LCG_E_ARG;
    const int n = p->A->length(), m = p->B->length();
    if (sigma)
        for (int i = 0; i <= n; ++i)
            for (int j = 0; j <= m; ++j)
                sigma[(size_t)i * (m + 1) + j] = (i == 0 || j == 0) ? 0 : p->sc->basematch(i, j);
    if (gapA)
        for (int i = 0; i <= n; ++i) gapA[i] = i ? p->sc->gapA(i) : 0;
    if (gapB)
        for (int j = 0; j <= m; ++j) gapB[j] = j ? p->sc->gapB(j) : 0;
    if (rowsA) *rowsA = p->A->num_rows();
    if (rowsB) *rowsB = p->B->num_rows();
    return LCG_OK;
}

extern "C" int lcg_pair_arcmatches(const lcg_pair *p, int32_t *al, int32_t *ar, int32_t *bl, int32_t *br,
                                   int64_t *score, int32_t cap) {
    if (!p) return LCG_E_ARG;
    if (p->deferred) return LCG_E_STATE;
    const int nam = p->am->num_arc_matches();
    if (cap < nam) return LCG_E_ARG;
    for (int k = 0; k < nam; ++k) {
        const ArcMatch &x = p->am->arcmatch(k);
        al[k] = p->bA->arc(x.arcA).left;
        ar[k] = p->bA->arc(x.arcA).right;
        bl[k] = p->bB->arc(x.arcB).left;
        br[k] = p->bB->arc(x.arcB).right;
        score[k] = p->sc->arcmatch(k);
    }
    return nam;
}

// ------------------------------------------------------------------ batch

// What the LocARNA-P tables of a pair depend on, captured from the prepared
// pair: reference scores and the derivative of every score with respect to
// the parameters theta (scores are affine in theta around the reference
// parameters; the reference's rounding is kept at theta = 0).
struct PfHost {
    int n = 0, m = 0, ncells = 0, nam = 0;
    double T = 300, ps = 1, thr = 0;
    int tau = 0;
    long alpha = 0;
    std::vector<int> codeA, codeB;
    std::vector<long> gapA, gapB;
    long sigma[LCG_NCODES][LCG_NCODES];
    std::vector<int> entry_of;     // am -> entry
    std::vector<long> am_score;    // by am
    std::vector<int> am_type;      // by am: (uA * 16 + uB), -1 if no RIBOSUM term
    std::vector<double> am_f;      // by am: f(arcA) + f(arcB), f = 1 - ln p / ln pe
    std::vector<int> am_ends;      // by am: al, ar, bl, br
    std::vector<int> min_col, max_col;
    double min_am_prob = 0, min_bm_prob = 0;  // cuts of the probability writers
};

struct lcg_batch {
    lcg_batch() = default;
    lcg_batch(const lcg_batch &) = delete;
    lcg_batch &operator=(const lcg_batch &) = delete;
    // releases every device resource the batch holds, also for a batch that
    // lcg_batch_create_shaped abandons on an error return
    ~lcg_batch();
    int device = 0;
    int npairs = 0;
    std::vector<PairDesc> descs;      // absolute offsets
    std::vector<int32_t> host_arena;  // image uploaded at creation (host packing)
    size_t arena_ints = 0;
    std::vector<int32_t> order;       // work queue order (longest first)
    std::vector<int64_t> work;        // DP cells per pair (work estimate)
    int32_t *d_arena = nullptr;
    PairDesc *d_descs = nullptr;
    int32_t *d_order = nullptr;
    int32_t *d_queue = nullptr;
    int32_t *d_barriers = nullptr;  // per-pair arrival counters (multi-workgroup fill)
    int wg_per_pair = 1;
    int32_t *d_scratch = nullptr;
    hipStream_t stream = nullptr;      // the stream every launch and copy of the batch goes to
    hipStream_t own_stream = nullptr;  // the batch's own stream (destroyed with it)
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_mid = nullptr;
    float fill_ms = -1, top_ms = -1;  // per-kernel times of the last mode-0 run
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_trace = -1;
    // launch shape: requested fields (0 = default), the effective shape and
    // what choosing one needs (CU count, LDS header ints, largest region)
    lcg_shape req{};
    lcg_shape eff{};
    int cus = 1, hdr = 0, gmax = 4, top_max = 4, max_len = 1;
    int sl_max = 0;  // --struct-local: ints of the largest group region (global scratch per wave)
    BatchParams bp;      // fill-D kernel
    int waves = 1, grid = 1;
    size_t lds_bytes = 0;
    BatchParams bp_top;  // top-level / traceback kernel
    int top_waves = 1, top_grid = 1;
    size_t top_lds_bytes = 0;
    bool have_results = false;
    bool fetched = false;
    bool run_invalid = false;  // the fetched run's status words reported an expired wait
    bool traced = false;
    // outputs of the last run: per pair the range [edges, out + 12) of its
    // image, packed at res_off[k] (gathered on the device, one copy back)
    std::vector<int32_t> result;
    std::vector<int32_t> res_off;
    int32_t *d_res = nullptr, *d_resoff = nullptr;
    // LocARNA-P
    std::vector<PfHost> pfh;
    // penalized / normalized alignment: per-pair position penalty of the top
    // level (device), the normalized scores of the last normalized run
    int32_t *d_lambda = nullptr;
    // k-best suboptimal alignments of the last lcg_batch_suboptimal call;
    // the per-pair restrictions (int4) and the top-level scratch of its runs
    struct SuboptAln {
        int64_t score;
        std::vector<int32_t> a, b;
        std::string sA, sB;
        int32_t restr[4];
    };
    std::vector<SuboptAln> subopt;
    int32_t *d_restr = nullptr;
    int32_t *d_kscratch = nullptr;
    size_t kscratch_ints = 0;
    std::vector<int64_t> norm_score;
    bool norm_valid = false;
    // names and sequences of every pair, for the alignment text
    std::vector<std::string> nameA, nameB, seqA, seqB;
    // inputs of every pair for the alignment and consensus writers (rows of
    // profiles, base-pair probabilities) and their expected probabilities
    std::vector<std::shared_ptr<const RnaData>> rdA, rdB;
    std::vector<double> pexpA, pexpB;
    std::vector<int64_t> score_bound;  // per pair, see score_bound()
    std::vector<PfDesc> pdescs;
    std::vector<double> pf_arena;
    double *d_pfa = nullptr;
    int32_t *d_pfsched = nullptr;  // LocARNA-P group schedule (rounds, LDS offsets)
    PfDesc *d_pdescs = nullptr;
    double *d_pfscratch = nullptr;
    double *d_pfstore = nullptr;  // LocARNA-P forward values of the fixed-column groups (inside -> outside)
    int pf_grid = 0, pf_waves = 8;
    size_t pf_lds = 0;
    PfBatchParams pbp;
    bool pf_have = false;
    std::vector<double> pf_result;
    // fragment match queries per pair (set for the next pf run), the
    // capacity of each pair's query block in the pf arena, and the
    // probabilities of the last run
    std::vector<std::vector<std::array<int32_t, 4>>> frags;
    std::vector<int32_t> frag_cap;
    std::vector<std::vector<double>> frag_prob;
    // device compaction (pairs from lcg_pair_prepare_device)
    bool device_built = false;
    std::vector<int32_t> cstats;      // per pair LCG_CS_INTS (counts, status)
    std::vector<CompactDesc> cdescs;  // absolute offsets
    int32_t *d_cin = nullptr, *d_ws = nullptr, *d_ws2 = nullptr, *d_cstats = nullptr;
    CompactDesc *d_cdescs = nullptr;
    CompactParams cp;
    int comp_lds = 0;
    float count_ms = -1, build_ms = -1;
};

namespace {
int hip_fail(hipError_t e, char *err = nullptr, int errlen = 0) {
    set_err(err, errlen, std::string("HIP: ") + hipGetErrorString(e));
    return LCG_E_HIP;
}
#define HIPCHECK(x)                          \
    do {                                     \
        hipError_t e_ = (x);                 \
        if (e_ != hipSuccess) return hip_fail(e_); \
    } while (0)
}  // namespace

static int nt4(char c) {
    switch (c) {
        case 'A': return 0;
        case 'C': return 1;
        case 'G': return 2;
        case 'U': return 3;
    }
    return -1;
}

static void capture_pf_host(lcg_batch &B, lcg_pair *const *pairs, int npairs) {
    B.pfh.resize(npairs);
    for (int k = 0; k < npairs; ++k) {
        const lcg_pair &P = *pairs[k];
        PfHost &h = B.pfh[k];
        h.n = P.desc.n;
        h.m = P.desc.m;
        h.ncells = P.desc.ncells;
        h.nam = P.desc.nam;
        h.T = (double)P.params.temperature_alipf;
        h.ps = P.params.pf_scale;
        h.thr = std::sqrt(P.params.min_am_prob);
        h.min_am_prob = P.params.min_am_prob;
        h.min_bm_prob = P.params.min_bm_prob;
        h.tau = P.params.tau;
        h.alpha = P.params.indel_opening;
        h.codeA.assign(h.n + 1, 0);
        h.codeB.assign(h.m + 1, 0);
        h.gapA.assign(h.n + 1, 0);
        h.gapB.assign(h.m + 1, 0);
        for (int i = 1; i <= h.n; ++i) {
            h.codeA[i] = P.sc->codeA(i);
            h.gapA[i] = P.sc->gapA(i);
        }
        for (int j = 1; j <= h.m; ++j) {
            h.codeB[j] = P.sc->codeB(j);
            h.gapB[j] = P.sc->gapB(j);
        }
        for (int a = 0; a < LCG_NCODES; ++a)
            for (int b = 0; b < LCG_NCODES; ++b) h.sigma[a][b] = P.sc->sigma_code(a, b);
        h.entry_of = P.entry_of;
        const BasePairs &bA = *P.bA, &bB = *P.bB;
        const double peA = P.params.exp_prob > 0 ? P.params.exp_prob : 1.0 / (2.0 * h.n);
        const double peB = P.params.exp_prob > 0 ? P.params.exp_prob : 1.0 / (2.0 * h.m);
        h.am_score.resize(h.nam);
        h.am_type.resize(h.nam);
        h.am_f.resize(h.nam);
        h.am_ends.resize(4 * (size_t)h.nam);
        for (int t = 0; t < h.nam; ++t) {
            const ArcMatch &x = P.am->arcmatch(t);
            const Arc &a = bA.arc(x.arcA), &b = bB.arc(x.arcB);
            h.am_score[t] = P.sc->arcmatch(t);
            const int u0 = nt4(P.A->nt(a.left)), u1 = nt4(P.A->nt(a.right));
            const int v0 = nt4(P.B->nt(b.left)), v1 = nt4(P.B->nt(b.right));
            const bool expl = P.am->explicit_scores();  // explicit scores: constants in theta
            h.am_type[t] = (!expl && P.params.use_ribosum && u0 >= 0 && u1 >= 0 && v0 >= 0 && v1 >= 0)
                               ? (u0 * 4 + u1) * 16 + (v0 * 4 + v1)
                               : -1;
            const double pa = P.A->arc_prob(a.left, a.right), pb = P.B->arc_prob(b.left, b.right);
            h.am_f[t] = expl ? 0.0 : (1 - std::log(pa) / std::log(peA)) + (1 - std::log(pb) / std::log(peB));
            h.am_ends[4 * t] = a.left;
            h.am_ends[4 * t + 1] = a.right;
            h.am_ends[4 * t + 2] = b.left;
            h.am_ends[4 * t + 3] = b.right;
        }
        h.min_col = P.tc->min_cols();
        h.max_col = P.tc->max_cols();
    }
}

namespace {
// LDS schedule of the fill of D for one pair (A: arena holding the pair at
// the offsets of d).  The groups of each row are dealt to K workgroup
// members (longest first to the least loaded member); each member packs its
// groups into rounds of at most `waves` regions that fit `pool` ints
// together, largest first.  Writes G[5] = round | wave<<12 | member<<16,
// G[6] = LDS offset, and the number of rounds of each row.  False if a row
// needs 4096 rounds or more.
bool schedule_groups(int32_t *A, const PairDesc &d, int waves, int pool, int K) {
    for (int al = d.n; al >= 1; --al) {
        const int g0 = A[d.rows + al + 1], g1 = A[d.rows + al];
        auto rows_of = [&](int g) {
            const int32_t *G = A + d.groups + g * LCG_GROUP_INTS;
            return (G[1] & 0xffff) - (G[0] & 0xffff);
        };
        std::vector<std::vector<int>> mem(K);
        {
            std::vector<int> by_len;
            for (int g = g0; g < g1; ++g) by_len.push_back(g);
            std::stable_sort(by_len.begin(), by_len.end(), [&](int x, int y) { return rows_of(x) > rows_of(y); });
            std::vector<long> load(K, 0);
            for (int g : by_len) {
                const int m = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                mem[m].push_back(g);
                load[m] += rows_of(g);
            }
        }
        int nr = 0;
        for (int m = 0; m < K; ++m) {
            auto &lst = mem[m];
            std::stable_sort(lst.begin(), lst.end(), [&](int x, int y) {
                return A[d.groups + x * LCG_GROUP_INTS + 7] > A[d.groups + y * LCG_GROUP_INTS + 7];
            });
            std::vector<int> used, cnt;
            for (int g : lst) {
                int32_t *G = A + d.groups + g * LCG_GROUP_INTS;
                size_t r = 0;
                while (r < used.size() && (cnt[r] >= waves || used[r] + G[7] > pool)) ++r;
                if (r == used.size()) {
                    used.push_back(0);
                    cnt.push_back(0);
                }
                G[5] = (int)r | (cnt[r] << 12) | (m << 16);
                G[6] = used[r];
                used[r] += G[7];
                cnt[r] += 1;
            }
            nr = std::max(nr, (int)used.size());
        }
        if (nr >= 4096) return false;
        A[d.rounds + al] = nr;
    }
    return true;
}
}  // namespace

extern "C" int lcg_pair_schedule(const lcg_pair *p, int32_t waves, int32_t pool_ints, int32_t members,
                                 int32_t *group_words, int32_t *rounds, int32_t cap) {
    if (!p || waves < 1 || waves > 16 || members < 1 || members > 64 || pool_ints < 4) return LCG_E_ARG;
    if (p->deferred) return LCG_E_STATE;
    const PairDesc &d = p->desc;
    if (cap < 4 * d.ngroups || cap < d.n + 2) return LCG_E_ARG;
    for (int g = 0; g < d.ngroups; ++g)
        if (p->blob[d.groups + g * LCG_GROUP_INTS + 7] > pool_ints) return LCG_E_CAPACITY;
    std::vector<int32_t> A = p->blob;
    if (!schedule_groups(A.data(), d, waves, pool_ints, members)) return LCG_E_CAPACITY;
    for (int g = 0; g < d.ngroups; ++g) {
        const int32_t *G = A.data() + d.groups + g * LCG_GROUP_INTS;
        group_words[4 * g] = G[5];
        group_words[4 * g + 1] = G[6];
        group_words[4 * g + 2] = G[7];
        group_words[4 * g + 3] = G[0];
    }
    for (int i = 0; i <= d.n + 1; ++i) rounds[i] = A[d.rounds + i];
    return d.ngroups;
}

namespace {
// Device compaction, pass 1: uploads the compaction inputs of the pairs,
// counts arc matches, groups and fill blocks on the device, and lays out the
// arena image of every pair from the counts (host parts written, arc-match
// structures left to pass 2).
int device_count_and_layout(lcg_batch &B, lcg_pair *const *pairs, int npairs, std::vector<int64_t> &work,
                            char *err, int errlen) {
    std::vector<int32_t> cin;
    size_t ws_ints = 0;
    int max_len = 1;
    B.cdescs.resize(npairs);
    for (int k = 0; k < npairs; ++k) {
        const lcg_pair *P = pairs[k];
        const int32_t base = (int32_t)cin.size();
        cin.insert(cin.end(), P->cimg.begin(), P->cimg.end());
        CompactDesc c = P->cdesc;
        int32_t *offs[] = {&c.arcsA, &c.arcsB, &c.wA, &c.wB, &c.ptA, &c.ptB, &c.lA, &c.lB,
                           &c.rpA, &c.rpB, &c.radjA, &c.radjB, &c.band, &c.tab, &c.hdr};
        for (int32_t *o : offs) *o += base;
        const size_t words = (size_t)c.NA * c.W64;
        ws_ints = (ws_ints + 1) & ~(size_t)1;
        c.ws_bits = (int32_t)ws_ints;
        ws_ints += 2 * words;
        c.ws_wpre = (int32_t)ws_ints;
        ws_ints += words;
        c.ws_rowcnt = (int32_t)ws_ints;
        ws_ints += c.NA;
        c.ws_sblk = (int32_t)ws_ints;
        ws_ints += c.NA + 1;
        c.ws_blka = (int32_t)ws_ints;
        ws_ints += c.NA;
        c.ws_blks = (int32_t)ws_ints;
        ws_ints += c.NA;
        if (ws_ints >= (size_t)0x7fffffff || cin.size() >= (size_t)0x7fffffff) {
            set_err(err, errlen, "batch too large for the compaction workspace");
            return LCG_E_CAPACITY;
        }
        B.cdescs[k] = c;
        max_len = std::max(max_len, std::max(c.n, c.m));
    }
    B.comp_lds = (3 * (max_len + 1) + 1) * 4;
    if (B.comp_lds > 160 * 1024 - 1024) {
        set_err(err, errlen, "sequence too long for the compaction kernels");
        return LCG_E_CAPACITY;
    }
    hipError_t e;
    if ((e = hipMalloc(&B.d_cin, std::max<size_t>(cin.size(), 4) * 4)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMalloc(&B.d_ws, std::max<size_t>(ws_ints, 4) * 4)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMalloc(&B.d_cdescs, sizeof(CompactDesc) * npairs)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMalloc(&B.d_cstats, 4 * (size_t)LCG_CS_INTS * npairs)) != hipSuccess)
        return hip_fail(e, err, errlen);
    if ((e = hipMemcpy(B.d_cin, cin.data(), cin.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, err, errlen);
    if ((e = hipMemcpy(B.d_cdescs, B.cdescs.data(), sizeof(CompactDesc) * npairs, hipMemcpyHostToDevice)) !=
        hipSuccess)
        return hip_fail(e, err, errlen);
    if ((e = hipEventRecord(B.ev0, B.stream)) != hipSuccess) return hip_fail(e, err, errlen);
    if (int rc = lcg_launch_compact_count(B.d_cdescs, B.d_cin, B.d_ws, B.d_cstats, npairs, B.comp_lds, B.stream))
        return hip_fail((hipError_t)rc, err, errlen);
    if ((e = hipEventRecord(B.ev1, B.stream)) != hipSuccess) return hip_fail(e, err, errlen);
    B.cstats.resize((size_t)LCG_CS_INTS * npairs);
    if ((e = hipMemcpyAsync(B.cstats.data(), B.d_cstats, B.cstats.size() * 4, hipMemcpyDeviceToHost, B.stream)) !=
        hipSuccess)
        return hip_fail(e, err, errlen);
    if ((e = hipStreamSynchronize(B.stream)) != hipSuccess) return hip_fail(e, err, errlen);
    hipEventElapsedTime(&B.count_ms, B.ev0, B.ev1);
    // arena layout from the counts
    B.descs.resize(npairs);
    size_t total = 0;
    std::vector<size_t> size(npairs);
    for (int k = 0; k < npairs; ++k) {
        const lcg_pair &P = *pairs[k];
        const int32_t *S = &B.cstats[(size_t)LCG_CS_INTS * k];
        PairDesc d;
        std::memset(&d, 0, sizeof d);
        d.n = P.desc.n;
        d.m = P.desc.m;
        d.nam = S[LCG_CS_NAM];
        d.ngroups = S[LCG_CS_NGROUPS];
        d.nfblk = S[LCG_CS_NBLK] + 1;
        int ncells = 0;
        for (int i = 0; i <= d.n; ++i) ncells += std::max(0, P.tc->max_col(i) - P.tc->min_col(i) + 1);
        d.ncells = ncells;
        size[k] = layout_pair(d, d.n, d.m, ncells, d.ngroups, d.nam, d.nfblk);
        d.gmax = S[LCG_CS_GMAX];
        int64_t cells = (int64_t)(((uint64_t)(uint32_t)S[LCG_CS_CELLS + 1] << 32) | (uint32_t)S[LCG_CS_CELLS]);
        top_region(P.tc->min_cols(), P.tc->max_cols(), d.n, d.m, d.wtop, d.top_ints, cells);
        work[k] = cells;
        B.descs[k] = d;
        total += size[k];
    }
    if (total >= (size_t)0x7fffffff) {
        set_err(err, errlen, "batch too large for one arena");
        return LCG_E_CAPACITY;
    }
    B.arena_ints = total;
    size_t base = 0;
    for (int k = 0; k < npairs; ++k) {
        rebase(B.descs[k], (int32_t)base);
        base += size[k];
    }
    return LCG_OK;
}

// Device compaction, pass 2 (after the arena image is on the device): the
// arc-match structures and the LDS schedule of every pair.
int device_build(lcg_batch &B, int npairs, int max_len, int waves, int pool, int K, char *err, int errlen) {
    size_t ws2 = 0;
    for (int k = 0; k < npairs; ++k) {
        CompactDesc &c = B.cdescs[k];
        const PairDesc &d = B.descs[k];
        c.ws_eof = (int32_t)ws2;
        ws2 += d.nam;
        c.ws_fof = (int32_t)ws2;
        ws2 += d.nam;
        c.ws_sched = (int32_t)ws2;
        ws2 += 3 * (size_t)d.ngroups + (size_t)K * (d.n + 2);
        if (ws2 >= (size_t)0x7fffffff) {
            set_err(err, errlen, "batch too large for the compaction workspace");
            return LCG_E_CAPACITY;
        }
    }
    hipError_t e;
    if (B.d_ws2) {  // a reshape with another number of workgroups per pair
        hipFree(B.d_ws2);
        B.d_ws2 = nullptr;
    }
    if ((e = hipMalloc(&B.d_ws2, std::max<size_t>(ws2, 4) * 4)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMemcpy(B.d_cdescs, B.cdescs.data(), sizeof(CompactDesc) * npairs, hipMemcpyHostToDevice)) !=
        hipSuccess)
        return hip_fail(e, err, errlen);
    B.cp.npairs = npairs;
    B.cp.max_len = max_len;
    B.cp.waves = waves;
    B.cp.pool = pool;
    B.cp.members = K;
    if ((e = hipEventRecord(B.ev0, B.stream)) != hipSuccess) return hip_fail(e, err, errlen);
    if (int rc = lcg_launch_compact_build(B.d_cdescs, B.d_descs, B.d_cin, B.d_arena, B.d_ws, B.d_ws2, B.d_cstats,
                                          B.cp, B.comp_lds, B.stream))
        return hip_fail((hipError_t)rc, err, errlen);
    if ((e = hipEventRecord(B.ev1, B.stream)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMemcpyAsync(B.cstats.data(), B.d_cstats, B.cstats.size() * 4, hipMemcpyDeviceToHost, B.stream)) !=
        hipSuccess)
        return hip_fail(e, err, errlen);
    if ((e = hipStreamSynchronize(B.stream)) != hipSuccess) return hip_fail(e, err, errlen);
    hipEventElapsedTime(&B.build_ms, B.ev0, B.ev1);
    for (int k = 0; k < npairs; ++k) {
        const int st = B.cstats[(size_t)LCG_CS_INTS * k + LCG_CS_STATUS];
        if (st == 1) {
            set_err(err, errlen, "too many LDS rounds in one row");
            return LCG_E_CAPACITY;
        }
        if (st != 0) {
            set_err(err, errlen, "internal: device arc-match compaction inconsistent with its counts");
            return LCG_E_STATE;
        }
    }
    return LCG_OK;
}
}  // namespace

namespace {
// A requested field (> 0) wins over the environment variable, which wins
// over the default.
int pick(int req, const char *env, int dflt, int lo, int hi) {
    if (req > 0) return std::min(hi, std::max(lo, req));
    if (const char *w = std::getenv(env)) return std::min(hi, std::max(lo, std::atoi(w)));
    return dflt;
}

constexpr int kLdsIntsCu = (160 * 1024 - 512) / 4;  // static LDS of the kernels included

struct FillShape {
    int waves, wgs_per_cu, K, pool;
};
struct TopShape {
    int waves, wgs_per_cu, pool;
};

// Fill of D: one workgroup of 16 waves per CU with the whole LDS as its
// pool by default: a larger pool packs more groups of a row into one round,
// which measured faster than two workgroups of 8 waves with half the pool
// each (allpairs64: 108 ms against 173 ms per step).  With fewer pairs than
// CUs, K workgroups share a pair (all co-resident).
int choose_fill(const lcg_batch &B, FillShape &s, char *err, int errlen) {
    s.waves = pick(B.req.waves, "LCG_WAVES", 16, 1, 16);
    s.wgs_per_cu = pick(B.req.wgs_per_cu, "LCG_WGS_PER_CU", 1, 1, 4);
    s.pool = kLdsIntsCu / s.wgs_per_cu - B.hdr;
    while (s.pool < B.gmax && s.wgs_per_cu > 1) {
        --s.wgs_per_cu;
        s.pool = kLdsIntsCu / s.wgs_per_cu - B.hdr;
    }
    s.pool &= ~3;
    if (s.pool < B.gmax) {
        set_err(err, errlen, "an arc-match group region exceeds the LDS of one workgroup");
        return LCG_E_CAPACITY;
    }
    const int np = B.npairs;
    s.K = pick(B.req.wg_per_pair, "LCG_WG_PER_PAIR", np < B.cus ? std::min(64, B.cus / np) : 1, 1, 64);
    s.K = std::min(s.K, std::max(1, B.cus / np));  // all workgroups of the fill stay co-resident
    return LCG_OK;
}

// The top-level / traceback kernel runs the top level on one wave: four
// workgroups of 4 waves per CU (a quarter of the LDS each) by default keep
// more pairs in flight, as long as every group region still fits the
// smaller pool (allpairs64: 3.36 ms against 4.57 ms with 2 x 8 and 7.42 ms
// with 1 x 16).
TopShape choose_top(const lcg_batch &B) {
    TopShape t;
    t.waves = pick(B.req.top_waves, "LCG_TOP_WAVES", 4, 1, 16);
    t.wgs_per_cu = pick(B.req.top_wgs_per_cu, "LCG_TOP_WGS_PER_CU", 4, 1, 4);
    t.pool = (kLdsIntsCu / t.wgs_per_cu - B.hdr) & ~3;
    while (t.pool < B.gmax && t.wgs_per_cu > 1) {
        --t.wgs_per_cu;
        t.pool = (kLdsIntsCu / t.wgs_per_cu - B.hdr) & ~3;
    }
    return t;
}

// LDS schedule of every pair of a host-packed batch into its host image.
bool schedule_all(lcg_batch &B, const FillShape &s) {
    for (int k = 0; k < B.npairs; ++k)
        if (!schedule_groups(B.host_arena.data(), B.descs[k], s.waves, s.pool, s.K)) return false;
    return true;
}

// Launch parameters of both kernels for the chosen shapes.
void apply_shape(lcg_batch &B, const FillShape &s, const TopShape &t) {
    const int np = B.npairs;
    B.wg_per_pair = s.K;
    B.waves = s.waves;
    B.lds_bytes = (size_t)(B.hdr + s.pool) * 4;
    B.bp.pool_ints = s.pool;
    B.bp.wg_per_pair = s.K;
    B.grid = s.K > 1 ? np * s.K : std::min(np, B.cus * s.wgs_per_cu);
    B.bp.sl_stride = B.sl_max;
    B.bp_top = B.bp;
    B.bp_top.pool_ints = t.pool;
    // the traceback's structure-local regions use the per-wave scratch too
    B.bp_top.scratch_stride = std::max(B.top_max, B.sl_max);
    B.top_waves = t.waves;
    B.top_lds_bytes = (size_t)(B.hdr + t.pool) * 4;
    B.top_grid = std::min(np, B.cus * t.wgs_per_cu);
    B.eff.waves = s.waves;
    B.eff.wgs_per_cu = s.wgs_per_cu;
    B.eff.wg_per_pair = s.K;
    B.eff.top_waves = t.waves;
    B.eff.top_wgs_per_cu = t.wgs_per_cu;
}

// global scratch of a batch: the top level's per-wave regions and, with
// --struct-local, the fill's and traceback's eight-state regions
size_t scratch_ints(const lcg_batch &B) {
    const size_t top = (size_t)B.top_grid * B.top_waves * std::max(B.top_max, B.sl_max);
    const size_t fill = (size_t)B.grid * B.waves * B.sl_max;
    return std::max(top, fill);
}
}  // namespace

extern "C" int lcg_batch_create(lcg_pair *const *pairs, int32_t npairs, int32_t device, lcg_batch **out, char *err,
                                int errlen) {
    return lcg_batch_create_shaped(pairs, npairs, device, nullptr, out, err, errlen);
}

extern "C" int lcg_batch_create_shaped(lcg_pair *const *pairs, int32_t npairs, int32_t device, const lcg_shape *shape,
                                       lcg_batch **out, char *err, int errlen) {
    if (!pairs || npairs <= 0 || !out) return LCG_E_ARG;
    *out = nullptr;
    auto B = std::make_unique<lcg_batch>();
    B->device = device;
    B->npairs = npairs;
    if (shape) B->req = *shape;
    const lcg_params &p0 = pairs[0]->params;
    const bool deferred = pairs[0]->deferred;
    for (int k = 0; k < npairs; ++k) {
        const lcg_pair *P = pairs[k];
        if (P->params.indel_opening != p0.indel_opening || P->params.sequ_local != p0.sequ_local ||
            P->params.free_endgaps != p0.free_endgaps || P->params.no_lonely_pairs != p0.no_lonely_pairs ||
            (P->params.stacking != 0) != (p0.stacking != 0) || P->params.struct_local != p0.struct_local ||
            (p0.struct_local && P->params.exclusion != p0.exclusion)) {
            set_err(err, errlen,
                    "all pairs of a batch must share indel_opening, sequ_local, free_endgaps, no_lonely_pairs, "
                    "stacking, struct_local and exclusion");
            return LCG_E_ARG;
        }
        if (P->deferred != deferred) {
            set_err(err, errlen, "a batch takes either host-packed or device-compacted pairs, not both");
            return LCG_E_ARG;
        }
        // the kernels read sigma either by symbol codes or, for profiles,
        // from a column-pair table: one kind per batch
        if (P->sc->profile() != pairs[0]->sc->profile()) {
            set_err(err, errlen, "a batch takes either pairs of alignment profiles or pairs of single sequences");
            return LCG_E_ARG;
        }
    }
    hipError_t e;
    int cus = 0;
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
        return hip_fail(e, err, errlen);
    if ((e = hipStreamCreate(&B->stream)) != hipSuccess) return hip_fail(e, err, errlen);
    B->own_stream = B->stream;
    if ((e = hipEventCreate(&B->ev0)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipEventCreate(&B->ev1)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipEventCreate(&B->ev_mid)) != hipSuccess) return hip_fail(e, err, errlen);

    std::vector<int64_t> work(npairs);
    if (deferred) {
        B->device_built = true;
        if (int rc = device_count_and_layout(*B, pairs, npairs, work, err, errlen)) return rc;
    } else {
        size_t total = 0;
        for (int k = 0; k < npairs; ++k) total += pairs[k]->blob.size();
        if (total >= (size_t)0x7fffffff) {
            set_err(err, errlen, "batch too large for one arena");
            return LCG_E_CAPACITY;
        }
        B->host_arena.reserve(total);
        B->arena_ints = total;
        for (int k = 0; k < npairs; ++k) {
            const lcg_pair *P = pairs[k];
            const int32_t base = (int32_t)B->host_arena.size();
            B->host_arena.insert(B->host_arena.end(), P->blob.begin(), P->blob.end());
            PairDesc d = P->desc;
            rebase(d, base);
            if (!image_in_place(d, base, (int32_t)P->blob.size(), B->host_arena, P->params.stacking != 0)) {
                set_err(err, errlen, "internal: image offsets");
                return LCG_E_STATE;
            }
            B->descs.push_back(d);
            work[k] = P->cells;
        }
    }
    int max_len = 1, gmax = 4, top_max = 4, nfblk_max = 1;
    for (const PairDesc &d : B->descs) B->sl_max = std::max(B->sl_max, d.slmax);
    for (const PairDesc &d : B->descs) {
        max_len = std::max(max_len, std::max(d.n, d.m));
        gmax = std::max(gmax, d.gmax);
        nfblk_max = std::max(nfblk_max, d.nfblk);
        top_max = std::max(top_max, d.top_ints);
    }
    // longest pairs first (greedy LPT over the persistent workgroups)
    B->work = work;
    B->order.resize(npairs);
    std::iota(B->order.begin(), B->order.end(), 0);
    std::stable_sort(B->order.begin(), B->order.end(), [&](int a, int b) { return work[a] > work[b]; });

    // LDS header: band, codes, gaps, sigma (7 L + 72 ints) of the longest
    // sequence, fill-list row pointers (L + 2) and blocks (2 per block); the
    // rest of a workgroup's LDS is the pool for the group regions of a row.
    // Per-row LDS schedule: groups of a row, largest first, into rounds of at
    // most `waves` regions that fit the pool together; with K workgroups per
    // pair the groups of a row are first dealt to the K members, longest
    // first to the least loaded member.
    const int fptr_lds = 7 * max_len + 72;
    const int fblk_lds = fptr_lds + max_len + 2;
    B->hdr = (fblk_lds + 2 * nfblk_max + 3) & ~3;
    B->cus = cus;
    B->gmax = gmax;
    B->top_max = top_max;
    B->max_len = max_len;
    FillShape fs;
    if (int rc = choose_fill(*B, fs, err, errlen)) return rc;
    const TopShape ts = choose_top(*B);
    if (!deferred && !schedule_all(*B, fs)) {
        set_err(err, errlen, "too many LDS rounds in one row");
        return LCG_E_CAPACITY;
    }
    B->bp.alpha = p0.indel_opening;
    B->bp.local = p0.sequ_local;
    B->bp.free_endgaps = p0.free_endgaps;
    B->bp.nolp = p0.no_lonely_pairs != 0;
    B->bp.stacking = p0.stacking != 0;
    B->bp.sigma_table = pairs[0]->sc->profile() ? 1 : 0;
    B->bp.struct_local = p0.struct_local != 0;
    B->bp.exclusion = p0.exclusion;
    if (B->bp.struct_local && B->bp.sigma_table) {
        set_err(err, errlen, "--struct-local: pairs of single sequences only");
        return LCG_E_ARG;
    }
    B->bp.trace = 1;
    B->bp.npairs = npairs;
    B->bp.hdr_ints = B->hdr;
    B->bp.fptr_lds = fptr_lds;
    B->bp.fblk_lds = fblk_lds;
    B->bp.max_len = max_len;
    B->bp.count = 0;
    B->bp.scratch_stride = top_max;
    B->bp.fixed_cols = 1;
    if (const char *f = std::getenv("LCG_FIXED_COLS")) B->bp.fixed_cols = std::atoi(f) != 0;
    apply_shape(*B, fs, ts);
    if ((e = hipMalloc(&B->d_arena, B->arena_ints * 4)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMalloc(&B->d_descs, sizeof(PairDesc) * npairs)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMalloc(&B->d_order, 4 * (size_t)npairs)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMalloc(&B->d_queue, 256)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMalloc(&B->d_barriers, 4 * (size_t)npairs)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMalloc(&B->d_scratch, scratch_ints(*B) * 4)) != hipSuccess) return hip_fail(e, err, errlen);
    if (deferred)  // the compaction writes the image over a cleared arena
        e = hipMemsetAsync(B->d_arena, 0, B->arena_ints * 4, B->stream);
    else
        e = hipMemcpy(B->d_arena, B->host_arena.data(), B->arena_ints * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMemcpy(B->d_descs, B->descs.data(), sizeof(PairDesc) * npairs, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, err, errlen);
    if ((e = hipMemcpy(B->d_order, B->order.data(), 4 * (size_t)npairs, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, err, errlen);
    B->res_off.resize(npairs);
    size_t res_ints = 0;
    for (int k = 0; k < npairs; ++k) {
        B->res_off[k] = (int32_t)res_ints;
        res_ints += (size_t)(B->descs[k].out + 12 - B->descs[k].edges);
    }
    B->result.resize(res_ints);
    if ((e = hipMalloc(&B->d_res, res_ints * 4)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMalloc(&B->d_resoff, 4 * (size_t)npairs)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipMemcpy(B->d_resoff, B->res_off.data(), 4 * (size_t)npairs, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, err, errlen);
    if (deferred) {
        if (int rc = device_build(*B, npairs, max_len, fs.waves, fs.pool, fs.K, err, errlen)) return rc;
    } else {
        capture_pf_host(*B, pairs, npairs);
    }
    for (int k = 0; k < npairs; ++k) {
        const lcg_pair &X = *pairs[k];
        B->score_bound.push_back(score_bound(X));
        B->nameA.push_back(X.A->name());
        B->nameB.push_back(X.B->name());
        std::string a, b;
        for (int i = 1; i <= X.A->length(); ++i) a += X.A->nt(i);
        for (int j = 1; j <= X.B->length(); ++j) b += X.B->nt(j);
        B->seqA.push_back(a);
        B->seqB.push_back(b);
        B->rdA.push_back(X.A);
        B->rdB.push_back(X.B);
        // locarna.cc:662-663
        B->pexpA.push_back(X.params.exp_prob > 0 ? X.params.exp_prob : 1.0 / (2.0 * X.A->length()));
        B->pexpB.push_back(X.params.exp_prob > 0 ? X.params.exp_prob : 1.0 / (2.0 * X.B->length()));
    }
    *out = B.release();
    return LCG_OK;
}

lcg_batch::~lcg_batch() {
    lcg_batch *B = this;
    // the stream is the first resource lcg_batch_create_shaped creates
    if (!B->stream) return;
    hipSetDevice(B->device);
    if (B->graph_exec) hipGraphExecDestroy(B->graph_exec);
    if (B->graph) hipGraphDestroy(B->graph);
    if (B->d_arena) hipFree(B->d_arena);
    if (B->d_descs) hipFree(B->d_descs);
    if (B->d_order) hipFree(B->d_order);
    if (B->d_queue) hipFree(B->d_queue);
    if (B->d_barriers) hipFree(B->d_barriers);
    if (B->d_scratch) hipFree(B->d_scratch);
    if (B->d_restr) hipFree(B->d_restr);
    if (B->d_kscratch) hipFree(B->d_kscratch);
    if (B->d_pfa) hipFree(B->d_pfa);
    if (B->d_pdescs) hipFree(B->d_pdescs);
    if (B->d_pfscratch) hipFree(B->d_pfscratch);
    if (B->d_pfsched) hipFree(B->d_pfsched);
    if (B->d_pfstore) hipFree(B->d_pfstore);
    if (B->ev0) hipEventDestroy(B->ev0);
    if (B->ev1) hipEventDestroy(B->ev1);
    if (B->ev_mid) hipEventDestroy(B->ev_mid);
    if (B->d_res) hipFree(B->d_res);
    if (B->d_resoff) hipFree(B->d_resoff);
    if (B->d_cin) hipFree(B->d_cin);
    if (B->d_ws) hipFree(B->d_ws);
    if (B->d_ws2) hipFree(B->d_ws2);
    if (B->d_cstats) hipFree(B->d_cstats);
    if (B->d_cdescs) hipFree(B->d_cdescs);
    if (B->d_lambda) hipFree(B->d_lambda);
    if (B->own_stream) hipStreamDestroy(B->own_stream);
}

extern "C" void lcg_batch_destroy(lcg_batch *B) { delete B; }

extern "C" int lcg_batch_pair_info(const lcg_batch *B, int32_t k, int32_t *nam, int32_t *ngroups,
                                   int64_t *arena_ints) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    const PairDesc &d = B->descs[k];
    if (nam) *nam = d.nam;
    if (ngroups) *ngroups = d.ngroups;
    if (arena_ints) {
        const size_t end = k + 1 < B->npairs ? (size_t)B->descs[k + 1].band : B->arena_ints;
        *arena_ints = (int64_t)(end - (size_t)d.band);
    }
    return LCG_OK;
}

extern "C" int lcg_batch_segment(const lcg_batch *B, int32_t k, int32_t *dst, int64_t cap) {
    if (!B || !dst || k < 0 || k >= B->npairs) return LCG_E_ARG;
    const size_t begin = (size_t)B->descs[k].band;
    const size_t end = k + 1 < B->npairs ? (size_t)B->descs[k + 1].band : B->arena_ints;
    if ((size_t)cap < end - begin) return LCG_E_ARG;
    HIPCHECK(hipSetDevice(B->device));
    HIPCHECK(hipStreamSynchronize(B->stream));
    HIPCHECK(hipMemcpy(dst, B->d_arena + begin, (end - begin) * 4, hipMemcpyDeviceToHost));
    return (int)(end - begin);
}

extern "C" int lcg_batch_compact_run(lcg_batch *B, float *count_ms, float *build_ms) {
    if (!B) return LCG_E_ARG;
    if (!B->device_built) return LCG_E_STATE;
    HIPCHECK(hipSetDevice(B->device));
    HIPCHECK(hipEventRecord(B->ev0, B->stream));
    if (int rc = lcg_launch_compact_count(B->d_cdescs, B->d_cin, B->d_ws, B->d_cstats, B->npairs, B->comp_lds,
                                          B->stream))
        return hip_fail((hipError_t)rc);
    HIPCHECK(hipEventRecord(B->ev_mid, B->stream));
    if (int rc = lcg_launch_compact_build(B->d_cdescs, B->d_descs, B->d_cin, B->d_arena, B->d_ws, B->d_ws2,
                                          B->d_cstats, B->cp, B->comp_lds, B->stream))
        return hip_fail((hipError_t)rc);
    HIPCHECK(hipEventRecord(B->ev1, B->stream));
    std::vector<int32_t> st(B->cstats.size());
    HIPCHECK(hipMemcpyAsync(st.data(), B->d_cstats, st.size() * 4, hipMemcpyDeviceToHost, B->stream));
    HIPCHECK(hipStreamSynchronize(B->stream));
    HIPCHECK(hipEventElapsedTime(&B->count_ms, B->ev0, B->ev_mid));
    HIPCHECK(hipEventElapsedTime(&B->build_ms, B->ev_mid, B->ev1));
    if (count_ms) *count_ms = B->count_ms;
    if (build_ms) *build_ms = B->build_ms;
    for (size_t t = 0; t < st.size(); ++t)
        if (st[t] != B->cstats[t]) return LCG_E_STATE;  // a rerun must reproduce the counts and status
    B->have_results = false;
    return LCG_OK;
}

extern "C" int lcg_batch_set_stream(lcg_batch *B, void *stream) {
    if (!B) return LCG_E_ARG;
    HIPCHECK(hipSetDevice(B->device));
    HIPCHECK(hipStreamSynchronize(B->stream));  // work queued on the previous stream completes first
    B->stream = stream ? static_cast<hipStream_t>(stream) : B->own_stream;
    return LCG_OK;
}

extern "C" int lcg_batch_set_spin_limit(lcg_batch *B, int32_t polls) {
    if (!B || polls < -1) return LCG_E_ARG;
    B->bp.spin_limit = polls;
    B->bp_top.spin_limit = polls;
    return LCG_OK;
}

extern "C" int lcg_batch_config(const lcg_batch *B, int32_t *waves, int32_t *workgroups, int32_t *lds_bytes) {
    if (!B) return LCG_E_ARG;
    if (waves) *waves = B->waves;
    if (workgroups) *workgroups = B->grid;
    if (lds_bytes) *lds_bytes = (int32_t)B->lds_bytes;
    return LCG_OK;
}

extern "C" int lcg_batch_run(lcg_batch *B, int32_t mode, int32_t trace, float *device_ms) {
    if (!B || mode < 0 || mode > 2) return LCG_E_ARG;
    HIPCHECK(hipSetDevice(B->device));
    BatchParams bp = B->bp;
    bp.trace = (trace & 1) ? 1 : 0;
    bp.count = (trace & 2) ? 1 : 0;
    if (mode == 2 && bp.count) return LCG_E_ARG;
    BatchParams top = B->bp_top;
    top.trace = bp.trace;
    top.count = bp.count;
    HIPCHECK(hipEventRecord(B->ev0, B->stream));
    if (mode == 0) {
        int rc = lcg_launch_align(B->d_descs, B->d_order, B->d_queue, B->d_arena, B->d_scratch, B->d_barriers, bp, B->grid, B->waves,
                                  B->lds_bytes, top, B->top_grid, B->top_waves, B->top_lds_bytes, nullptr, 0, B->ev_mid, B->stream);
        if (rc) return hip_fail((hipError_t)rc);
    } else if (mode == 1) {
        BatchParams one = bp, one_top = top;
        one.npairs = 1;
        one_top.npairs = 1;
        for (int k = 0; k < B->npairs; ++k) {
            int rc = lcg_launch_align(B->d_descs, B->d_order + k, B->d_queue, B->d_arena, B->d_scratch, B->d_barriers, one, B->wg_per_pair, B->waves,
                                      B->lds_bytes, one_top, 1, B->top_waves, B->top_lds_bytes, nullptr, 0, nullptr, B->stream);
            if (rc) return hip_fail((hipError_t)rc);
        }
    } else {
        if (!B->graph_exec || B->graph_trace != bp.trace) {
            if (B->graph_exec) hipGraphExecDestroy(B->graph_exec);
            if (B->graph) hipGraphDestroy(B->graph);
            B->graph_exec = nullptr;
            B->graph = nullptr;
            hipStream_t cs;
            HIPCHECK(hipStreamCreate(&cs));
            HIPCHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            int rc = lcg_launch_align(B->d_descs, B->d_order, B->d_queue, B->d_arena, B->d_scratch, B->d_barriers, bp, B->grid, B->waves,
                                      B->lds_bytes, top, B->top_grid, B->top_waves, B->top_lds_bytes, nullptr, 0, nullptr, cs);
            hipGraph_t g = nullptr;
            hipError_t ec = hipStreamEndCapture(cs, &g);
            hipStreamDestroy(cs);
            if (rc) return hip_fail((hipError_t)rc);
            if (ec != hipSuccess) return hip_fail(ec);
            B->graph = g;
            HIPCHECK(hipGraphInstantiate(&B->graph_exec, g, nullptr, nullptr, 0));
            B->graph_trace = bp.trace;
        }
        HIPCHECK(hipGraphLaunch(B->graph_exec, B->stream));
    }
    HIPCHECK(hipEventRecord(B->ev1, B->stream));
    HIPCHECK(hipEventSynchronize(B->ev1));
    if (device_ms) HIPCHECK(hipEventElapsedTime(device_ms, B->ev0, B->ev1));
    if (mode == 0) {
        HIPCHECK(hipEventElapsedTime(&B->fill_ms, B->ev0, B->ev_mid));
        HIPCHECK(hipEventElapsedTime(&B->top_ms, B->ev_mid, B->ev1));
    } else {
        B->fill_ms = B->top_ms = -1;
    }
    B->have_results = true;
    B->fetched = false;
    B->traced = bp.trace != 0;
    B->norm_valid = false;
    return LCG_OK;
}

extern "C" int lcg_noalloc_scratch(int device, int **ptr, long long *ints);

// The fill of D, top level and traceback of every pair (as lcg_batch_run
// mode 0), enqueued on the batch's stream with the work-queue counters, the
// arrival counters and the top-level per-wave regions in the library's
// static device array instead of the batch's own buffers.  Calls nothing
// that allocates, records no event and does not wait: the getters
// synchronize with the stream.  Fewer top-level workgroups than the batch's
// own launch when the array holds fewer per-wave regions (results are the
// same for any number).
extern "C" int lcg_batch_run_noalloc(lcg_batch *B, int32_t trace) {
    if (!B || (trace & ~1)) return LCG_E_ARG;
    if (B->bp.struct_local) return LCG_E_ARG;  // its regions need the batch's own scratch
    HIPCHECK(hipSetDevice(B->device));
    int *S = nullptr;
    long long cap = 0;
    if (int rc = lcg_noalloc_scratch(B->device, &S, &cap)) return hip_fail((hipError_t)rc);
    const long long bar_off = 64;                                 // queues at 0 and 16
    const long long top_off = (bar_off + B->npairs + 63) & ~63LL;  // 256-byte aligned
    const long long per_wg = (long long)B->top_waves * B->top_max;
    if (top_off + per_wg > cap) return LCG_E_CAPACITY;
    const int top_grid = (int)std::min<long long>(B->top_grid, (cap - top_off) / per_wg);
    BatchParams bp = B->bp, top = B->bp_top;
    bp.trace = top.trace = trace & 1;
    bp.count = top.count = 0;
    if (int rc = lcg_launch_align(B->d_descs, B->d_order, S, B->d_arena, S + top_off, S + bar_off, bp, B->grid,
                                  B->waves, B->lds_bytes, top, top_grid, B->top_waves, B->top_lds_bytes, nullptr, 0,
                                  nullptr, B->stream))
        return hip_fail((hipError_t)rc);
    B->fill_ms = B->top_ms = -1;
    B->have_results = true;
    B->fetched = false;
    B->traced = (trace & 1) != 0;
    B->norm_valid = false;
    return LCG_OK;
}

namespace {
int fetch(const lcg_batch *Bc);
inline int32_t res(const lcg_batch *B, int k, int32_t off);

// One run with the per-pair position penalties in d_lambda (top_only: the
// top level and traceback again over the D of the last run).
int run_with_penalty(lcg_batch *B, bool top_only, float *device_ms) {
    HIPCHECK(hipSetDevice(B->device));
    BatchParams bp = B->bp, top = B->bp_top;
    bp.trace = top.trace = 1;
    bp.count = top.count = 0;
    HIPCHECK(hipEventRecord(B->ev0, B->stream));
    int rc = lcg_launch_align(B->d_descs, B->d_order, B->d_queue, B->d_arena, B->d_scratch, B->d_barriers, bp,
                              B->grid, B->waves, B->lds_bytes, top, B->top_grid, B->top_waves, B->top_lds_bytes,
                              B->d_lambda, top_only ? 1 : 0, nullptr, B->stream);
    if (rc) return hip_fail((hipError_t)rc);
    HIPCHECK(hipEventRecord(B->ev1, B->stream));
    HIPCHECK(hipEventSynchronize(B->ev1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, B->ev0, B->ev1));
    if (device_ms) *device_ms += ms;
    B->fill_ms = B->top_ms = -1;
    B->have_results = true;
    B->fetched = false;
    B->traced = true;
    B->norm_valid = false;
    return LCG_OK;
}

int upload_lambda(lcg_batch *B, const std::vector<int32_t> &lam) {
    HIPCHECK(hipSetDevice(B->device));
    if (!B->d_lambda) HIPCHECK(hipMalloc(&B->d_lambda, 4 * (size_t)B->npairs));
    HIPCHECK(hipMemcpy(B->d_lambda, lam.data(), 4 * (size_t)B->npairs, hipMemcpyHostToDevice));
    return LCG_OK;
}
}  // namespace

// The top level charged a position penalty lam stays in the kernels' score
// range: lam shifts the score of each aligned position once (a base match
// twice, scoring.cc:77-90), so by at most 2 |lam| (n + m).
static bool penalty_in_range(const lcg_batch *B, int k, int64_t lam) {
    const PairDesc &d = B->descs[k];
    return B->score_bound[k] + 2 * std::llabs(lam) * (int64_t)(d.n + d.m) <= LCG_SCORE_BOUND;
}

extern "C" int lcg_batch_run_penalized(lcg_batch *B, int64_t position_penalty, float *device_ms) {
    if (!B || position_penalty < -100000 || position_penalty > 100000) return LCG_E_ARG;
    for (int k = 0; k < B->npairs; ++k)
        if (!penalty_in_range(B, k, position_penalty)) return LCG_E_CAPACITY;
    if (int rc = upload_lambda(B, std::vector<int32_t>(B->npairs, (int32_t)position_penalty))) return rc;
    if (device_ms) *device_ms = 0;
    return run_with_penalty(B, false, device_ms);
}

extern "C" int lcg_batch_run_normalized(lcg_batch *B, int64_t L, int32_t max_iterations, int32_t *iterations,
                                        float *device_ms) {
    if (!B || max_iterations < 1 || L < 0) return LCG_E_ARG;
    if (!B->bp.local) return LCG_E_ARG;  // normalized alignment is sequence-local (src/locarna.cc:428-440)
    const int np = B->npairs;
    std::vector<int32_t> lam(np, 0);
    std::vector<char> done(np, 0);
    if (device_ms) *device_ms = 0;
    if (int rc = upload_lambda(B, lam)) return rc;
    if (int rc = run_with_penalty(B, false, device_ms)) return rc;  // D, then lambda = 0
    std::vector<int32_t> ea, eb;
    for (int it = 1;; ++it) {
        if (int rc = fetch(B)) return rc;
        bool all = true;
        for (int k = 0; k < np; ++k) {
            if (done[k]) continue;
            const PairDesc &d = B->descs[k];
            const int64_t score = res(B, k, d.out);
            const int max_i = res(B, k, d.out + 1), max_j = res(B, k, d.out + 2);
            // the cell where the local trace stopped (M = 0): before the
            // first edge, a match edge for the recursion's gap costs; the
            // empty alignment stops at (0, 0)
            int min_i = 0, min_j = 0;
            ea.assign(d.n + d.m + 2, 0);
            eb.assign(d.n + d.m + 2, 0);
            const int ne = lcg_batch_alignment(B, k, ea.data(), eb.data(), (int32_t)ea.size(), nullptr, nullptr, 0);
            if (ne < 0) return ne;
            if (ne > 0) {
                for (int t = 0; t < ne; ++t)
                    if (ea[t] > 0) {
                        min_i = ea[t] - 1;
                        break;
                    }
                for (int t = 0; t < ne; ++t)
                    if (eb[t] > 0) {
                        min_j = eb[t] - 1;
                        break;
                    }
            }
            // aligner.cc:1562-1575
            const int64_t length = (int64_t)(max_i - min_i + 1) + (max_j - min_j + 1);
            const int64_t next = (score + length * lam[k]) / (length + L);
            if (next == lam[k]) {
                done[k] = 1;
                continue;
            }
            if (next < -100000 || next > 100000) return LCG_E_STATE;
            if (!penalty_in_range(B, k, next)) return LCG_E_CAPACITY;
            lam[k] = (int32_t)next;
            all = false;
        }
        if (all) {
            B->norm_score.assign(lam.begin(), lam.end());
            B->norm_valid = true;
            if (iterations) *iterations = it;
            return LCG_OK;
        }
        if (it >= max_iterations) return LCG_E_STATE;
        if (int rc = upload_lambda(B, lam)) return rc;
        if (int rc = run_with_penalty(B, true, device_ms)) return rc;
    }
}

namespace {
// Copies the outputs of a run back once, on the first result request.
int fetch(const lcg_batch *Bc) {
    lcg_batch *B = const_cast<lcg_batch *>(Bc);
    if (!B->have_results) return LCG_E_STATE;
    if (B->fetched) return B->run_invalid ? LCG_E_STATE : LCG_OK;
    HIPCHECK(hipSetDevice(B->device));
    if (int rc = lcg_launch_gather(B->d_descs, B->d_resoff, B->d_arena, B->d_res, B->npairs, B->stream))
        return hip_fail((hipError_t)rc);
    HIPCHECK(hipMemcpyAsync(B->result.data(), B->d_res, B->result.size() * 4, hipMemcpyDeviceToHost, B->stream));
    HIPCHECK(hipStreamSynchronize(B->stream));
    // a bounded wait of the multi-workgroup fill that expired invalidates the
    // run: every later result request of this run fails too
    B->run_invalid = false;
    for (int k = 0; k < B->npairs; ++k)
        if (B->result[B->res_off[k] + (B->descs[k].out + 3 - B->descs[k].edges)] != 0) B->run_invalid = true;
    B->fetched = true;
    return B->run_invalid ? LCG_E_STATE : LCG_OK;
}

// output word at arena offset `off` (in [edges, out + 12)) of pair k
inline int32_t res(const lcg_batch *B, int k, int32_t off) {
    return B->result[B->res_off[k] + (off - B->descs[k].edges)];
}
}  // namespace

// k-best suboptimal alignments by interval splitting (Aligner::suboptimal,
// aligner.cc:1382-1514, not normalized).  The fill of D runs once for the
// batch; every restriction of pair k is then one top-level launch over that
// pair only (one workgroup), with the traceback for the enumerated ones.
// Tasks are ordered as the reference orders them: a max-heap on the score
// (greater_second, aligner.hh:34-45, compares with <; std::priority_queue).
namespace {
int restricted_top(lcg_batch *B, int k, const int32_t r[4], bool trace, int64_t *score) {
    HIPCHECK(hipSetDevice(B->device));
    const PairDesc &d = B->descs[k];
    // the stored width of a restricted region is at most m + 1
    const size_t need = std::max<size_t>(((size_t)(d.m + 1) * 2 + (size_t)(d.n + 1) * (d.m + 1) + 3) & ~(size_t)3,
                                         (size_t)B->sl_max);
    if (need > B->kscratch_ints) {
        if (B->d_kscratch) HIPCHECK(hipFree(B->d_kscratch));
        B->d_kscratch = nullptr;
        HIPCHECK(hipMalloc(&B->d_kscratch, need * B->top_waves * 4));
        B->kscratch_ints = need;
    }
    if (!B->d_restr) HIPCHECK(hipMalloc(&B->d_restr, 16 * (size_t)B->npairs));
    // the restriction goes into the pair's slot (indexed by pair, like the
    // position penalties)
    HIPCHECK(hipMemcpyAsync(B->d_restr + 4 * (size_t)k, r, 16, hipMemcpyHostToDevice, B->stream));
    // queue position of pair k (order[pos] == k)
    int pos = -1;
    for (int t = 0; t < B->npairs && pos < 0; ++t)
        if (B->order[t] == k) pos = t;
    if (pos < 0) return LCG_E_STATE;
    BatchParams bp = B->bp, top = B->bp_top;
    bp.npairs = top.npairs = 1;
    bp.trace = top.trace = trace ? 1 : 0;
    bp.count = top.count = 0;
    top.scratch_stride = (int32_t)B->kscratch_ints;
    int rc = lcg_launch_align(B->d_descs, B->d_order + pos, B->d_queue, B->d_arena, B->d_kscratch, B->d_barriers, bp,
                              1, B->waves, B->lds_bytes, top, 1, B->top_waves, B->top_lds_bytes, nullptr, 1, nullptr,
                              B->stream, B->d_restr);
    if (rc) return hip_fail((hipError_t)rc);
    HIPCHECK(hipStreamSynchronize(B->stream));
    B->fill_ms = B->top_ms = -1;
    B->have_results = true;
    B->fetched = false;
    B->traced = trace;
    B->norm_valid = false;
    if (int rc2 = fetch(B)) return rc2;
    *score = (int64_t)res(B, k, B->descs[k].out);
    return LCG_OK;
}
}  // namespace

extern "C" int lcg_batch_suboptimal(lcg_batch *B, int32_t k, int32_t kbest, int64_t threshold, int64_t *scores,
                                    int32_t cap) {
    if (!B || k < 0 || k >= B->npairs || kbest == 0 || cap < 0 || (cap > 0 && !scores)) return LCG_E_ARG;
    B->subopt.clear();
    // the fill of D of every pair (and the unrestricted alignment)
    if (int rc = lcg_batch_run(B, 0, 1, nullptr)) return rc;
    if (int rc = fetch(B)) return rc;
    const PairDesc &d = B->descs[k];
    const int n = d.n, m = d.m;
    struct Task {
        int32_t r[4];  // startA, endA, startB, endB
        int64_t score;
    };
    struct Less {
        bool operator()(const Task &x, const Task &y) const { return x.score < y.score; }
    };
    std::priority_queue<Task, std::vector<Task>, Less> tasks;
    Task whole{{1, n, 1, m}, 0};
    if (int rc = restricted_top(B, k, whole.r, false, &whole.score)) return rc;
    tasks.push(whole);
    std::vector<int32_t> ea(n + m + 2), eb(n + m + 2);
    std::string sa(std::max(n, m) + 2, '\0'), sb(std::max(n, m) + 2, '\0');
    size_t i = 1;
    while (kbest < 0 || i <= (size_t)kbest) {
        if (tasks.empty()) break;
        const Task task = tasks.top();
        tasks.pop();
        // a restriction without an alignment (minus infinity) ends the
        // enumeration like one below the threshold
        if (task.score < threshold + 1 || task.score < LCG_NEG_LIMIT) break;
        int64_t sc = 0;
        if (int rc = restricted_top(B, k, task.r, true, &sc)) return rc;
        const int ne = lcg_batch_alignment(B, k, ea.data(), eb.data(), (int32_t)ea.size(), &sa[0], &sb[0],
                                           (int32_t)sa.size());
        if (ne < 0) return ne;
        if (ne == 0) continue;  // empty alignment
        lcg_batch::SuboptAln al;
        al.score = task.score;
        al.a.assign(ea.begin(), ea.begin() + ne);
        al.b.assign(eb.begin(), eb.begin() + ne);
        al.sA = sa.c_str();
        al.sB = sb.c_str();
        std::copy(task.r, task.r + 4, al.restr);
        B->subopt.push_back(std::move(al));
        if (kbest >= 0 && i == (size_t)kbest) break;
        // start / end positions (Alignment::start_positions / end_positions,
        // alignment.cc:170-195) and the split of the longer side
        int fa = -1, fb = -1, la = 0, lb = 0;
        for (int t = 0; t < ne; ++t) {
            if (fa < 0 && ea[t] > 0) fa = ea[t];
            if (fb < 0 && eb[t] > 0) fb = eb[t];
            if (ea[t] > 0) la = ea[t];
            if (eb[t] > 0) lb = eb[t];
        }
        if (fa < 0) fa = n;
        if (fb < 0) fb = m;
        Task r1 = task, r2 = task;
        if (task.r[1] - task.r[0] > task.r[3] - task.r[2]) {
            const int split = (fa + la) / 2;
            r1.r[1] = split;
            r2.r[0] = split;
        } else {
            const int split = (fb + lb) / 2;
            r1.r[3] = split;
            r2.r[2] = split;
        }
        // a restriction the split leaves empty aligns nothing (its region
        // would have no rows or columns): score 0 locally, minus infinity
        // globally, without a launch
        for (Task *t : {&r1, &r2}) {
            if (t->r[1] < t->r[0] || t->r[3] < t->r[2] || t->r[0] < 1 || t->r[2] < 1 || t->r[1] > n ||
                t->r[3] > m) {
                t->score = B->bp.local ? 0 : LCG_NEG;
                continue;
            }
            if (int rc = restricted_top(B, k, t->r, false, &t->score)) return rc;
        }
        tasks.push(r1);
        tasks.push(r2);
        ++i;
    }
    const int cnt = (int)std::min<size_t>(B->subopt.size(), (size_t)cap);
    for (int t = 0; t < cnt; ++t) scores[t] = B->subopt[t].score;
    return (int)B->subopt.size();
}

extern "C" int lcg_batch_suboptimal_alignment(const lcg_batch *B, int32_t t, int32_t *a, int32_t *b, int32_t cap,
                                              char *structA, char *structB, int32_t scap, int32_t *restriction) {
    if (!B || t < 0 || t >= (int32_t)B->subopt.size()) return LCG_E_ARG;
    const auto &al = B->subopt[t];
    if (cap < (int32_t)al.a.size()) return LCG_E_ARG;
    std::copy(al.a.begin(), al.a.end(), a);
    std::copy(al.b.begin(), al.b.end(), b);
    if (structA && structB) {
        if (scap <= (int32_t)al.sA.size() || scap <= (int32_t)al.sB.size()) return LCG_E_ARG;
        std::memcpy(structA, al.sA.c_str(), al.sA.size() + 1);
        std::memcpy(structB, al.sB.c_str(), al.sB.size() + 1);
    }
    if (restriction) std::copy(al.restr, al.restr + 4, restriction);
    return (int32_t)al.a.size();
}

extern "C" int lcg_batch_score(const lcg_batch *B, int32_t k, int64_t *score) {
    if (!B || k < 0 || k >= B->npairs || !score) return LCG_E_ARG;
    if (int rc = fetch(B)) return rc;
    *score = B->norm_valid ? B->norm_score[k] : (int64_t)res(B, k, B->descs[k].out);
    return LCG_OK;
}

extern "C" int lcg_batch_alignment(const lcg_batch *B, int32_t k, int32_t *a, int32_t *b, int32_t cap,
                                   char *structA, char *structB, int32_t scap) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    if (!B->traced) return LCG_E_STATE;
    if (int rc = fetch(B)) return rc;
    const PairDesc &d = B->descs[k];
    int cnt = 0;
    for (int s = 0; s <= d.n + d.m; ++s) {
        const int32_t v = res(B, k, d.edges + s);
        if (v == -1) continue;
        if (cnt >= cap) return LCG_E_ARG;
        const int x = v >> 16, y = v & 0xffff;
        a[cnt] = x ? x : -1;
        b[cnt] = y ? y : -1;
        ++cnt;
    }
    if (structA && structB) {
        if (scap <= d.n || scap <= d.m) return LCG_E_ARG;
        static const char sym[3] = {'.', '(', ')'};
        for (int i = 0; i < d.n; ++i) structA[i] = sym[res(B, k, d.structs + i)];
        for (int j = 0; j < d.m; ++j) structB[j] = sym[res(B, k, d.structs + d.n + j)];
        structA[d.n] = 0;
        structB[d.m] = 0;
    }
    return cnt;
}

namespace {
// Alignment::alignment_edges (alignment.cc:121-167): the traced edges of
// pair k as (a, b) columns (-1 for a gap), with the positions a local or
// free-end-gap alignment leaves out as gap columns unless only the local
// part is asked for.  Returns LCG_OK or a status.
int full_edges(const lcg_batch *B, int32_t k, bool only_local, std::vector<std::pair<int, int>> &cols,
               std::string *strA = nullptr, std::string *strB = nullptr) {
    const int n = (int)B->seqA[k].size(), m = (int)B->seqB[k].size();
    std::vector<int32_t> ea(n + m + 2), eb(n + m + 2);
    std::string stA(n + 2, 0), stB(m + 2, 0);
    const int ne = lcg_batch_alignment(B, k, ea.data(), eb.data(), (int32_t)ea.size(), stA.data(), stB.data(),
                                       (int32_t)std::max(n, m) + 2);
    if (ne < 0) return ne;
    cols.clear();
    auto col = [&](int a, int b) {
        cols.push_back({a > 0 ? a : -1, b > 0 ? b : -1});
        if (strA) *strA += a > 0 ? stA[a - 1] : '-';
        if (strB) *strB += b > 0 ? stB[b - 1] : '-';
    };
    int lastA = 1, lastB = 1;
    for (int t = 0; t < ne; ++t) {
        if (ea[t] > 0) {
            for (; lastA < ea[t]; ++lastA)
                if (!only_local) col(lastA, -1);
        }
        if (eb[t] > 0) {
            for (; lastB < eb[t]; ++lastB)
                if (!only_local) col(-1, lastB);
        }
        if (ea[t] > 0) ++lastA;
        if (eb[t] > 0) ++lastB;
        col(ea[t], eb[t]);
    }
    if (!only_local) {
        for (; lastA <= n; ++lastA) col(lastA, -1);
        for (; lastB <= m; ++lastB) col(-1, lastB);
    }
    return LCG_OK;
}
}  // namespace

extern "C" int64_t lcg_batch_write_clustal(const lcg_batch *B, int32_t k, int32_t width, int32_t flags, char *buf,
                                           int64_t cap) {
    if (!B || k < 0 || k >= B->npairs || width < 1) return LCG_E_ARG;
    if (!B->traced) return LCG_E_STATE;
    int64_t score = 0;
    if (int rc = lcg_batch_score(B, k, &score)) return rc;
    std::vector<std::pair<int, int>> cols;
    std::string strA, strB;
    if (int rc = full_edges(B, k, (flags & LCG_CLUSTAL_LOCAL_ONLY) != 0, cols, &strA, &strB)) return rc;
    // rows of both inputs with the alignment's gaps, "A." / "B." prefixes
    // on a name clash (MultipleAlignment::init, multiple_alignment.cc:186-247)
    const std::vector<std::pair<std::string, std::string>> ali = aligned_rows(*B->rdA[k], *B->rdB[k], cols);
    // header (main_helper.icc:557-569: the score for pairs of single
    // sequences only) and blocks of `width` columns (MultipleAlignment::write,
    // multiple_alignment.cc:994-1085), with --write-structure the structure
    // of A first and that of B last
    std::string text = "CLUSTAL W --- LocARNA 2.0.1";
    if (B->rdA[k]->num_rows() == 1 && B->rdB[k]->num_rows() == 1) {
        text += " --- Score: ";
        text += score < LCG_NEG_LIMIT ? std::string("-inf") : std::to_string(score);
    }
    text += "\n\n";
    std::vector<std::pair<std::string, const std::string *>> rows;
    const bool st = (flags & LCG_CLUSTAL_STRUCTURES) != 0;  // --write-structure
    if (st) rows.push_back({"", &strA});
    for (const auto &r : ali) rows.push_back({r.first, &r.second});
    if (st) rows.push_back({"", &strB});
    text += clustal_blocks(rows, (size_t)width);
    if (buf && cap > (int64_t)text.size()) std::memcpy(buf, text.c_str(), text.size() + 1);
    return (int64_t)text.size();
}

extern "C" int64_t lcg_batch_write_consensus_pp(const lcg_batch *B, int32_t k, char *buf, int64_t cap) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    if (!B->traced) return LCG_E_STATE;
    std::vector<std::pair<int, int>> cols;
    if (int rc = full_edges(B, k, false, cols)) return rc;
    const std::string text =
        consensus_pp_text(*B->rdA[k], *B->rdB[k], cols, B->pexpA[k], B->pexpB[k], B->bp.stacking != 0);
    if (buf && cap > (int64_t)text.size()) std::memcpy(buf, text.c_str(), text.size() + 1);
    return (int64_t)text.size();
}

extern "C" int lcg_batch_D(const lcg_batch *B, int32_t k, int64_t *D, int32_t cap) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    if (int rc = fetch(B)) return rc;
    const PairDesc &d = B->descs[k];
    if (cap < d.nam) return LCG_E_ARG;
    std::vector<int32_t> v(std::max(d.nam, 1));
    HIPCHECK(hipMemcpy(v.data(), B->d_arena + d.D, 4 * (size_t)d.nam, hipMemcpyDeviceToHost));
    for (int t = 0; t < d.nam; ++t) D[t] = v[t];
    return d.nam;
}

extern "C" int lcg_batch_counters(const lcg_batch *B, int32_t k, int64_t *cells, int64_t *candidates) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    if (int rc = fetch(B)) return rc;
    const PairDesc &d = B->descs[k];
    uint64_t c[4];
    for (int t = 0; t < 4; ++t) {
        const uint32_t lo = (uint32_t)res(B, k, d.out + 4 + 2 * t), hi = (uint32_t)res(B, k, d.out + 5 + 2 * t);
        c[t] = ((uint64_t)hi << 32) | lo;
    }
    if (cells) *cells = (int64_t)(c[0] + c[2]);
    if (candidates) *candidates = (int64_t)(c[1] + c[3]);
    return LCG_OK;
}

extern "C" int lcg_batch_counters_split(const lcg_batch *B, int32_t k, int64_t *fill_cells, int64_t *fill_candidates,
                                        int64_t *top_cells, int64_t *top_candidates) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    if (int rc = fetch(B)) return rc;
    const PairDesc &d = B->descs[k];
    uint64_t c[4];
    for (int t = 0; t < 4; ++t) {
        const uint32_t lo = (uint32_t)res(B, k, d.out + 4 + 2 * t), hi = (uint32_t)res(B, k, d.out + 5 + 2 * t);
        c[t] = ((uint64_t)hi << 32) | lo;
    }
    if (fill_cells) *fill_cells = (int64_t)c[0];
    if (fill_candidates) *fill_candidates = (int64_t)c[1];
    if (top_cells) *top_cells = (int64_t)c[2];
    if (top_candidates) *top_candidates = (int64_t)c[3];
    return LCG_OK;
}

extern "C" int lcg_batch_kernel_times(const lcg_batch *B, float *fill_ms, float *top_ms) {
    if (!B) return LCG_E_ARG;
    if (B->fill_ms < 0) return LCG_E_STATE;
    if (fill_ms) *fill_ms = B->fill_ms;
    if (top_ms) *top_ms = B->top_ms;
    return LCG_OK;
}

// ------------------------------------------------------------ LocARNA-P

extern "C" int lcg_launch_pf(const PairDesc *descs, const PfDesc *pdescs, const int *order, int *queue, int *arena,
                             double *pfa, double *scratch, const int *sched, int *barriers, double *fstore,
                             PfBatchParams bp, int grid, int waves, size_t lds_bytes, hipStream_t stream);

namespace {
int pf_setup(lcg_batch &B) {
    if (B.d_pfa) return LCG_OK;
    const int np = B.npairs;
    B.pdescs.resize(np);
    size_t off = 0;
    int max_len = 1;
    size_t stride = 0;
    for (int k = 0; k < np; ++k) {
        const PfHost &h = B.pfh[k];
        PfDesc &f = B.pdescs[k];
        std::memset(&f, 0, sizeof f);
        auto take = [&](size_t cnt) {
            int32_t o = (int32_t)off;
            off += (cnt + 1) & ~(size_t)1;
            return o;
        };
        f.egA = take(h.n + 1);
        f.egB = take(h.m + 1);
        f.es = take(LCG_NCODES * LCG_NCODES);
        f.eam = take(h.nam);
        f.Dc = take(h.nam);
        f.Dbar = take(h.nam);
        f.edge_full = take(h.ncells);
        f.edge_ref = take(h.ncells);
        f.scal = take(4);
        f.out = take(4);
        f.fD = take(h.nam);
        const int cap = B.frags.empty() ? 0 : (int)B.frags[k].size();
        f.frag = take(8 * (size_t)cap);
        f.nfrag = 0;
        max_len = std::max(max_len, std::max(h.n, h.m));
        const PairDesc &d = B.descs[k];
        const size_t top_cells = (size_t)(d.n + 1) * d.wtop;
        stride = std::max(stride, 4 * top_cells + 2 * (size_t)(d.m + 1) + 8);
        // groups are no larger than the top level region
    }
    B.pf_arena.assign(off, 0.0);
    B.frag_cap.assign(np, 0);
    for (int k = 0; k < np && !B.frags.empty(); ++k) B.frag_cap[k] = (int32_t)B.frags[k].size();
    B.pf_waves = pick(B.req.pf_waves, "LCG_PF_WAVES", 8, 1, 8);
    int nfblk_max = 1;
    for (int k = 0; k < np; ++k) nfblk_max = std::max(nfblk_max, B.descs[k].nfblk);
    B.pbp.npairs = np;
    B.pbp.max_len = max_len;
    B.pbp.nfblk_max = nfblk_max;
    // LDS: int header (band, cell offsets, codes, fill blocks), double
    // header (exp gap scores, exp sigma), then the pool of group regions
    const int hdr_ints = (6 * max_len + 8 + 2 * nfblk_max + 1) & ~1;
    const int hdr_doubles = 2 * (max_len + 1) + LCG_NCODES * LCG_NCODES;
    const int lds_total = 160 * 1024 - 512;
    const int pool = (lds_total - 4 * hdr_ints - 8 * hdr_doubles) / 8;
    if (pool < 64) return LCG_E_CAPACITY;
    B.pbp.pool = pool;
    B.pf_lds = (size_t)4 * hdr_ints + (size_t)8 * (hdr_doubles + pool);
    // workgroups per pair: with fewer pairs than CUs, K workgroups share a
    // pair (all resident, one per CU)
    int cus = 0;
    HIPCHECK(hipSetDevice(B.device));
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, B.device));
    const bool forced = B.req.pf_wg_per_pair > 0 || std::getenv("LCG_PF_WG_PER_PAIR");
    int K = pick(B.req.pf_wg_per_pair, "LCG_PF_WG_PER_PAIR", np < cus ? std::min(64, cus / np) : 1, 1, 64);
    K = std::min(K, std::max(1, cus / np));
    // the first nst pairs of the work order get K workgroups each, statically;
    // the others are pulled from the queue by the remaining workgroups.  With
    // at most half as many pairs as CUs all pairs are static.  Otherwise the largest pairs
    // get two workgroups where that shortens the estimated makespan (a pair
    // on two workgroups measured about 1.75x faster; the rest by longest
    // processing time first over the other workgroups).
    int nst = K > 1 ? np : 0;
    if (K == 1 && cus >= 4 && !forced) {
        auto makespan = [&](int ns) {
            double t = 0;
            for (int q = 0; q < ns; ++q) t = std::max(t, (double)B.work[B.order[q]] / 1.75);
            const int rest = cus - 2 * ns;
            if (ns < np && rest < 1) return 1e300;
            std::vector<double> load(std::max(rest, 1), 0.0);
            for (int q = ns; q < np; ++q) {
                auto it = std::min_element(load.begin(), load.end());
                *it += (double)B.work[B.order[q]];
            }
            for (double x : load) t = std::max(t, x);
            return t;
        };
        double best = makespan(0);
        for (int ns = 1; 2 * ns <= cus && ns <= np; ++ns) {
            const double t = makespan(ns);
            if (t < best) {
                best = t;
                nst = ns;
                K = 2;
            }
        }
    }
    B.pbp.wg_per_pair = K;
    B.pbp.nstatic = nst;
    // schedule: per row, the groups dealt to the K members (longest first to
    // the least loaded), then per member in rounds of at most pf_waves;
    // narrow groups (columns and rows at most 128) get LDS regions that fit
    // the pool together, the others run in per-wave global scratch (offset
    // -1).  Word: round | wave<<12 | member<<16.
    std::vector<int32_t> sched;
    size_t max_fixed = 0;
    std::vector<int> Kp(np, 1);  // workgroups of each pair
    for (int q = 0; q < nst; ++q) Kp[B.order[q]] = K;
    for (int k = 0; k < np; ++k) {
        const PairDesc &d = B.descs[k];
        PfDesc &f = B.pdescs[k];
        const int32_t *A = B.host_arena.data();
        const int KK = Kp[k];
        f.sched = (int32_t)sched.size();
        sched.resize(sched.size() + 4 * (size_t)d.ngroups);
        f.rin = (int32_t)sched.size();
        sched.resize(sched.size() + d.n + 2, 0);
        f.rout = (int32_t)sched.size();
        sched.resize(sched.size() + d.n + 2, 0);
        const int gbase = d.groups;
        for (int al = d.n; al >= 1; --al) {
            const int g0 = A[d.rows + al + 1], g1 = A[d.rows + al];
            for (int pass = 0; pass < 2; ++pass) {
                auto size_of = [&](int g) -> long {
                    const int32_t *G = A + gbase + g * LCG_GROUP_INTS;
                    const int gal = G[0] & 0xffff, gbl = G[0] >> 16, gar = G[1] & 0xffff, gbr = G[1] >> 16;
                    if (gbr - gbl - 1 > 128 || gar - gal - 1 > 128) return -1;
                    const long cells = (long)(gar - gal) * G[2];
                    const long sz = (gbr - gbl) + (pass == 0 ? 1 : 2) * cells;
                    return sz <= pool ? ((sz + 1) & ~1L) : -1;
                };
                auto rows_of = [&](int g) {
                    const int32_t *G = A + gbase + g * LCG_GROUP_INTS;
                    return (G[1] & 0xffff) - (G[0] & 0xffff);
                };
                std::vector<std::vector<int>> mem(KK);
                {
                    std::vector<int> by_len;
                    for (int g = g0; g < g1; ++g) by_len.push_back(g);
                    std::stable_sort(by_len.begin(), by_len.end(), [&](int x, int y) { return rows_of(x) > rows_of(y); });
                    std::vector<long> load(KK, 0);
                    for (int g : by_len) {
                        const int mm = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                        mem[mm].push_back(g);
                        load[mm] += rows_of(g);
                    }
                }
                size_t nr = 0;
                for (int mm = 0; mm < KK; ++mm) {
                    std::vector<int> &idx = mem[mm];
                    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return size_of(x) > size_of(y); });
                    std::vector<long> used;
                    std::vector<int> cnt;
                    for (int g : idx) {
                        const long sz = size_of(g);
                        size_t r = 0;
                        while (r < used.size() && (cnt[r] >= B.pf_waves || (sz > 0 && used[r] + sz > pool))) ++r;
                        if (r == used.size()) {
                            used.push_back(0);
                            cnt.push_back(0);
                        }
                        int32_t *S = &sched[f.sched + 4 * (size_t)g + 2 * pass];
                        S[0] = (int32_t)r | (cnt[r] << 12) | (mm << 16);
                        S[1] = sz > 0 ? (int32_t)used[r] : -1;
                        if (sz > 0) {
                            used[r] += sz;
                            const int32_t *G = A + gbase + g * LCG_GROUP_INTS;
                            max_fixed = std::max(max_fixed, (size_t)2 * ((G[1] & 0xffff) - (G[0] & 0xffff)) * G[2]);
                        }
                        cnt[r] += 1;
                    }
                    nr = std::max(nr, used.size());
                }
                sched[(pass == 0 ? f.rin : f.rout) + al] = (int32_t)nr;
            }
        }
        // group indices in sched are relative to the pair's group array
        (void)gbase;
    }
    // forward-value store: every group that runs the fixed-column kernels in
    // both passes gets 3 (ar - al) W doubles (E, s, M) while the store stays
    // within its budget; the others are recomputed by the outside pass
    const size_t store_budget = (size_t)1 << 31;  // doubles (16 GiB)
    size_t store = 0;
    for (int k = 0; k < np; ++k) {
        const PairDesc &d = B.descs[k];
        PfDesc &f = B.pdescs[k];
        const int32_t *A = B.host_arena.data();
        f.fso = (int32_t)sched.size();
        sched.resize(sched.size() + (size_t)d.ngroups, -1);
        for (int g = 0; g < d.ngroups; ++g) {
            const int32_t *S = &sched[f.sched + 4 * (size_t)g];
            if (S[1] < 0 || S[3] < 0) continue;
            const int32_t *G = A + d.groups + g * LCG_GROUP_INTS;
            const size_t need = (3 * (size_t)((G[1] & 0xffff) - (G[0] & 0xffff)) * G[2] + 3) & ~(size_t)3;
            if (store + need > store_budget) continue;
            sched[f.fso + (size_t)g] = (int32_t)(store / 4);
            store += need;
        }
    }
    stride = std::max(stride, max_fixed);
    B.pbp.scratch_stride = (int32_t)((stride + 31) & ~(size_t)31);
    hipError_t e;
    // all workgroups resident at once (one per CU): static members + the rest
    B.pf_grid = nst == np ? np * K : (nst > 0 ? cus : std::min(np, cus));
    B.eff.pf_waves = B.pf_waves;
    B.eff.pf_wg_per_pair = forced ? K : 0;  // 0: chosen per pair by the makespan rule above
    HIPCHECK(hipMalloc(&B.d_pfa, std::max<size_t>(1, off) * 8));
    HIPCHECK(hipMalloc(&B.d_pdescs, sizeof(PfDesc) * np));
    HIPCHECK(hipMalloc(&B.d_pfscratch, (size_t)B.pf_grid * B.pf_waves * B.pbp.scratch_stride * 8));
    HIPCHECK(hipMalloc(&B.d_pfsched, std::max<size_t>(1, sched.size()) * 4));
    HIPCHECK(hipMalloc(&B.d_pfstore, std::max<size_t>(4, store) * 8));
    HIPCHECK(hipMemcpy(B.d_pfsched, sched.data(), sched.size() * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(B.d_pdescs, B.pdescs.data(), sizeof(PfDesc) * np, hipMemcpyHostToDevice));
    (void)e;
    return LCG_OK;
}

// Fragment match queries of every pair into the pf arena: i, j, k, l, the
// stored row width W of the block's region (i-1, k-1; j+1, l+1), and in /
// out cleared.  A block whose corner cell (i-1, k-1) or (j, l) lies outside
// the band has probability 0 (no alignment in the band passes through it):
// W = 0, the kernels skip it.
int frag_write(lcg_batch &B) {
    for (int k = 0; k < B.npairs; ++k) {
        PfDesc &f = B.pdescs[k];
        const int nq = B.frags.empty() ? 0 : (int)B.frags[k].size();
        if (nq > B.frag_cap[k]) return LCG_E_STATE;  // set_fragments drops the arena first
        f.nfrag = nq;
        const PfHost &h = B.pfh[k];
        for (int q = 0; q < nq; ++q) {
            const auto &x = B.frags[k][q];
            const int i = x[0], j = x[1], kk = x[2], l = x[3];
            auto valid = [&](int r, int c) { return h.min_col[r] <= c && c <= h.max_col[r]; };
            int W = 0;
            if (valid(i - 1, kk - 1) && valid(j, l)) {
                W = std::min(l, h.max_col[i - 1]) - (kk - 1) + 1;
                for (int r = i; r <= j; ++r)
                    W = std::max(W, std::min(l, h.max_col[r]) - std::max(kk - 1, h.min_col[r]) + 1);
                const size_t need = 4 * (size_t)(j - i + 2) * W + 2 * (size_t)(l - kk + 2);
                if (need > (size_t)B.pbp.scratch_stride) return LCG_E_CAPACITY;
            }
            double *d = B.pf_arena.data() + f.frag + 8 * q;
            d[0] = i;
            d[1] = j;
            d[2] = kk;
            d[3] = l;
            d[4] = W;
            d[5] = d[6] = d[7] = 0.0;
        }
    }
    return LCG_OK;
}

// Boltzmann tables of pair k for parameters theta0 + dtheta.
void pf_tables(const lcg_batch &B, int k, const double *dth, double *A) {
    const PfHost &h = B.pfh[k];
    const PfDesc &f = B.pdescs[k];
    auto dt = [&](int idx) { return dth ? dth[idx] : 0.0; };
    const double T = h.T;
    const double dindel = dt(272), dopen = dt(273), dsw = dt(274);
    for (int i = 1; i <= h.n; ++i) A[f.egA + i] = std::exp((h.gapA[i] + dindel) / T);
    for (int j = 1; j <= h.m; ++j) A[f.egB + j] = std::exp((h.gapB[j] + dindel) / T);
    for (int a = 0; a < LCG_NCODES; ++a)
        for (int b = 0; b < LCG_NCODES; ++b)
            A[f.es + a * LCG_NCODES + b] = std::exp((h.sigma[a][b] + ((a < 4 && b < 4) ? dt(a * 4 + b) : 0.0)) / T);
    for (int am = 0; am < h.nam; ++am) {
        double s = (double)h.am_score[am];
        if (h.am_type[am] >= 0) s += (h.tau / 100.0) * dt(16 + h.am_type[am]);
        s += dsw * h.am_f[am];
        A[f.eam + h.entry_of[am]] = std::exp(s / T);
    }
    A[f.scal + 0] = std::exp((h.alpha + dopen) / T);
    A[f.scal + 1] = h.ps;
    A[f.scal + 2] = h.thr;
}
}  // namespace

extern "C" int lcg_batch_pf_run(lcg_batch *B, const double *dtheta, float *device_ms) {
    if (!B) return LCG_E_ARG;
    B->pf_have = false;  // the results of an earlier run are overwritten from here on
    if (B->bp.free_endgaps) return LCG_E_ARG;  // free end gaps: integer path only
    if (B->bp.nolp) return LCG_E_ARG;          // --noLP: integer path only
    if (B->bp.stacking) return LCG_E_ARG;      // --stacking: integer path only
    if (B->bp.sigma_table) return LCG_E_ARG;   // profiles: integer path only
    if (B->bp.struct_local) return LCG_E_ARG;  // --struct-local: integer path only
    if (B->device_built) return LCG_E_STATE;  // LocARNA-P reads the host arc-match lists
    if (int rc = pf_setup(*B)) return rc;
    for (int k = 0; k < B->npairs; ++k) pf_tables(*B, k, dtheta, B->pf_arena.data());
    if (int rc = frag_write(*B)) return rc;
    HIPCHECK(hipSetDevice(B->device));
    HIPCHECK(hipMemcpy(B->d_pfa, B->pf_arena.data(), B->pf_arena.size() * 8, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(B->d_pdescs, B->pdescs.data(), sizeof(PfDesc) * B->npairs, hipMemcpyHostToDevice));
    B->pbp.spin_limit = B->bp.spin_limit;
    HIPCHECK(hipEventRecord(B->ev0, B->stream));
    int rc = lcg_launch_pf(B->d_descs, B->d_pdescs, B->d_order, B->d_queue, B->d_arena, B->d_pfa, B->d_pfscratch,
                           B->d_pfsched, B->d_barriers, B->d_pfstore, B->pbp, B->pf_grid, B->pf_waves, B->pf_lds,
                           B->stream);
    if (rc) return hip_fail((hipError_t)rc);
    HIPCHECK(hipEventRecord(B->ev1, B->stream));
    HIPCHECK(hipEventSynchronize(B->ev1));
    if (device_ms) HIPCHECK(hipEventElapsedTime(device_ms, B->ev0, B->ev1));
    B->pf_result.resize(B->pf_arena.size());
    HIPCHECK(hipMemcpy(B->pf_result.data(), B->d_pfa, B->pf_result.size() * 8, hipMemcpyDeviceToHost));
    // a bounded wait of the multi-workgroup path that expired invalidates the run
    for (const PfDesc &f : B->pdescs)
        if (B->pf_result[f.out + 2] != 0.0) return LCG_E_STATE;
    B->frag_prob.assign(B->npairs, {});
    for (int k = 0; k < B->npairs; ++k) {
        const PfDesc &f = B->pdescs[k];
        const double Z = B->pf_result[f.out], ps = B->pfh[k].ps;
        for (int q = 0; q < f.nfrag; ++q) {
            const double *x = B->pf_result.data() + f.frag + 8 * q;
            B->frag_prob[k].push_back(x[4] > 0.0 ? x[5] * x[6] * ps / Z : 0.0);
        }
    }
    B->pf_have = true;
    return LCG_OK;
}

namespace {
void drop_pf_state(lcg_batch &B);
}

extern "C" int lcg_batch_pf_set_fragments(lcg_batch *B, int32_t k, int32_t nq, const int32_t *ijkl) {
    if (!B || k < 0 || k >= B->npairs || nq < 0 || (nq > 0 && !ijkl)) return LCG_E_ARG;
    if (B->device_built) return LCG_E_STATE;  // LocARNA-P needs host-packed pairs
    const PfHost &h = B->pfh[k];
    std::vector<std::array<int32_t, 4>> q((size_t)nq);
    for (int t = 0; t < nq; ++t) {
        const int32_t *x = ijkl + 4 * t;
        if (x[0] < 1 || x[1] < x[0] || x[1] > h.n || x[2] < 1 || x[3] < x[2] || x[3] > h.m) return LCG_E_ARG;
        q[t] = {x[0], x[1], x[2], x[3]};
    }
    if (B->frags.empty()) B->frags.resize(B->npairs);
    // a larger query block than the pf arena holds: lay it out again
    if (B->d_pfa && (B->frag_cap.empty() || nq > B->frag_cap[k])) drop_pf_state(*B);
    B->frags[k] = std::move(q);
    return LCG_OK;
}

extern "C" int lcg_batch_pf_fragment_probs(const lcg_batch *B, int32_t k, double *prob, int32_t cap) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    if (!B->pf_have) return LCG_E_STATE;
    const std::vector<double> &p = B->frag_prob[k];
    if (cap < (int32_t)p.size() || (!prob && !p.empty())) return LCG_E_ARG;
    std::copy(p.begin(), p.end(), prob);
    return (int)p.size();
}

extern "C" int lcg_batch_pf_config(lcg_batch *B, int32_t *wg_per_pair, int32_t *nstatic, int32_t *workgroups) {
    if (!B) return LCG_E_ARG;
    if (B->device_built) return LCG_E_STATE;
    if (int rc = pf_setup(*B)) return rc;
    if (wg_per_pair) *wg_per_pair = B->pbp.wg_per_pair;
    if (nstatic) *nstatic = B->pbp.nstatic;
    if (workgroups) *workgroups = B->pf_grid;
    return LCG_OK;
}

extern "C" int lcg_batch_pf_result(const lcg_batch *B, int32_t k, double *Z, double *openings) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    if (!B->pf_have) return LCG_E_STATE;
    const PfDesc &f = B->pdescs[k];
    const double z = B->pf_result[f.out];
    if (Z) *Z = z;
    if (openings) *openings = B->pf_result[f.out + 1] / z;
    return LCG_OK;
}

extern "C" int lcg_batch_am_probs(const lcg_batch *B, int32_t k, double *prob, int32_t cap) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    if (!B->pf_have) return LCG_E_STATE;
    const PfHost &h = B->pfh[k];
    const PfDesc &f = B->pdescs[k];
    if (cap < h.nam) return LCG_E_ARG;
    const double Z = B->pf_result[f.out];
    for (int am = 0; am < h.nam; ++am) {
        const int t = h.entry_of[am];
        prob[am] = B->pf_result[f.Dc + t] * B->pf_result[f.Dbar + t] / Z;
    }
    return h.nam;
}

extern "C" int lcg_batch_pf_D(const lcg_batch *B, int32_t k, double *D, double *Dprime, int32_t cap) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    if (!B->pf_have) return LCG_E_STATE;
    const PfHost &h = B->pfh[k];
    const PfDesc &f = B->pdescs[k];
    if (cap < h.nam) return LCG_E_ARG;
    for (int am = 0; am < h.nam; ++am) {
        const int t = h.entry_of[am];
        if (D) D[am] = B->pf_result[f.Dc + t];
        if (Dprime) Dprime[am] = B->pf_result[f.Dbar + t] * B->pf_result[f.eam + t] / h.ps;
    }
    return h.nam;
}

extern "C" int lcg_batch_bm_probs(const lcg_batch *B, int32_t k, int32_t mode, double *dense, int32_t cap) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    if (!B->pf_have) return LCG_E_STATE;
    const PfHost &h = B->pfh[k];
    const PfDesc &f = B->pdescs[k];
    const int W = h.m + 1;
    if (cap < (h.n + 1) * W) return LCG_E_ARG;
    const double Z = B->pf_result[f.out];
    std::fill(dense, dense + (size_t)(h.n + 1) * W, 0.0);
    const int32_t edge = (mode == 2) ? f.edge_full : f.edge_ref;
    int cell = 0;
    for (int i = 0; i <= h.n; ++i) {
        for (int j = h.min_col[i]; j <= h.max_col[i]; ++j, ++cell) {
            if (i == 0 || j == 0) continue;
            const bool vm = h.min_col[i - 1] <= j - 1 && j - 1 <= h.max_col[i - 1];
            if (mode == 2 || vm) dense[(size_t)i * W + j] = B->pf_result[edge + cell] / Z;
        }
    }
    if (mode == 1) {  // --include-am-in-bm (aligner_p.icc:1382-1397)
        for (int am = 0; am < h.nam; ++am) {
            const int t = h.entry_of[am];
            const double p = B->pf_result[f.Dc + t] * B->pf_result[f.Dbar + t] / Z;
            dense[(size_t)h.am_ends[4 * am] * W + h.am_ends[4 * am + 2]] += p;
            dense[(size_t)h.am_ends[4 * am + 1] * W + h.am_ends[4 * am + 3]] += p;
        }
    }
    return (h.n + 1) * W;
}

namespace {
int64_t put_text(const std::string &text, char *buf, int64_t cap) {
    if (buf && cap > (int64_t)text.size()) std::memcpy(buf, text.c_str(), text.size() + 1);
    return (int64_t)text.size();
}
}  // namespace

extern "C" int64_t lcg_batch_write_am_probs(const lcg_batch *B, int32_t k, char *buf, int64_t cap) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    if (!B->pf_have) return LCG_E_STATE;
    const PfHost &h = B->pfh[k];
    std::vector<double> prob(std::max(h.nam, 1));
    if (int rc = lcg_batch_am_probs(B, k, prob.data(), (int32_t)prob.size()); rc < 0) return rc;
    std::string text;
    char line[128];
    for (int am = 0; am < h.nam; ++am) {
        if (!(prob[am] >= h.min_am_prob)) continue;
        // ostream << double: %g, precision 6
        std::snprintf(line, sizeof line, "%d %d %d %d %g\n", h.am_ends[4 * am], h.am_ends[4 * am + 1],
                      h.am_ends[4 * am + 2], h.am_ends[4 * am + 3], prob[am]);
        text += line;
    }
    return put_text(text, buf, cap);
}

extern "C" int64_t lcg_batch_write_bm_probs(const lcg_batch *B, int32_t k, int32_t include_am_in_bm, char *buf,
                                            int64_t cap) {
    if (!B || k < 0 || k >= B->npairs) return LCG_E_ARG;
    if (!B->pf_have) return LCG_E_STATE;
    const PfHost &h = B->pfh[k];
    const int W = h.m + 1;
    std::vector<double> dense((size_t)(h.n + 1) * W);
    if (int rc = lcg_batch_bm_probs(B, k, include_am_in_bm ? 1 : 0, dense.data(), (int32_t)dense.size()); rc < 0)
        return rc;
    std::string text;
    char line[96];
    for (int i = 1; i <= h.n; ++i) {
        for (int j = 1; j <= h.m; ++j) {
            const double p = dense[(size_t)i * W + j];
            if (!(p >= h.min_bm_prob)) continue;
            std::snprintf(line, sizeof line, "%d %d %g\n", i, j, p);
            text += line;
        }
    }
    return put_text(text, buf, cap);
}

namespace {
// ------------------------------------------------------------ sampling
// Stochastic traceback of LocARNA-P: alignments drawn from the Boltzmann
// distribution of the last lcg_batch_pf_run.  Host code: the inside
// matrices of the top level and of every group a sample enters are
// recomputed from the device's D values with the same tables and the same
// recurrence as the kernels, then sampled backwards with the affine states
// (total, s = not ending with an insertion, E = deletion, F = insertion):
//   E(i,j) = E(i-1,j) gA + (M(i-1,j) - E(i-1,j)) gA o
//   s(i,j) = M(i-1,j-1) es(i,j) + E(i,j) + ps sum_arcs M(p-1,q-1) D(am)
//   F(i,j) = F(i,j-1) gB + s(i,j-1) gB o,   M = s + F
// Row al and column bl of a region hold the initial gap runs.  Global top
// level only.
struct SampleRegion {
    int al = 0, bl = 0, ar = 0, br = 0, w = 0;
    std::vector<double> M, E, F, S;  // dense (ar - al) x (br - bl + 1), 0 outside
    std::vector<uint8_t> st;         // stored cell
    size_t at(int i, int j) const { return (size_t)(i - al) * w + (j - bl); }
    bool in(int i, int j) const { return i >= al && i < ar && j >= bl && j <= br - 1 && st[at(i, j)]; }
    double get(const std::vector<double> &X, int i, int j) const { return in(i, j) ? X[at(i, j)] : 0.0; }
};

struct Sampler {
    const lcg_batch &B;
    int k;
    const PfHost &h;
    const PfDesc &f;
    const PairDesc &d;
    const int32_t *A;        // host arena
    const double *tab;       // Boltzmann tables of the last run
    const double *res;       // device results of the last run (Dc)
    double eo, ps;
    std::vector<int> c2am;   // candidate entry -> arc match
    std::map<std::pair<int, int>, int> group_of;
    std::map<int, SampleRegion> cache;  // region by group (-1: top level)
    uint64_t rng;

    Sampler(const lcg_batch &B_, int k_, uint64_t seed)
        : B(B_), k(k_), h(B_.pfh[k_]), f(B_.pdescs[k_]), d(B_.descs[k_]), A(B_.host_arena.data()),
          tab(B_.pf_arena.data()), res(B_.pf_result.data()), rng(seed) {
        eo = tab[f.scal + 0];
        ps = tab[f.scal + 1];
        c2am.assign(h.nam, -1);
        for (int am = 0; am < h.nam; ++am) c2am[h.entry_of[am]] = am;
        for (int g = 0; g < d.ngroups; ++g) {
            const int32_t *G = A + d.groups + g * LCG_GROUP_INTS;
            group_of[{G[0] & 0xffff, G[0] >> 16}] = g;
        }
    }
    double uniform() {  // splitmix64, u = (x >> 11) 2^-53
        uint64_t z = (rng += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
    double gA(int i) const { return tab[f.egA + i]; }
    double gB(int j) const { return tab[f.egB + j]; }
    double es(int i, int j) const { return tab[f.es + h.codeA[i] * LCG_NCODES + h.codeB[j]]; }
    int mn(int i) const { return h.min_col[i]; }
    int mx(int i) const { return h.max_col[i]; }
    int cell(int i, int j) const { return A[d.band + 2 * (h.n + 1) + i] + (j - mn(i)); }

    const SampleRegion &region(int g) {
        auto it = cache.find(g);
        if (it != cache.end()) return it->second;
        SampleRegion R;
        if (g < 0) {
            R.al = 0, R.bl = 0, R.ar = h.n + 1, R.br = h.m + 1;
        } else {
            const int32_t *G = A + d.groups + g * LCG_GROUP_INTS;
            R.al = G[0] & 0xffff, R.bl = G[0] >> 16, R.ar = G[1] & 0xffff, R.br = G[1] >> 16;
        }
        R.w = R.br - R.bl + 1;
        const size_t cells = (size_t)(R.ar - R.al) * R.w;
        R.M.assign(cells, 0.0);
        R.E.assign(cells, 0.0);
        R.F.assign(cells, 0.0);
        R.S.assign(cells, 0.0);
        R.st.assign(cells, 0);
        const int al = R.al, bl = R.bl, ar = R.ar, br = R.br;
        // row al: M(al, bl) = 1/ps, M(al, j) = o/ps prod gB
        const int hi0 = std::min(br - 1, mx(al));
        R.st[R.at(al, bl)] = 1;
        R.M[R.at(al, bl)] = R.S[R.at(al, bl)] = 1.0 / ps;
        double run = eo / ps;
        for (int j = bl + 1; j <= hi0; ++j) {
            run *= gB(j);
            R.st[R.at(al, j)] = 1;
            R.M[R.at(al, j)] = R.S[R.at(al, j)] = run;
        }
        std::vector<double> Ev(br - bl + 1, 0.0);
        double colv = eo / ps;
        bool colopen = true;
        for (int i = al + 1; i < ar; ++i) {
            if (colopen && mn(i) <= bl) {
                colv *= gA(i);
                R.st[R.at(i, bl)] = 1;
                R.M[R.at(i, bl)] = R.S[R.at(i, bl)] = colv;
            } else {
                colopen = false;
            }
            const int jlo = std::max(bl + 1, mn(i)), hi = std::min(br - 1, mx(i));
            double Fprev = 0.0, sprev = colopen ? colv : 0.0;
            for (int j = jlo; j <= hi; ++j) {
                const double up = R.get(R.M, i - 1, j), dg = R.get(R.M, i - 1, j - 1);
                const double ep = Ev[j - bl];
                const double e = ep * gA(i) + (up - ep) * gA(i) * eo;
                Ev[j - bl] = e;
                double sv = dg * es(i, j) + e;
                const int c = cell(i, j);
                double sa = 0.0;
                for (int t = A[d.cptr + c]; t < A[d.cptr + c + 1]; ++t) {
                    const int p = A[d.cand + 4 * t] & 0xffff, q = A[d.cand + 4 * t] >> 16;
                    if (p <= al) break;
                    if (q <= bl) continue;
                    sa += R.get(R.M, p - 1, q - 1) * res[f.Dc + t];
                }
                sv += sa * ps;
                const double Fv = Fprev * gB(j) + sprev * gB(j) * eo;
                const size_t x = R.at(i, j);
                R.st[x] = 1;
                R.E[x] = e;
                R.S[x] = sv;
                R.F[x] = Fv;
                R.M[x] = sv + Fv;
                Fprev = Fv;
                sprev = sv;
            }
        }
        return cache.emplace(g, std::move(R)).first->second;
    }

    // features of one alignment (score = reference score + theta . f)
    std::vector<double> feat;
    int nmatch = 0;
    void open_run() { feat[273] += 1.0; }
    void match_edge(int i, int j) {
        ++nmatch;
        const int a = h.codeA[i], b = h.codeB[j];
        if (a < 4 && b < 4) feat[a * 4 + b] += 1.0;
    }

    // samples region R backwards from (i, j) in the total state
    void sample(int g, int i, int j) {
        const SampleRegion *Rp = &region(g);
        enum { TOT, SS, SNE, EE, FF } state = TOT;  // SNE: s without the deletion term
        for (;;) {
            const SampleRegion &R = *Rp;
            if (state == TOT || state == SS || state == SNE) {
                if (i == R.al) {  // initial insertion run of row al
                    if (j > R.bl) open_run();
                    return;
                }
                if (j == R.bl) {  // initial deletion run of column bl
                    open_run();
                    return;
                }
            }
            const size_t x = R.at(i, j);
            if (state == TOT) {
                state = uniform() * R.M[x] < R.F[x] ? FF : SS;
                continue;
            }
            if (state == FF) {  // insertion of j
                const double ext = R.get(R.F, i, j - 1) * gB(j);
                const double tot = R.F[x];
                j -= 1;
                if (uniform() * tot < ext) {
                    state = FF;
                } else {
                    open_run();
                    state = SS;
                }
                continue;
            }
            if (state == EE) {  // deletion of i
                const double Ep = R.get(R.E, i - 1, j), Mp = R.get(R.M, i - 1, j);
                const double ext = Ep * gA(i), opn = std::max(0.0, Mp - Ep) * gA(i) * eo;
                i -= 1;
                if (uniform() * (ext + opn) < ext) {
                    state = EE;
                } else {
                    open_run();
                    // total without the deletion states: F or s without E
                    const SampleRegion &Q = *Rp;
                    if (i == Q.al || j == Q.bl) {
                        state = TOT;
                    } else {
                        const size_t y = Q.at(i, j);
                        const double sne = std::max(0.0, Q.S[y] - Q.E[y]);
                        state = uniform() * (sne + Q.F[y]) < Q.F[y] ? FF : SNE;
                    }
                }
                continue;
            }
            // SS / SNE: match, (deletion), arc matches
            const double mt = R.get(R.M, i - 1, j - 1) * es(i, j);
            const double de = state == SS ? R.E[x] : 0.0;
            double u = uniform() * (state == SS ? R.S[x] : std::max(0.0, R.S[x] - R.E[x]));
            if (u < mt) {
                match_edge(i, j);
                i -= 1;
                j -= 1;
                state = TOT;
                continue;
            }
            u -= mt;
            if (u < de) {
                state = EE;
                continue;
            }
            u -= de;
            const int c = cell(i, j);
            int pick = -1, last = -1;
            for (int t = A[d.cptr + c]; t < A[d.cptr + c + 1]; ++t) {
                const int p = A[d.cand + 4 * t] & 0xffff, q = A[d.cand + 4 * t] >> 16;
                if (p <= R.al) break;
                if (q <= R.bl) continue;
                const double v = R.get(R.M, p - 1, q - 1) * res[f.Dc + t] * ps;
                if (v <= 0.0) continue;
                last = t;
                if (u < v) {
                    pick = t;
                    break;
                }
                u -= v;
            }
            if (pick < 0) pick = last;  // rounding at the end of the list
            if (pick < 0) {  // no arc term: rounding in the choice between match and deletion
                if (mt > 0.0) {
                    match_edge(i, j);
                    i -= 1;
                    j -= 1;
                    state = TOT;
                } else {
                    state = EE;
                }
                continue;
            }
            const int am = c2am[pick];
            const int p = A[d.cand + 4 * pick] & 0xffff, q = A[d.cand + 4 * pick] >> 16;
            nmatch += 2;
            if (h.am_type[am] >= 0) feat[16 + h.am_type[am]] += h.tau / 100.0;
            feat[274] += h.am_f[am];
            // inside of the arc match: its group from (i-1, j-1) back to (p, q)
            sample(group_of.at({p, q}), i - 1, j - 1);
            Rp = &region(g);
            i = p - 1;
            j = q - 1;
            state = TOT;
        }
    }

    std::vector<double> one() {
        feat.assign(LCG_NTHETA, 0.0);
        nmatch = 0;
        sample(-1, h.n, h.m);
        feat[272] = h.n + h.m - 2.0 * nmatch;
        return feat;
    }
};
}  // namespace

extern "C" int lcg_batch_pf_sample(const lcg_batch *B, int32_t k, uint64_t seed, int32_t nsamples, double *mean,
                                   double *var, int32_t cap) {
    if (!B || k < 0 || k >= B->npairs || nsamples < 1 || !mean || cap < LCG_NTHETA) return LCG_E_ARG;
    if (!B->pf_have) return LCG_E_STATE;
    if (B->bp.local) return LCG_E_ARG;  // the sampler follows the global top level
    try {
        Sampler S(*B, k, seed);
        std::vector<double> sum(LCG_NTHETA, 0.0), sum2(LCG_NTHETA, 0.0);
        for (int r = 0; r < nsamples; ++r) {
            const std::vector<double> x = S.one();
            for (int t = 0; t < LCG_NTHETA; ++t) {
                sum[t] += x[t];
                sum2[t] += x[t] * x[t];
            }
        }
        for (int t = 0; t < LCG_NTHETA; ++t) {
            mean[t] = sum[t] / nsamples;
            if (var) var[t] = std::max(0.0, sum2[t] / nsamples - mean[t] * mean[t]);
        }
        return LCG_NTHETA;
    } catch (std::exception &) {
        return LCG_E_STATE;
    }
}

extern "C" int lcg_batch_pf_grad(const lcg_batch *B, int32_t k, double *grad, int32_t cap) {
    if (!B || k < 0 || k >= B->npairs || cap < LCG_NTHETA) return LCG_E_ARG;
    if (!B->pf_have) return LCG_E_STATE;
    const PfHost &h = B->pfh[k];
    const PfDesc &f = B->pdescs[k];
    const double Z = B->pf_result[f.out];
    const double T = h.T;
    std::fill(grad, grad + LCG_NTHETA, 0.0);
    double matches = 0.0;
    int cell = 0;
    for (int i = 0; i <= h.n; ++i) {
        for (int j = h.min_col[i]; j <= h.max_col[i]; ++j, ++cell) {
            if (i == 0 || j == 0) continue;
            const double p = B->pf_result[f.edge_full + cell] / Z;
            matches += p;
            const int a = h.codeA[i], b = h.codeB[j];
            if (a < 4 && b < 4) grad[a * 4 + b] += p / T;
        }
    }
    double dsw = 0.0;
    for (int am = 0; am < h.nam; ++am) {
        const int t = h.entry_of[am];
        const double p = B->pf_result[f.Dc + t] * B->pf_result[f.Dbar + t] / Z;
        matches += 2 * p;
        if (h.am_type[am] >= 0) grad[16 + h.am_type[am]] += (h.tau / 100.0) * p / T;
        dsw += p * h.am_f[am];
    }
    grad[272] = (h.n + h.m - 2 * matches) / T;
    grad[273] = (B->pf_result[f.out + 1] / Z) / T;
    grad[274] = dsw / T;
    return LCG_NTHETA;
}

namespace {
// Fingerprint of the integer results of the last run: score, alignment
// edges and structure strings of every pair (FNV-1a over the words).
int result_digest(const lcg_batch *B, uint64_t &h) {
    h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) {
        h ^= v;
        h *= 1099511628211ull;
    };
    std::vector<int32_t> a, b;
    std::vector<char> sa, sb;
    for (int k = 0; k < B->npairs; ++k) {
        int64_t sc = 0;
        if (int rc = lcg_batch_score(B, k, &sc)) return rc;
        mix((uint64_t)sc);
        const PairDesc &d = B->descs[k];
        const int cap = d.n + d.m + 2;
        a.assign(cap, 0);
        b.assign(cap, 0);
        sa.assign(cap, 0);
        sb.assign(cap, 0);
        const int cnt = lcg_batch_alignment(B, k, a.data(), b.data(), cap, sa.data(), sb.data(), cap);
        if (cnt < 0) return cnt;
        for (int t = 0; t < cnt; ++t) mix(((uint64_t)(uint32_t)a[t] << 32) | (uint32_t)b[t]);
        for (int t = 0; t < d.n; ++t) mix((uint64_t)sa[t]);
        for (int t = 0; t < d.m; ++t) mix((uint64_t)sb[t]);
    }
    return LCG_OK;
}

void drop_pf_state(lcg_batch &B) {
    if (B.d_pfa) hipFree(B.d_pfa);
    if (B.d_pdescs) hipFree(B.d_pdescs);
    if (B.d_pfscratch) hipFree(B.d_pfscratch);
    if (B.d_pfsched) hipFree(B.d_pfsched);
    if (B.d_pfstore) hipFree(B.d_pfstore);
    B.d_pfa = nullptr;
    B.d_pdescs = nullptr;
    B.d_pfscratch = nullptr;
    B.d_pfsched = nullptr;
    B.d_pfstore = nullptr;
    B.pf_have = false;
}

std::string shape_text(const lcg_shape &s) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "waves %d, wgs/CU %d, wgs/pair %d, top waves %d, top wgs/CU %d, pf waves %d, pf wgs/pair %d",
                  s.waves, s.wgs_per_cu, s.wg_per_pair, s.top_waves, s.top_wgs_per_cu, s.pf_waves, s.pf_wg_per_pair);
    return buf;
}
}  // namespace

extern "C" int lcg_batch_get_shape(const lcg_batch *B, lcg_shape *shape) {
    if (!B || !shape) return LCG_E_ARG;
    *shape = B->eff;
    if (!B->d_pfa) {  // LocARNA-P not set up yet: what it would take
        shape->pf_waves = pick(B->req.pf_waves, "LCG_PF_WAVES", 8, 1, 8);
        shape->pf_wg_per_pair = B->req.pf_wg_per_pair;
    }
    return LCG_OK;
}

extern "C" int lcg_batch_reshape(lcg_batch *B, const lcg_shape *shape, char *err, int errlen) {
    if (!B || !shape) return LCG_E_ARG;
    HIPCHECK(hipSetDevice(B->device));
    HIPCHECK(hipStreamSynchronize(B->stream));
    const lcg_shape old = B->req;
    FillShape fs0;
    if (int rc = choose_fill(*B, fs0, err, errlen)) return rc;
    B->req = *shape;
    FillShape fs;
    if (int rc = choose_fill(*B, fs, err, errlen)) {
        B->req = old;
        return rc;
    }
    const TopShape ts = choose_top(*B);
    if (fs.waves != fs0.waves || fs.pool != fs0.pool || fs.K != fs0.K) {
        if (B->device_built) {
            // as at creation: both compaction passes (the build checks the
            // counts of the count pass), the image over a cleared arena
            auto rebuild = [&](const FillShape &f, char *e, int elen) -> int {
                HIPCHECK(hipMemsetAsync(B->d_arena, 0, B->arena_ints * 4, B->stream));
                if (int rc = lcg_launch_compact_count(B->d_cdescs, B->d_cin, B->d_ws, B->d_cstats, B->npairs,
                                                      B->comp_lds, B->stream))
                    return hip_fail((hipError_t)rc, e, elen);
                return device_build(*B, B->npairs, B->max_len, f.waves, f.pool, f.K, e, elen);
            };
            if (int rc = rebuild(fs, err, errlen)) {
                // back to the shape that was built before
                B->req = old;
                char e2[64];
                if (rebuild(fs0, e2, sizeof e2)) return LCG_E_STATE;
                return rc;
            }
        } else {
            if (!schedule_all(*B, fs)) {
                B->req = old;
                schedule_all(*B, fs0);  // succeeded before
                set_err(err, errlen, "too many LDS rounds in one row");
                return LCG_E_CAPACITY;
            }
            HIPCHECK(hipMemcpy(B->d_arena, B->host_arena.data(), B->arena_ints * 4, hipMemcpyHostToDevice));
        }
    }
    const size_t old_scratch = scratch_ints(*B);
    apply_shape(*B, fs, ts);
    const size_t new_scratch = scratch_ints(*B);
    if (new_scratch > old_scratch) {
        HIPCHECK(hipFree(B->d_scratch));
        B->d_scratch = nullptr;
        HIPCHECK(hipMalloc(&B->d_scratch, new_scratch * 4));
    }
    if (B->graph_exec) hipGraphExecDestroy(B->graph_exec);
    if (B->graph) hipGraphDestroy(B->graph);
    B->graph_exec = nullptr;
    B->graph = nullptr;
    B->graph_trace = -1;
    B->have_results = false;
    B->fetched = false;
    if (B->d_pfa && (shape->pf_waves != old.pf_waves || shape->pf_wg_per_pair != old.pf_wg_per_pair))
        drop_pf_state(*B);  // set up again with the new shape by the next LocARNA-P run
    return LCG_OK;
}

extern "C" int lcg_batch_autotune(lcg_batch *B, int32_t path, int32_t reps, lcg_shape *best, float *best_ms,
                                  float *default_ms, char *err, int errlen) {
    if (!B || (path != 0 && path != 1) || reps < 1 || reps > 100) return LCG_E_ARG;
    const lcg_shape start = B->req;
    lcg_shape cur;
    lcg_batch_get_shape(B, &cur);
    const int np = B->npairs, cus = B->cus;
    auto fail = [&](int rc, const std::string &msg) {
        set_err(err, errlen, msg.c_str());
        char e2[64];
        lcg_batch_reshape(B, &start, e2, sizeof e2);
        return rc;
    };
    if (path == 0) {
        // integer path: fill shapes first (timed by the fill kernel alone),
        // then top-level shapes with the best fill; every candidate's scores
        // and alignments must equal those of the starting shape
        float g0 = 1e30f;
        for (int r = 0; r <= reps; ++r) {
            float ms = 0;
            if (int rc = lcg_batch_run(B, 2, 1, &ms)) return fail(rc, "run failed at the starting shape");
            if (r > 0) g0 = std::min(g0, ms);
        }
        uint64_t ref = 0;
        if (int rc = result_digest(B, ref)) return fail(rc, "results of the starting shape unavailable");
        auto time_shape = [&](const lcg_shape &s, float &fill, float &top, bool &skip) -> int {
            char e2[128];
            skip = false;
            const int rc = lcg_batch_reshape(B, &s, e2, sizeof e2);
            if (rc == LCG_E_CAPACITY) {  // shape does not fit these pairs
                skip = true;
                return LCG_OK;
            }
            if (rc) return rc;
            fill = top = 1e30f;
            for (int r = 0; r <= reps; ++r) {
                if (int rc2 = lcg_batch_run(B, 0, 1, nullptr)) return rc2;
                if (r > 0) {
                    fill = std::min(fill, B->fill_ms);
                    top = std::min(top, B->top_ms);
                }
            }
            uint64_t h = 0;
            if (int rc2 = result_digest(B, h)) return rc2;
            if (h != ref) return LCG_E_STATE;
            return LCG_OK;
        };
        std::vector<lcg_shape> fills;
        auto add_fill = [&](int w, int g, int k) {
            lcg_shape s = cur;
            s.waves = w;
            s.wgs_per_cu = g;
            s.wg_per_pair = k;
            fills.push_back(s);
        };
        if (np >= cus) {
            const int wg[][2] = {{16, 1}, {12, 1}, {8, 1}, {8, 2}, {4, 2}, {4, 4}};
            for (auto &x : wg) add_fill(x[0], x[1], 1);
        } else {
            const int kmax = std::min(64, cus / np);
            std::vector<int> ks = {kmax, kmax / 2, kmax / 4, 1};
            ks.erase(std::remove_if(ks.begin(), ks.end(), [](int k) { return k < 1; }), ks.end());
            std::sort(ks.begin(), ks.end(), std::greater<int>());
            ks.erase(std::unique(ks.begin(), ks.end()), ks.end());
            for (int k : ks) add_fill(16, 1, k);
            add_fill(8, 1, kmax);
        }
        lcg_shape bestf = cur;
        float bf = 1e30f;
        for (const lcg_shape &s : fills) {
            float fill = 0, top = 0;
            bool skip = false;
            const int rc = time_shape(s, fill, top, skip);
            if (rc == LCG_E_STATE) return fail(rc, "results differ at shape " + shape_text(s));
            if (rc) return fail(rc, "run failed at shape " + shape_text(s));
            if (!skip && fill < bf) {
                bf = fill;
                lcg_batch_get_shape(B, &bestf);
            }
        }
        const int tw[][2] = {{4, 4}, {8, 2}, {2, 4}, {16, 1}, {8, 1}, {4, 2}};
        lcg_shape bestt = bestf;
        float bt = 1e30f;
        for (auto &x : tw) {
            lcg_shape s = bestf;
            s.top_waves = x[0];
            s.top_wgs_per_cu = x[1];
            float fill = 0, top = 0;
            bool skip = false;
            const int rc = time_shape(s, fill, top, skip);
            if (rc == LCG_E_STATE) return fail(rc, "results differ at shape " + shape_text(s));
            if (rc) return fail(rc, "run failed at shape " + shape_text(s));
            if (!skip && top < bt) {
                bt = top;
                lcg_batch_get_shape(B, &bestt);
            }
        }
        // the winner, timed as the starting shape was (graph replay); kept
        // only if it beats the starting shape
        char e2[128];
        if (int rc = lcg_batch_reshape(B, &bestt, e2, sizeof e2)) return fail(rc, e2);
        float g1 = 1e30f;
        for (int r = 0; r <= reps; ++r) {
            float ms = 0;
            if (int rc = lcg_batch_run(B, 2, 1, &ms)) return fail(rc, "run failed at the chosen shape");
            if (r > 0) g1 = std::min(g1, ms);
        }
        if (g1 > g0) {
            if (int rc = lcg_batch_reshape(B, &start, e2, sizeof e2)) return rc;
            g1 = g0;
        }
        if (default_ms) *default_ms = g0;
        if (best_ms) *best_ms = g1;
        if (best) lcg_batch_get_shape(B, best);
        return LCG_OK;
    }
    // LocARNA-P: waves and workgroups per pair of the inside/outside kernel;
    // every candidate's log Z must agree with the starting shape's to 1e-12
    if (B->device_built) return LCG_E_STATE;
    auto time_pf = [&](float &t) -> int {
        t = 1e30f;
        for (int r = 0; r <= reps; ++r) {
            float ms = 0;
            if (int rc = lcg_batch_pf_run(B, nullptr, &ms)) return rc;
            if (r > 0) t = std::min(t, ms);
        }
        return LCG_OK;
    };
    float p0 = 0;
    if (int rc = time_pf(p0)) return fail(rc, "LocARNA-P run failed at the starting shape");
    std::vector<double> z0(np);
    for (int k = 0; k < np; ++k) z0[k] = B->pf_result[B->pdescs[k].out];
    std::vector<lcg_shape> cands;
    std::vector<int> ks = {0, 1};
    if (np < cus) {
        const int kmax = std::min(64, cus / np);
        for (int k : {kmax, kmax / 2, 2})
            if (k > 1) ks.push_back(k);
    }
    std::sort(ks.begin(), ks.end());
    ks.erase(std::unique(ks.begin(), ks.end()), ks.end());
    for (int w : {8, 4})
        for (int k : ks) {
            lcg_shape s = B->req;
            s.pf_waves = w;
            s.pf_wg_per_pair = k;
            cands.push_back(s);
        }
    lcg_shape bests = B->req;
    float bp = p0;
    for (const lcg_shape &s : cands) {
        char e2[128];
        if (int rc = lcg_batch_reshape(B, &s, e2, sizeof e2)) return fail(rc, e2);
        float t = 0;
        if (int rc = time_pf(t)) return fail(rc, "LocARNA-P run failed at shape " + shape_text(s));
        for (int k = 0; k < np; ++k) {
            const double z = B->pf_result[B->pdescs[k].out];
            if (!(std::fabs(z - z0[k]) <= 1e-12 * std::fabs(z0[k])))
                return fail(LCG_E_STATE, "partition functions differ at shape " + shape_text(s));
        }
        if (t < bp) {
            bp = t;
            bests = s;
        }
    }
    char e2[128];
    if (int rc = lcg_batch_reshape(B, &bests, e2, sizeof e2)) return fail(rc, e2);
    if (default_ms) *default_ms = p0;
    if (best_ms) *best_ms = bp;
    if (best) lcg_batch_get_shape(B, best);
    return LCG_OK;
}

extern "C" size_t lcg_band_lds_bytes(int lmax);
extern "C" int lcg_launch_band(const BandDesc *descs, const int *order, int *queue, int *iarena,
                               const double *darena, double *fwd, BandParams bp, int grid, hipStream_t stream);

namespace {
// fn(k) for k in [0, n) on `threads` host threads.  Returns the first k
// whose call threw (its message in msg), or -1.
template <class Fn>
int parallel_for(int n, int threads, Fn fn, std::string &msg, bool *capacity = nullptr) {
    std::atomic<int> next{0}, bad{-1};
    std::mutex mu;
    auto work = [&]() {
        for (;;) {
            const int k = next.fetch_add(1);
            if (k >= n || bad.load() >= 0) return;
            try {
                fn(k);
            } catch (std::exception &e) {
                std::lock_guard<std::mutex> g(mu);
                if (bad.load() < 0) {
                    msg = e.what();
                    if (capacity) *capacity = dynamic_cast<capacity_failure *>(&e) != nullptr;
                    bad.store(k);
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < std::min(threads, n); ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    return bad.load();
}

// Device memory of one band launch, freed on every path.
struct BandBuffers {
    BandDesc *descs = nullptr;
    int32_t *order = nullptr, *queue = nullptr, *iarena = nullptr;
    double *darena = nullptr, *fwd = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ~BandBuffers() {
        if (descs) hipFree(descs);
        if (order) hipFree(order);
        if (queue) hipFree(queue);
        if (iarena) hipFree(iarena);
        if (darena) hipFree(darena);
        if (fwd) hipFree(fwd);
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
        if (stream) hipStreamDestroy(stream);
    }
};

// Trace bands of the pairs `todo` (all with the same parameters) on the
// device, in chunks of at most 4 GiB of forward values.  A pair the kernel
// flags keeps its unrestricted band and is appended to `recheck` (the host
// bands it); the others get the kernel's rows.
int device_bands(std::vector<std::unique_ptr<lcg_pair>> &P, std::vector<int> todo, const lcg_params &p,
                 int device, double eps, int threads, std::vector<int> &recheck, float &ms_total, char *err,
                 int errlen) {
    if (todo.empty()) return LCG_OK;
    auto cells = [&](int k) { return (int64_t)(P[k]->A->length() + 1) * (P[k]->B->length() + 1); };
    std::stable_sort(todo.begin(), todo.end(), [&](int a, int b) { return cells(a) > cells(b); });
    hipError_t e;
    int cus = 0;
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, err, errlen);
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
        return hip_fail(e, err, errlen);
    BandParams bp;
    std::memset(&bp, 0, sizeof bp);
    const ScoringParams &sp = P[todo[0]]->sp;
    // as the host band (edge_probs.icc:76-162 with main_helper.icc:386-396)
    bp.temp = sp.temperature_alipf / 100.0;
    bp.g_open = std::exp((sp.indel_opening / 100.0) / bp.temp);
    bp.g_ext = std::exp((sp.indel / 100.0) / bp.temp);
    bp.sw = sp.struct_weight / 100.0;
    bp.thr = p.min_trace_prob;
    bp.eps = eps;
    band_similarity(sp, bp.sim);
    bp.local = p.sequ_local != 0;
    bp.extended = p.band_extended != 0;
    bp.free_endgaps = p.free_endgaps;
    const int64_t budget = (int64_t)1 << 29;  // doubles (4 GiB)
    size_t c0 = 0;
    while (c0 < todo.size()) {
        size_t c1 = c0;
        int64_t fw = 0;
        while (c1 < todo.size() && (c1 == c0 || fw + 3 * cells(todo[c1]) <= budget)) fw += 3 * cells(todo[c1++]);
        const int np = (int)(c1 - c0);
        // layout first (offsets from the lengths), then every pair's inputs
        // on the host threads (the structure profiles sum over the pairs of
        // each position)
        std::vector<BandDesc> descs(np);
        size_t ni = 0, nd = 0;
        int lmax = 1;
        int64_t foff = 0;
        for (int q = 0; q < np; ++q) {
            const lcg_pair &X = *P[todo[c0 + q]];
            const int n = X.A->length(), m = X.B->length();
            BandDesc &d = descs[q];
            d.n = n;
            d.m = m;
            lmax = std::max(lmax, std::max(n, m));
            d.codes = (int32_t)ni;
            d.band = (int32_t)(ni + n + m + 2);
            d.rows = (int32_t)(ni + n + m + 2 + 2 * (n + 1));
            ni += (size_t)(n + m + 2) + 2 * (n + 1) + 2 * (n + 1) + 1;
            d.prof = (int32_t)nd;
            nd += 2 * (size_t)(n + 1) + 2 * (size_t)(m + 1);
            d.fwd = foff;
            foff += 3 * (int64_t)(n + 1) * (m + 1);
            if (ni >= (size_t)INT_MAX || nd >= (size_t)INT_MAX) break;
        }
        if (ni >= (size_t)INT_MAX || nd >= (size_t)INT_MAX) {
            set_err(err, errlen, "band batch too large");
            return LCG_E_CAPACITY;
        }
        std::vector<int32_t> ia(ni, 0);
        std::vector<double> da(nd, 0.0);
        std::string msg;
        parallel_for(
            np, threads,
            [&](int q) {
                const lcg_pair &X = *P[todo[c0 + q]];
                const BandDesc &d = descs[q];
                const int n = d.n, m = d.m;
                int32_t *c = ia.data() + d.codes;
                c[0] = -1;
                for (int i = 1; i <= n; ++i) c[i] = nt_code(X.A->nt(i));
                c[n + 1] = -1;
                for (int j = 1; j <= m; ++j) c[n + 1 + j] = nt_code(X.B->nt(j));
                int32_t *b = ia.data() + d.band;
                for (int i = 0; i <= n; ++i) {
                    b[i] = X.tc->min_col(i);
                    b[n + 1 + i] = X.tc->max_col(i);
                }
                double *f = da.data() + d.prof;
                for (int i = 1; i <= n; ++i) {
                    f[i] = X.A->prob_paired_upstream(i);
                    f[n + 1 + i] = X.A->prob_paired_downstream(i);
                }
                double *g = f + 2 * (n + 1);
                for (int j = 1; j <= m; ++j) {
                    g[j] = X.B->prob_paired_upstream(j);
                    g[m + 1 + j] = X.B->prob_paired_downstream(j);
                }
            },
            msg);
        bp.npairs = np;
        bp.lmax = lmax;
        std::vector<int32_t> order(np);
        std::iota(order.begin(), order.end(), 0);
        const size_t lds = lcg_band_lds_bytes(lmax);
        const int per_cu = std::max(1, std::min(8, (int)((160 * 1024 - 256) / lds)));
        const int grid = std::min(np, cus * per_cu);
        BandBuffers bb;
        if ((e = hipStreamCreate(&bb.stream)) != hipSuccess) return hip_fail(e, err, errlen);
        if ((e = hipEventCreate(&bb.ev0)) != hipSuccess) return hip_fail(e, err, errlen);
        if ((e = hipEventCreate(&bb.ev1)) != hipSuccess) return hip_fail(e, err, errlen);
        if ((e = hipMalloc(&bb.descs, sizeof(BandDesc) * np)) != hipSuccess) return hip_fail(e, err, errlen);
        if ((e = hipMalloc(&bb.order, 4 * (size_t)np)) != hipSuccess) return hip_fail(e, err, errlen);
        if ((e = hipMalloc(&bb.queue, 256)) != hipSuccess) return hip_fail(e, err, errlen);
        if ((e = hipMalloc(&bb.iarena, 4 * ia.size())) != hipSuccess) return hip_fail(e, err, errlen);
        if ((e = hipMalloc(&bb.darena, 8 * std::max<size_t>(1, da.size()))) != hipSuccess)
            return hip_fail(e, err, errlen);
        if ((e = hipMalloc(&bb.fwd, 8 * (size_t)foff)) != hipSuccess) return hip_fail(e, err, errlen);
        if ((e = hipMemcpy(bb.descs, descs.data(), sizeof(BandDesc) * np, hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, err, errlen);
        if ((e = hipMemcpy(bb.order, order.data(), 4 * (size_t)np, hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, err, errlen);
        if ((e = hipMemcpy(bb.iarena, ia.data(), 4 * ia.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, err, errlen);
        if ((e = hipMemcpy(bb.darena, da.data(), 8 * da.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, err, errlen);
        if ((e = hipEventRecord(bb.ev0, bb.stream)) != hipSuccess) return hip_fail(e, err, errlen);
        if (int rc = lcg_launch_band(bb.descs, bb.order, bb.queue, bb.iarena, bb.darena, bb.fwd, bp, grid,
                                     bb.stream))
            return hip_fail((hipError_t)rc, err, errlen);
        if ((e = hipEventRecord(bb.ev1, bb.stream)) != hipSuccess) return hip_fail(e, err, errlen);
        if ((e = hipMemcpyAsync(ia.data(), bb.iarena, 4 * ia.size(), hipMemcpyDeviceToHost, bb.stream)) !=
            hipSuccess)
            return hip_fail(e, err, errlen);
        if ((e = hipStreamSynchronize(bb.stream)) != hipSuccess) return hip_fail(e, err, errlen);
        float ms = 0;
        hipEventElapsedTime(&ms, bb.ev0, bb.ev1);
        ms_total += ms;
        for (int q = 0; q < np; ++q) {
            lcg_pair &X = *P[todo[c0 + q]];
            const BandDesc &d = descs[q];
            const int n = d.n;
            if (ia[d.rows + 2 * (n + 1)] != 0) {
                recheck.push_back(todo[c0 + q]);
                continue;
            }
            std::vector<int> first(n + 1), last(n + 1);
            for (int i = 0; i <= n; ++i) {
                const int lo = ia[d.rows + i], hi = ia[d.rows + n + 1 + i];
                first[i] = lo == INT_MAX ? -1 : lo;
                last[i] = hi;
            }
            X.tc->restrict_by_rows(first, last);
        }
        c0 = c1;
    }
    return LCG_OK;
}
}  // namespace

extern "C" int lcg_pairs_prepare(const char *const *ppA, const char *const *ppB, int32_t npairs,
                                 const lcg_params *p, const lcg_prepare_opts *opts, lcg_pair **out,
                                 lcg_prepare_stats *stats, char *err, int errlen) {
    if (!ppA || !ppB || !p || !out || npairs < 0) return LCG_E_ARG;
    for (int k = 0; k < npairs; ++k) {
        out[k] = nullptr;
        if (!ppA[k] || !ppB[k]) return LCG_E_ARG;
    }
    lcg_prepare_opts o;
    std::memset(&o, 0, sizeof o);
    if (opts) o = *opts;
    const int threads = o.threads > 0 ? o.threads : std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    const double eps = o.band_eps > 0 ? o.band_eps : 1e-9;
    const bool deferred = (o.flags & LCG_PREPARE_DEVICE_COMPACTION) != 0;
    const bool dev_band = (o.flags & LCG_PREPARE_DEVICE_BAND) != 0;
    if (deferred && p->no_lonely_pairs) {  // the device compaction builds no --noLP groups
        set_err(err, errlen, "--noLP needs the arc matches enumerated on the host");
        return LCG_E_ARG;
    }
    if (deferred && p->stacking) {  // nor the inner arc matches and stacked scores
        set_err(err, errlen, "--stacking needs the arc matches enumerated on the host");
        return LCG_E_ARG;
    }
    if (deferred && p->struct_local) {  // nor the eight-state regions
        set_err(err, errlen, "--struct-local needs the arc matches enumerated on the host");
        return LCG_E_ARG;
    }
    std::vector<std::unique_ptr<lcg_pair>> P(npairs);
    std::string msg;
    using clk = std::chrono::steady_clock;
    auto since = [](clk::time_point t) { return std::chrono::duration<float, std::milli>(clk::now() - t).count(); };
    auto t0 = clk::now();
    float parse_ms = 0, band_wall_ms = 0;
    auto input_error = [&](int k) {
        set_err(err, errlen, "pair " + std::to_string(k) + ": " + msg);
        return LCG_E_INPUT;
    };
    // every distinct input text is parsed once (all-pairs batches name each
    // sequence's text n - 1 times; texts are told apart by address)
    std::unordered_map<const char *, int> uid;
    std::vector<const char *> texts;
    std::vector<int> ua(npairs), ub(npairs), first_pair;
    for (int k = 0; k < npairs; ++k) {
        for (int side = 0; side < 2; ++side) {
            const char *t = side ? ppB[k] : ppA[k];
            auto it = uid.find(t);
            int u;
            if (it == uid.end()) {
                u = (int)texts.size();
                uid.emplace(t, u);
                texts.push_back(t);
                first_pair.push_back(k);
            } else {
                u = it->second;
            }
            (side ? ub : ua)[k] = u;
        }
    }
    std::vector<std::shared_ptr<const RnaData>> rd(texts.size());
    std::vector<std::shared_ptr<const BasePairs>> bps(texts.size());
    int bad = parallel_for(
        (int)texts.size(), threads,
        [&](int u) {
            rd[u] = std::make_shared<const RnaData>(texts[u], p->min_prob, p->max_bp_span, p->stacking != 0);
            bps[u] = std::make_shared<const BasePairs>(rd[u].get(), p->min_prob);
        },
        msg);
    if (bad >= 0) return input_error(first_pair[bad]);
    bad = parallel_for(
        npairs, threads,
        [&](int k) {
            P[k].reset(new lcg_pair);
            prepare_parse(ppA[k], ppB[k], p, *P[k], rd[ua[k]], rd[ub[k]]);
        },
        msg);
    if (bad >= 0) return input_error(bad);
    parse_ms = since(t0);
    t0 = clk::now();
    std::vector<int> host_band;
    float band_ms = 0;
    int on_device = 0;
    if (p->min_trace_prob > 0.0) {
        if (dev_band) {
            std::vector<int> todo, recheck;
            for (int k = 0; k < npairs; ++k) {
                // the band kernel reads one symbol per position: profiles
                // are banded on the host
                if (P[k]->A->length() <= LCG_BAND_MAX_LEN && P[k]->B->length() <= LCG_BAND_MAX_LEN &&
                    P[k]->A->num_rows() == 1 && P[k]->B->num_rows() == 1)
                    todo.push_back(k);
                else
                    host_band.push_back(k);
            }
            if (int rc = device_bands(P, todo, *p, o.device, eps, threads, recheck, band_ms, err, errlen))
                return rc;
            on_device = (int)(todo.size() - recheck.size());
            host_band.insert(host_band.end(), recheck.begin(), recheck.end());
        } else {
            host_band.resize(npairs);
            std::iota(host_band.begin(), host_band.end(), 0);
        }
        bad = parallel_for((int)host_band.size(), threads, [&](int t) { prepare_band_host(*P[host_band[t]]); },
                           msg);
        if (bad >= 0) return input_error(host_band[bad]);
    }
    band_wall_ms = since(t0);
    t0 = clk::now();
    bool capacity = false;
    bad = parallel_for(
        npairs, threads, [&](int k) { prepare_finish(*P[k], deferred, bps[ua[k]], bps[ub[k]]); }, msg, &capacity);
    if (bad >= 0) {
        const int rc = input_error(bad);
        return capacity ? LCG_E_CAPACITY : rc;
    }
    for (int k = 0; k < npairs; ++k) out[k] = P[k].release();
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->band_ms = band_ms;
        stats->device_banded = on_device;
        stats->host_banded = (int32_t)host_band.size();
        stats->parse_ms = parse_ms;
        stats->band_wall_ms = band_wall_ms;
        stats->finish_ms = since(t0);
    }
    return LCG_OK;
}
