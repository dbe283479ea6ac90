// gfx950 kernels of LocARNA-P (sum-product DP in FP64): inside partition
// function, and its adjoint (reverse mode), which yields the outside values
// of every arc match, the arc-match and base-match probabilities and the
// expected number of gap openings.
//
// Reference: AlignerP, src/LocARNA/aligner_p.icc: inside :148-438;
// outside :450-1139; probabilities :1150-1399.  Here the outside values come
// from the adjoint of the inside recursion: Dbar(am) = dZ/dD(am), so that
// P(am) = D(am) Dbar(am) / Z (the reference's D D' pf_scale / (Z e^{s/T}));
// the base-match probability of (i, j) is the weight of alignments with a
// match edge into (i, j) directly inside an arc match with P(am) >
// sqrt(min_am_prob), or at the top level (aligner_p.icc:1201-1399), divided
// by Z.  The unrestricted match-edge weights and the gap-opening weight give
// the gradient of log Z with respect to the score parameters.
//
// Layout per group (origin al, bl; rows al..ar-1; bound br): rows stored from
// column max(bl, min_col(i)) with stride W, as in the integer kernel.  The
// scratch of a wave holds, for the group being processed, the matrices M,
// E, F and Mbar (doubles) plus E / Ebar vectors over the columns.
//
// Cell recurrence (inside), lane l owns column j of row i:
//   E(i,j) = E(i-1,j) gA (1 - o) + M(i-1,j) gA o
//   s(i,j) = M(i-1,j-1) es(i,j) + E(i,j) + ps * sum_arcs M(p-1,q-1) D
//   F(i,j) = F(i,j-1) gB + s(i,j-1) gB o              (s(i,jlo-1) = M(i,jlo-1))
// (the reference's F gB + (M(i,j-1) - F) gB o with M = s + F; the E chain
// keeps its (1 - o) since M(i-1,j) contains E(i-1,j))
//   M(i,j) = s(i,j) + F(i,j)
// The F chain is a linear recurrence along the row, solved by a wave scan
// of (multiplier, offset) pairs; its adjoint is the mirrored suffix scan.
#include <hip/hip_runtime.h>

#include "device_layout.h"
#include "pf_layout.h"

namespace {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// DPP move of a double (both halves): lane l receives the lane CTRL names;
// lanes without a source, and rows outside ROWMASK, receive `old`
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d(double old, double v) {
    const unsigned long long o = (unsigned long long)__double_as_longlong(old);
    const unsigned long long x = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)x, CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(x >> 32), CTRL, ROWMASK,
                                               0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ double readlane_d(double v, int l) {
    const unsigned long long x = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)x, l);
    const int hi = __builtin_amdgcn_readlane((int)(unsigned)(x >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// one step of the prefix composition: (a, b) after the map from the source
// lane (identity where there is none)
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void affine_step(double &a, double &b) {
    const double pa = dpp_d<CTRL, ROWMASK>(1.0, a), pb = dpp_d<CTRL, ROWMASK>(0.0, b);
    b = a * pb + b;
    a = a * pa;
}

// inclusive scan of the affine maps x -> a x + b over the lanes (lane 0
// first): result (A_l, B_l) with F_l = A_l F_in + B_l.  DPP row shifts
// within rows of 16, then row broadcasts of lanes 15 and 31.
__device__ __forceinline__ void wave_scan_affine(double &a, double &b) {
    affine_step<0x111, 0xf>(a, b);  // row_shr:1
    affine_step<0x112, 0xf>(a, b);  // row_shr:2
    affine_step<0x114, 0xf>(a, b);  // row_shr:4
    affine_step<0x118, 0xf>(a, b);  // row_shr:8
    affine_step<0x142, 0xa>(a, b);  // row_bcast:15 -> rows 1, 3
    affine_step<0x143, 0xc>(a, b);  // row_bcast:31 -> rows 2, 3
}

// suffix version: x_l = g_l + c_l x_{l+1} (last lane first); returns the
// pair (C_l, G_l) with x_l = C_l x_in + G_l where x_in enters at lane 63.
// DPP row shifts to the left within rows of 16; then each row is composed
// with the rows after it (their totals read from lanes 16, 32, 48).
__device__ __forceinline__ void wave_suffix_affine(double &c, double &g) {
    affine_step<0x101, 0xf>(c, g);  // row_shl:1
    affine_step<0x102, 0xf>(c, g);  // row_shl:2
    affine_step<0x104, 0xf>(c, g);  // row_shl:4
    affine_step<0x108, 0xf>(c, g);  // row_shl:8
    const double c1 = readlane_d(c, 16), g1 = readlane_d(g, 16);
    const double c2 = readlane_d(c, 32), g2 = readlane_d(g, 32);
    const double c3 = readlane_d(c, 48), g3 = readlane_d(g, 48);
    const double s1c = c2 * c3, s1g = c2 * g3 + g2;      // rows 2..3
    const double s0c = c1 * s1c, s0g = c1 * s1g + g1;    // rows 1..3
    const int row = lane_id() >> 4;
    const double sc = row == 0 ? s0c : row == 1 ? s1c : row == 2 ? c3 : 1.0;
    const double sg = row == 0 ? s0g : row == 1 ? s1g : row == 2 ? g3 : 0.0;
    g = c * sg + g;
    c = c * sc;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

struct PfPair {
    const int *mn, *mx, *coff, *codeA, *codeB;  // LDS
    const double *egA, *egB, *es;                // LDS: exp gap scores, 8x8 exp sigma
    const int *cptr;
    const int4 *cand;
    double *Dc;     // inside D per candidate entry
    double *Dbar;   // outside dZ/dD per candidate entry
    const double *eam;  // exp(s(am)/T) per candidate entry
    double *edge_full, *edge_ref;  // per band cell
    double eo, ps;
    int n, m;
    const int *fbp, *fbk;  // LDS: fill blocks of row i (left end p, first entry)
    const int *fcand;      // int arena: fill entries (p | q<<16, j | min_col(p-1)<<16, -)
    const int *f2c;        // fill entry -> candidate entry
    double *fD;            // inside D per fill entry
    double *fq;            // fragment match queries (8 doubles each)
    int nfq;
};

// Fragment match queries (A[i..j] with B[k..l]): the reverse sweep of every
// region holding the cells (i-1, k-1) and (j, l) adds M(i-1, k-1) times the
// adjoint of the summand s(j, l) (the weight of every completion from (j,
// l), the block being a summand of s that is not part of E) to the query's
// out; the block's inside is a separate fill.  P = in out ps / Z.
__device__ __forceinline__ void frag_query(const PfPair &P, int q, int &i, int &j, int &k, int &l) {
    const double *x = P.fq + 8 * q;
    i = (int)x[0];
    j = (int)x[1];
    k = (int)x[2];
    l = (int)x[3];
}

__device__ __forceinline__ int row_lo(const PfPair &P, int i, int bl) { return max(bl, P.mn[i]); }
__device__ __forceinline__ int row_hi(const PfPair &P, int i, int br) { return min(br - 1, P.mx[i]); }

struct Region {
    int al, bl, ar, br, W;
    double *M, *E, *F, *Mb;  // matrices (R x W)
    double *Ev, *Eb;         // vectors over columns bl..br-1
    __device__ __forceinline__ double *at(double *X, const PfPair &P, int i, int j) const {
        const int lo = (i == al) ? bl : row_lo(P, i, bl);
        return X + (size_t)(i - al) * W + (j - lo);
    }
    __device__ __forceinline__ bool stored(const PfPair &P, int i, int j) const {
        if (i < al || i >= ar) return false;
        const int lo = (i == al) ? bl : row_lo(P, i, bl);
        return j >= lo && j <= row_hi(P, i, br);
    }
    __device__ __forceinline__ double get(double *X, const PfPair &P, int i, int j) const {
        return stored(P, i, j) ? *at(X, P, i, j) : 0.0;
    }
};

// Orders a workgroup's accesses to its global scratch and to the pair's
// arrays.  One workgroup owns a pair, its waves run on one CU and share that
// CU's vector L1 (write-through), so a workgroup-scope fence (wait for the
// outstanding accesses) suffices; an agent-scope fence would write back and
// invalidate the L2 on every call.
__device__ __forceinline__ void scratch_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

// Inside fill of a region (aligner_p.icc:150-312).  Stores M, E, F.
__device__ void pf_fill(const PfPair &P, const Region &R) {
    const int lane = lane_id();
    const int al = R.al, bl = R.bl, ar = R.ar, br = R.br, W = R.W;
    const double eo = P.eo, ps = P.ps;
    for (int c = lane; c < br - bl; c += 64) R.Ev[c] = 0.0;
    // row al: M(al, bl) = 1/ps, M(al, j) = eo/ps prod gB
    const int hi0 = row_hi(P, al, br);
    {
        double carry = eo / ps;
        for (int j0 = bl + 1; j0 <= hi0; j0 += 64) {
            const int j = j0 + lane;
            const bool act = j <= hi0;
            double a = act ? P.egB[j] : 1.0, b = 0.0;
            wave_scan_affine(a, b);  // prefix products
            if (act) {
                R.M[j - bl] = carry * a;
                R.E[j - bl] = 0.0;
                R.F[j - bl] = 0.0;
            }
            carry *= readlane_d(a, 63);
        }
        if (lane == 0) {
            R.M[0] = 1.0 / ps;
            R.E[0] = 0.0;
            R.F[0] = 0.0;
        }
    }
    scratch_sync();
    double colv = eo / ps;
    bool colopen = true;
    for (int i = al + 1; i < ar; ++i) {
        const int mni = P.mn[i];
        if (colopen && mni <= bl) {
            colv *= P.egA[i];
        } else {
            colopen = false;
        }
        const int lo = max(bl, mni), hi = min(br - 1, P.mx[i]);
        double *row = R.M + (size_t)(i - al) * W;
        if (colopen && lane == 0) {
            row[0] = colv;
            R.E[(size_t)(i - al) * W] = 0.0;
            R.F[(size_t)(i - al) * W] = 0.0;
        }
        const int jlo = max(bl + 1, mni);
        double Fcarry = 0.0;
        double sprev = colopen ? colv : 0.0;  // M(i, jlo-1)
        const double gA = P.egA[i];
        const int cb = P.coff[i] - mni;
        for (int j0 = jlo; j0 <= hi; j0 += 64) {
            const int j = j0 + lane;
            const bool act = j <= hi;
            double s = 0.0, gB = 1.0, e = 0.0;
            if (act) {
                const double up = R.get(R.M, P, i - 1, j), dg = R.get(R.M, P, i - 1, j - 1);
                const double ep = R.Ev[j - bl];
                e = ep * gA + (up - ep) * gA * eo;
                R.Ev[j - bl] = e;
                s = dg * P.es[P.codeA[i] * LCG_NCODES + P.codeB[j]] + e;
                const int t0 = P.cptr[cb + j], t1 = P.cptr[cb + j + 1];
                double sa = 0.0;
                for (int t = t0; t < t1; ++t) {
                    const int4 en = P.cand[t];
                    const int p = en.x & 0xffff, q = en.x >> 16;
                    if (p <= al) break;
                    if (q <= bl) continue;
                    sa += R.get(R.M, P, p - 1, q - 1) * P.Dc[t];
                }
                s += sa * ps;
                gB = P.egB[j];
            }
            // F chain: F_j = F_{j-1} gB + s_{j-1} gB o
            double sl = dpp_d<0x138, 0xf>(s, s);  // lane l-1 (lane 0: its own)
            if (lane == 0) sl = sprev;
            double a = act ? gB : 1.0, b = act ? sl * gB * eo : 0.0;
            wave_scan_affine(a, b);
            const double F = a * Fcarry + b;
            if (act) {
                row[j - lo] = s + F;
                R.E[(size_t)(i - al) * W + (j - lo)] = e;
                R.F[(size_t)(i - al) * W + (j - lo)] = F;
            }
            Fcarry = readlane_d(F, 63);
            sprev = readlane_d(s, 63);
        }
        scratch_sync();
    }
}

// Adjoint of pf_fill for a region whose Mb holds the seeds.  Accumulates
// Dbar (per candidate entry), match-edge weights into `edge` (and into
// `edge2` when given) and the gap-opening weight.  Mb is consumed.
__device__ void pf_reverse(const PfPair &P, const Region &R, bool want_dbar, double *edge, double *openings,
                           double *edge2 = nullptr) {
    const int lane = lane_id();
    const int al = R.al, bl = R.bl, ar = R.ar, br = R.br, W = R.W;
    const double eo = P.eo, ps = P.ps;
    for (int c = lane; c < br - bl; c += 64) R.Eb[c] = 0.0;
    scratch_sync();
    double open_acc = 0.0;
    for (int i = ar - 1; i > al; --i) {
        const int mni = P.mn[i];
        const int lo = max(bl, mni), hi = min(br - 1, P.mx[i]);
        const int jlo = max(bl + 1, mni);
        const double gA = P.egA[i];
        const int cb = P.coff[i] - mni;
        const bool colopen_row = (mni <= bl);  // column bl stored for this row
        // columns of the region not computed in row i end their E chain
        for (int c = lane; c < br - bl; c += 64) {
            const int j = bl + c;
            if (j < jlo || j > hi) R.Eb[c] = 0.0;
        }
        if (jlo > hi) {
            scratch_sync();
            continue;
        }
        // chunks from the right: F adjoint flows right to left
        const int nch = (hi - jlo) / 64;
        double Fb_in = 0.0;        // Fbar of the column right of the chunk
        double a_in = 0.0;         // gB of that column
        for (int ch = nch; ch >= 0; --ch) {
            const int j0 = jlo + ch * 64;
            const int j = j0 + lane;
            const bool act = j <= hi;
            const size_t idx = act ? (size_t)(i - al) * W + (j - lo) : 0;
            const double g = act ? R.Mb[idx] : 0.0;
            const double gB = act ? P.egB[j] : 1.0;
            // Fbar_j = g_j + gB(j+1) Fbar_{j+1}
            double a_next = dpp_d<0x130, 0xf>(act ? gB : 0.0, act ? gB : 0.0);  // lane l+1
            if (lane == 63 || !(j + 1 <= hi)) a_next = (lane == 63) ? a_in : 0.0;
            double c = act ? a_next : 0.0, gg = act ? g : 0.0;
            wave_suffix_affine(c, gg);
            const double Fb = c * Fb_in + gg;  // Fbar_j
            // sbar_j = g_j + Fbar_{j+1} gB(j+1) o
            double Fb_next = dpp_d<0x130, 0xf>(Fb, Fb);
            double gB_next = dpp_d<0x130, 0xf>(gB, gB);
            if (lane == 63) {
                Fb_next = Fb_in;
                gB_next = a_in;
            }
            if (!(j + 1 <= hi) && lane != 63) {
                Fb_next = 0.0;
                gB_next = 0.0;
            }
            const double sb = g + Fb_next * gB_next * eo;
            if (act && want_dbar) {
                for (int q = 0; q < P.nfq; ++q) {
                    int fi, fj, fk, fl;
                    frag_query(P, q, fi, fj, fk, fl);
                    if (fj == i && fl == j && P.fq[8 * q + 4] > 0.0 && R.stored(P, fi - 1, fk - 1))
                        atomicAdd(&P.fq[8 * q + 6], sb * *R.at(R.M, P, fi - 1, fk - 1));
                }
            }
            if (act) {
                const double F = R.F[idx];
                // F_j = F_{j-1} gB (1-o) + s_{j-1} gB o ; opening part = F_j - F_{j-1} gB
                const double Fprev = (j > jlo) ? R.F[idx - 1] : 0.0;
                open_acc += Fb * (F - Fprev * gB);
                // boundary: F_jlo uses M(i, jlo-1)
                if (j == jlo) {
                    if (colopen_row) {
                        double *mcol = R.at(R.Mb, P, i, jlo - 1);
                        *mcol += Fb * gB * eo;
                    }
                }
                // s_j = M(i-1,j-1) es + E + ps sum arcs
                const double es = P.es[P.codeA[i] * LCG_NCODES + P.codeB[j]];
                if (R.stored(P, i - 1, j - 1)) {
                    const double dg = *R.at(R.M, P, i - 1, j - 1);
                    atomicAdd(R.at(R.Mb, P, i - 1, j - 1), sb * es);
                    atomicAdd(&edge[P.coff[i] + (j - mni)], sb * dg * es);
                    if (edge2) atomicAdd(&edge2[P.coff[i] + (j - mni)], sb * dg * es);
                }
                // E_j = Eprev gA (1-o) + M(i-1,j) gA o
                const double Eb = R.Eb[j - bl] + sb;
                const double E = R.E[idx];
                const double Eprev = R.stored(P, i - 1, j) && (i - 1 > al) ? *R.at(R.E, P, i - 1, j) : 0.0;
                if (R.stored(P, i - 1, j)) atomicAdd(R.at(R.Mb, P, i - 1, j), Eb * gA * eo);
                open_acc += Eb * (E - Eprev * gA);
                R.Eb[j - bl] = Eb * gA * (1.0 - eo);
                // arcs
                const int t0 = P.cptr[cb + j], t1 = P.cptr[cb + j + 1];
                for (int t = t0; t < t1; ++t) {
                    const int4 en = P.cand[t];
                    const int p = en.x & 0xffff, q = en.x >> 16;
                    if (p <= al) break;
                    if (q <= bl) continue;
                    const double Dv = P.Dc[t];
                    if (Dv == 0.0) continue;
                    // M(p-1, q-1) outside the band reads as 0 (aligner_p.icc:150-189 zeroes
                    // every such cell the recursion can reach)
                    if (!R.stored(P, p - 1, q - 1)) continue;
                    atomicAdd(R.at(R.Mb, P, p - 1, q - 1), sb * Dv * ps);
                    if (want_dbar) atomicAdd(&P.Dbar[t], sb * (*R.at(R.M, P, p - 1, q - 1)) * ps);
                }
            }
            Fb_in = readlane_d(Fb, 0);
            a_in = readlane_d(act ? gB : 0.0, 0);
        }
        scratch_sync();
    }
    // openings of the init row al and column bl (one each per cell value)
    const int hi0 = row_hi(P, al, br);
    for (int j = bl + 1 + lane; j <= hi0; j += 64) open_acc += R.Mb[j - bl] * R.M[j - bl];
    for (int i = al + 1 + lane; i < ar; i += 64) {
        if (P.mn[i] > bl) continue;
        bool open_col = true;
        for (int k = al + 1; k <= i; ++k)
            if (P.mn[k] > bl) open_col = false;
        if (open_col) open_acc += *R.at(R.Mb, P, i, bl) * *R.at(R.M, P, i, bl);
    }
    if (openings) {
        const double tot = wave_sum(open_acc);
        if (lane == 0) atomicAdd(openings, tot);  // waves of the workgroup share it
    }
}

// ---- fixed-column group kernels (columns bl+1 .. br-1 in NC chunks of 64
// lanes, lane l of chunk c owns column bl+1+64c+l in every row; rows
// al+1 .. ar-1 at most 128).  Region in LDS (doubles): [S: br-bl] [M:
// (ar-al) W] [Mbar: (ar-al) W], rows stored from max(bl, min_col(i)) as in
// the general layout.  The arc candidates of row i are the fill-list prefix
// p > al (left ends descending), located from the LDS block table.

struct PfEnt {
    int pq, jm;
    double D;
};

// lane l receives lane l-1's value, lane 0 `first` (DPP wave_shr:1)
__device__ __forceinline__ double shift_up_d(double v, double first) { return dpp_d<0x138, 0xf>(first, v); }
// lane l receives lane l+1's value, lane 63 `last` (DPP wave_shl:1)
__device__ __forceinline__ double shift_down_d(double v, double last) { return dpp_d<0x130, 0xf>(last, v); }

// candidate range [t0, t1) of rows al+1+r, r = lane + 64 h (RT tables: at
// most 64 RT rows)
template <int RT>
__device__ __forceinline__ void pf_row_table(const PfPair &P, int al, int bl, int br, int nrows, int rt0[RT],
                                             int rt1[RT]) {
#pragma unroll
    for (int h = 0; h < RT; ++h) {
        const int r = lane_id() + 64 * h;
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
}
template <int RT>
__device__ __forceinline__ void pf_row_cands(const int rt0[RT], const int rt1[RT], int r, int &t0, int &t1) {
    t0 = t1 = 0;
#pragma unroll
    for (int h = 0; h < RT; ++h)
        if ((r >> 6) == h) {
            t0 = __builtin_amdgcn_readlane(rt0[h], r & 63);
            t1 = __builtin_amdgcn_readlane(rt1[h], r & 63);
        }
}
// Unpredicated loads: lanes past the range read its last entry (or entry 0
// of the pair; a group implies arc matches) and are ignored by the caller's
// lane flag.
__device__ __forceinline__ PfEnt pf_fetch(const PfPair &P, int t, int t1) {
    const int tt = t < t1 ? t : max(t1 - 1, 0);
    PfEnt e;
    struct PJ {
        int x, y;
    };  // 4-byte aligned pair (one 8-byte load)
    const PJ pj = *reinterpret_cast<const PJ *>(P.fcand + 3 * tt);
    e.pq = pj.x;
    e.jm = pj.y;
    e.D = P.fD[tt];
    return e;
}
__device__ __forceinline__ void lds_fence() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); }

// Inside fill of a group (aligner_p.icc:150-312), values as pf_fill.  When
// Es/Ss are given (global scratch, layout of M), E and s = M - F of every
// computed cell are stored for the adjoint.
template <int NC, int RT = 2>
__device__ __forceinline__ void pf_fwd_fixed(const PfPair &P, int al, int bl, int ar, int br, int W, double *base, double *Es,
                             double *Ss) {
    const int lane = lane_id();
    const double eo = P.eo, ps = P.ps;
    double *Srow = base;
    double *Mst = base + (br - bl);
    const int nrows = ar - al - 1;
    int rt0[RT], rt1[RT];
    pf_row_table<RT>(P, al, bl, br, nrows, rt0, rt1);
    for (int c = lane; c < br - bl; c += 64) Srow[c] = 0.0;
    double Mp[NC], Ep[NC], gBc[NC];
    int cBc[NC];
    const int hi0 = min(br - 1, P.mx[al]);
    {
        double carry = eo / ps;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int j = bl + 1 + 64 * c + lane;
            const bool in = j <= hi0, col = j <= br - 1;
            double a = in ? P.egB[j] : 1.0, b = 0.0;
            wave_scan_affine(a, b);  // prefix products
            Mp[c] = in ? carry * a : 0.0;
            if (in) Mst[j - bl] = carry * a;
            carry *= readlane_d(a, 63);
            Ep[c] = 0.0;
            gBc[c] = col ? P.egB[j] : 1.0;
            cBc[c] = col ? P.codeB[j] : 0;
        }
        if (lane == 0) Mst[0] = 1.0 / ps;
    }
    lds_fence();
    double colv = eo / ps;
    bool colopen = true;
    double Mcol_prev = 1.0 / ps;  // M(i-1, bl)
    PfEnt cur[4];
    {
        int u0 = 0, u1 = 0;
        if (nrows > 0) pf_row_cands<RT>(rt0, rt1, 0, u0, u1);
#pragma unroll
        for (int k = 0; k < 4; ++k) cur[k] = pf_fetch(P, u0 + 64 * k + lane, u1);
    }
    for (int i = al + 1; i < ar; ++i) {
        const int r = i - al - 1;
        const int mni = P.mn[i];
        double Mcol = 0.0;  // M(i, bl)
        if (colopen && mni <= bl) {
            colv *= P.egA[i];
            Mcol = colv;
        } else {
            colopen = false;
        }
        const int lo = max(bl, mni), hi = min(br - 1, P.mx[i]);
        double *row = Mst + (size_t)(r + 1) * W;
        if (colopen && lane == 0) row[0] = colv;
        const int jlo = max(bl + 1, mni);
        int t0, t1;
        pf_row_cands<RT>(rt0, rt1, r, t0, t1);
        // arc terms ps * M(p-1, q-1) * D, candidate-parallel, LDS atomic adds
        auto arc = [&](const PfEnt &e, bool in) {
            const int p = e.pq & 0xffff, q = e.pq >> 16, j = e.jm & 0xffff;
            const int rl = max(bl, e.jm >> 16);
            const bool ok = in && q > bl && j >= jlo && j <= hi && q - 1 >= rl;
            const double v = Mst[ok ? (p - 1 - al) * W + (q - 1 - rl) : 0] * e.D * ps;
            if (ok) atomicAdd(&Srow[j - bl - 1], v);
        };
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (t0 + 64 * k < t1) arc(cur[k], t0 + 64 * k + lane < t1);
        for (int t = t0 + 256 + lane; t < t1; t += 64) arc(pf_fetch(P, t, t1), true);
        if (r + 1 < nrows) {  // the next row's chunks
            int u0, u1;
            pf_row_cands<RT>(rt0, rt1, r + 1, u0, u1);
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (u0 + 64 * k < u1) cur[k] = pf_fetch(P, u0 + 64 * k + lane, u1);
        }
        lds_fence();
        double Fcarry = 0.0, sprev = Mcol, dgin = Mcol_prev;
        const double gA = P.egA[i];
        const double *esrow = P.es + P.codeA[i] * LCG_NCODES;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int j = bl + 1 + 64 * c + lane;
            const bool act = j >= jlo && j <= hi;
            const double up = Mp[c];
            const double dg = shift_up_d(Mp[c], dgin);
            dgin = readlane_d(Mp[c], 63);
            double s = 0.0, e = 0.0;
            if (act) {
                const double ep = Ep[c];
                e = ep * gA + (up - ep) * gA * eo;
                Ep[c] = e;
                const double sv = Srow[j - bl - 1];
                Srow[j - bl - 1] = 0.0;
                s = dg * esrow[cBc[c]] + e + sv;
            }
            const double sl = shift_up_d(s, sprev);
            double a = act ? gBc[c] : 1.0, b = act ? sl * gBc[c] * eo : 0.0;
            wave_scan_affine(a, b);
            const double F = a * Fcarry + b;
            const double Mv = s + F;
            if (act) {
                const size_t idx = (size_t)(r + 1) * W + (j - lo);
                row[j - lo] = Mv;
                if (Es) {
                    Es[idx] = e;
                    Ss[idx] = s;
                }
            }
            Mp[c] = act ? Mv : 0.0;
            Fcarry = readlane_d(F, 63);
            sprev = readlane_d(s, 63);
        }
        Mcol_prev = Mcol;
        lds_fence();
    }
}

// Adjoint of pf_fwd_fixed for a group whose Mbar rows (LDS, after M) hold the
// seeds; Es/Ss from the forward recompute.  FULL: also Dbar (dZ/dD per
// candidate entry) and the gap-opening weight (returned); edge receives the
// match-edge weights.  Same terms as pf_reverse.
//
// The sweep visits rows ar-1 .. al+1.  Once a row is done no later row reads
// its Mbar, so each computed cell's Mbar is replaced by its sbar there; the
// arc terms of the row read sbar from LDS.  The global accumulations (Dbar +=
// sbar M(p-1,q-1) ps per arc candidate, edge += sbar M(i-1,j-1) es per cell)
// are issued inside the sweep, next to the values they use (a separate pass
// after the sweep that re-reads sbar and M measured slower).  The sweep's
// Ss/Es loads (FULL) are issued a row ahead.
template <int NC, bool FULL, int RT = 2>
__device__ __forceinline__ double pf_rev_fixed(const PfPair &P, int al, int bl, int ar, int br, int W, double *base, const double *Es,
                               const double *Ss, double *edge, double *edge2 = nullptr) {
    const int lane = lane_id();
    const double eo = P.eo, ps = P.ps;
    const double *Mst = base + (br - bl);
    double *Mb = base + (br - bl) + (size_t)(ar - al) * W;
    const int nrows = ar - al - 1;
    int rt0[RT], rt1[RT];
    pf_row_table<RT>(P, al, bl, br, nrows, rt0, rt1);
    double carry[NC], Eb[NC], gBc[NC];
    int cBc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int j = bl + 1 + 64 * c + lane;
        carry[c] = 0.0;
        Eb[c] = 0.0;
        gBc[c] = (j <= br - 1) ? P.egB[j] : 1.0;
        cBc[c] = (j <= br - 1) ? P.codeB[j] : 0;
    }
    // first stored column of row i of the region (row al from bl)
    auto lo_of = [&](int i) { return (i == al) ? bl : max(bl, P.mn[i]); };
    // Ss of row i at column j - 1 and Es of row i - 1 at column j (clamped
    // into the stored rows; the caller's flags decide whether they are used)
    double ssc[NC], esc[NC];
    auto load_se = [&](int i, double *ss, double *es) {
        const int r = i - al - 1;
        const int lo = lo_of(i), lo_u = lo_of(i - 1);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int j = bl + 1 + 64 * c + lane;
            ss[c] = Ss[(size_t)(r + 1) * W + min(max(j - 1 - lo, 0), W - 1)];
            es[c] = Es[(size_t)r * W + min(max(j - lo_u, 0), W - 1)];
        }
    };
    if (FULL && nrows > 0) load_se(ar - 1, ssc, esc);
    double colcarry = 0.0;  // contribution of row i+1 to Mbar(i, bl)
    double open_acc = 0.0;
    // first 64 fill entries of the row, fetched one row ahead; rows are
    // visited downwards
    PfEnt ce;
    {
        int u0 = 0, u1 = 0;
        if (nrows > 0) pf_row_cands<RT>(rt0, rt1, nrows - 1, u0, u1);
        ce = pf_fetch(P, u0 + lane, u1);
    }
    lds_fence();
    for (int i = ar - 1; i > al; --i) {
        const int r = i - al - 1;
        PfEnt ne = ce;
        double ssn[NC], esn[NC];
        if (r > 0) {
            int u0, u1;
            pf_row_cands<RT>(rt0, rt1, r - 1, u0, u1);
            if (u1 > u0) ne = pf_fetch(P, u0 + lane, u1);
            if (FULL) load_se(i - 1, ssn, esn);
        }
        const int mni = P.mn[i];
        const int lo = max(bl, mni), hi = min(br - 1, P.mx[i]), jlo = max(bl + 1, mni);
        const bool colopen_i = mni <= bl;
        const int lo_u = lo_of(i - 1), hi_u = min(br - 1, P.mx[i - 1]);
        const int jlo_u = max(bl + 1, P.mn[i - 1]);
        const bool upc_row = i - 1 > al;  // row i-1 computed (has E)
        const double *rowM = Mst + (size_t)(r + 1) * W;
        const double *rowU = Mst + (size_t)r * W;
        double *rowB = Mb + (size_t)(r + 1) * W;
        const double gA = P.egA[i];
        const double *esrow = P.es + P.codeA[i] * LCG_NCODES;
        double Fb_in = 0.0, gB_in = 0.0, d_in = 0.0;
        double newcarry[NC];
        double colcarry_next = 0.0, colb = 0.0;
#pragma unroll
        for (int cc = NC - 1; cc >= 0; --cc) {
            const int j = bl + 1 + 64 * cc + lane;
            const bool act = j >= jlo && j <= hi;
            const double g = act ? rowB[j - lo] + carry[cc] : 0.0;
            const double gB = act ? gBc[cc] : 0.0;
            const double gB_next = shift_down_d(gB, gB_in);
            double cm = act ? gB_next : 0.0, gg = g;
            wave_suffix_affine(cm, gg);
            const double Fb = cm * Fb_in + gg;  // Fbar_j
            const double Fb_next = shift_down_d(Fb, Fb_in);
            const double sb = act ? g + Fb_next * gB_next * eo : 0.0;
            if (act) rowB[j - lo] = sb;  // this row's Mbar is not read again
            const bool bcol = act && j == jlo && jlo == bl + 1 && colopen_i;  // s_{j-1} = M(i, bl)
            if (FULL && act) {
                const double sl = (j == jlo) ? (bcol ? rowM[0] : 0.0) : ssc[cc];
                open_acc += Fb * sl * gB * eo;
            }
            if (cc == 0) colb = readlane_d(bcol ? Fb * gB * eo : 0.0, 0);
            // E chain: E(i,j) = E(i-1,j) gA (1-o) + M(i-1,j) gA o
            const double eb = act ? sb + Eb[cc] : 0.0;
            const bool upst = j >= lo_u && j <= hi_u;
            if (FULL && act) {
                const double up = upst ? rowU[j - lo_u] : 0.0;
                const bool upc = upc_row && j >= jlo_u && j <= hi_u;
                const double Ep = upc ? esc[cc] : 0.0;
                open_acc += eb * (up - Ep) * gA * eo;
            }
            Eb[cc] = act ? eb * gA * (1.0 - eo) : 0.0;
            // match: s_j uses M(i-1, j-1) es
            const bool dgst = act && j - 1 >= lo_u && j - 1 <= hi_u;
            const double d = dgst ? sb * esrow[cBc[cc]] : 0.0;  // to Mbar(i-1, j-1)
            if (dgst) {  // match-edge weight sbar M(i-1, j-1) es
                const double w = d * rowU[j - 1 - lo_u];
                atomicAdd(&edge[P.coff[i] - mni + j], w);
                if (edge2) atomicAdd(&edge2[P.coff[i] - mni + j], w);
            }
            const double d_right = shift_down_d(d, d_in);
            newcarry[cc] = upst ? (act ? eb * gA * eo : 0.0) + d_right : 0.0;
            if (cc == 0) colcarry_next = readlane_d(d, 0);
            d_in = readlane_d(d, 0);
            Fb_in = readlane_d(Fb, 0);
            gB_in = readlane_d(gB, 0);
        }
        // the init column cell (i, bl): one opening per value
        if (FULL && colopen_i && lane == 0) open_acc += (rowB[0] + colcarry + colb) * rowM[0];
        lds_fence();
        // arc terms: Mbar(p-1, q-1) += sbar D ps
        int t0, t1;
        pf_row_cands<RT>(rt0, rt1, r, t0, t1);
        // arc terms: Mbar(p-1, q-1) += sbar D ps and (FULL) Dbar += sbar
        // M(p-1, q-1) ps of the candidate entry tt
        auto arc = [&](const PfEnt &e, int tt, bool in) {
            const int p = e.pq & 0xffff, q = e.pq >> 16, j = e.jm & 0xffff;
            const int rl = max(bl, e.jm >> 16);
            if (in && q > bl && j >= jlo && j <= hi && q - 1 >= rl && e.D != 0.0) {
                atomicAdd(&Mb[(p - 1 - al) * W + (q - 1 - rl)], rowB[j - lo] * e.D * ps);
                if (FULL) atomicAdd(&P.Dbar[P.f2c[tt]], rowB[j - lo] * Mst[(p - 1 - al) * W + (q - 1 - rl)] * ps);
            }
        };
        if (t1 > t0) arc(ce, t0 + lane, t0 + lane < t1);
        for (int t = t0 + 64 + lane; t < t1; t += 64) arc(pf_fetch(P, t, t1), t, true);
        ce = ne;
        if (FULL && r > 0) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                ssc[c] = ssn[c];
                esc[c] = esn[c];
            }
        }
        lds_fence();
#pragma unroll
        for (int c = 0; c < NC; ++c) carry[c] = newcarry[c];
        colcarry = (lo_u == bl) ? colcarry_next : 0.0;
    }
    // init row al: one opening per value M(al, j), j > bl
    if (FULL) {
        const int hi0 = min(br - 1, P.mx[al]);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int j = bl + 1 + 64 * c + lane;
            if (j <= hi0) open_acc += (Mb[j - bl] + carry[c]) * Mst[j - bl];
        }
    }
    // fragment queries: the computed cells hold sbar now (lanes over queries)
    if (FULL && P.nfq > 0) {
        lds_fence();
        for (int q = lane; q < P.nfq; q += 64) {
            int fi, fj, fk, fl;
            frag_query(P, q, fi, fj, fk, fl);
            if (!(P.fq[8 * q + 4] > 0.0) || fj <= al || fj >= ar || fi - 1 < al || fk - 1 < bl) continue;
            const int clo = lo_of(fi - 1), chi = min(br - 1, P.mx[fi - 1]);
            const int jlo = max(bl + 1, P.mn[fj]), jhi = min(br - 1, P.mx[fj]);
            if (fk - 1 < clo || fk - 1 > chi || fl < jlo || fl > jhi) continue;
            const double mv = Mst[(size_t)(fi - 1 - al) * W + (fk - 1 - clo)];
            const double sbv = Mb[(size_t)(fj - al) * W + (fl - lo_of(fj))];
            atomicAdd(&P.fq[8 * q + 6], mv * sbv);
        }
    }
    return wave_sum(open_acc);
}

__device__ __forceinline__ void zero_region_bar(const Region &R, const PfPair &P) {
    const int lane = lane_id();
    const size_t total = (size_t)(R.ar - R.al) * R.W;
    for (size_t c = lane; c < total; c += 64) R.Mb[c] = 0.0;
    scratch_sync();
}

}  // namespace

// One workgroup per pair (persistent, pairs from a queue).  Inside: rows
// al = n .. 1, the groups of a row in the host's LDS rounds; narrow groups
// (columns and rows at most 128) run the fixed-column kernels on LDS
// regions, others the general kernels on per-wave global scratch.  Top
// level (wave 0, global scratch): inside, full and restricted adjoint.
// Outside: rows al = 1 .. n, each group's forward values recomputed, then
// the full adjoint (Dbar, edge weights, openings) and, if any of its arc
// matches has P(am) > sqrt(min_am_prob), the restricted one.
#define LCG_PF_SPIN_LIMIT (1 << 22)

// The group of rows g0..g1-1 whose schedule word (S[4 g + slot]) is `want`,
// -1 if none: the lanes compare 64 words at a time (a wave holds at most
// one group per round).
__device__ __forceinline__ int find_group(const int *S, int slot, int g0, int g1, int want) {
    for (int gb = g0; gb < g1; gb += 64) {
        const int g = gb + lane_id();
        const unsigned long long hit = __ballot(g < g1 && S[4 * g + slot] == want);
        if (hit) return gb + __ffsll((long long)hit) - 1;
    }
    return -1;
}

extern "C" __global__ void __launch_bounds__(512) lcg_pf_pairs(const PairDesc *__restrict__ descs,
                                                                const PfDesc *__restrict__ pdescs,
                                                                const int *__restrict__ order, int *queue,
                                                                int *arena, double *pfa, double *scratch,
                                                                const int *__restrict__ sched, int *barriers,
                                                                double *fstore, PfBatchParams bp) {
    extern __shared__ int lds[];
    __shared__ int s_pair;
    __shared__ double s_open;
    const int nwaves = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = lane_id();
    const int L = bp.max_len;
    int *mn = lds, *mx = lds + (L + 1), *coff = lds + 2 * (L + 1);
    int *codeA = lds + 3 * L + 4, *codeB = lds + 4 * L + 5;
    int *fbp = lds + 5 * L + 6, *fbk = lds + 6 * L + 8;
    double *dl = reinterpret_cast<double *>(lds + ((6 * L + 8 + 2 * bp.nfblk_max + 1) & ~1));
    double *egA = dl, *egB = dl + (L + 1), *es = dl + 2 * (L + 1);
    double *pool = es + LCG_NCODES * LCG_NCODES;
    double *wscr = scratch + (size_t)(blockIdx.x * nwaves + wave) * bp.scratch_stride;
    // The first nstatic pairs of the queue order have wg_per_pair workgroups
    // (members) each, blocks K k .. K k + K - 1, all resident; the groups of a
    // row are dealt to the members, which meet at a barrier after every row
    // (arrival counter per pair, bounded wait).  The other workgroups take
    // the remaining pairs from the queue, one workgroup per pair.
    const bool stat = (int)blockIdx.x < bp.wg_per_pair * bp.nstatic;
    const int K = stat ? bp.wg_per_pair : 1;
    const int member = stat ? (int)blockIdx.x % K : 0;
    const int spin_lim = bp.spin_limit ? bp.spin_limit : LCG_PF_SPIN_LIMIT;
    for (int iter = 0;; ++iter) {
        if (stat) {
            if (iter > 0) break;
            if (threadIdx.x == 0) s_pair = (int)blockIdx.x / K;
        } else if (threadIdx.x == 0) {
            s_pair = bp.nstatic + atomicAdd(queue, 1);
        }
        __syncthreads();
        const int k = s_pair;
        if (k >= bp.npairs) break;
        const int pid = order[k];
        const PairDesc d = descs[pid];
        const PfDesc f = pdescs[pid];
        const int n = d.n, m = d.m;
        for (int t = threadIdx.x; t <= n; t += blockDim.x) {
            mn[t] = arena[d.band + t];
            mx[t] = arena[d.band + n + 1 + t];
            codeA[t] = arena[d.codes + t];
            egA[t] = pfa[f.egA + t];
        }
        for (int t = threadIdx.x; t <= m; t += blockDim.x) {
            codeB[t] = arena[d.codes + n + 1 + t];
            egB[t] = pfa[f.egB + t];
        }
        for (int t = threadIdx.x; t <= n + 1; t += blockDim.x) {
            coff[t] = arena[d.band + 2 * (n + 1) + t];
            fbp[t] = arena[d.fptr + t];
        }
        for (int t = threadIdx.x; t < 2 * d.nfblk; t += blockDim.x) fbk[t] = arena[d.fblk + t];
        for (int t = threadIdx.x; t < LCG_NCODES * LCG_NCODES; t += blockDim.x) es[t] = pfa[f.es + t];
        if (member == 0) {  // the pair's accumulators (then a barrier when K > 1)
            for (int t = threadIdx.x; t < d.nam; t += blockDim.x) {
                pfa[f.Dc + t] = 0.0;
                pfa[f.Dbar + t] = 0.0;
                pfa[f.fD + t] = 0.0;
            }
            for (int t = threadIdx.x; t < d.ncells; t += blockDim.x) {
                pfa[f.edge_full + t] = 0.0;
                pfa[f.edge_ref + t] = 0.0;
            }
            if (threadIdx.x == 0) {
                pfa[f.out + 1] = 0.0;  // gap openings (members add theirs)
                pfa[f.out + 2] = 0.0;  // status: 1 if a bounded wait expired
            }
        }
        if (threadIdx.x == 0) s_open = 0.0;
        int passed = 0;  // barrier generations passed
        // all members' global writes before any member goes on
        auto pair_barrier = [&]() {
            __threadfence();
            __syncthreads();
            ++passed;
            if (threadIdx.x == 0) {
                atomicAdd(&barriers[k], 1);
                for (int spin = 0; spin_lim < 0 || atomicAdd(&barriers[k], 0) < K * passed; ++spin) {
                    if (spin > spin_lim) {
                        pfa[f.out + 2] = 1.0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        };
        PfPair P;
        P.mn = mn;
        P.mx = mx;
        P.coff = coff;
        P.codeA = codeA;
        P.codeB = codeB;
        P.egA = egA;
        P.egB = egB;
        P.es = es;
        P.cptr = arena + d.cptr;
        P.cand = reinterpret_cast<const int4 *>(arena + d.cand);
        P.Dc = pfa + f.Dc;
        P.Dbar = pfa + f.Dbar;
        P.eam = pfa + f.eam;
        P.edge_full = pfa + f.edge_full;
        P.edge_ref = pfa + f.edge_ref;
        P.eo = pfa[f.scal + 0];
        P.ps = pfa[f.scal + 1];
        P.n = n;
        P.m = m;
        P.fbp = fbp;
        P.fbk = fbk;
        P.fcand = arena + d.fcand;
        P.f2c = arena + d.f2c;
        P.fD = pfa + f.fD;
        P.fq = pfa + f.frag;
        P.nfq = f.nfrag;
        if (K > 1) {
            pair_barrier();
        } else {
            scratch_sync();
            __syncthreads();
        }
        const int mine = member << 16;
        const int *rows = arena + d.rows;
        const int *groups = arena + d.groups;
        const int4 *gams = reinterpret_cast<const int4 *>(arena + d.gams);
        const int *S = sched + f.sched;
        auto region = [&](int al, int bl, int ar, int br, int W) {
            Region R;
            R.al = al;
            R.bl = bl;
            R.ar = ar;
            R.br = br;
            R.W = W;
            const size_t cells = (size_t)(ar - al) * W;
            R.M = wscr;
            R.E = wscr + cells;
            R.F = wscr + 2 * cells;
            R.Mb = wscr + 3 * cells;
            R.Ev = wscr + 4 * cells;
            R.Eb = R.Ev + (br - bl);
            return R;
        };
        // stored value of (i, j) in a fixed region (row al from bl)
        auto fixed_at = [&](int al, int bl, int br, int W, int i, int j) {
            const int lo = (i == al) ? bl : max(bl, P.mn[i]);
            const int hi = min(br - 1, P.mx[i]);
            return (j >= lo && j <= hi) ? (i - al) * W + (j - lo) : -1;
        };
        // ---- inside: rows al = n .. 1
        for (int al = n; al >= 1; --al) {
            const int g0 = rows[al + 1], g1 = rows[al];
            if (g0 == g1) continue;
            const int nr = sched[f.rin + al];
            for (int rnd = 0; rnd < nr; ++rnd) {
                const int g = find_group(S, 0, g0, g1, rnd | (wave << 12) | mine);
                if (g >= 0) {
                    const int *G = groups + g * LCG_GROUP_INTS;
                    const int gal = G[0] & 0xffff, gbl = G[0] >> 16, gar = G[1] & 0xffff, gbr = G[1] >> 16;
                    const int W = G[2];
                    const int off = S[4 * g + 1];
                    if (off >= 0) {
                        double *base = pool + off;
                        // the forward values the outside pass reads (E, s, M), if stored
                        const int so = sched[f.fso + g];
                        const size_t cells = (size_t)(gar - gal) * W;
                        double *Es = so >= 0 ? fstore + 4 * (size_t)so : nullptr;
                        double *Ss = so >= 0 ? Es + cells : nullptr;
                        if (gbr - gbl - 1 <= 64)
                            pf_fwd_fixed<1>(P, gal, gbl, gar, gbr, W, base, Es, Ss);
                        else
                            pf_fwd_fixed<2>(P, gal, gbl, gar, gbr, W, base, Es, Ss);
                        const double *Mst = base + (gbr - gbl);
                        if (so >= 0) {
                            double *Ms = Ss + cells;
                            for (size_t c = lane; c < cells; c += 64) Ms[c] = Mst[c];
                        }
                        for (int t = G[3] + lane; t < G[4]; t += 64) {
                            const int4 ga = gams[t];
                            const int ar = ga.x & 0xffff, br = ga.x >> 16;
                            const int x = fixed_at(gal, gbl, gbr, W, ar - 1, br - 1);
                            const double v = (x >= 0 ? Mst[x] : 0.0) * P.eam[ga.y];
                            P.Dc[ga.y] = v;
                            P.fD[ga.w] = v;
                        }
                    } else {
                        const Region R = region(gal, gbl, gar, gbr, W);
                        pf_fill(P, R);
                        for (int t = G[3] + lane; t < G[4]; t += 64) {
                            const int4 ga = gams[t];
                            const int ar = ga.x & 0xffff, br = ga.x >> 16;
                            const double v = R.get(R.M, P, ar - 1, br - 1) * P.eam[ga.y];
                            P.Dc[ga.y] = v;
                            P.fD[ga.w] = v;
                        }
                    }
                }
                __threadfence_block();
                __syncthreads();
            }
            if (K > 1) pair_barrier();
        }
        // ---- top level and its adjoint (wave 0 of member 0)
        // (the restricted base-match weights count every top-level match edge,
        // aligner_p.icc:1201-1399: the seeds are the same, so one adjoint
        // sweep gives both)
        if (wave == 0 && member == 0) {
            const int W = d.wtop;
            const size_t cells = (size_t)(n + 1) * W;
            if (n <= 192 && m <= 192 && (size_t)(m + 1) + 2 * cells <= (size_t)bp.pool) {
                // the whole LDS pool is free while the other waves wait: the
                // fixed-column kernels (E, s of every cell in the wave's scratch)
                double *base = pool;
                double *Es = wscr, *Ss = wscr + cells;
                pf_fwd_fixed<3, 3>(P, 0, 0, n + 1, m + 1, W, base, Es, Ss);
                const int xz = fixed_at(0, 0, m + 1, W, n, m);
                const double Z = xz >= 0 ? base[(m + 1) + xz] : 0.0;
                if (lane == 0) pfa[f.out + 0] = Z;
                double *Mb = base + (m + 1) + cells;
                for (size_t c = lane; c < cells; c += 64) Mb[c] = 0.0;
                lds_fence();
                if (lane == 0 && xz >= 0) Mb[xz] = 1.0;
                scratch_sync();  // Es/Ss (global) and the seed (LDS) before the sweep
                const double o = pf_rev_fixed<3, true, 3>(P, 0, 0, n + 1, m + 1, W, base, Es, Ss, P.edge_full,
                                                          P.edge_ref);
                if (lane == 0) atomicAdd(&s_open, o);
            } else {
                const Region R = region(0, 0, n + 1, m + 1, W);
                pf_fill(P, R);
                const double Z = *R.at(R.M, P, n, m);
                if (lane == 0) pfa[f.out + 0] = Z;
                zero_region_bar(R, P);
                if (lane == 0) *R.at(R.Mb, P, n, m) = 1.0;
                scratch_sync();
                pf_reverse(P, R, true, P.edge_full, &s_open, P.edge_ref);
            }
        }
        if (K > 1) {
            pair_barrier();
        } else {
            scratch_sync();
            __syncthreads();
        }
        const double Z = pfa[f.out + 0];
        const double thr = pfa[f.scal + 2];
        // ---- outside over the groups, rows al = 1 .. n
        for (int al = 1; al <= n; ++al) {
            const int g0 = rows[al + 1], g1 = rows[al];
            if (g0 == g1) continue;
            const int nr = sched[f.rout + al];
            for (int rnd = 0; rnd < nr; ++rnd) {
                const int g = find_group(S, 2, g0, g1, rnd | (wave << 12) | mine);
                if (g >= 0) {
                    const int *G = groups + g * LCG_GROUP_INTS;
                    const int gal = G[0] & 0xffff, gbl = G[0] >> 16, gar = G[1] & 0xffff, gbr = G[1] >> 16;
                    const int W = G[2];
                    const int off = S[4 * g + 3];
                    bool any = false;
                    for (int t = G[3] + lane; t < G[4]; t += 64) {
                        const int4 ga = gams[t];
                        if (P.Dc[ga.y] * P.Dbar[ga.y] / Z > thr) any = true;
                    }
                    const bool sig = __ballot(any) != 0;
                    if (off >= 0) {
                        double *base = pool + off;
                        const size_t cells = (size_t)(gar - gal) * W;
                        double *Mb = base + (gbr - gbl) + cells;
                        const bool nc1 = gbr - gbl - 1 <= 64;
                        const int so = sched[f.fso + g];
                        double *Es = wscr, *Ss = wscr + cells;
                        if (so >= 0) {
                            // the inside pass's forward values: M into the region
                            Es = fstore + 4 * (size_t)so;
                            Ss = Es + cells;
                            const double *Ms = Ss + cells;
                            double *Mst = base + (gbr - gbl);
                            for (size_t c = lane; c < cells; c += 64) Mst[c] = Ms[c];
                            lds_fence();
                        } else if (nc1) {
                            pf_fwd_fixed<1>(P, gal, gbl, gar, gbr, W, base, Es, Ss);
                        } else {
                            pf_fwd_fixed<2>(P, gal, gbl, gar, gbr, W, base, Es, Ss);
                        }
                        for (int pass = 0; pass < (sig ? 2 : 1); ++pass) {
                            for (size_t c = lane; c < cells; c += 64) Mb[c] = 0.0;
                            lds_fence();
                            for (int t = G[3] + lane; t < G[4]; t += 64) {
                                const int4 ga = gams[t];
                                if (pass == 1 && !(P.Dc[ga.y] * P.Dbar[ga.y] / Z > thr)) continue;
                                const int ar = ga.x & 0xffff, br = ga.x >> 16;
                                const int x = fixed_at(gal, gbl, gbr, W, ar - 1, br - 1);
                                if (x >= 0) atomicAdd(&Mb[x], P.Dbar[ga.y] * P.eam[ga.y]);
                            }
                            scratch_sync();  // Es/Ss (global) and the seeds (LDS) before the sweep
                            if (pass == 0) {
                                const double o = nc1 ? pf_rev_fixed<1, true>(P, gal, gbl, gar, gbr, W, base, Es, Ss,
                                                                                P.edge_full)
                                                     : pf_rev_fixed<2, true>(P, gal, gbl, gar, gbr, W, base, Es, Ss,
                                                                                P.edge_full);
                                if (lane == 0) atomicAdd(&s_open, o);
                            } else {
                                if (nc1)
                                    pf_rev_fixed<1, false>(P, gal, gbl, gar, gbr, W, base, Es, Ss, P.edge_ref);
                                else
                                    pf_rev_fixed<2, false>(P, gal, gbl, gar, gbr, W, base, Es, Ss, P.edge_ref);
                            }
                        }
                    } else {
                        const Region R = region(gal, gbl, gar, gbr, W);
                        pf_fill(P, R);
                        zero_region_bar(R, P);
                        for (int t = G[3] + lane; t < G[4]; t += 64) {
                            const int4 ga = gams[t];
                            const int ar = ga.x & 0xffff, br = ga.x >> 16;
                            if (R.stored(P, ar - 1, br - 1))
                                atomicAdd(R.at(R.Mb, P, ar - 1, br - 1), P.Dbar[ga.y] * P.eam[ga.y]);
                        }
                        scratch_sync();
                        pf_reverse(P, R, true, P.edge_full, &s_open);
                        if (sig) {
                            zero_region_bar(R, P);
                            for (int t = G[3] + lane; t < G[4]; t += 64) {
                                const int4 ga = gams[t];
                                const int ar = ga.x & 0xffff, br = ga.x >> 16;
                                if (P.Dc[ga.y] * P.Dbar[ga.y] / Z > thr && R.stored(P, ar - 1, br - 1))
                                    atomicAdd(R.at(R.Mb, P, ar - 1, br - 1), P.Dbar[ga.y] * P.eam[ga.y]);
                            }
                            scratch_sync();
                            pf_reverse(P, R, false, P.edge_ref, nullptr);
                        }
                    }
                }
                scratch_sync();
                __syncthreads();
            }
            if (K > 1) pair_barrier();
        }
        // fragment queries: the inside of each block (the virtual arc match
        // (i-1, j+1; k-1, l+1), AlignerP::compute_fragment_match_prob,
        // aligner_p.icc:1440-1479), one wave per query in its global scratch
        for (int q = member * nwaves + wave; q < P.nfq; q += K * nwaves) {
            int fi, fj, fk, fl;
            frag_query(P, q, fi, fj, fk, fl);
            const int Wq = (int)P.fq[8 * q + 4];
            if (Wq <= 0) continue;
            const Region R = region(fi - 1, fk - 1, fj + 1, fl + 1, Wq);
            pf_fill(P, R);
            if (lane == 0) P.fq[8 * q + 5] = R.get(R.M, P, fj, fl);
        }
        if (threadIdx.x == 0) {
            if (K > 1)
                atomicAdd(&pfa[f.out + 1], s_open);
            else
                pfa[f.out + 1] = s_open;
        }
        __syncthreads();
    }
}

extern "C" __global__ void lcg_pf_reset_barriers(int *barriers, int n) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) barriers[t] = 0;
}

extern "C" __global__ void lcg_pf_reset_queue(int *queue) {
    if (threadIdx.x == 0) *queue = 0;
}

extern "C" int lcg_launch_pf(const PairDesc *descs, const PfDesc *pdescs, const int *order, int *queue, int *arena,
                             double *pfa, double *scratch, const int *sched, int *barriers, double *fstore,
                             PfBatchParams bp, int grid, int waves, size_t lds_bytes, hipStream_t stream) {
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)lcg_pf_pairs, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           160 * 1024 - 256);
        if (e != hipSuccess) return (int)e;
        attr_set = true;
    }
    hipLaunchKernelGGL(lcg_pf_reset_queue, dim3(1), dim3(64), 0, stream, queue);
    if (bp.wg_per_pair > 1)
        hipLaunchKernelGGL(lcg_pf_reset_barriers, dim3(1), dim3(256), 0, stream, barriers, bp.npairs);
    hipLaunchKernelGGL(lcg_pf_pairs, dim3(grid), dim3(waves * 64), lds_bytes, stream, descs, pdescs, order, queue,
                       arena, pfa, scratch, sched, barriers, fstore, bp);
    return (int)hipGetLastError();
}
