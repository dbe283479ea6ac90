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

// inclusive scan of the affine maps x -> a x + b over the lanes (lane 0
// first): result (A_l, B_l) with F_l = A_l F_in + B_l
__device__ __forceinline__ void wave_scan_affine(double &a, double &b) {
    const int lane = lane_id();
    for (int d = 1; d < 64; d <<= 1) {
        const double pa = __shfl_up(a, d), pb = __shfl_up(b, d);
        if (lane >= d) {
            b = a * pb + b;
            a = a * pa;
        }
    }
}

// suffix version: x_l = g_l + c_l x_{l+1} (last lane first); returns the
// pair (C_l, G_l) with x_l = C_l x_in + G_l where x_in enters at lane 63
__device__ __forceinline__ void wave_suffix_affine(double &c, double &g) {
    const int lane = lane_id();
    for (int d = 1; d < 64; d <<= 1) {
        const double nc = __shfl_down(c, d), ng = __shfl_down(g, d);
        if (lane + d < 64) {
            g = c * ng + g;
            c = c * nc;
        }
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double readlane_d(double v, int l) {
    return __shfl(v, l);
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
};

__device__ __forceinline__ int row_lo(const PfPair &P, int i, int bl) { return max(bl, P.mn[i]); }
__device__ __forceinline__ int row_hi(const PfPair &P, int i, int br) { return min(br - 1, P.mx[i]); }

struct Region {
    int al, bl, ar, br, W;
    double *M, *E, *F, *Mb;  // matrices (R x W)
    double *Ev, *Eb;         // vectors over columns bl..br-1
    __device__ double *at(double *X, const PfPair &P, int i, int j) const {
        const int lo = (i == al) ? bl : row_lo(P, i, bl);
        return X + (size_t)(i - al) * W + (j - lo);
    }
    __device__ bool stored(const PfPair &P, int i, int j) const {
        if (i < al || i >= ar) return false;
        const int lo = (i == al) ? bl : row_lo(P, i, bl);
        return j >= lo && j <= row_hi(P, i, br);
    }
    __device__ double get(double *X, const PfPair &P, int i, int j) const {
        return stored(P, i, j) ? *at(X, P, i, j) : 0.0;
    }
};

__device__ __forceinline__ void scratch_sync() { __threadfence(); }

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
            double sl = __shfl_up(s, 1);
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
// Dbar (per candidate entry), match-edge weights into `edge` and the
// gap-opening weight.  Mb is consumed.
__device__ void pf_reverse(const PfPair &P, const Region &R, bool want_dbar, double *edge, double *openings) {
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
            double a_next = __shfl_down(act ? gB : 0.0, 1);
            if (lane == 63 || !(j + 1 <= hi)) a_next = (lane == 63) ? a_in : 0.0;
            double c = act ? a_next : 0.0, gg = act ? g : 0.0;
            wave_suffix_affine(c, gg);
            const double Fb = c * Fb_in + gg;  // Fbar_j
            // sbar_j = g_j + Fbar_{j+1} gB(j+1) o
            double Fb_next = __shfl_down(Fb, 1);
            double gB_next = __shfl_down(gB, 1);
            if (lane == 63) {
                Fb_next = Fb_in;
                gB_next = a_in;
            }
            if (!(j + 1 <= hi) && lane != 63) {
                Fb_next = 0.0;
                gB_next = 0.0;
            }
            const double sb = g + Fb_next * gB_next * eo;
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

__device__ void zero_region_bar(const Region &R, const PfPair &P) {
    const int lane = lane_id();
    const size_t total = (size_t)(R.ar - R.al) * R.W;
    for (size_t c = lane; c < total; c += 64) R.Mb[c] = 0.0;
    scratch_sync();
}

}  // namespace

extern "C" __global__ void __launch_bounds__(512) lcg_pf_pairs(const PairDesc *__restrict__ descs,
                                                               const PfDesc *__restrict__ pdescs,
                                                               const int *__restrict__ order, int *queue,
                                                               int *arena, double *pfa, double *scratch,
                                                               PfBatchParams bp) {
    extern __shared__ int lds[];
    __shared__ int s_pair;
    __shared__ double s_open;
    const int nwaves = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = lane_id();
    const int L = bp.max_len;
    int *mn = lds, *mx = lds + (L + 1), *coff = lds + 2 * (L + 1);
    int *codeA = lds + 3 * L + 4, *codeB = lds + 4 * L + 5;
    double *dl = reinterpret_cast<double *>(lds + ((5 * L + 6 + 1) & ~1));
    double *egA = dl, *egB = dl + (L + 1), *es = dl + 2 * (L + 1);
    double *wscr = scratch + (size_t)(blockIdx.x * nwaves + wave) * bp.scratch_stride;
    for (;;) {
        if (threadIdx.x == 0) s_pair = atomicAdd(queue, 1);
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
        for (int t = threadIdx.x; t <= n + 1; t += blockDim.x) coff[t] = arena[d.band + 2 * (n + 1) + t];
        for (int t = threadIdx.x; t < LCG_NCODES * LCG_NCODES; t += blockDim.x) es[t] = pfa[f.es + t];
        for (int t = threadIdx.x; t < d.nam; t += blockDim.x) {
            pfa[f.Dc + t] = 0.0;
            pfa[f.Dbar + t] = 0.0;
        }
        for (int t = threadIdx.x; t < d.ncells; t += blockDim.x) {
            pfa[f.edge_full + t] = 0.0;
            pfa[f.edge_ref + t] = 0.0;
        }
        if (threadIdx.x == 0) s_open = 0.0;
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
        __threadfence();
        __syncthreads();
        const int *rows = arena + d.rows;
        const int *groups = arena + d.groups;
        const int4 *gams = reinterpret_cast<const int4 *>(arena + d.gams);
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
        // ---- inside: rows al = n .. 1
        for (int al = n; al >= 1; --al) {
            const int g0 = rows[al + 1], g1 = rows[al];
            for (int g = g0 + wave; g < g1; g += nwaves) {
                const int *G = groups + g * LCG_GROUP_INTS;
                const Region R = region(G[0] & 0xffff, G[0] >> 16, G[1] & 0xffff, G[1] >> 16, G[2]);
                pf_fill(P, R);
                for (int t = G[3] + lane; t < G[4]; t += 64) {
                    const int4 ga = gams[t];
                    const int ar = ga.x & 0xffff, br = ga.x >> 16;
                    P.Dc[ga.y] = R.get(R.M, P, ar - 1, br - 1) * P.eam[ga.y];
                }
            }
            __threadfence();
            __syncthreads();
        }
        // ---- top level and its adjoint (wave 0)
        if (wave == 0) {
            const Region R = region(0, 0, n + 1, m + 1, d.wtop);
            pf_fill(P, R);
            const double Z = *R.at(R.M, P, n, m);
            if (lane == 0) pfa[f.out + 0] = Z;
            // full adjoint, seed dZ/dM(n,m) = 1
            zero_region_bar(R, P);
            if (lane == 0) *R.at(R.Mb, P, n, m) = 1.0;
            scratch_sync();
            pf_reverse(P, R, true, P.edge_full, &s_open);
            zero_region_bar(R, P);
            if (lane == 0) *R.at(R.Mb, P, n, m) = 1.0;
            scratch_sync();
            pf_reverse(P, R, false, P.edge_ref, nullptr);
        }
        __threadfence();
        __syncthreads();
        const double Z = pfa[f.out + 0];
        const double thr = pfa[f.scal + 2];
        // ---- outside over the groups, rows al = 1 .. n
        for (int al = 1; al <= n; ++al) {
            const int g0 = rows[al + 1], g1 = rows[al];
            for (int g = g0 + wave; g < g1; g += nwaves) {
                const int *G = groups + g * LCG_GROUP_INTS;
                const Region R = region(G[0] & 0xffff, G[0] >> 16, G[1] & 0xffff, G[1] >> 16, G[2]);
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
                // restricted pass for the base-match probabilities
                bool any = false;
                for (int t = G[3] + lane; t < G[4]; t += 64) {
                    const int4 ga = gams[t];
                    if (P.Dc[ga.y] * P.Dbar[ga.y] / Z > thr) any = true;
                }
                if (__ballot(any)) {
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
            __threadfence();
            __syncthreads();
        }
        if (threadIdx.x == 0) pfa[f.out + 1] = s_open;
        __syncthreads();
    }
}

extern "C" __global__ void lcg_pf_reset_queue(int *queue) {
    if (threadIdx.x == 0) *queue = 0;
}

extern "C" int lcg_launch_pf(const PairDesc *descs, const PfDesc *pdescs, const int *order, int *queue, int *arena,
                             double *pfa, double *scratch, PfBatchParams bp, int grid, int waves, size_t lds_bytes,
                             hipStream_t stream) {
    hipLaunchKernelGGL(lcg_pf_reset_queue, dim3(1), dim3(64), 0, stream, queue);
    hipLaunchKernelGGL(lcg_pf_pairs, dim3(grid), dim3(waves * 64), lds_bytes, stream, descs, pdescs, order, queue,
                       arena, pfa, scratch, bp);
    return (int)hipGetLastError();
}
