// gfx950 kernel of the trace band: the Gotoh partition function over the
// StralScore similarity, forward and on the reversed sequences, and the
// trace probability of every cell of the band, reduced to the first and
// last column of each row whose probability reaches the cut.
//
// Reference: PFGotoh::pf_gotoh (edge_probs.icc:76-162; initialization and
// the three recursions), PFTraceProbs (:212-268; the combination of forward
// and reverse values), StralScore::sigma (stral_score.cc:29-60), and
// TraceController::restrict_by_trace_probabilities
// (trace_controller.cc:565-597), whose per-row first / last passing column
// the host applies.
//
// One workgroup of 256 threads per pair (a queue over the pairs, longest
// first).  The workgroup sweeps anti-diagonals d = i + j; the three
// matrices of the last three diagonals live in LDS, the forward values of
// every cell in HBM for the probabilities.  Every diagonal is renormalized
// by an exact power of two (the largest exponent of the diagonal before), so
// the values stay near 1 at any length: the true value of a cell on
// diagonal d is v * 2^E[d].  The probabilities are computed in double.  A
// pair is flagged, and the host recomputes its band in its own precision
// (so the band always equals the host's), when a cell's probability lies
// within eps (relative) of the cut or is not finite, when a true value lies
// outside the range of the host's precision (double: 2^-1000 .. 2^1000,
// long double: 2^-16000 .. 2^16000), or when a nonzero contribution falls
// below 2^-1000 of its diagonal's scale (not kept exactly here).
#include <hip/hip_runtime.h>

#include <climits>

#include "band_layout.h"

namespace {

constexpr int BT = 256;

__device__ __forceinline__ int fexp(double v) { return v != 0.0 ? ilogb(v) + 1 : INT_MIN; }

struct BandLds {
    double *M[3], *A[3], *B[3];  // last three diagonals, indexed by i
    int *EF, *ER;                // diagonal exponents, forward and reverse passes
    int *lo, *hi;                // first / last passing column per row
    double *upA, *dnA, *upB, *dnB;
    int *cA, *cB;
    int *emax;       // [3] largest exponent of a diagonal (ring)
    double *zsum;    // [3] per-diagonal part of Z (ring): M (local mode) or the
                     // free-end-gap terms of the last column / row (global)
    double *chain;   // [3] gap chain of column 0, [3] of row 0 (rings)
    int *misc;       // [0] flags, [1] ze
    double *zmisc;   // [0] zm, [1] global Z mantissa
};

// Z accumulator (mantissa, exponent): z += t * 2^e
__device__ void zadd(double &zm, int &ze, double t, int e) {
    if (t == 0.0) return;
    if (zm == 0.0) {
        zm = t;
        ze = e;
        return;
    }
    if (ilogb(t) + e > ilogb(zm) + ze) {
        zm = ldexp(zm, ze - e) + t;
        ze = e;
    } else {
        zm += ldexp(t, e - ze);
    }
}

}  // namespace

extern "C" __global__ void __launch_bounds__(BT) lcg_band_pairs(const BandDesc *__restrict__ descs,
                                                                const int *__restrict__ order, int *queue,
                                                                int *iarena, const double *__restrict__ darena,
                                                                double *fwd, BandParams bp) {
    extern __shared__ double lds_d[];
    __shared__ int s_pair;
    const int L1 = bp.lmax + 1;
    BandLds S;
    {
        double *p = lds_d;
        for (int r = 0; r < 3; ++r) {
            S.M[r] = p;
            p += L1;
            S.A[r] = p;
            p += L1;
            S.B[r] = p;
            p += L1;
        }
        S.upA = p;
        p += L1;
        S.dnA = p;
        p += L1;
        S.upB = p;
        p += L1;
        S.dnB = p;
        p += L1;
        S.zsum = p;
        p += 4;
        S.zmisc = p;
        p += 2;
        S.chain = p;
        p += 6;
        int *q = reinterpret_cast<int *>(p);
        S.EF = q;
        q += 2 * L1;
        S.ER = q;
        q += 2 * L1;
        S.lo = q;
        q += L1;
        S.hi = q;
        q += L1;
        S.cA = q;
        q += L1;
        S.cB = q;
        q += L1;
        S.emax = q;
        q += 4;
        S.misc = q;
    }
    const int tid = threadIdx.x;
    const double go = bp.g_open, ge = bp.g_ext, goge = go * ge;
    const int rlim = bp.extended ? 16000 : 1000;
    // free end gaps: the forward pass takes the flags as given, the reverse
    // pass those of the reversed sequences (left and right swapped); the
    // global Z gains the last column's A (second sequence's left flag) and
    // the last row's B (first sequence's left flag) of the forward pass
    // (edge_probs.icc:39-53, :120-131)
    const int fe_fwd = bp.free_endgaps;
    const int fe_rev = ((fe_fwd & 2) >> 1) | ((fe_fwd & 1) << 1) | ((fe_fwd & 8) >> 1) | ((fe_fwd & 4) << 1);
    const bool zextra = !bp.local && (fe_fwd & 5);
    for (;;) {
        if (tid == 0) s_pair = atomicAdd(queue, 1);
        __syncthreads();
        const int k = s_pair;
        if (k >= bp.npairs) break;
        const BandDesc D = descs[order[k]];
        const int n = D.n, m = D.m;
        const int *mn = iarena + D.band, *mx = iarena + D.band + n + 1;
        for (int t = tid; t <= n; t += BT) {
            S.cA[t] = iarena[D.codes + t];
            S.upA[t] = darena[D.prof + t];
            S.dnA[t] = darena[D.prof + n + 1 + t];
            S.lo[t] = INT_MAX;
            S.hi[t] = -1;
        }
        for (int t = tid; t <= m; t += BT) {
            S.cB[t] = iarena[D.codes + n + 1 + t];
            S.upB[t] = darena[D.prof + 2 * (n + 1) + t];
            S.dnB[t] = darena[D.prof + 2 * (n + 1) + m + 1 + t];
        }
        if (tid == 0) S.misc[0] = 0;
        double *F = fwd + D.fwd;
        // Boltzmann weight of the similarity of cell (i, j), 1 <= i <= n,
        // 1 <= j <= m (stral_score.cc:29-60, one row per sequence)
        auto weight = [&](int i, int j) {
            const int a = S.cA[i], b = S.cB[j];
            double seq = 0.0;
            if (a >= 0 && b >= 0) seq = bp.sim[a * 4 + b];
            const double s = bp.sw * (sqrt(S.dnA[i] * S.dnB[j]) + sqrt(S.upA[i] * S.upB[j])) + seq;
            return exp(s / bp.temp);
        };
        double zm = 0.0;  // thread 0: Z of the forward pass
        int ze = 0;
        for (int pass = 0; pass < 2; ++pass) {
            const bool rev = pass == 1;
            int *E = rev ? S.ER : S.EF;
            const int fe = rev ? fe_rev : fe_fwd;
            const bool zacc = !rev && (bp.local || zextra);
            // band of this orientation: reversed cell (i', j') is cell
            // (n - i', m - j'), reversed row i' spans m - max_col(n - i')
            // .. m - min_col(n - i')
            auto bmn = [&](int i) { return rev ? m - mx[n - i] : mn[i]; };
            auto bmx = [&](int i) { return rev ? m - mn[n - i] : mx[i]; };
            if (tid < 3) {
                S.emax[tid] = INT_MIN;
                S.zsum[tid] = 0.0;
            }
            __syncthreads();
            for (int d = 0; d <= n + m; ++d) {
                const int r0 = d % 3, r1 = (d + 2) % 3, r2 = (d + 1) % 3;  // slots of d, d-1, d-2
                int Ed = 0;
                if (d > 0) {
                    const int e1 = S.emax[r1];
                    Ed = E[d - 1] + (e1 == INT_MIN ? 0 : e1);
                }
                if (tid == 0) {
                    E[d] = Ed;
                    if (d > 0 && zacc) zadd(zm, ze, S.zsum[r1], E[d - 1]);
                    S.emax[r2] = INT_MIN;  // slot of d + 1
                    S.zsum[r2] = 0.0;
                }
                const double s0 = ldexp(1.0, -Ed);
                const double f1 = d >= 1 ? ldexp(1.0, E[d - 1] - Ed) : 0.0;
                const double f2 = d >= 2 ? ldexp(1.0, E[d - 2] - Ed) : 0.0;
                const int i0 = d - m > 0 ? d - m : 0, i1 = d < n ? d : n;
                int xmax = INT_MIN;
                double zpart = 0.0;
                int flags = 0;
                // a nonzero contribution t whose scaled result v fell below
                // 2^-1000 (relative to the diagonal's scale) is not kept
                // exactly: the pair goes back to the host
                auto under = [&](double t, double v) {
                    if (t != 0.0 && (v == 0.0 || ilogb(v) < -1000)) flags |= LCG_BAND_RANGE;
                };
                // true value outside the host precision's range
                auto range = [&](double v) {
                    if (v != 0.0) {
                        const int x = fexp(v) + Ed;
                        if (x > rlim || x < -rlim) flags |= LCG_BAND_RANGE;
                    }
                };
                for (int i = i0 + tid; i <= i1; i += BT) {
                    const int j = d - i;
                    double vM = 0.0, vA = 0.0, vB = 0.0;
                    if (i >= 1 && j >= 1) {
                        const int lo = bmn(i) > 1 ? bmn(i) : 1, hi = bmx(i) < m ? bmx(i) : m;
                        if (j >= lo && j <= hi) {
                            const double w = rev ? weight(n + 1 - i, m + 1 - j) : weight(i, j);
                            const double tM = S.M[r2][i - 1] * w + S.A[r2][i - 1] * w + S.B[r2][i - 1] * w;
                            const double tA = S.A[r1][i - 1] * ge + S.M[r1][i - 1] * goge + S.B[r1][i - 1] * goge;
                            const double tB = S.B[r1][i] * ge + S.M[r1][i] * goge + S.A[r1][i] * goge;
                            vM = tM * f2;
                            under(tM, vM);
                            if (bp.local) {
                                const double lw = w * s0;
                                under(w, lw);
                                vM += lw;
                            }
                            vA = tA * f1;
                            under(tA, vA);
                            vB = tB * f1;
                            under(tB, vB);
                        }
                    } else if (i == 0 && j == 0) {
                        if (bmn(0) <= 0 && 0 <= bmx(0)) vM = bp.local ? 0.0 : s0;
                    } else if (j == 0) {  // i >= 1
                        // gap chain of column 0 (kept apart from the free
                        // start, which the reference adds after the chain)
                        double ch = 0.0;
                        if (i == 1) {
                            if (bmn(1) <= 0 && 0 <= bmx(1)) {
                                ch = goge * s0;
                                under(goge, ch);
                            }
                        } else if (bmn(i) <= 0) {
                            const double t = S.chain[r1] * ge;
                            ch = t * f1;
                            under(t, ch);
                        }
                        S.chain[r0] = ch;
                        vA = ch;
                        if (fe & 4) {  // free left end of the second sequence: 1 at A(i, 0)
                            under(1.0, s0);
                            vA += s0;
                        }
                    } else {  // i == 0, j >= 1
                        double ch = 0.0;
                        if (j == 1) {
                            if (bmn(0) <= 1 && 1 <= bmx(0)) {
                                ch = goge * s0;
                                under(goge, ch);
                            }
                        } else if (j >= (bmn(0) > 2 ? bmn(0) : 2) && j <= (bmx(0) < m ? bmx(0) : m)) {
                            const double t = S.chain[3 + r1] * ge;
                            ch = t * f1;
                            under(t, ch);
                        }
                        S.chain[3 + r0] = ch;
                        vB = ch;
                        if (fe & 1) {  // free left end of the first sequence: 1 at B(0, j)
                            under(1.0, s0);
                            vB += s0;
                        }
                    }
                    S.M[r0][i] = vM;
                    S.A[r0][i] = vA;
                    S.B[r0][i] = vB;
                    range(vM);
                    range(vA);
                    range(vB);
                    const int xm = fexp(vM), xa = fexp(vA), xb = fexp(vB);
                    const int x = xm > xa ? (xm > xb ? xm : xb) : (xa > xb ? xa : xb);
                    xmax = x > xmax ? x : xmax;
                    if (!rev) {
                        double *c = F + 3 * ((size_t)i * (m + 1) + j);
                        c[0] = vM;
                        c[1] = vA;
                        c[2] = vB;
                        if (bp.local) zpart += vM;
                        if (zextra) {
                            if (j == m && (fe_fwd & 4)) zpart += vA;
                            if (i == n && (fe_fwd & 1)) zpart += vB;
                        }
                    } else {
                        // trace probability of cell (n - i, m - j)
                        const int ci = n - i, cj = m - j;
                        if (cj >= mn[ci] && cj <= mx[ci]) {
                            const double *c = F + 3 * ((size_t)ci * (m + 1) + cj);
                            const double fM = c[0], fA = c[1], fB = c[2];
                            const double la = bp.local ? s0 : 0.0;
                            const double rs = vM + vA + vB + la;
                            const double z = fM * rs + fA * (vM + vA / go + vB) + fB * (vM + vA + vB / go);
                            const int eF = S.EF[ci + cj];
                            const double P = ldexp(z / S.zmisc[0], eF + Ed - S.misc[1]);
                            if (!bp.extended && rs != 0.0) {
                                // products the host forms in double
                                const int xs = fexp(rs) + Ed;
                                const double fv[3] = {fM, fA, fB};
                                for (int q = 0; q < 3; ++q) {
                                    if (fv[q] == 0.0) continue;
                                    const int xp = xs + fexp(fv[q]) + eF;
                                    if (xp > rlim || xp < -rlim) flags |= LCG_BAND_RANGE;
                                }
                            }
                            if (!isfinite(P) || fabs(P - bp.thr) <= bp.eps * bp.thr) {
                                flags |= LCG_BAND_AMBIGUOUS;
                            } else if (P >= bp.thr) {
                                atomicMin(&S.lo[ci], cj);
                                atomicMax(&S.hi[ci], cj);
                            }
                        }
                    }
                }
                if (xmax != INT_MIN) atomicMax(&S.emax[r0], xmax);
                if (flags) atomicOr(&S.misc[0], flags);
                if (zacc && zpart != 0.0) atomicAdd(&S.zsum[r0], zpart);
                if (!rev && !bp.local && tid == 0 && d == n + m) {
                    // the thread of cell (n, m) is the one with i = n on the
                    // last diagonal: i0 = i1 = n there, so thread 0
                    S.zmisc[1] = S.M[r0][n] + S.A[r0][n] + S.B[r0][n];
                }
                __syncthreads();
            }
            if (!rev && tid == 0) {
                if (bp.local) {
                    zadd(zm, ze, S.zsum[(n + m) % 3], E[n + m]);
                    zadd(zm, ze, 1.0, 0);  // the empty alignment
                } else {
                    if (zextra) zadd(zm, ze, S.zsum[(n + m) % 3], E[n + m]);
                    zadd(zm, ze, S.zmisc[1], E[n + m]);  // cell (n, m)
                }
                S.zmisc[0] = zm;
                S.misc[1] = ze;
                if (!(zm > 0.0) || !isfinite(zm)) S.misc[0] |= LCG_BAND_AMBIGUOUS;
            }
            __syncthreads();
        }
        int *rows = iarena + D.rows;
        for (int t = tid; t <= n; t += BT) {
            rows[t] = S.lo[t];
            rows[n + 1 + t] = S.hi[t];
        }
        if (tid == 0) rows[2 * (n + 1)] = S.misc[0];
        __syncthreads();
    }
}

namespace {
int set_band_lds_limit() {
    static bool done = false;
    if (done) return 0;
    hipError_t e = hipFuncSetAttribute((const void *)lcg_band_pairs, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       160 * 1024 - 256);
    if (e != hipSuccess) return (int)e;
    done = true;
    return 0;
}
}  // namespace

// Dynamic LDS bytes of lcg_band_pairs for sequences of at most lmax.
extern "C" size_t lcg_band_lds_bytes(int lmax) {
    const size_t L1 = (size_t)lmax + 1;
    return (13 * L1 + 12) * sizeof(double) + (8 * L1 + 8) * sizeof(int);
}

extern "C" int lcg_launch_band(const BandDesc *descs, const int *order, int *queue, int *iarena,
                               const double *darena, double *fwd, BandParams bp, int grid, hipStream_t stream) {
    if (int e = set_band_lds_limit()) return e;
    const size_t lds = lcg_band_lds_bytes(bp.lmax);
    if (hipError_t e = hipMemsetAsync(queue, 0, sizeof(int), stream)) return (int)e;
    hipLaunchKernelGGL(lcg_band_pairs, dim3(grid), dim3(BT), lds, stream, descs, order, queue, iarena, darena, fwd,
                       bp);
    return (int)hipGetLastError();
}
