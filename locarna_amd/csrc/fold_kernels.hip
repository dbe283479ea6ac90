// Base-pair probabilities of single sequences (McCaskill's inside and
// outside algorithm) over the energy model of fold_model.h: the input step
// that mlocarna runs once per sequence (locarna_rnafold_pp ->
// RnaEnsemble, rna_ensemble.cc:141-193; mlocarna:3148-3188).
//
// One workgroup of 16 waves folds one sequence.  The inside pass walks the
// diagonals d = j - i upward; the cells of a diagonal are independent, a
// quarter wave (16 lanes) per cell, and the lanes of a cell split its terms
// (interior loops, multiloop splits) and sum them with a butterfly; 16
// lanes per cell measured faster than 8, 32 or 64 on 64 sequences of 60-200 nt
// (the work is latency-bound: more cells in flight per wave).  The exterior
// prefix sums Z5 (forward) and Z5b (their adjoints) run on one wave.  The
// outside pass walks the diagonals downward and gathers, for every cell, the
// terms of the enclosing cells that use it (the pull form of the adjoint of
// the inside recursions), so that no cell is written by more than one wave
// and the result does not depend on the schedule.  The probabilities
// P = Qb Qbb / Z and the stacked joint probabilities are written as the
// outside value of a cell is known.
//
// Recursions (1-based, t = pair type of (i, j), 0 when (i, j) cannot pair):
//   Qb(i,j)  = H(i,j) + sum_{k,l} L(i,j,k,l) Qb(k,l)
//              + MLC TP(t) sum_u Qm(i+1,u-1) Qm1(u,j-1)
//   Qm1(i,j) = sum_{l <= j} Qb(i,l) BR(i,l)
//   Qm(i,j)  = sum_{u >= i} (1 + Qm(i,u-1)) Qm1(u,j)
//   Z5(j)    = Z5(j-1) + sum_k Z5(k-1) Qb(k,j) EXT(k,j),  Z = Z5(n)
// Every value of a segment of L bases is stored times sigma^-L (sigma per
// sequence, FoldDesc.log_sigma; ViennaRNA's per-nucleotide pf_scale), so
// each term carries sigma^-(bases it adds): SC[l1 + l2 + 2] on an interior
// loop, SC[2] on a multiloop's closing pair, SC[d + 1] on a hairpin, SC[j - l]
// on the unpaired tail of a branch, SC[1] per exterior base.  The
// probabilities do not depend on sigma; log Z = log Z5(n) + n log sigma.
#include <hip/hip_runtime.h>

#include "fold_layout.h"

namespace {

constexpr int FT = LCG_FOLD_THREADS;
#ifndef LCG_FOLD_CELL_LANES
#define LCG_FOLD_CELL_LANES 16
#endif
// lanes per cell: a wave works on 64 / CL cells of a diagonal at once
constexpr int CL = LCG_FOLD_CELL_LANES;

// sum over the CL lanes of a cell
__device__ __forceinline__ double cell_sum(double v) {
#pragma unroll
    for (int o = CL / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ int weak(int t) { return (0x72 >> t) & 1; }

struct Seq {
    const signed char *cs;  // codes in LDS, position i at cs[i]
    const int *ptab;        // pair type of (code a, code b) at a * 5 + b
    int n, W;
    // pair type of (i, j) with the hairpin and span limits, 0 if none
    __device__ __forceinline__ int pt(int i, int j) const {
        const int d = j - i;
        return (d >= 4 && d < W) ? ptab[cs[i] * 5 + cs[j]] : 0;
    }
};

__device__ __forceinline__ double loopf(const double *T, int l1, int l2, int to, int ti) {
    const int s = l1 + l2;
    if (s == 0) return T[LCG_FOLD_TAB_STACK + to * 8 + ti];
    if (s == 1) return T[LCG_FOLD_TAB_BULGE1 + to * 8 + ti];
    return T[LCG_FOLD_TAB_LOOP + (l1 * 31 + l2) * 4 + weak(to) * 2 + weak(ti)];
}

// number of (l1, l2) entries with l1 + l2 <= s
__device__ __forceinline__ int loop_count(int s) { return s < 0 ? 0 : (s + 1) * (s + 2) / 2; }

}  // namespace

extern "C" __global__ __launch_bounds__(FT) void lcg_fold_seqs(const FoldDesc *descs, const signed char *codes,
                                                               const double *tabs, double *mats, double *outs,
                                                               double *zout, int nseq) {
    extern __shared__ double lds[];
    const int sq = blockIdx.x;
    if (sq >= nseq) return;
    const FoldDesc D = descs[sq];
    const int n = D.n, W = D.W;
    double *T = lds;                       // fixed factor table
    double *Z5 = T + LCG_FOLD_TAB_FIXED;   // n + 1
    double *Z5b = Z5 + (n + 1);            // n + 1
    double *SC = Z5b + (n + 1);            // n + 2: SC[L] = sigma^-L
    int *ptab = (int *)(SC + (n + 2));     // 25
    unsigned short *LL = (unsigned short *)(ptab + 28);  // LCG_FOLD_NLOOP: l1 | l2 << 8
    signed char *cs = (signed char *)(LL + LCG_FOLD_NLOOP + 8);  // n + 2
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = FT / 64;
    const int cl = lane % CL, cg = wave * (64 / CL) + lane / CL, ncg = nw * (64 / CL);

    for (int x = tid; x < LCG_FOLD_TAB_FIXED; x += FT) T[x] = tabs[x];
    for (int x = tid; x < n + 2; x += FT) cs[x] = (x >= 1 && x <= n) ? codes[D.codes + x] : (signed char)4;
    for (int x = tid; x < n + 2; x += FT) SC[x] = exp(-(double)x * D.log_sigma);
    if (tid < 25) {
        // canonical pairs AU, CG, GC, UA, GU, UG as types 1..6
        const int a = tid / 5, b = tid % 5;
        int t = 0;
        if (a == 0 && b == 3) t = 1;
        else if (a == 1 && b == 2) t = 2;
        else if (a == 2 && b == 1) t = 3;
        else if (a == 3 && b == 0) t = 4;
        else if (a == 2 && b == 3) t = 5;
        else if (a == 3 && b == 2) t = 6;
        ptab[tid] = t;
    }
    for (int e = tid; e < LCG_FOLD_NLOOP; e += FT) {
        int s = 0;
        while (loop_count(s) <= e) ++s;
        const int l1 = e - loop_count(s - 1);
        LL[e] = (unsigned short)(l1 | ((s - l1) << 8));
    }
    __syncthreads();

    const Seq S{cs, ptab, n, W};
    const size_t sz = (size_t)(n + 2) * W;
    double *Qb = mats + D.mat, *Qm = Qb + sz, *Qm1 = Qm + sz;
    double *Qbb = Qm1 + sz, *Qmb = Qbb + sz, *Qm1b = Qmb + sz;
    double *P = outs + D.out, *P2 = P + sz;
    const double *hp = tabs + LCG_FOLD_TAB_HP;
    const double mlc = T[LCG_FOLD_TAB_MLC];

    // ---------------- inside
    for (int d = 0; d < W; ++d) {
        for (int c = cg; c < n - d; c += ncg) {
            const int i = 1 + c, j = i + d;
            double qb = 0, qm1 = 0, qm = 0;
            if (d >= 4) {
                const int t = S.pt(i, j);
                if (t) {
                    double s = 0;
                    const int cnt = loop_count(min(LCG_FOLD_MAXLOOP, d - 6));
#pragma unroll 4
                    for (int e = cl; e < cnt; e += CL) {
                        const int l1 = LL[e] & 255, l2 = LL[e] >> 8;
                        const int k = i + 1 + l1, l = j - 1 - l2;
                        const int ti = S.pt(k, l);
                        // the load does not wait for the pair test: (k, l)
                        // lies in the band and holds 0 where it cannot pair
                        const double q = Qb[(size_t)k * W + (l - k)];
                        s += (ti ? loopf(T, l1, l2, t, ti) * SC[l1 + l2 + 2] : 0.0) * q;
                    }
                    double mq = 0;
                    for (int u = i + 6 + cl; u <= j - 5; u += CL)
                        mq += Qm[(size_t)(i + 1) * W + (u - 2 - i)] * Qm1[(size_t)u * W + (j - 1 - u)];
                    s = cell_sum(s);
                    mq = cell_sum(mq);
                    qb = hp[(d - 1) * 2 + weak(t)] * SC[d + 1] + s + mq * mlc * T[LCG_FOLD_TAB_TP + weak(t)] * SC[2];
                }
                double a1 = 0;
                for (int l = i + 4 + cl; l < j; l += CL) {
                    const int tl = S.pt(i, l);
                    const double q = Qb[(size_t)i * W + (l - i)];  // 0 where (i, l) cannot pair
                    a1 += (tl ? T[LCG_FOLD_TAB_BRANCH + weak(tl)] * SC[j - l] : 0.0) * q;
                }
                a1 = cell_sum(a1);
                if (t) a1 += qb * T[LCG_FOLD_TAB_BRANCH + weak(t)];
                qm1 = a1;
                double a2 = 0;
                for (int u = i + 1 + cl; u <= j - 4; u += CL)
                    a2 += (SC[u - i] + Qm[(size_t)i * W + (u - 1 - i)]) * Qm1[(size_t)u * W + (j - u)];
                qm = cell_sum(a2) + qm1;
            }
            if (cl == 0) {
                const size_t x = (size_t)i * W + d;
                Qb[x] = qb;
                Qm1[x] = qm1;
                Qm[x] = qm;
            }
        }
        __syncthreads();
    }

    // ---------------- exterior loop: Z5 forward, then its adjoints.  The
    // sum of column j reads Z5 up to j - 5 only, so four consecutive columns
    // are summed at once (a quarter wave each) and chained in order.
    if (wave == 0) {
        const int g = lane >> 4, gl = lane & 15;
        double zlast = 1.0;
        if (lane == 0) Z5[0] = 1.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        for (int j0 = 1; j0 <= n; j0 += 4) {
            const int j = j0 + g;
            double f = 0;
            if (j <= n)
                for (int k = max(1, j - W + 1) + gl; k <= j - 4; k += 16) {
                    const int t = S.pt(k, j);
                    const double q = Qb[(size_t)k * W + (j - k)];
                    f += (t ? Z5[k - 1] * T[LCG_FOLD_TAB_EXT + weak(t)] : 0.0) * q;
                }
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) f += __shfl_xor(f, o);
            const double z0 = zlast * SC[1] + __shfl(f, 0), z1 = z0 * SC[1] + __shfl(f, 16),
                         z2 = z1 * SC[1] + __shfl(f, 32), z3 = z2 * SC[1] + __shfl(f, 48);
            if (lane == 0) {
                Z5[j0] = z0;
                if (j0 + 1 <= n) Z5[j0 + 1] = z1;
                if (j0 + 2 <= n) Z5[j0 + 2] = z2;
                if (j0 + 3 <= n) Z5[j0 + 3] = z3;
            }
            zlast = j0 + 3 <= n ? z3 : j0 + 2 <= n ? z2 : j0 + 1 <= n ? z1 : z0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        }
        // Z5b(m) = Z5b(m+1) + sum_j Z5b(j) Qb(m+1,j) EXT reads Z5b from m + 5 up
        double glast = 1.0;
        if (lane == 0) Z5b[n] = 1.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        for (int m0 = n - 1; m0 >= 0; m0 -= 4) {
            const int m = m0 - g;
            double f = 0;
            if (m >= 0)
                for (int j = m + 5 + gl; j <= min(n, m + W); j += 16) {
                    const int t = S.pt(m + 1, j);
                    const double q = Qb[(size_t)(m + 1) * W + (j - m - 1)];
                    f += (t ? Z5b[j] * T[LCG_FOLD_TAB_EXT + weak(t)] : 0.0) * q;
                }
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) f += __shfl_xor(f, o);
            const double z0 = glast * SC[1] + __shfl(f, 0), z1 = z0 * SC[1] + __shfl(f, 16),
                         z2 = z1 * SC[1] + __shfl(f, 32), z3 = z2 * SC[1] + __shfl(f, 48);
            if (lane == 0) {
                Z5b[m0] = z0;
                if (m0 - 1 >= 0) Z5b[m0 - 1] = z1;
                if (m0 - 2 >= 0) Z5b[m0 - 2] = z2;
                if (m0 - 3 >= 0) Z5b[m0 - 3] = z3;
            }
            glast = m0 - 3 >= 0 ? z3 : m0 - 2 >= 0 ? z2 : m0 - 1 >= 0 ? z1 : z0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        }
        if (lane == 0) zout[sq] = Z5[n];
    }
    __syncthreads();
    const double Z = Z5[n];

    // ---------------- outside (gathered adjoints), probabilities
    for (int d = W - 1; d >= 0; --d) {
        for (int c = cg; c < n - d; c += ncg) {
            const int a = 1 + c, e = a + d;
            double qmb = 0, qm1b = 0, qbb = 0, p = 0, p2 = 0;
            if (d >= 4) {
                // Qm(a, e): inside Qb(a-1, j) as Qm(i+1, u-1), inside Qm(a, j) as Qm(i, u-1)
                double s = 0;
                if (a >= 2)
                    for (int j = e + 6 + cl; j <= min(n, a + W - 2); j += CL) {
                        const int t = S.pt(a - 1, j);
                        const double o = Qbb[(size_t)(a - 1) * W + (j - a + 1)], q = Qm1[(size_t)(e + 1) * W + (j - e - 2)];
                        s += (t ? mlc * T[LCG_FOLD_TAB_TP + weak(t)] * SC[2] : 0.0) * o * q;
                    }
                for (int j = e + 5 + cl; j <= min(n, a + W - 1); j += CL)
                    s += Qmb[(size_t)a * W + (j - a)] * Qm1[(size_t)(e + 1) * W + (j - e - 1)];
                qmb = cell_sum(s);
                // Qm1(a, e): inside Qb(i, e+1) as Qm1(u, j-1), inside Qm(i, e) as Qm1(u, j)
                s = 0;
                if (e + 1 <= n)
                    for (int i = max(1, e + 2 - W) + cl; i <= a - 6; i += CL) {
                        const int t = S.pt(i, e + 1);
                        const double o = Qbb[(size_t)i * W + (e + 1 - i)], q = Qm[(size_t)(i + 1) * W + (a - 2 - i)];
                        s += (t ? mlc * T[LCG_FOLD_TAB_TP + weak(t)] * SC[2] : 0.0) * o * q;
                    }
                for (int i = max(1, e - W + 1) + cl; i <= a - 1; i += CL)
                    s += Qmb[(size_t)i * W + (e - i)] * (SC[a - i] + Qm[(size_t)i * W + (a - 1 - i)]);
                qm1b = cell_sum(s) + qmb;
                // Qb(a, e): exterior loop, enclosing pairs, multiloop branches
                const int t = S.pt(a, e);
                if (t) {
                    double si = 0, sb = 0;
                    const int cnt = loop_count(min(LCG_FOLD_MAXLOOP, W - 3 - d));
#pragma unroll 4
                    for (int q = cl; q < cnt; q += CL) {
                        const int l1 = LL[q] & 255, l2 = LL[q] >> 8;
                        const int i = a - 1 - l1, j = e + 1 + l2;
                        const bool in = i >= 1 && j <= n;
                        // (i, j) in range lies in the band (cnt bounds l1 + l2); the load
                        // does not wait for the pair test, cells that cannot pair hold 0
                        const double o = in ? Qbb[(size_t)i * W + (j - i)] : 0.0;
                        const int to = in ? S.pt(i, j) : 0;
                        si += (to ? loopf(T, l1, l2, to, t) * SC[l1 + l2 + 2] : 0.0) * o;
                    }
                    for (int j = e + 1 + cl; j <= min(n, a + W - 1); j += CL)
                        sb += Qm1b[(size_t)a * W + (j - a)] * SC[j - e];
                    si = cell_sum(si);
                    sb = cell_sum(sb);
                    qbb = Z5b[e] * Z5[a - 1] * T[LCG_FOLD_TAB_EXT + weak(t)] + si +
                          T[LCG_FOLD_TAB_BRANCH + weak(t)] * (sb + qm1b);
                    const double qb = Qb[(size_t)a * W + d];
                    p = qb * qbb / Z;
                    if (d >= 6) {
                        const int ti = S.pt(a + 1, e - 1);
                        if (ti)
                            p2 = qbb * T[LCG_FOLD_TAB_STACK + t * 8 + ti] * SC[2] * Qb[(size_t)(a + 1) * W + (d - 2)] / Z;
                    }
                }
            }
            if (cl == 0) {
                const size_t x = (size_t)a * W + d;
                Qmb[x] = qmb;
                Qm1b[x] = qm1b;
                Qbb[x] = qbb;
                P[x] = p;
                P2[x] = p2;
            }
        }
        __syncthreads();
    }
}

extern "C" size_t lcg_fold_lds_bytes(int nmax) {
    return (size_t)(LCG_FOLD_TAB_FIXED + 2 * (nmax + 1) + (nmax + 2)) * sizeof(double) + 28 * sizeof(int) +
           (LCG_FOLD_NLOOP + 8) * sizeof(unsigned short) + (size_t)(nmax + 2);
}

extern "C" int lcg_launch_fold(const FoldDesc *descs, const signed char *codes, const double *tabs, double *mats,
                               double *outs, double *zout, int nseq, int nmax, hipStream_t stream) {
    const size_t lds = lcg_fold_lds_bytes(nmax);
    if (hipError_t e = hipFuncSetAttribute((const void *)lcg_fold_seqs, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds))
        return (int)e;
    if (nseq <= 0) return 0;
    hipLaunchKernelGGL(lcg_fold_seqs, dim3(nseq), dim3(FT), lds, stream, descs, codes, tabs, mats, outs, zout, nseq);
    return (int)hipGetLastError();
}
