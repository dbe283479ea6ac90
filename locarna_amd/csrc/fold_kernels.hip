// Base-pair probabilities of single sequences (McCaskill's inside and
// outside algorithm) over the energy model of fold_model.h: the input step
// that mlocarna runs once per sequence (locarna_rnafold_pp ->
// RnaEnsemble, rna_ensemble.cc:141-193; mlocarna:3148-3188).
//
// One workgroup of 16 waves folds one sequence.  The inside pass walks the
// diagonals d = j - i upward; the cells of a diagonal are independent, one
// wave per cell, and the lanes of the wave split the cell's terms (interior
// loops, multiloop splits) and sum them with a butterfly.  The exterior
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
#include <hip/hip_runtime.h>

#include "fold_layout.h"

namespace {

constexpr int FT = LCG_FOLD_THREADS;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
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
    int *ptab = (int *)(Z5b + (n + 1));    // 25
    unsigned short *LL = (unsigned short *)(ptab + 28);  // LCG_FOLD_NLOOP: l1 | l2 << 8
    signed char *cs = (signed char *)(LL + LCG_FOLD_NLOOP + 8);  // n + 2
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = FT / 64;

    for (int x = tid; x < LCG_FOLD_TAB_FIXED; x += FT) T[x] = tabs[x];
    for (int x = tid; x < n + 2; x += FT) cs[x] = (x >= 1 && x <= n) ? codes[D.codes + x] : (signed char)4;
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
        for (int c = wave; c < n - d; c += nw) {
            const int i = 1 + c, j = i + d;
            double qb = 0, qm1 = 0, qm = 0;
            if (d >= 4) {
                const int t = S.pt(i, j);
                if (t) {
                    double s = 0;
                    const int cnt = loop_count(min(LCG_FOLD_MAXLOOP, d - 6));
                    for (int e = lane; e < cnt; e += 64) {
                        const int l1 = LL[e] & 255, l2 = LL[e] >> 8;
                        const int k = i + 1 + l1, l = j - 1 - l2;
                        const int ti = S.pt(k, l);
                        if (ti) s += loopf(T, l1, l2, t, ti) * Qb[(size_t)k * W + (l - k)];
                    }
                    double mq = 0;
                    for (int u = i + 6 + lane; u <= j - 5; u += 64)
                        mq += Qm[(size_t)(i + 1) * W + (u - 2 - i)] * Qm1[(size_t)u * W + (j - 1 - u)];
                    s = wave_sum(s);
                    mq = wave_sum(mq);
                    qb = hp[(d - 1) * 2 + weak(t)] + s + mq * mlc * T[LCG_FOLD_TAB_TP + weak(t)];
                }
                double a1 = 0;
                for (int l = i + 4 + lane; l < j; l += 64) {
                    const int tl = S.pt(i, l);
                    if (tl) a1 += Qb[(size_t)i * W + (l - i)] * T[LCG_FOLD_TAB_BRANCH + weak(tl)];
                }
                a1 = wave_sum(a1);
                if (t) a1 += qb * T[LCG_FOLD_TAB_BRANCH + weak(t)];
                qm1 = a1;
                double a2 = 0;
                for (int u = i + 1 + lane; u <= j - 4; u += 64)
                    a2 += (1.0 + Qm[(size_t)i * W + (u - 1 - i)]) * Qm1[(size_t)u * W + (j - u)];
                qm = wave_sum(a2) + qm1;
            }
            if (lane == 0) {
                const size_t x = (size_t)i * W + d;
                Qb[x] = qb;
                Qm1[x] = qm1;
                Qm[x] = qm;
            }
        }
        __syncthreads();
    }

    // ---------------- exterior loop: Z5 forward, then its adjoints
    if (wave == 0) {
        double zprev = 1.0;
        if (lane == 0) Z5[0] = 1.0;
        for (int j = 1; j <= n; ++j) {
            double z = 0;
            for (int k = max(1, j - W + 1) + lane; k <= j - 4; k += 64) {
                const int t = S.pt(k, j);
                if (t) z += Z5[k - 1] * Qb[(size_t)k * W + (j - k)] * T[LCG_FOLD_TAB_EXT + weak(t)];
            }
            zprev += wave_sum(z);
            if (lane == 0) Z5[j] = zprev;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        }
        double gprev = 1.0;
        if (lane == 0) Z5b[n] = 1.0;
        for (int m = n - 1; m >= 0; --m) {
            double g = 0;
            for (int j = m + 5 + lane; j <= min(n, m + W); j += 64) {
                const int t = S.pt(m + 1, j);
                if (t) g += Z5b[j] * Qb[(size_t)(m + 1) * W + (j - m - 1)] * T[LCG_FOLD_TAB_EXT + weak(t)];
            }
            gprev += wave_sum(g);
            if (lane == 0) Z5b[m] = gprev;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        }
        if (lane == 0) zout[sq] = Z5[n];
    }
    __syncthreads();
    const double Z = Z5[n];

    // ---------------- outside (gathered adjoints), probabilities
    for (int d = W - 1; d >= 0; --d) {
        for (int c = wave; c < n - d; c += nw) {
            const int a = 1 + c, e = a + d;
            double qmb = 0, qm1b = 0, qbb = 0, p = 0, p2 = 0;
            if (d >= 4) {
                // Qm(a, e): inside Qb(a-1, j) as Qm(i+1, u-1), inside Qm(a, j) as Qm(i, u-1)
                double s = 0;
                if (a >= 2)
                    for (int j = e + 6 + lane; j <= min(n, a + W - 2); j += 64) {
                        const int t = S.pt(a - 1, j);
                        if (t)
                            s += Qbb[(size_t)(a - 1) * W + (j - a + 1)] * mlc * T[LCG_FOLD_TAB_TP + weak(t)] *
                                 Qm1[(size_t)(e + 1) * W + (j - e - 2)];
                    }
                for (int j = e + 5 + lane; j <= min(n, a + W - 1); j += 64)
                    s += Qmb[(size_t)a * W + (j - a)] * Qm1[(size_t)(e + 1) * W + (j - e - 1)];
                qmb = wave_sum(s);
                // Qm1(a, e): inside Qb(i, e+1) as Qm1(u, j-1), inside Qm(i, e) as Qm1(u, j)
                s = 0;
                if (e + 1 <= n)
                    for (int i = max(1, e + 2 - W) + lane; i <= a - 6; i += 64) {
                        const int t = S.pt(i, e + 1);
                        if (t)
                            s += Qbb[(size_t)i * W + (e + 1 - i)] * mlc * T[LCG_FOLD_TAB_TP + weak(t)] *
                                 Qm[(size_t)(i + 1) * W + (a - 2 - i)];
                    }
                for (int i = max(1, e - W + 1) + lane; i <= a - 1; i += 64)
                    s += Qmb[(size_t)i * W + (e - i)] * (1.0 + Qm[(size_t)i * W + (a - 1 - i)]);
                qm1b = wave_sum(s) + qmb;
                // Qb(a, e): exterior loop, enclosing pairs, multiloop branches
                const int t = S.pt(a, e);
                if (t) {
                    double si = 0, sb = 0;
                    const int cnt = loop_count(min(LCG_FOLD_MAXLOOP, W - 3 - d));
                    for (int q = lane; q < cnt; q += 64) {
                        const int l1 = LL[q] & 255, l2 = LL[q] >> 8;
                        const int i = a - 1 - l1, j = e + 1 + l2;
                        if (i >= 1 && j <= n) {
                            const int to = S.pt(i, j);
                            if (to) si += Qbb[(size_t)i * W + (j - i)] * loopf(T, l1, l2, to, t);
                        }
                    }
                    for (int j = e + 1 + lane; j <= min(n, a + W - 1); j += 64) sb += Qm1b[(size_t)a * W + (j - a)];
                    si = wave_sum(si);
                    sb = wave_sum(sb);
                    qbb = Z5b[e] * Z5[a - 1] * T[LCG_FOLD_TAB_EXT + weak(t)] + si +
                          T[LCG_FOLD_TAB_BRANCH + weak(t)] * (sb + qm1b);
                    const double qb = Qb[(size_t)a * W + d];
                    p = qb * qbb / Z;
                    if (d >= 6) {
                        const int ti = S.pt(a + 1, e - 1);
                        if (ti)
                            p2 = qbb * T[LCG_FOLD_TAB_STACK + t * 8 + ti] * Qb[(size_t)(a + 1) * W + (d - 2)] / Z;
                    }
                }
            }
            if (lane == 0) {
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
    return (size_t)(LCG_FOLD_TAB_FIXED + 2 * (nmax + 1)) * sizeof(double) + 28 * sizeof(int) +
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
