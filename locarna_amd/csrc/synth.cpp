// Seeded synthetic RNA inputs: sequences and base-pair probabilities.
//
// LocARNA reads base-pair probabilities from ViennaRNA (RNAfold -p) or from
// `#PP 2.0` files (reference: src/LocARNA/rna_data.cc:52-61 for folding,
// :984-1124 for the pp reader).  ViennaRNA is not available here, so this
// file provides the input side of the benchmark workloads:
//
//  * a splitmix64 stream that draws lengths, GC contents and nucleotides;
//  * a McCaskill-style base-pair partition function (inside + outside, FP64)
//    over a reduced nearest-neighbour energy model: canonical pairs
//    (AU, CG, GU), minimum hairpin of 3, stacking energies, size-dependent
//    hairpin / bulge / interior loop initiation (loops up to 30 unpaired),
//    a linear multiloop model and terminal AU/GU penalties, with a maximum
//    base-pair span (the model is fold_model.h, shared with the device
//    folding of fold_kernels.hip, which computes the same probabilities).
//
// The energies are a simplification of the Turner 2004 set; the probabilities
// are therefore McCaskill-like, not RNAfold-identical.  They only serve as
// realistic, reproducible input: the aligner consumes whatever `.pp` text the
// generator writes, and the same text is fed to every implementation that is
// compared.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "fold_model.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

using namespace locarna_amd::fold_model;

// Triangular (i<j) storage as dense (n+2)x(n+2) for simplicity.
struct Tri {
    int n;
    std::vector<double> v;
    explicit Tri(int n_) : n(n_), v((size_t)(n_ + 2) * (n_ + 2), 0.0) {}
    double &operator()(int i, int j) { return v[(size_t)i * (n + 2) + j]; }
};

}  // namespace

extern "C" {

// Draws `count` sequences from a splitmix64 stream seeded with `seed`.
// Per sequence: x = next(); len = len_min + x % (len_max - len_min + 1);
// gc = gc_min + (gc_max - gc_min) * u; then len nucleotides with
// P(A) = P(U) = (1-gc)/2, P(C) = P(G) = gc/2, cumulative order A, C, G, U.
// Sequence k is written to out_seqs + k * stride (NUL-terminated);
// out_lens[k] receives its length.  Returns 0, or -1 if stride is too small.
int lcg_synth_sequences(uint64_t seed, int count, int len_min, int len_max,
                        double gc_min, double gc_max, char *out_seqs,
                        int stride, int *out_lens) {
    if (len_max < len_min || stride <= len_max) return -1;
    SplitMix64 rng(seed);
    for (int k = 0; k < count; ++k) {
        uint64_t x = rng.next();
        int len = len_min + (int)(x % (uint64_t)(len_max - len_min + 1));
        double gc = gc_min + (gc_max - gc_min) * rng.uniform();
        char *s = out_seqs + (size_t)k * stride;
        for (int i = 0; i < len; ++i) {
            double u = rng.uniform();
            char c;
            if (u < (1.0 - gc) / 2) c = 'A';
            else if (u < 0.5) c = 'C';
            else if (u < (1.0 + gc) / 2) c = 'G';
            else c = 'U';
            s[i] = c;
        }
        s[len] = 0;
        out_lens[k] = len;
    }
    return 0;
}

// McCaskill-style base-pair probabilities of `seq` (ACGU, length len) with
// base pairs restricted to span j-i+1 <= max_span (max_span <= 0: no limit).
// Writes every pair with probability > cutoff, in order i ascending, j
// ascending (1-based), into out_i/out_j/out_p (capacity cap).  Returns the
// number of pairs written, -1 on bad input, -2 if cap is too small.
int lcg_synth_bpp(const char *seq, int len, int max_span, double cutoff,
                  int *out_i, int *out_j, double *out_p, int cap) {
    if (len < 0) return -1;
    int n = len;
    if (max_span <= 0 || max_span > n) max_span = n;
    std::string s(" ");
    s.append(seq, seq + n);
    for (int i = 1; i <= n; ++i) {
        char c = s[i];
        if (c == 'T') c = 'U';
        if (c >= 'a' && c <= 'z') c = c - 'a' + 'A';
        s[i] = c;
    }
    auto ptype = [&](int i, int j) -> int {
        if (j - i - 1 < 3 || j - i + 1 > max_span) return 0;
        return pair_type(s[i], s[j]);
    };

    Tri Qb(n), Qm(n), Qm1(n);
    const double ml_close = bf(ML_CLOSING + ML_BRANCH);

    // interior-loop Boltzmann factor for (i,j) closing (k,l)
    auto loop_bf = [&](int i, int j, int k, int l) -> double {
        int to = ptype(i, j), ti = pair_type(s[k], s[l]);
        int l1 = k - i - 1, l2 = j - l - 1;
        double e;
        if (l1 == 0 && l2 == 0) e = stack_energy(s[i], s[k], s[l], s[j]);
        else if (l1 == 0 || l2 == 0)
            e = bulge_energy(l1 + l2, to, ti, stack_energy(s[i], s[k], s[l], s[j]));
        else e = interior_energy(l1, l2, to, ti);
        return bf(e);
    };
    auto branch_bf = [&](int i, int j) -> double {
        return bf(ML_BRANCH + terminal_penalty(pair_type(s[i], s[j])));
    };

    // ---------------- inside
    for (int d = 4; d <= n - 1; ++d) {
        for (int i = 1; i + d <= n; ++i) {
            int j = i + d;
            int t = ptype(i, j);
            if (t) {
                double q = bf(hairpin_energy(j - i - 1, t));
                for (int k = i + 1; k <= i + MAX_LOOP + 1 && k < j - 4; ++k) {
                    int l1 = k - i - 1;
                    for (int l = j - 1; l > k + 3; --l) {
                        int l2 = j - l - 1;
                        if (l1 + l2 > MAX_LOOP) break;
                        if (!ptype(k, l)) continue;
                        q += loop_bf(i, j, k, l) * Qb(k, l);
                    }
                }
                // multiloop: Qm(i+1,u-1) * Qm1(u,j-1)
                double mq = 0;
                for (int u = i + 6; u <= j - 5; ++u) mq += Qm(i + 1, u - 1) * Qm1(u, j - 1);
                q += mq * ml_close * bf(terminal_penalty(t));
                Qb(i, j) = q;
            }
            if (d < max_span) {
                // Qm1(i,j) = sum_l Qb(i,l) * branch(i,l)   (unpaired cost 0)
                double m1 = 0;
                for (int l = i + 4; l <= j; ++l)
                    if (Qb(i, l) != 0) m1 += Qb(i, l) * branch_bf(i, l);
                Qm1(i, j) = m1;
                // Qm(i,j) = sum_u (1 + Qm(i,u-1)) * Qm1(u,j)
                double m = 0;
                for (int u = i; u + 4 <= j; ++u)
                    m += (1.0 + (u - 1 >= i ? Qm(i, u - 1) : 0.0)) * Qm1(u, j);
                Qm(i, j) = m;
            }
        }
    }
    std::vector<double> Z5(n + 1, 0.0), Z5b(n + 1, 0.0);
    Z5[0] = 1.0;
    auto ext_bf = [&](int k, int j) { return bf(terminal_penalty(pair_type(s[k], s[j]))); };
    for (int j = 1; j <= n; ++j) {
        double z = Z5[j - 1];
        for (int k = (j - max_span + 1 > 1 ? j - max_span + 1 : 1); k + 4 <= j; ++k)
            if (Qb(k, j) != 0) z += Z5[k - 1] * Qb(k, j) * ext_bf(k, j);
        Z5[j] = z;
    }
    const double Z = Z5[n];

    // ---------------- outside (adjoints of the inside assignments)
    Tri Qbb(n), Qmb(n), Qm1b(n);
    Z5b[n] = 1.0;
    for (int j = n; j >= 1; --j) {
        double g = Z5b[j];
        Z5b[j - 1] += g;
        for (int k = (j - max_span + 1 > 1 ? j - max_span + 1 : 1); k + 4 <= j; ++k)
            if (Qb(k, j) != 0) {
                double e = ext_bf(k, j);
                Z5b[k - 1] += g * Qb(k, j) * e;
                Qbb(k, j) += g * Z5[k - 1] * e;
            }
    }
    for (int d = n - 1; d >= 4; --d) {
        for (int i = 1; i + d <= n; ++i) {
            int j = i + d;
            if (d < max_span) {
                double gm = Qmb(i, j);
                if (gm != 0) {
                    for (int u = i; u + 4 <= j; ++u) {
                        double left = 1.0 + (u - 1 >= i ? Qm(i, u - 1) : 0.0);
                        Qm1b(u, j) += gm * left;
                        if (u - 1 >= i) Qmb(i, u - 1) += gm * Qm1(u, j);
                    }
                }
                double gm1 = Qm1b(i, j);
                if (gm1 != 0) {
                    for (int l = i + 4; l <= j; ++l)
                        if (Qb(i, l) != 0) Qbb(i, l) += gm1 * branch_bf(i, l);
                }
            }
            int t = ptype(i, j);
            if (t && Qb(i, j) != 0) {
                double g = Qbb(i, j);
                if (g == 0) continue;
                for (int k = i + 1; k <= i + MAX_LOOP + 1 && k < j - 4; ++k) {
                    int l1 = k - i - 1;
                    for (int l = j - 1; l > k + 3; --l) {
                        int l2 = j - l - 1;
                        if (l1 + l2 > MAX_LOOP) break;
                        if (!ptype(k, l)) continue;
                        Qbb(k, l) += g * loop_bf(i, j, k, l);
                    }
                }
                double c = g * ml_close * bf(terminal_penalty(t));
                for (int u = i + 6; u <= j - 5; ++u) {
                    Qmb(i + 1, u - 1) += c * Qm1(u, j - 1);
                    Qm1b(u, j - 1) += c * Qm(i + 1, u - 1);
                }
            }
        }
    }

    int cnt = 0;
    for (int i = 1; i <= n; ++i) {
        for (int j = i + 4; j <= n; ++j) {
            if (!ptype(i, j) || Qb(i, j) == 0) continue;
            double p = Qb(i, j) * Qbb(i, j) / Z;
            if (p > cutoff) {
                if (cnt >= cap) return -2;
                out_i[cnt] = i;
                out_j[cnt] = j;
                out_p[cnt] = p > 1.0 ? 1.0 : p;
                ++cnt;
            }
        }
    }
    return cnt;
}

}  // extern "C"
