/* CPU oracle for the base-pair probability step: test infrastructure only
 * (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg), never
 * linked by the product.
 *
 * The reference folds with ViennaRNA 2.x (RnaEnsemble,
 * src/LocARNA/rna_ensemble.cc:141-193: vrna_mfe, vrna_exp_params_rescale,
 * vrna_pf; Turner 2004 parameters, dangles 2), a dependency absent here,
 * as are its parameter files.  ViennaRNA's published algorithm is
 * McCaskill's (J. S. McCaskill, Biopolymers 29:1105-1119, 1990): inside
 * partition functions of the loop decomposition, outside partition
 * functions, P(i,j) = Qb(i,j) Qb^(i,j) / Z.  This file restates that
 * algorithm over the library's reduced energy model (stated below); the
 * joint probability of (i,j) and (i+1,j-1) follows
 * RnaEnsembleImpl::arc_2_prob_noali (rna_ensemble.cc:415-436).
 *
 * Pinning: Z and every P(i,j) and P2(i,j) equal the sums over all secondary
 * structures, enumerated with an independent loop-energy evaluator, on
 * small sequences (tests/test_oracle_fold.py), with span limits; the
 * probabilities obey sum_j P(i,j) <= 1 and P2 <= min(P(i,j), P(i+1,j-1)).
 * Parity unpinned: the energy model is this library's, not the Turner 2004
 * set ViennaRNA uses, so the probabilities are not RNAfold's.
 *
 * Energy model (kcal/mol, 37 C, kT = 0.0019872 * 310.15): canonical pairs
 * AU, CG, GU with j - i >= 4 and j - i + 1 <= max span; stack: Turner 2004
 * Watson-Crick stacks, -0.5 for GU on GU, -1.4 for GU on CG, -1.0 for GU on
 * AU; hairpin of L unpaired: 5.4 5.6 5.7 5.4 6.0 6.1 6.2 for L = 3..9, else
 * 6.2 + 1.75 kT ln(L/9), plus the terminal penalty (0.5 for AU and GU
 * pairs); bulge of L: 3.8 + stack for L = 1, 2.8 3.2 3.6 4.0 4.4 for
 * L = 2..6, else 4.4 + 1.75 kT ln(L/6), plus both terminal penalties;
 * interior l1 x l2 (both > 0, l1 + l2 <= 30): 0.5 1.6 1.1 2.0 2.0 for
 * l1 + l2 = 2..6, else 2.0 + 1.08 ln((l1+l2)/6), plus min(3, 0.6 |l1 - l2|)
 * and 0.7 per AU or GU pair; multiloop: 3.4 + 0.4 per pair (the closing one
 * included) plus the terminal penalties of its pairs; exterior loop: the
 * terminal penalties of its pairs.
 */
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

const double KT = 0.0019872 * 310.15;

int ptype_of(char a, char b) {
    const std::string p = std::string(1, a) + b;
    static const char *names[] = {"AU", "CG", "GC", "UA", "GU", "UG"};
    for (int t = 0; t < 6; ++t)
        if (p == names[t]) return t + 1;
    return 0;
}
bool weak_pair(int t) { return t == 1 || t == 4 || t == 5 || t == 6; }
bool gu_pair(int t) { return t == 5 || t == 6; }
double term_pen(int t) { return weak_pair(t) ? 0.5 : 0.0; }

// outer pair (x, xr) closing the inner pair (y, yr): 5' x y 3' / 3' xr yr 5'
double stack_dg(char x, char y, char yr, char xr) {
    const int to = ptype_of(x, xr), ti = ptype_of(y, yr);
    if (gu_pair(to) && gu_pair(ti)) return -0.5;
    if (gu_pair(to) || gu_pair(ti)) {
        const int o = gu_pair(to) ? ti : to;
        return (o == 2 || o == 3) ? -1.4 : -1.0;
    }
    // Turner 2004 Watson-Crick stacks, 5'AB3'/3'CD5' as "ABCD"
    static const struct {
        const char *k;
        double g;
    } wc[] = {{"AAUU", -0.93}, {"AUUA", -1.10}, {"UAAU", -1.33}, {"CUGA", -2.08}, {"CAGU", -2.11},
              {"GUCA", -2.24}, {"GACU", -2.35}, {"CGGC", -2.36}, {"GGCC", -3.26}, {"GCCG", -3.42}};
    const char a[5] = {x, y, xr, yr, 0}, b[5] = {yr, xr, y, x, 0};
    for (const auto &e : wc)
        if (!std::strcmp(e.k, a) || !std::strcmp(e.k, b)) return e.g;
    std::abort();
}

double hairpin_dg(int L, int t) {
    static const double h[10] = {0, 0, 0, 5.4, 5.6, 5.7, 5.4, 6.0, 6.1, 6.2};
    return (L <= 9 ? h[L] : 6.2 + 1.75 * KT * std::log(L / 9.0)) + term_pen(t);
}

// loop closed by (i,j) with the inner pair (k,l) (1-based positions of s)
double inner_loop_dg(const std::string &s, int i, int j, int k, int l) {
    const int to = ptype_of(s[i], s[j]), ti = ptype_of(s[k], s[l]);
    const int l1 = k - i - 1, l2 = j - l - 1, L = l1 + l2;
    if (L == 0) return stack_dg(s[i], s[k], s[l], s[j]);
    if (l1 == 0 || l2 == 0) {
        if (L == 1) return 3.8 + stack_dg(s[i], s[k], s[l], s[j]);
        static const double b[7] = {0, 0, 2.8, 3.2, 3.6, 4.0, 4.4};
        return (L <= 6 ? b[L] : 4.4 + 1.75 * KT * std::log(L / 6.0)) + term_pen(to) + term_pen(ti);
    }
    static const double in[7] = {0, 0, 0.5, 1.6, 1.1, 2.0, 2.0};
    double g = L <= 6 ? in[L] : 2.0 + 1.08 * std::log(L / 6.0);
    g += std::min(3.0, 0.6 * std::abs(l1 - l2));
    if (weak_pair(to)) g += 0.7;
    if (weak_pair(ti)) g += 0.7;
    return g;
}

double boltz(double g) { return std::exp(-g / KT); }

}  // namespace

extern "C" int oracle_fold(const char *seq, int max_span, double *P, double *P2, double *log_z) {
    if (!seq) return -1;
    std::string s = " ";
    for (const char *c = seq; *c; ++c) {
        char x = (char)std::toupper((unsigned char)*c);
        s += x == 'T' ? 'U' : x;
    }
    const int n = (int)s.size() - 1;
    const int span = max_span < 0 ? n : max_span;
    const int N = n + 2;
    auto pairable = [&](int i, int j) { return j - i >= 4 && j - i + 1 <= span ? ptype_of(s[i], s[j]) : 0; };
    const double ML_A = 3.4, ML_B = 0.4;
    const double mlc = boltz(ML_A + ML_B);
    std::vector<double> Qb(N * N, 0.0), Qm(N * N, 0.0), Qm1(N * N, 0.0);
    std::vector<double> Qbb(N * N, 0.0), Qmb(N * N, 0.0), Qm1b(N * N, 0.0);
#define AT(M, i, j) M[(size_t)(i) * N + (j)]
    // inside (McCaskill 1990, eq. for Qb, Qm, Qm1)
    for (int d = 4; d < n; ++d)
        for (int i = 1; i + d <= n; ++i) {
            const int j = i + d;
            const int t = pairable(i, j);
            if (t) {
                double q = boltz(hairpin_dg(j - i - 1, t));
                for (int k = i + 1; k < j && k - i - 1 <= 30; ++k)
                    for (int l = j - 1; l > k && (k - i - 1) + (j - l - 1) <= 30; --l) {
                        if (!pairable(k, l)) continue;
                        q += boltz(inner_loop_dg(s, i, j, k, l)) * AT(Qb, k, l);
                    }
                double m = 0;
                for (int u = i + 6; u <= j - 5; ++u) m += AT(Qm, i + 1, u - 1) * AT(Qm1, u, j - 1);
                q += m * mlc * boltz(term_pen(t));
                AT(Qb, i, j) = q;
            }
            if (d + 1 <= span) {
                double m1 = 0;
                for (int l = i + 4; l <= j; ++l) {
                    const int tl = pairable(i, l);
                    if (tl) m1 += AT(Qb, i, l) * boltz(ML_B + term_pen(tl));
                }
                AT(Qm1, i, j) = m1;
                double m = 0;
                for (int u = i; u <= j - 4; ++u) m += (1.0 + (u > i ? AT(Qm, i, u - 1) : 0.0)) * AT(Qm1, u, j);
                AT(Qm, i, j) = m;
            }
        }
    std::vector<double> Z5(n + 1, 0.0), Z5b(n + 1, 0.0);
    Z5[0] = 1.0;
    for (int j = 1; j <= n; ++j) {
        double z = Z5[j - 1];
        for (int k = 1; k + 4 <= j; ++k) {
            const int t = pairable(k, j);
            if (t) z += Z5[k - 1] * AT(Qb, k, j) * boltz(term_pen(t));
        }
        Z5[j] = z;
    }
    const double Z = Z5[n];
    // outside: derivative of Z with respect to every inside value, gathered
    // over the terms that use it
    Z5b[n] = 1.0;
    for (int m = n - 1; m >= 0; --m) {
        double g = Z5b[m + 1];
        for (int j = m + 5; j <= n; ++j) {
            const int t = pairable(m + 1, j);
            if (t) g += Z5b[j] * AT(Qb, m + 1, j) * boltz(term_pen(t));
        }
        Z5b[m] = g;
    }
    for (int d = n - 1; d >= 4; --d)
        for (int a = 1; a + d <= n; ++a) {
            const int e = a + d;
            if (d + 1 <= span) {
                double qmb = 0;
                // Qm(a,e) = Qm(i+1,u-1) in Qb(a-1,j), u = e+1
                for (int j = e + 6; j <= n && a >= 2; ++j) {
                    const int t = pairable(a - 1, j);
                    if (t) qmb += AT(Qbb, a - 1, j) * mlc * boltz(term_pen(t)) * AT(Qm1, e + 1, j - 1);
                }
                // Qm(a,e) = Qm(i,u-1) in Qm(a,j), u = e+1
                for (int j = e + 5; j <= n && j - a + 1 <= span; ++j) qmb += AT(Qmb, a, j) * AT(Qm1, e + 1, j);
                AT(Qmb, a, e) = qmb;
                double qm1b = 0;
                // Qm1(a,e) = Qm1(u,j-1) in Qb(i,e+1), u = a
                for (int i = 1; i <= a - 6 && e + 1 <= n; ++i) {
                    const int t = pairable(i, e + 1);
                    if (t) qm1b += AT(Qbb, i, e + 1) * mlc * boltz(term_pen(t)) * AT(Qm, i + 1, a - 1);
                }
                // Qm1(a,e) = Qm1(u,j) in Qm(i,e), u = a
                for (int i = 1; i <= a; ++i)
                    if (e - i + 1 <= span) qm1b += AT(Qmb, i, e) * (1.0 + (i < a ? AT(Qm, i, a - 1) : 0.0));
                AT(Qm1b, a, e) = qm1b;
            }
            const int t = pairable(a, e);
            if (!t) continue;
            double qbb = Z5b[e] * Z5[a - 1] * boltz(term_pen(t));
            for (int i = a - 1; i >= 1 && a - i - 1 <= 30; --i)
                for (int j = e + 1; j <= n && (a - i - 1) + (j - e - 1) <= 30; ++j) {
                    if (!pairable(i, j)) continue;
                    qbb += AT(Qbb, i, j) * boltz(inner_loop_dg(s, i, j, a, e));
                }
            double br = 0;
            for (int j = e; j <= n && j - a + 1 <= span; ++j) br += AT(Qm1b, a, j);
            qbb += br * boltz(ML_B + term_pen(t));
            AT(Qbb, a, e) = qbb;
        }
    const size_t M = (size_t)(n + 1) * (n + 1);
    if (P) std::memset(P, 0, M * sizeof(double));
    if (P2) std::memset(P2, 0, M * sizeof(double));
    for (int i = 1; i <= n; ++i)
        for (int j = i + 4; j <= n; ++j) {
            if (!pairable(i, j)) continue;
            if (P) P[(size_t)i * (n + 1) + j] = AT(Qb, i, j) * AT(Qbb, i, j) / Z;
            if (P2 && j - i >= 6 && pairable(i + 1, j - 1))
                P2[(size_t)i * (n + 1) + j] =
                    AT(Qbb, i, j) * boltz(inner_loop_dg(s, i, j, i + 1, j - 1)) * AT(Qb, i + 1, j - 1) / Z;
        }
#undef AT
    if (log_z) *log_z = std::log(Z);
    return n;
}
