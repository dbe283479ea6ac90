// The folding energy model: loop energies of a reduced nearest-neighbour
// model (kcal/mol at 37 C) and their Boltzmann factors.
//
// LocARNA takes base-pair probabilities from ViennaRNA's McCaskill
// implementation (rna_ensemble.cc:141-193: vrna_mfe, vrna_exp_params_rescale,
// vrna_pf) with the Turner 2004 parameter set.  Neither is available here;
// this model keeps the loop decomposition of that energy model (stacks,
// hairpins, bulges, interior loops up to 30 unpaired bases, a linear
// multiloop and terminal AU/GU penalties, canonical pairs AU, CG, GU,
// hairpins of at least 3 bases) with a reduced parameter set: the Watson-Crick
// stacks are the Turner 2004 values, every other term a size-dependent
// simplification without sequence-dependent mismatch, tetraloop or dangle
// terms.  Energies of a structure are the sum over its loops; the exterior
// loop adds only the terminal penalties of its pairs.
//
// Used by the seeded input generator (synth.cpp) and by the folding kernels'
// factor tables (fold.cpp), so that both compute the same ensemble.
#ifndef LOCARNA_AMD_FOLD_MODEL_H
#define LOCARNA_AMD_FOLD_MODEL_H

#include <cmath>
#include <cstdlib>
#include <cstring>

namespace locarna_amd {
namespace fold_model {

const double kT = 0.0019872 * 310.15;  // kcal/mol at 37 C

// pair types: 0 none, 1 AU, 2 CG, 3 GC, 4 UA, 5 GU, 6 UG
inline int pair_type(char a, char b) {
    if (a == 'A' && b == 'U') return 1;
    if (a == 'C' && b == 'G') return 2;
    if (a == 'G' && b == 'C') return 3;
    if (a == 'U' && b == 'A') return 4;
    if (a == 'G' && b == 'U') return 5;
    if (a == 'U' && b == 'G') return 6;
    return 0;
}

inline bool is_weak(int t) { return t == 1 || t == 4 || t == 5 || t == 6; }
inline bool is_gu(int t) { return t == 5 || t == 6; }
inline double terminal_penalty(int t) { return is_weak(t) ? 0.5 : 0.0; }

// Watson-Crick stacks, Turner 2004 (kcal/mol); key: top strand 5'XY3',
// bottom strand 3'X'Y'5' (X pairs X', Y pairs Y').
inline double wc_stack(char X, char Y, char Yp, char Xp) {
    struct Entry {
        const char *k;
        double e;
    };
    static const Entry tab[] = {{"AA/UU", -0.93}, {"AU/UA", -1.10}, {"UA/AU", -1.33}, {"CU/GA", -2.08},
                                {"CA/GU", -2.11}, {"GU/CA", -2.24}, {"GA/CU", -2.35}, {"CG/GC", -2.36},
                                {"GG/CC", -3.26}, {"GC/CG", -3.42}};
    char key1[6] = {X, Y, '/', Xp, Yp, 0};
    char key2[6] = {Yp, Xp, '/', Y, X, 0};  // same stack read from the other strand
    for (const auto &t : tab) {
        if (!std::strcmp(t.k, key1) || !std::strcmp(t.k, key2)) return t.e;
    }
    return -1.0;
}

// stack of the outer pair (X, X') on the inner pair (Y, Y')
inline double stack_energy(char X, char Y, char Yp, char Xp) {
    int to = pair_type(X, Xp), ti = pair_type(Y, Yp);
    if (!is_gu(to) && !is_gu(ti)) return wc_stack(X, Y, Yp, Xp);
    if (is_gu(to) && is_gu(ti)) return -0.5;
    int other = is_gu(to) ? ti : to;
    return (other == 2 || other == 3) ? -1.4 : -1.0;
}

inline double hairpin_energy(int L, int t) {
    static const double init[] = {0, 0, 0, 5.4, 5.6, 5.7, 5.4, 6.0, 6.1, 6.2};
    double e = (L <= 9) ? init[L] : 6.2 + 1.75 * kT * std::log(L / 9.0);
    return e + terminal_penalty(t);
}

inline double bulge_energy(int L, int to, int ti, double stack_e) {
    if (L == 1) return 3.8 + stack_e;
    static const double init[] = {0, 0, 2.8, 3.2, 3.6, 4.0, 4.4};
    double e = (L <= 6) ? init[L] : 4.4 + 1.75 * kT * std::log(L / 6.0);
    return e + terminal_penalty(to) + terminal_penalty(ti);
}

inline double interior_energy(int l1, int l2, int to, int ti) {
    int n = l1 + l2;
    static const double init[] = {0, 0, 0.5, 1.6, 1.1, 2.0, 2.0};
    double e = (n <= 6) ? init[n] : 2.0 + 1.08 * std::log(n / 6.0);
    double asym = 0.6 * std::abs(l1 - l2);
    if (asym > 3.0) asym = 3.0;
    e += asym;
    if (is_weak(to)) e += 0.7;
    if (is_weak(ti)) e += 0.7;
    return e;
}

const double ML_CLOSING = 3.4;
const double ML_BRANCH = 0.4;
const int MAX_LOOP = 30;

inline double bf(double e) { return std::exp(-e / kT); }

// the nucleotides of a pair type (inverse of pair_type)
inline void pair_bases(int t, char &a, char &b) {
    static const char L[] = "?ACGUGU", R[] = "?UGCAUG";
    a = L[t];
    b = R[t];
}

}  // namespace fold_model
}  // namespace locarna_amd

#endif
