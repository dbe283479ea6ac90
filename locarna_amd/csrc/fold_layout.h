// Device layout of the base-pair probability step (McCaskill inside and
// outside over the energy model of fold_model.h), shared by fold.cpp and
// fold_kernels.hip.
//
// One workgroup folds one sequence.  Every matrix of a sequence is a band:
// entry (i, j), 1 <= i <= n, 0 <= j - i < W, at i * W + (j - i), with
// W = min(n, max span) (a pair (i, j) needs j - i + 1 <= max span).  Six
// bands hold the inside values Qb (pair (i, j) closes the segment), Qm (at
// least one multiloop branch in i..j), Qm1 (exactly one branch, starting at
// i) and their outside values; two more, in a separate output arena, hold the
// pair probabilities P(i, j) and the joint probabilities P2(i, j) of (i, j)
// and (i + 1, j - 1).
#ifndef LOCARNA_AMD_FOLD_LAYOUT_H
#define LOCARNA_AMD_FOLD_LAYOUT_H

#define LCG_FOLD_MAXLOOP 30
// (l1, l2) with l1 + l2 <= 30, ordered by l1 + l2
#define LCG_FOLD_NLOOP 496
// factor table (doubles): Boltzmann factors of the loop terms
#define LCG_FOLD_TAB_LOOP 0       // [l1 * 31 + l2][weak(outer) * 2 + weak(inner)], l1 + l2 >= 2
#define LCG_FOLD_TAB_STACK 3844   // [outer type * 8 + inner type], l1 = l2 = 0
#define LCG_FOLD_TAB_BULGE1 3908  // [outer type * 8 + inner type], l1 + l2 = 1
#define LCG_FOLD_TAB_EXT 3972     // [weak]: a pair in the exterior loop
#define LCG_FOLD_TAB_BRANCH 3974  // [weak]: a branch of a multiloop
#define LCG_FOLD_TAB_TP 3976      // [weak]: the closing pair's terminal penalty
#define LCG_FOLD_TAB_MLC 3978     // multiloop closing (closing + one branch)
#define LCG_FOLD_TAB_FIXED 3980
#define LCG_FOLD_TAB_HP 3980      // [L * 2 + weak]: hairpin of L unpaired bases
#define LCG_FOLD_MAX_LEN 4096     // longest sequence (Z5, Z5b and the scale powers live in LDS)
#define LCG_FOLD_THREADS 1024

struct FoldDesc {
    int n;           // sequence length
    int W;           // band width
    long long codes; // offset of the codes: position i at codes + i (0 A, 1 C, 2 G, 3 U, 4 other)
    long long mat;   // offset (doubles) of the six bands Qb, Qm, Qm1, Qbb, Qmb, Qm1b
    long long out;   // offset (doubles) of the two bands P, P2
    double log_sigma; // scale per base: values of L bases are stored times exp(-L log_sigma)
};

#endif
