"""Pin of the band step's Gotoh partition function (PFGotoh + PFTraceProbs,
edge_probs.icc:7-162, :212-268) by enumeration: for tiny pairs every
alignment is listed with its Boltzmann weight (match steps
exp(sigma(i, j) / T), a gap run of length k w_open * w_ext^k, a switch
between the two gap directions opening a new run), and the probability that
an alignment passes through cell (i, j) is compared with the oracle's trace
probability.  Global: paths from (0, 0) to (n, m) through every point they
reach.  Sequence-local, as the reference combines its local forward and reverse
recursions: the value of cell (i, j) is the sum over prefixes ending in
(i, j) with last step X and suffixes leaving (i, j) with first step Y of
their weights (divided by w_open when X = Y is the same gap direction, the
run continuing), plus the prefixes ending with a match alone; prefixes start
with a match step anywhere or at (0, 0) with a gap step (the recursion
initializes the gap states of row 0 and column 0 in both modes,
edge_probs.icc:82-95), suffixes end with a match step anywhere or with a
gap step into (n, m); the normalization is 1 + the prefixes that end with a
match anywhere.  The similarity sigma(i, j) itself (StralScore) is the oracle's;
what is pinned is the recursion, its reverse pass and their combination.
Runs on the CPU."""
import math

import pytest

from oracle import oracle
from tiny import tiny_pair


def _paths(n, m):
    """Every global alignment as its list of steps ('M', 'A', 'B')."""
    out = []

    def walk(i, j, steps):
        if i == n and j == m:
            out.append(list(steps))
        for s, di, dj in (("M", 1, 1), ("A", 1, 0), ("B", 0, 1)):
            if i + di <= n and j + dj <= m:
                steps.append(s)
                walk(i + di, j + dj, steps)
                steps.pop()

    walk(0, 0, [])
    return out


def _brute(S, w, n, m):
    """Global trace probabilities: weight of the alignments through (i, j)."""
    w_open, w_ext, T = w
    P = [[0.0] * (m + 1) for _ in range(n + 1)]
    Z = 0.0
    for steps in _paths(n, m):
        i = j = 0
        weight = 1.0
        visited = [(0, 0)]
        prev = None
        for s in steps:
            if s == "M":
                i, j = i + 1, j + 1
                weight *= math.exp(S[i][j] / T)
            elif s == "A":
                i += 1
                weight *= w_ext * (w_open if prev != "A" else 1.0)
            else:
                j += 1
                weight *= w_ext * (w_open if prev != "B" else 1.0)
            prev = s
            visited.append((i, j))
        Z += weight
        for (a, b) in visited:
            P[a][b] += weight
    return [[x / Z for x in row] for row in P]


def _local_brute(S, w, n, m):
    """Trace probabilities of the local recursion from enumerated prefixes
    and suffixes."""
    w_open, w_ext, T = w
    steps3 = (("M", 1, 1), ("A", 1, 0), ("B", 0, 1))

    def step_weight(s, prev, i, j):  # weight of step s arriving at (i, j)
        if s == "M":
            return math.exp(S[i][j] / T)
        return w_ext * (w_open if prev != s else 1.0)

    F = {}  # (i, j, X) -> summed weight of prefixes ending in (i, j) with step X

    def grow(i, j, last, weight):
        F[(i, j, last)] = F.get((i, j, last), 0.0) + weight
        for s, di, dj in steps3:
            if i + di <= n and j + dj <= m:
                grow(i + di, j + dj, s, weight * step_weight(s, last, i + di, j + dj))

    for i0 in range(n):
        for j0 in range(m):
            grow(i0 + 1, j0 + 1, "M", step_weight("M", None, i0 + 1, j0 + 1))
    if n > 0:
        grow(1, 0, "A", step_weight("A", None, 1, 0))
    if m > 0:
        grow(0, 1, "B", step_weight("B", None, 0, 1))

    def suffixes(i, j):
        """{first step: summed weight} of suffixes from (i, j)"""
        out = {}

        def walk(a, b, first, last, weight):
            for s, di, dj in steps3:
                if a + di > n or b + dj > m:
                    continue
                wgt = weight * step_weight(s, last, a + di, b + dj)
                f = first or s
                if s == "M" or (a + di == n and b + dj == m):
                    out[f] = out.get(f, 0.0) + wgt
                walk(a + di, b + dj, f, s, wgt)

        walk(i, j, None, None, 1.0)
        return out

    Z = 1.0 + sum(v for (i, j, x), v in F.items() if x == "M")
    P = [[0.0] * (m + 1) for _ in range(n + 1)]
    for i in range(n + 1):
        for j in range(m + 1):
            B = suffixes(i, j)
            z = 0.0
            for X in ("M", "A", "B"):
                fx = F.get((i, j, X), 0.0)
                if fx == 0.0:
                    continue
                tail = sum(v / w_open if (Y == X and X != "M") else v for Y, v in B.items())
                z += fx * (tail + (1.0 if X == "M" else 0.0))
            P[i][j] = z / Z
    return P


@pytest.mark.parametrize("local", [0, 1])
@pytest.mark.parametrize("extended", [1, 0])
def test_trace_probs_match_enumeration(local, extended):
    for seed in range(6):
        a, b = tiny_pair(9100 + seed, nA=4 + seed % 3, nB=4 + (seed + 1) % 3)
        params = oracle.locarna_params(min_prob=0.01, max_bp_span=-1, sequ_local=local, band_extended=extended,
                                       indel=-150, indel_opening=-200)
        P, S, w = oracle.trace_probs(a, b, params)
        n, m = len(P) - 1, len(P[0]) - 1
        ref = _local_brute(S, w, n, m) if local else _brute(S, w, n, m)
        for i in range(n + 1):
            for j in range(m + 1):
                assert abs(P[i][j] - ref[i][j]) <= 1e-12, (seed, local, i, j, P[i][j], ref[i][j])


def test_enumeration_catches_a_wrong_gap_rule():
    # the enumeration distinguishes affine from linear gap weights
    a, b = tiny_pair(9100, nA=5, nB=5)
    params = oracle.locarna_params(min_prob=0.01, max_bp_span=-1, indel=-150, indel_opening=-200)
    P, S, w = oracle.trace_probs(a, b, params)
    n, m = len(P) - 1, len(P[0]) - 1
    linear = _brute(S, (1.0, w[1] * w[0] ** 0.5, w[2]), n, m)
    assert max(abs(P[i][j] - linear[i][j]) for i in range(n + 1) for j in range(m + 1)) > 1e-6


def _prefix_sums(S, w, n, m, fe):
    """Forward quantities of the band's Gotoh recursion with free end gaps,
    by enumeration of prefix paths: F[(i, j, X)] sums the weights of paths
    ending in (i, j) with step X.  A path starts at (0, 0) (weight 1, state
    M) or, with the second sequence's left end free, at (i, 0) in state A
    with weight 1 (i >= 1), or, with the first's, at (0, j) in state B; it
    continues down column 0 / along row 0 only from (0, 0)."""
    w_open, w_ext, T = w
    F = {}

    def walk(i, j, last, weight, chain):
        F[(i, j, last)] = F.get((i, j, last), 0.0) + weight
        if i < n and j < m:
            walk(i + 1, j + 1, "M", weight * math.exp(S[i + 1][j + 1] / T), False)
        if i < n and (j >= 1 or chain):
            walk(i + 1, j, "A", weight * w_ext * (w_open if last != "A" else 1.0), chain and j == 0)
        if j < m and (i >= 1 or chain):
            walk(i, j + 1, "B", weight * w_ext * (w_open if last != "B" else 1.0), chain and i == 0)

    walk(0, 0, "M", 1.0, True)
    if fe & 4:
        for i in range(1, n + 1):
            walk(i, 0, "A", 1.0, False)
    if fe & 1:
        for j in range(1, m + 1):
            walk(0, j, "B", 1.0, False)
    return F


def _free_probs(S, w, n, m, fe):
    w_open = w[0]
    F = _prefix_sums(S, w, n, m, fe)
    # reversed sequences: similarity mirrored, left and right ends swapped
    Sr = [[0.0] * (m + 1) for _ in range(n + 1)]
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            Sr[i][j] = S[n + 1 - i][m + 1 - j]
    rev = ((fe & 2) >> 1) | ((fe & 1) << 1) | ((fe & 8) >> 1) | ((fe & 4) << 1)
    R = _prefix_sums(Sr, w, n, m, rev)
    f = lambda d, i, j, x: d.get((i, j, x), 0.0)
    Z = sum(f(F, n, m, x) for x in "MAB")
    if fe & 4:
        Z += sum(f(F, i, m, "A") for i in range(n + 1))
    if fe & 1:
        Z += sum(f(F, n, j, "B") for j in range(m + 1))
    P = [[0.0] * (m + 1) for _ in range(n + 1)]
    for i in range(n + 1):
        for j in range(m + 1):
            rM, rA, rB = (f(R, n - i, m - j, x) for x in "MAB")
            z = f(F, i, j, "M") * (rM + rA + rB) + f(F, i, j, "A") * (rM + rA / w_open + rB) + \
                f(F, i, j, "B") * (rM + rA + rB / w_open)
            P[i][j] = z / Z
    return P


@pytest.mark.parametrize("fe", [1, 2, 4, 8, 5, 10, 15])
def test_trace_probs_free_endgaps_match_enumeration(fe):
    # edge_probs.icc:76-162 with free end gaps: 1 added to A(i, 0) / B(0, j)
    # under the left flags, the reverse pass with the flags of the reversed
    # sequences, and Z with the last column's A and last row's B entries
    # (:39-53)
    for seed in range(4):
        a, b = tiny_pair(9500 + seed, nA=4 + seed % 2, nB=4 + (seed + 1) % 3)
        params = oracle.locarna_params(min_prob=0.01, max_bp_span=-1, free_endgaps=fe, indel=-150,
                                       indel_opening=-200)
        P, S, w = oracle.trace_probs(a, b, params)
        n, m = len(P) - 1, len(P[0]) - 1
        ref = _free_probs(S, w, n, m, fe)
        for i in range(n + 1):
            for j in range(m + 1):
                assert abs(P[i][j] - ref[i][j]) <= 1e-12 * max(1.0, abs(ref[i][j])), (fe, seed, i, j)


def test_free_enumeration_reduces_to_global():
    a, b = tiny_pair(9100, nA=5, nB=4)
    params = oracle.locarna_params(min_prob=0.01, max_bp_span=-1, indel=-150, indel_opening=-200)
    P, S, w = oracle.trace_probs(a, b, params)
    n, m = len(P) - 1, len(P[0]) - 1
    ref = _free_probs(S, w, n, m, 0)
    glob = _brute(S, w, n, m)
    for i in range(n + 1):
        for j in range(m + 1):
            assert abs(ref[i][j] - glob[i][j]) <= 1e-12 and abs(P[i][j] - ref[i][j]) <= 1e-12
