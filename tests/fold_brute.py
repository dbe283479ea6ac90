"""Exhaustive check of the folding partition function on short sequences.

Every secondary structure (canonical pairs, hairpins of at least 3 bases,
pair span j - i + 1 <= max span, interior loops of at most 30 unpaired
bases) is enumerated and scored loop by loop with an evaluator written
independently of the recursions; Z, the pair probabilities and the joint
probabilities of stacked pairs are the Boltzmann-weighted sums.  The
energies are the library's reduced nearest-neighbour model (Turner 2004
Watson-Crick stacks; the other loop terms size-dependent).
"""
import math

KT = 0.0019872 * 310.15
PAIRS = {"AU": 1, "CG": 2, "GC": 3, "UA": 4, "GU": 5, "UG": 6}
WC = {"AAUU": -0.93, "AUUA": -1.10, "UAAU": -1.33, "CUGA": -2.08, "CAGU": -2.11,
      "GUCA": -2.24, "GACU": -2.35, "CGGC": -2.36, "GGCC": -3.26, "GCCG": -3.42}


def _weak(a, b):
    return a + b in ("AU", "UA", "GU", "UG")


def _gu(a, b):
    return a + b in ("GU", "UG")


def _tp(a, b):
    return 0.5 if _weak(a, b) else 0.0


def _stack(x, y, yr, xr):
    """outer pair (x, xr) on inner pair (y, yr)"""
    if _gu(x, xr) and _gu(y, yr):
        return -0.5
    if _gu(x, xr) or _gu(y, yr):
        other = (y, yr) if _gu(x, xr) else (x, xr)
        return -1.4 if other[0] + other[1] in ("CG", "GC") else -1.0
    key = x + y + xr + yr
    if key in WC:
        return WC[key]
    return WC[yr + xr + y + x]


def _loop_energy(s, i, j, children):
    a, b = s[i], s[j]
    if not children:
        L = j - i - 1
        h = {3: 5.4, 4: 5.6, 5: 5.7, 6: 5.4, 7: 6.0, 8: 6.1, 9: 6.2}
        return (h[L] if L <= 9 else 6.2 + 1.75 * KT * math.log(L / 9.0)) + _tp(a, b)
    if len(children) == 1:
        k, l = children[0]
        l1, l2 = k - i - 1, j - l - 1
        if l1 + l2 == 0:
            return _stack(a, s[k], s[l], b)
        if l1 == 0 or l2 == 0:
            L = l1 + l2
            if L == 1:
                return 3.8 + _stack(a, s[k], s[l], b)
            g = {2: 2.8, 3: 3.2, 4: 3.6, 5: 4.0, 6: 4.4}.get(L)
            if g is None:
                g = 4.4 + 1.75 * KT * math.log(L / 6.0)
            return g + _tp(a, b) + _tp(s[k], s[l])
        L = l1 + l2
        g = {2: 0.5, 3: 1.6, 4: 1.1, 5: 2.0, 6: 2.0}.get(L)
        if g is None:
            g = 2.0 + 1.08 * math.log(L / 6.0)
        g += min(3.0, 0.6 * abs(l1 - l2))
        g += (0.7 if _weak(a, b) else 0.0) + (0.7 if _weak(s[k], s[l]) else 0.0)
        return g
    g = 3.4 + 0.4 * (len(children) + 1) + _tp(a, b)
    for (k, l) in children:
        g += _tp(s[k], s[l])
    return g


def _children(partner, i, j):
    out, p = [], i + 1
    while p < j:
        q = partner[p]
        if q > p:
            out.append((p, q))
            p = q + 1
        else:
            p += 1
    return out


def structures(s, max_span):
    """All pair lists of the 1-based sequence s[1..n] (s[0] a pad)."""
    n = len(s) - 1
    span = n if max_span < 0 else max_span
    memo = {}

    def rec(i, j):
        if i > j:
            return [[]]
        if (i, j) in memo:
            return memo[(i, j)]
        res = [st for st in rec(i + 1, j)]
        for k in range(i + 4, j + 1):
            if s[i] + s[k] in PAIRS and k - i + 1 <= span:
                for a in rec(i + 1, k - 1):
                    for b in rec(k + 1, j):
                        res.append([(i, k)] + a + b)
        memo[(i, j)] = res
        return res

    return rec(1, n)


def energy(s, pairs):
    """Free energy of a structure, or None if it has an interior loop of more
    than 30 unpaired bases."""
    n = len(s) - 1
    partner = [0] * (n + 2)
    for (i, j) in pairs:
        partner[i], partner[j] = j, i
    g = 0.0
    for (k, l) in _children(partner, 0, n + 1):
        g += _tp(s[k], s[l])
    for (i, j) in pairs:
        ch = _children(partner, i, j)
        if len(ch) == 1 and (ch[0][0] - i - 1) + (j - ch[0][1] - 1) > 30:
            return None
        g += _loop_energy(s, i, j, ch)
    return g


def fold(seq, max_span=-1):
    """(Z, P, P2) by enumeration; P and P2 dicts keyed by (i, j)."""
    s = " " + seq.upper().replace("T", "U")
    Z, P, P2 = 0.0, {}, {}
    for st in structures(s, max_span):
        g = energy(s, st)
        if g is None:
            continue
        w = math.exp(-g / KT)
        Z += w
        ps = set(st)
        for (i, j) in st:
            P[(i, j)] = P.get((i, j), 0.0) + w
            if (i + 1, j - 1) in ps:
                P2[(i, j)] = P2.get((i, j), 0.0) + w
    return Z, {k: v / Z for k, v in P.items()}, {k: v / Z for k, v in P2.items()}
