"""Exhaustive enumeration of LocARNA structure alignments on tiny inputs.

Independent of any DP: every alignment of A and B (a monotone path of match /
deletion / insertion edges through valid cells of the trace band) is
enumerated together with every nested set of matched arc pairs whose end
edges are matches of the path.  The score of a structure alignment is

    sum over match edges that are not arc ends  sigma(i, j)
  + sum over maximal gap runs                   alpha + sum gamma
  + sum over matched arc pairs                  s(am)

(the model of the reference DP, src/LocARNA/aligner.cc:153-233 with the arc
match decomposition of :574-610; a group's leading run along its first row /
column is its init, :264-368).  This pins the oracle's optimum, its
partition function, its arc-match probabilities and its traceback tie rule
(aligner.cc:1082-1343: at each cell, backwards, the first of match >
deletion of length 1, 2, ... > insertion of length 1, 2, ... > arc matches in
(al desc, bl desc) order that continues an optimal alignment).
"""
import math


class Problem:
    def __init__(self, lenA, lenB, sigma, gapA, gapB, alpha, arcmatches, am_score,
                 min_col, max_col, local=False, free=0, no_lonely_pairs=False, am_score_stacked=None,
                 struct_local=False, exclusion=0):
        self.n, self.m = lenA, lenB
        # --struct-local: inside every arc match (not at the top level) the
        # alignment may exclude one contiguous block of A positions and one
        # of B positions; an excluded block costs `exclusion`, pays no gap
        # costs, ends any gap run it touches and takes no edges.  (Modelled
        # for gap costs <= 0 and exclusion <= 0, where excluding a deleted
        # prefix is never worse and an empty block never helps.)
        self.struct_local = struct_local
        self.exclusion = exclusion
        # --stacking: the stacked score of each stackable arc match (None
        # elsewhere); an arc match whose interior is exactly its inner arc
        # match (al+1, ar-1, bl+1, br-1) scores the better of its two scores
        # (aligner.cc:574-610)
        self.am_score_stacked = am_score_stacked
        # --noLP: only structure alignments whose every matched arc pair has
        # its directly inner or directly outer arc pair matched as well
        # (stacks of two or more, aligner.cc:612-657)
        self.no_lonely_pairs = no_lonely_pairs
        # free end gaps (locarna --free-endgaps as bits: 1 left end of the
        # first sequence, 2 its right end, 4 left end of the second, 8 its
        # right end): free leading runs of the top level cost nothing and
        # are not written; with a right end free the alignment may end in
        # the last column (second sequence) or the last row (first)
        self.free = free
        self.sigma = sigma                  # dict (i, j) -> score
        self.gapA, self.gapB = gapA, gapB   # lists indexed 1..n / 1..m
        self.alpha = alpha
        self.min_col, self.max_col = min_col, max_col
        self.local = local
        self.arcmatches = arcmatches        # list of (al, ar, bl, br)
        self.am_score = am_score
        # arc matches by right ends, sorted (al desc, bl desc)
        self.by_right = {}
        for k, (al, ar, bl, br) in enumerate(arcmatches):
            self.by_right.setdefault((ar, br), []).append((al, bl, k))
        for v in self.by_right.values():
            v.sort(key=lambda t: (-t[0], -t[1]))

    def valid(self, i, j):
        return 0 <= i <= self.n and self.min_col[i] <= j <= self.max_col[i]


# An alignment of a group is a tuple of steps, LAST step first:
#   ('M', i, j)          match edge into (i, j)
#   ('D', k, i, j)       k deletions ending at (i, j)
#   ('I', k, i, j)       k insertions ending at (i, j)
#   ('A', rank, k, inner) arc match k (rank in its right-end list) whose
#                        inner alignment is `inner`
#   ('pad', i, j)        the group's init run from its origin to (i, j)

def _live(P, r0, c0, i, j):
    """A cell whose M the group with origin (r0, c0) defines: a band cell,
    or a cell of its init row (columns up to the row's max_col) or init
    column (rows while min_col <= c0), which the reference's init_state
    sets whether or not the band holds them (aligner.cc:264-368).  They
    differ from band cells only for a restricted top level whose origin the
    band excludes."""
    if P.valid(i, j) or (i, j) == (r0, c0):
        return True
    if i == r0:
        return c0 <= j <= P.max_col[r0]
    if j == c0:
        return i > r0 and all(P.min_col[x] <= c0 for x in range(r0 + 1, i + 1))
    return False


def _sl_alignments(P, r0, c0, i, j, memo):
    """Structure-local alignments of the arc-match region with origin (r0,
    c0) ending at (i, j): list of (score, steps, last, usedA, usedB); steps
    as group_alignments plus ('XA', k, i, j) / ('XB', k, i, j) for an
    excluded block of k positions of A / B ending at (i, j)."""
    key = ('sl', r0, c0, i, j)
    if key in memo:
        return memo[key]
    out = []
    ok = lambda x, y: (x, y) == (r0, c0) or P.valid(x, y)
    if (i, j) == (r0, c0):
        out.append((0, (), None, False, False))
    elif ok(i, j):
        if i > r0 and j > c0 and ok(i - 1, j - 1):
            for sc, st, _, ua, ub in _sl_alignments(P, r0, c0, i - 1, j - 1, memo):
                out.append((sc + P.sigma[(i, j)], (('M', i, j),) + st, 'M', ua, ub))
        for kind, di, dj in (('D', 1, 0), ('I', 0, 1)):
            k = 1
            while i - k * di >= r0 and j - k * dj >= c0 and ok(i - k * di, j - k * dj):
                x, y = i - k * di, j - k * dj
                cost = P.alpha + (sum(P.gapA[r] for r in range(x + 1, i + 1)) if di else
                                  sum(P.gapB[c] for c in range(y + 1, j + 1)))
                for sc, st, last, ua, ub in _sl_alignments(P, r0, c0, x, y, memo):
                    if last != kind:
                        out.append((sc + cost, ((kind, k, i, j),) + st, kind, ua, ub))
                k += 1
        for kind, di, dj in (('XA', 1, 0), ('XB', 0, 1)):
            k = 1
            while i - k * di >= r0 and j - k * dj >= c0 and ok(i - k * di, j - k * dj):
                x, y = i - k * di, j - k * dj
                for sc, st, last, ua, ub in _sl_alignments(P, r0, c0, x, y, memo):
                    if (ua if kind == 'XA' else ub):
                        continue
                    out.append((sc + P.exclusion, ((kind, k, i, j),) + st, kind,
                                ua or kind == 'XA', ub or kind == 'XB'))
                k += 1
        if i > r0 and j > c0:
            for rank, (al, bl, kk) in enumerate(P.by_right.get((i, j), [])):
                if al <= r0 or bl <= c0:
                    continue
                inner = group_alignments(P, al, bl, i - 1, j - 1, False, memo)
                for isc, ist, _ in inner:
                    for psc, pst, _, ua, ub in _sl_alignments(P, r0, c0, al - 1, bl - 1, memo):
                        out.append((psc + isc + P.am_score[kk], (('A', rank, kk, ist),) + pst, 'A', ua, ub))
    memo[key] = out
    return out


def group_alignments(P, r0, c0, i, j, top, memo):
    """All structure alignments of the group with origin (r0, c0) ending at
    cell (i, j): list of (score, steps, last_edge_type)."""
    key = (r0, c0, i, j, top)
    if key in memo:
        return memo[key]
    if P.struct_local and not top:
        memo[key] = [(sc, st, last) for sc, st, last, _, _ in _sl_alignments(P, r0, c0, i, j, memo)]
        return memo[key]
    out = []
    free_row = top and (P.local or P.free & 1)  # insertions before the first position of A
    free_col = top and (P.local or P.free & 4)  # deletions before the first position of B
    if (i, j) == (r0, c0):
        out.append((0, (), None))
    elif i == r0:
        if all(_live(P, r0, c0, r0, c) for c in range(c0, j + 1)):
            sc = 0 if free_row else P.alpha + sum(P.gapB[c] for c in range(c0 + 1, j + 1))
            out.append((sc, (('pad', i, j),), 'I'))
    elif j == c0:
        if all(_live(P, r0, c0, r, c0) for r in range(r0, i + 1)):
            sc = 0 if free_col else P.alpha + sum(P.gapA[r] for r in range(r0 + 1, i + 1))
            out.append((sc, (('pad', i, j),), 'D'))
    elif P.valid(i, j):
        if _live(P, r0, c0, i - 1, j - 1):
            for sc, st, _ in group_alignments(P, r0, c0, i - 1, j - 1, top, memo):
                out.append((sc + P.sigma[(i, j)], (('M', i, j),) + st, 'M'))
        # a gap run starts at any cell with a value and runs through band cells
        k = 1
        while i - k >= r0 and _live(P, r0, c0, i - k, j):
            cost = P.alpha + sum(P.gapA[r] for r in range(i - k + 1, i + 1))
            for sc, st, last in group_alignments(P, r0, c0, i - k, j, top, memo):
                if last != 'D':
                    out.append((sc + cost, (('D', k, i, j),) + st, 'D'))
            if not P.valid(i - k, j):
                break
            k += 1
        k = 1
        while j - k >= c0 and _live(P, r0, c0, i, j - k):
            cost = P.alpha + sum(P.gapB[c] for c in range(j - k + 1, j + 1))
            for sc, st, last in group_alignments(P, r0, c0, i, j - k, top, memo):
                if last != 'I':
                    out.append((sc + cost, (('I', k, i, j),) + st, 'I'))
            if not P.valid(i, j - k):
                break
            k += 1
        for rank, (al, bl, kk) in enumerate(P.by_right.get((i, j), [])):
            if al <= r0 or bl <= c0:
                continue
            inner = group_alignments(P, al, bl, i - 1, j - 1, False, memo)
            prefix = group_alignments(P, r0, c0, al - 1, bl - 1, top, memo)
            for isc, ist, _ in inner:
                s_am = P.am_score[kk]
                if P.am_score_stacked is not None and len(ist) == 1 and ist[0][0] == 'A' and \
                        P.arcmatches[ist[0][2]] == (al + 1, i - 1, bl + 1, j - 1):
                    st = P.am_score_stacked[kk]
                    if P.no_lonely_pairs:
                        # --noLP with stacking (aligner.cc:612-657): an arc
                        # match around its inner one always scores stacked,
                        # and only a stackable one can
                        if st is None:
                            continue
                        s_am = st
                    elif st is not None:
                        s_am = max(s_am, st)
                for psc, pst, _ in prefix:
                    out.append((psc + isc + s_am, (('A', rank, kk, ist),) + pst, 'A'))
    memo[key] = out
    return out


def _decision(step):
    t = step[0]
    if t == 'M':
        return (0,)
    if t == 'D':
        return (1, step[1])
    if t == 'I':
        return (2, step[1])
    if t == 'A':
        return (3, step[1], trace_key(step[3]))
    return (9,)


def trace_key(steps):
    """Backward decision sequence; the traceback picks the minimum."""
    return tuple(_decision(s) for s in steps)


def edges(P, steps, r0, c0, top):
    """Forward edge list (a, b), -1 for gaps, and the matched arc indices."""
    out, arcs = [], []
    for s in reversed(steps):
        t = s[0]
        if t == 'pad':
            if s[1] == r0:
                if not (top and (P.local or P.free & 1)):
                    out += [(-1, c) for c in range(c0 + 1, s[2] + 1)]
            else:
                if not (top and (P.local or P.free & 4)):
                    out += [(r, -1) for r in range(r0 + 1, s[1] + 1)]
        elif t == 'M':
            out.append((s[1], s[2]))
        elif t == 'D':
            out += [(r, -1) for r in range(s[2] - s[1] + 1, s[2] + 1)]
        elif t == 'I':
            out += [(-1, c) for c in range(s[3] - s[1] + 1, s[3] + 1)]
        elif t == 'A':
            al, ar, bl, br = P.arcmatches[s[2]]
            arcs.append(s[2])
            out.append((al, bl))
            e2, a2 = edges(P, s[3], al, bl, False)
            out += e2
            arcs += a2
            out.append((ar, br))
    return out, arcs


def matched_arcs(P, steps):
    """Arc-match indices of a structure alignment (all nesting levels)."""
    out = []
    for s in steps:
        if s[0] == 'A':
            out.append(s[2])
            out += matched_arcs(P, s[3])
    return out


def stacked(P, steps):
    """True unless --noLP forbids the alignment: every matched arc pair
    (al, ar, bl, br) needs (al+1, ar-1, bl+1, br-1) or (al-1, ar+1, bl-1,
    br+1) matched too."""
    if not P.no_lonely_pairs:
        return True
    S = {P.arcmatches[k] for k in matched_arcs(P, steps)}
    return all((al + 1, ar - 1, bl + 1, br - 1) in S or (al - 1, ar + 1, bl - 1, br + 1) in S
               for (al, ar, bl, br) in S)


def best_global(P):
    """(max score, traced edges, traced arc-match indices, #alignments)."""
    memo = {}
    ends = [(P.n, P.m)]
    if P.free & (2 | 8):
        # aligner.cc:775-816: the last column for the second sequence's
        # right end (rows ascending), then the last row for the first's
        # (columns ascending); the first strict maximum
        ends = []
        if P.free & 8:
            ends += [(i, P.m) for i in range(1, P.n + 1) if P.max_col[i] >= P.m]
        if P.free & 2:
            ends += [(P.n, j) for j in range(max(1, P.min_col[P.n]), min(P.m, P.max_col[P.n]) + 1)]
    best, cell, count = None, None, 0
    for (i, j) in ends:
        alns = [a for a in group_alignments(P, 0, 0, i, j, True, memo) if stacked(P, a[1])]
        count += len(alns)
        if not alns:
            continue
        sc = max(x for x, _, _ in alns)
        if best is None or sc > best:
            best, cell = sc, (i, j)
    alns = [a for a in group_alignments(P, 0, 0, cell[0], cell[1], True, memo) if stacked(P, a[1])]
    chosen = min((st for sc, st, _ in alns if sc == best), key=trace_key)
    e, a = edges(P, chosen, 0, 0, True)
    return best, e, a, count


def best_local(P):
    """Sequence-local top level (aligner.cc:823-880): alignments start at any
    valid top-level cell with score 0 (the top level's init row/column are 0)
    and end at any valid cell; the empty alignment scores 0.  Returns the
    optimum and the set of optimal end cells."""
    memo = {}
    best, ends = 0, []
    for i in range(P.n + 1):
        for j in range(P.m + 1):
            if not P.valid(i, j):
                continue
            for r0 in range(i + 1):
                for c0 in range(j + 1):
                    if not P.valid(r0, c0):
                        continue
                    if (r0, c0) != (0, 0) and (r0 == 0 or c0 == 0):
                        continue  # starts on the zero init row / column are covered by (0,0)
                    for sc, st, _ in group_alignments(P, r0, c0, i, j, True, memo):
                        if not stacked(P, st):
                            continue
                        if sc > best:
                            best, ends = sc, [(i, j)]
                        elif sc == best and (i, j) not in ends:
                            ends.append((i, j))
    return best, ends


def partition_function(P, T):
    """Z = sum exp(score / T) over all structure alignments and the
    probability of every arc match and of every base-match edge that is not
    an arc end."""
    alns = group_alignments(P, 0, 0, P.n, P.m, True, {})
    Z = 0.0
    am = [0.0] * len(P.arcmatches)
    bm = {}
    for sc, st, _ in alns:
        w = math.exp(sc / T)
        Z += w
        e, arcs = edges(P, st, 0, 0, True)
        ends = set()
        for k in arcs:
            am[k] += w
            al, ar, bl, br = P.arcmatches[k]
            ends.add((al, bl))
            ends.add((ar, br))
        for (a, b) in e:
            if a > 0 and b > 0 and (a, b) not in ends:
                bm[(a, b)] = bm.get((a, b), 0.0) + w
    return Z, [x / Z for x in am], {k: v / Z for k, v in bm.items()}, len(alns)


def _walk(P, steps, r0, c0, top, owner, out):
    """Collects (kind, payload, owner) for the structure alignment: match
    edges ('M', (i, j), owner arc match or None), arc matches ('A', k,
    owner) and gap runs ('G', length, owner)."""
    for s in reversed(steps):
        t = s[0]
        if t == 'pad':
            length = (s[2] - c0) if s[1] == r0 else (s[1] - r0)
            if length > 0:
                out.append(('G', length, owner))
        elif t == 'M':
            out.append(('M', (s[1], s[2]), owner))
        elif t in ('D', 'I'):
            out.append(('G', s[1], owner))
        elif t == 'A':
            al, ar, bl, br = P.arcmatches[s[2]]
            out.append(('A', s[2], owner))
            _walk(P, s[3], al, bl, False, s[2], out)


def pf_details(P, T, min_am_prob):
    """Z, P(am), unrestricted match-edge marginals, the reference's
    base-match probabilities (match edges directly inside arc matches with
    P(am) > sqrt(min_am_prob) or at the top level), and the expected number
    of gap openings."""
    alns = group_alignments(P, 0, 0, P.n, P.m, True, {})
    Z = 0.0
    am = [0.0] * len(P.arcmatches)
    full, items = {}, []
    open_w = 0.0
    for sc, st, _ in alns:
        w = math.exp(sc / T)
        Z += w
        out = []
        _walk(P, st, 0, 0, True, None, out)
        items.append((w, out))
        for kind, x, owner in out:
            if kind == 'A':
                am[x] += w
            elif kind == 'M':
                full[x] = full.get(x, 0.0) + w
            elif kind == 'G':
                open_w += w
    am = [x / Z for x in am]
    thr = math.sqrt(min_am_prob)
    ref = {}
    for w, out in items:
        for kind, x, owner in out:
            if kind == 'M' and (owner is None or am[owner] > thr):
                ref[x] = ref.get(x, 0.0) + w
    return (Z, am, {k: v / Z for k, v in full.items()}, {k: v / Z for k, v in ref.items()}, open_w / Z)


def optimal_global_edges(P):
    """Edge lists of every optimal global structure alignment (top level
    ending in (n, m))."""
    alns = [a for a in group_alignments(P, 0, 0, P.n, P.m, True, {}) if stacked(P, a[1])]
    best = max(sc for sc, _, _ in alns)
    return best, [edges(P, st, 0, 0, True)[0] for sc, st, _ in alns if sc == best]


def _group_paths(P, steps, r0, c0, out):
    """Collects, for the group with origin (r0, c0) and every group nested in
    it, the cells its path visits in order and the interior cells of its gap
    runs: out gets (cells, run_interior) per group."""
    cells = [(r0, c0)]
    inner_cells = set()
    for s in reversed(steps):
        t = s[0]
        if t == 'pad':
            i, j = s[1], s[2]
            if i == r0:
                path = [(r0, c) for c in range(c0 + 1, j + 1)]
            else:
                path = [(r, c0) for r in range(r0 + 1, i + 1)]
            inner_cells.update(path[:-1])
            cells.extend(path)
        elif t == 'M':
            cells.append((s[1], s[2]))
        elif t == 'D':
            k, i, j = s[1], s[2], s[3]
            path = [(r, j) for r in range(i - k + 1, i + 1)]
            inner_cells.update(path[:-1])
            cells.extend(path)
        elif t == 'I':
            k, i, j = s[1], s[2], s[3]
            path = [(i, c) for c in range(j - k + 1, j + 1)]
            inner_cells.update(path[:-1])
            cells.extend(path)
        elif t == 'A':
            al, ar, bl, br = P.arcmatches[s[2]]
            cells.append((ar, br))
            _group_paths(P, s[3], al, bl, out)
    out.append((cells, inner_cells))


def fragment_probabilities(P, T, frags, split=True):
    """Probabilities that A[i..j] and B[k..l] are matched as one block, for
    each (i, j, k, l) of frags (AlignerP::compute_fragment_match_prob,
    aligner_p.icc:1440-1479), by enumeration: the weight of the structure
    alignments one of whose groups (the top level or an arc match's
    interior) passes through the cells (i-1, k-1) and (j, l), over Z.  The
    block starts and ends its own gap runs, as in the reference's
    decomposition into the block's inside and its outside: a run through
    either cell (the cell inside the run) counts twice, its opening score is
    added once more (split=False leaves that out, for a check that the rule
    is observed)."""
    alns = group_alignments(P, 0, 0, P.n, P.m, True, {})
    Z = 0.0
    hit = [0.0] * len(frags)
    for sc, st, _ in alns:
        Z += math.exp(sc / T)
        groups = []
        _group_paths(P, st, 0, 0, groups)
        for cells, run_inner in groups:
            cs = set(cells)
            for q, (i, j, k, l) in enumerate(frags):
                a, b = (i - 1, k - 1), (j, l)
                if a in cs and b in cs:
                    splits = ((a in run_inner) + (b in run_inner)) if split else 0
                    hit[q] += math.exp((sc + splits * P.alpha) / T)
    return [h / Z for h in hit]


def best_restricted(P, sA, eA, sB, eB):
    """Top level restricted to A[sA..eA] x B[sB..eB] (AlignerRestriction,
    aligner.cc:736-880 with r_): global, the best alignment from the
    origin (sA-1, sB-1) to (eA, eB) and its traced edges (the trace's tie
    rule); sequence-local, alignments that start at any valid cell of the
    rectangle with score 0 (the origin's row and column are the zero init)
    and end at any cell, and the optimum's first end cell in row-major
    order.  Returns (best, edges or None, end cell or None)."""
    memo = {}
    if not P.local:
        alns = group_alignments(P, sA - 1, sB - 1, eA, eB, True, memo)
        if not alns:
            return None, None, None
        best = max(sc for sc, _, _ in alns)
        chosen = min((st for sc, st, _ in alns if sc == best), key=trace_key)
        return best, edges(P, chosen, sA - 1, sB - 1, True)[0], (eA, eB)
    best, end = 0, None
    for i in range(sA - 1, eA + 1):
        for j in range(sB - 1, eB + 1):
            if not P.valid(i, j):
                continue
            for r0 in range(sA - 1, i + 1):
                for c0 in range(sB - 1, j + 1):
                    if (r0, c0) != (sA - 1, sB - 1) and not P.valid(r0, c0):
                        continue
                    if (r0, c0) != (sA - 1, sB - 1) and (r0 == sA - 1 or c0 == sB - 1):
                        continue  # starts on the zero init row / column are covered by the origin
                    for sc, st, _ in group_alignments(P, r0, c0, i, j, True, memo):
                        if sc > best:
                            best, end = sc, (i, j)
    return best, None, end
