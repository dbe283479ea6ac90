"""Tiny random problems (sequences of a few nucleotides with a handful of
arcs) written as #PP 2.0 text, for exhaustive checks."""
import random

from locarna_amd.synth import pp_text, pp_text_rows


def tiny_pp(rng, n, narcs, name="t"):
    seq = "".join(rng.choice("ACGU") for _ in range(n))
    cand = [(i, j) for i in range(1, n + 1) for j in range(i + 3, n + 1)]
    rng.shuffle(cand)
    arcs = sorted(cand[:narcs])
    pairs = [(i, j, round(rng.uniform(0.02, 0.95), 4)) for (i, j) in arcs]
    return pp_text(name, seq, pairs, bpcut=0.0005)


def tiny_pair(seed, nA=None, nB=None, arcs=(1, 3)):
    rng = random.Random(seed)
    nA = nA or rng.randint(4, 7)
    nB = nB or rng.randint(4, 7)
    return (tiny_pp(rng, nA, rng.randint(*arcs), "a"),
            tiny_pp(rng, nB, rng.randint(*arcs), "b"))


def tiny_stacked_pp(rng, n, name="t"):
    """A tiny sequence whose arcs include stacks: one or two stacks of two
    or three arcs (i, j), (i+1, j-1), ... and possibly a lonely arc."""
    seq = "".join(rng.choice("ACGU") for _ in range(n))
    arcs = set()
    for _ in range(rng.randint(1, 2)):
        depth = rng.choice((2, 3)) if n >= 9 else 2
        i = rng.randint(1, n - 3 - 2 * (depth - 1))
        j = rng.randint(i + 3 + 2 * (depth - 1), n)
        for d in range(depth):
            arcs.add((i + d, j - d))
    if rng.random() < 0.7:
        i = rng.randint(1, n - 3)
        j = rng.randint(i + 3, n)
        arcs.add((i, j))
    pairs = [(i, j, round(rng.uniform(0.05, 0.95), 4)) for (i, j) in sorted(arcs)]
    return pp_text(name, seq, pairs, bpcut=0.0005)


def tiny_stacked_pair(seed, nA=None, nB=None):
    rng = random.Random(seed)
    nA = nA or rng.randint(6, 9)
    nB = nB or rng.randint(6, 9)
    return tiny_stacked_pp(rng, nA, "a"), tiny_stacked_pp(rng, nB, "b")


def tiny_stacking_pp(rng, n, name="t"):
    """A tiny stacked sequence (tiny_stacked_pp's arcs) whose stacked arcs
    carry joint probabilities with their inner arcs (#STACK), some of them
    at or below the cut."""
    seq = "".join(rng.choice("ACGU") for _ in range(n))
    arcs = set()
    for _ in range(rng.randint(1, 2)):
        depth = rng.choice((2, 3)) if n >= 9 else 2
        i = rng.randint(1, n - 3 - 2 * (depth - 1))
        j = rng.randint(i + 3 + 2 * (depth - 1), n)
        for d in range(depth):
            arcs.add((i + d, j - d))
    if rng.random() < 0.7:
        i = rng.randint(1, n - 3)
        j = rng.randint(i + 3, n)
        arcs.add((i, j))
    prob = {a: round(rng.uniform(0.05, 0.95), 4) for a in sorted(arcs)}
    pairs = []
    for (i, j) in sorted(arcs):
        q = prob.get((i + 1, j - 1))
        if q is not None and rng.random() < 0.85:
            p2 = round(min(prob[(i, j)], q) * rng.uniform(0.2, 1.0), 4)
            if rng.random() < 0.1:
                p2 = 0.0004  # at the cut: not a joint probability
            pairs.append((i, j, prob[(i, j)], p2))
        else:
            pairs.append((i, j, prob[(i, j)]))
    return pp_text(name, seq, pairs, bpcut=0.0005)


def tiny_stacking_pair(seed, nA=None, nB=None):
    rng = random.Random(seed)
    nA = nA or rng.randint(6, 9)
    nB = nB or rng.randint(6, 9)
    return tiny_stacking_pp(rng, nA, "a"), tiny_stacking_pp(rng, nB, "b")


def tiny_profile_pp(rng, n, nrows, narcs, name="p", gap_rate=0.25, n_rate=0.05):
    """A tiny alignment of nrows rows over n columns (gaps '-', now and then
    an N) with arcs on its columns; rows are named name0, name1, ..."""
    rows = []
    for r in range(nrows):
        row = "".join("-" if rng.random() < gap_rate else ("N" if rng.random() < n_rate else rng.choice("ACGU"))
                      for _ in range(n))
        rows.append(("%s%d" % (name, r), row))
    cand = [(i, j) for i in range(1, n + 1) for j in range(i + 3, n + 1)]
    rng.shuffle(cand)
    arcs = sorted(cand[:narcs])
    pairs = [(i, j, round(rng.uniform(0.02, 0.95), 4)) for (i, j) in arcs]
    return pp_text_rows(rows, pairs, bpcut=0.0005)


def tiny_profile_pair(seed, rows=(1, 3), arcs=(1, 3)):
    """Two tiny profiles (a single-row side now and then)."""
    rng = random.Random(seed)
    return (tiny_profile_pp(rng, rng.randint(4, 7), rng.randint(*rows), rng.randint(*arcs), "a"),
            tiny_profile_pp(rng, rng.randint(4, 7), rng.randint(*rows), rng.randint(*arcs), "b"))


def replicate_rows(pp, k):
    """The pp text with its single sequence row repeated k times (rows
    named name.0 .. name.k-1): a profile that scores as the sequence."""
    lines = pp.split("\n")
    end = lines.index("#END")
    seqlines = [i for i in range(end) if lines[i] and not lines[i].startswith("#")]
    assert len(seqlines) == 1
    name, seq = lines[seqlines[0]].split()
    rows = ["%s.%d %s" % (name, t, seq) for t in range(k)]
    return "\n".join(lines[:seqlines[0]] + rows + lines[seqlines[0] + 1:])
