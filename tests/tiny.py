"""Tiny random problems (sequences of a few nucleotides with a handful of
arcs) written as #PP 2.0 text, for exhaustive checks."""
import random

from locarna_amd.synth import pp_text


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
