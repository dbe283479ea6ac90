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
