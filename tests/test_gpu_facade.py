"""The C++ facade (include/locarna_amd.hpp: Aligner, AlignerP,
AlignerParams in the reference's named-argument style) driven from a
compiled program, locarna_gpu (locarna_amd/csrc/locarna_gpu.cpp), with the
call sequence of the reference's locarna / locarna_p main programs: scores,
alignments and structures equal the oracle's; the partition function is
within 1e-9 relative and the probability files equal the ctypes path's."""
import os
import subprocess

import pytest

from locarna_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "locarna_amd", "_build", "locarna_gpu")

FLAGS = ["--indel=-350", "--indel-opening=-500", "--struct-weight=200", "--tau=0", "--min-prob=0.0005",
         "--maxBPspan=100"]


def _files(tmp_path, seed, n1, n2):
    seqs = synth.sequences(seed, 2, min(n1, n2), max(n1, n2), 0.5, 0.5)
    ppA, ppB = synth.make_pp("a", seqs[0], 100), synth.make_pp("b", seqs[1], 100)
    fa, fb = tmp_path / "a.pp", tmp_path / "b.pp"
    fa.write_text(ppA)
    fb.write_text(ppB)
    return ppA, ppB, str(fa), str(fb)


def _run(args):
    print("locarna_gpu", " ".join(args), flush=True)
    r = subprocess.run([EXE] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    return r.returncode, r.stdout, r.stderr


def _parse(out):
    res = {}
    for line in out.splitlines():
        if line.startswith("Score: "):
            res["score"] = int(line.split()[1])
        elif line.startswith("EDGES"):
            res["edges"] = [tuple(int(x) for x in e.split(":")) for e in line.split()[1:]]
        elif line.startswith("STRUCTURE_A "):
            res["sA"] = line.split()[1] if len(line.split()) > 1 else ""
        elif line.startswith("STRUCTURE_B "):
            res["sB"] = line.split()[1] if len(line.split()) > 1 else ""
        elif line.startswith("Z_EXACT "):
            res["Z"] = float(line.split()[1])
        elif line.startswith("FRAGMENT_MATCH_PROB "):
            f = line.split()
            res.setdefault("frags", []).append((tuple(int(x) for x in f[1:5]), float(f[6])))
    return res


@pytest.mark.parametrize("extra,kw", [
    ([], {}),
    (["--sequ-local"], {"sequ_local": 1}),
    (["--free-endgaps=++--"], {"free_endgaps": 1 | 2}),
])
def test_aligner_matches_oracle(tmp_path, extra, kw):
    ppA, ppB, fa, fb = _files(tmp_path, 41, 90, 120)
    rc, out, err = _run([fa, fb] + FLAGS + extra + ["--write-structure", "--clustal"])
    assert rc == 0, err
    res = _parse(out)
    o = oracle.align(ppA, ppB, oracle.locarna_params(**kw))
    assert res["score"] == o.score
    assert res["edges"] == o.edges
    assert res["sA"] == o.structA and res["sB"] == o.structB
    assert "CLUSTAL W --- LocARNA 2.0.1 --- Score: %d" % o.score in out


def test_aligner_penalized_matches_oracle(tmp_path):
    ppA, ppB, fa, fb = _files(tmp_path, 42, 70, 80)
    rc, out, err = _run([fa, fb] + FLAGS + ["--sequ-local", "--penalized=30"])
    assert rc == 0, err
    res = _parse(out)
    o = oracle.align_variant(ppA, ppB, oracle.locarna_params(sequ_local=1), penalized=30)
    assert res["score"] == o.score and res["edges"] == o.edges


def test_alignerp_matches_oracle_and_ctypes_path(tmp_path):
    ppA, ppB, fa, fb = _files(tmp_path, 44, 80, 100)
    am, bm = tmp_path / "am.txt", tmp_path / "bm.txt"
    pflags = ["--indel=-350", "--indel-opening=-500", "--struct-weight=200", "--min-prob=0.0005",
              "--maxBPspan=100"]
    rc, out, err = _run(["--pf", fa, fb] + pflags + ["--write-arcmatch-probs=%s" % am,
                                                     "--write-basematch-probs=%s" % bm])
    assert rc == 0, err
    res = _parse(out)
    o = oracle.pf(ppA, ppB, oracle.locarna_p_params(min_am_prob=0.001, min_bm_prob=0.001))
    assert res["Z"] == pytest.approx(o.pf, rel=1e-9)
    pa = api.locarna_p_params(min_am_prob=0.001, min_bm_prob=0.001)
    batch = api.Batch([api.Pair(ppA, ppB, pa)])
    batch.pf_run()
    assert am.read_text() == batch.am_probs_text(0)
    assert bm.read_text() == batch.bm_probs_text(0)
    batch.destroy()


def test_aligner_no_lonely_pairs_matches_oracle(tmp_path):
    ppA, ppB, fa, fb = _files(tmp_path, 45, 80, 90)
    rc, out, err = _run([fa, fb] + FLAGS + ["--noLP", "--write-structure"])
    assert rc == 0, err
    res = _parse(out)
    o = oracle.align(ppA, ppB, oracle.locarna_params(no_lonely_pairs=1))
    assert res["score"] == o.score and res["edges"] == o.edges
    assert res["sA"] == o.structA and res["sB"] == o.structB


@pytest.mark.parametrize("flags,bits", [(["--stacking", "--exp-prob=0.01"], 1), (["--new-stacking"], 2),
                                        (["--stacking", "--new-stacking", "--exp-prob=0.01"], 3),
                                        (["--stacking"], 0)])
def test_aligner_stacking_matches_oracle(tmp_path, flags, bits):
    """--stacking / --new-stacking on #STACK files; --stacking without
    --exp-prob is turned off with a warning (src/locarna.cc:407-414)."""
    seqs = synth.sequences(47, 2, 70, 90, 0.5, 0.5)
    pps = [synth.pp_text(n, s, synth.with_stacking(synth.base_pair_probs(s, 100), seed=k))
           for k, (n, s) in enumerate(zip("ab", seqs))]
    fa, fb = tmp_path / "a.pp", tmp_path / "b.pp"
    fa.write_text(pps[0])
    fb.write_text(pps[1])
    rc, out, err = _run([str(fa), str(fb)] + FLAGS + flags + ["--write-structure"])
    assert rc == 0, err
    assert ("stacking turned off" in err) == (bits == 0 and "--new-stacking" not in flags)
    res = _parse(out)
    exp = 0.01 if any(f.startswith("--exp-prob") for f in flags) else -1.0
    o = oracle.align(pps[0], pps[1], oracle.locarna_params(stacking=bits, exp_prob=exp))
    assert res["score"] == o.score and res["edges"] == o.edges
    assert res["sA"] == o.structA and res["sB"] == o.structB


def test_fragment_match_probs_match_oracle(tmp_path):
    """locarna_p --fragment-match-probs: one FRAGMENT_MATCH_PROB line per
    fragment (src/locarna_p.cc:537-561), printed to six significant digits."""
    ppA, ppB, fa, fb = _files(tmp_path, 46, 60, 70)
    frags = [(1, 10, 1, 12), (5, 30, 8, 35), (20, 21, 22, 23)]
    spec = "; ".join("%d %d %d %d" % f for f in frags)
    pflags = ["--indel=-350", "--indel-opening=-500", "--struct-weight=200", "--min-prob=0.0005",
              "--maxBPspan=100"]
    rc, out, err = _run(["--pf", fa, fb] + pflags + ["--fragment-match-probs=" + spec])
    assert rc == 0, err
    got = _parse(out)["frags"]
    want = oracle.pf_fragments(ppA, ppB, frags, oracle.locarna_p_params(min_am_prob=0.001, min_bm_prob=0.001))
    assert [g[0] for g in got] == frags
    for (f, x), y in zip(got, want):
        assert abs(x - y) <= 1e-5 * abs(y) + 1e-300, (f, x, y)


def test_kbest_matches_oracle(tmp_path):
    """locarna --kbest: one Score / EDGES block per enumerated alignment."""
    ppA, ppB, fa, fb = _files(tmp_path, 47, 60, 70)
    rc, out, err = _run([fa, fb] + FLAGS + ["--sequ-local", "--kbest=4"])
    assert rc == 0, err
    blocks = [b for b in out.strip().split("\n\n") if b.strip()]
    got = [(_parse(b)["score"], _parse(b)["edges"]) for b in blocks]
    want = oracle.suboptimal(ppA, ppB, 4, params=oracle.locarna_params(sequ_local=1))
    assert got == [(w[0], w[1]) for w in want]
