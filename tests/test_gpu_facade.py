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
