"""The pairwise alignment text of locarna --clustal (lcg_batch_write_clustal),
read back: the header carries the score mlocarna extracts
(src/Utils/mlocarna:3516-3527) and equals the oracle's; names are padded to
18 columns and blocks hold `width` columns; each row without gaps is the
whole sequence (the aligned part with local-only output); the match columns
are the oracle's match edges; structure rows when asked.  Global,
sequence-local and free-end-gap alignments.  Calls the kernels through the
C ABI."""
import re

import pytest

from locarna_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def _parse(text, nrows):
    lines = text.split("\n")
    m = re.fullmatch(r"CLUSTAL W --- LocARNA 2\.0\.1 --- Score: (\S+)", lines[0])
    assert m and lines[1] == "", lines[:2]
    body = lines[2:]
    assert body[-1] == ""  # text ends with a newline
    body = body[:-1]
    blocks, cur = [], []
    for l in body:
        if l == "":
            blocks.append(cur)
            cur = []
        else:
            cur.append(l)
    blocks.append(cur)
    rows = [("", "")] * nrows
    widths = []
    for b in blocks:
        assert len(b) == nrows, b
        for r, l in enumerate(b):
            name, seq = l[:18].rstrip(), l[19:]
            assert l[18] == " "
            rows[r] = (name, rows[r][1] + seq)
        widths.append(len(b[0]) - 19)
    return m.group(1), rows, widths


def _match_edges(rowA, rowB):
    i = j = 0
    out = []
    for a, b in zip(rowA, rowB):
        i += a != "-"
        j += b != "-"
        if a != "-" and b != "-":
            out.append((i, j))
    return out


def _check(pp, kw, label, width=120):
    pa, po = api.locarna_params(**kw), oracle.locarna_params(**kw)
    P = [api.Pair(a, b, pa) for a, b in pp]
    B = api.Batch(P)
    try:
        B.run(mode=2)
        for k, (a, b) in enumerate(pp):
            o = oracle.align(a, b, po)
            text = B.clustal(k, width=width)
            score, rows, widths = _parse(text, 2)
            assert score == ("-inf" if o.score < -(1 << 60) else str(o.score)), (label, k)
            assert all(w == width for w in widths[:-1]) and 0 < widths[-1] <= width, (label, k)
            (nA, rA), (nB, rB) = rows
            assert len(rA) == len(rB)
            assert len(rA.replace("-", "")) == o.lenA and len(rB.replace("-", "")) == o.lenB, (label, k)
            want = [(x, y) for x, y in o.edges if x > 0 and y > 0]
            assert _match_edges(rA, rB) == want, (label, k)
            # local-only output: only the aligned part
            lo = _parse(B.clustal(k, width=width, local_only=True), 2)[1]
            if o.edges:
                a_pos = [x for x, _ in o.edges if x > 0]
                assert len(lo[0][1].replace("-", "")) == (max(a_pos) - min(a_pos) + 1 if a_pos else 0), (label, k)
            # structure rows first and last, over the same columns
            _, srows, _ = _parse(B.clustal(k, width=width, structures=True), 4)
            assert srows[0][0] == "" and srows[3][0] == "" and srows[1][1] == rA and srows[2][1] == rB
            assert set(srows[0][1]) <= set("().-") and len(srows[0][1]) == len(rA)
            sA = "".join(c for c, x in zip(srows[0][1], rA) if x != "-")
            assert sA == o.structA or o.structA == "", (label, k)
    finally:
        B.destroy()


def test_global_pairs():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=30)
    _check([(pps[a], pps[b]) for a, b in pairs], dict(max_bp_span=span), "global")
    _check([(pps[a], pps[b]) for a, b in pairs[:5]], dict(max_bp_span=span), "global width 37", width=37)


def test_local_pairs():
    pps, pairs, span = synth.workload("long500", limit_pairs=2)
    _check([(pps[a], pps[b]) for a, b in pairs], dict(max_bp_span=span, sequ_local=1), "local")


def test_free_endgap_pairs():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=20)
    _check([(pps[a], pps[b]) for a, b in pairs], dict(max_bp_span=span, free_endgaps=api.free_endgaps("++++")),
           "free")
