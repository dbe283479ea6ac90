"""Argument checks of the folding entry points that return before any
device work (CPU): parameter ranges, the length limit, the pipeline's
minimum of two sequences, and the locarna_rnafold_pp program's refusals."""
import os
import subprocess

import pytest

from locarna_amd import api


@pytest.mark.parametrize("kw", [dict(min_prob=-0.1), dict(min_prob=1.0), dict(max_bp_span=0),
                                dict(max_bp_span=-2)])
def test_fold_parameter_ranges(kw):
    with pytest.raises(api.LcgError, match="fold parameters"):
        api.Fold(["ACGUACGUACGU"], **kw)


def test_fold_length_limit():
    with pytest.raises(api.LcgError, match="folding limit of 4096"):
        api.Fold(["ACGU" * 1024 + "A"])


def test_mlocarna_needs_two_sequences():
    with pytest.raises(api.LcgError):
        api.Mlocarna(["a"], ["ACGUACGUACGU"])


@pytest.mark.parametrize("args,stdin,msg", [(["--noLP"], "x ACGU\n", "not supported"),
                                            (["--in-loop"], "x ACGU\n", "not supported"),
                                            ([], ">a\nACGU\n>b\nACGU\n", "more than one sequence"),
                                            (["--bogus"], "x ACGU\n", "unknown option")])
def test_rnafold_pp_refusals(args, stdin, msg):
    exe = os.path.join(os.path.dirname(api.__file__), "_build", "locarna_rnafold_pp")
    r = subprocess.run([exe] + args, input=stdin, text=True, capture_output=True, timeout=60)
    assert r.returncode != 0 and msg in r.stderr
