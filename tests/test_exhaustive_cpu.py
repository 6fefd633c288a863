"""Exhaustive CPU checks of arithmetic identities the kernels rely on.

small_div.c: the binary distance's quotient of two counts, taken on the GPU
as a reciprocal (v_rcp_f32, within 1 ulp) followed by one correction from
the exact remainder, equals the correctly rounded quotient for every pair of
integers 0 <= a <= b <= 131072 (the library uses it up to 131072 features),
whichever float within one ulp of 1/b the reciprocal returns.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BOUND = 131072          # SMALL_DIV_MAX in gt_binary.hip


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_small_quotient_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "small_div")
    src = os.path.join(HERE, "exhaustive", "small_div.c")
    base = ["gcc", "-O3", "-ffp-contract=off", "-fopenmp", src, "-o", exe, "-lm"]
    # hardware fma where the host has one (the check is 10x slower through libm's)
    if subprocess.run(base + ["-march=native"], capture_output=True).returncode != 0:
        subprocess.run(base, check=True)
    r = subprocess.run([exe, str(BOUND)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failures=0" in r.stdout


def test_bound_matches_the_kernel():
    src = open(os.path.join(HERE, "..", "gputools_amd", "csrc", "gt_binary.hip")).read()
    assert "constexpr int64_t SMALL_DIV_MAX = %d;" % BOUND in src
