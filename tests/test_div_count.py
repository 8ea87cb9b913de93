"""The cross sweep divides each support sum by its integer count with a
reciprocal, one fused residual correction and an IEEE fallback for
NaN/infinite/denormal/zero quotients (rtstereo_amd/csrc/rts_internal.h, div_count).
This runs the same arithmetic on the CPU for every count the short-arm sweep
can produce (1..1225) and a sample of counts up to 2^19, with every reciprocal
within two ulp of 1/c, and requires the IEEE quotient bit for bit
(seq/CrossBasedAggregation.cpp:165 divides with `costSum /= cnt`)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("divc") / "div_count_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "div_count_check.c"), "-lm"],
                   check=True)
    return exe


def test_all_short_arm_counts(checker):
    r = subprocess.run([checker, "1225", "4000", "7"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0


def test_long_arm_counts_sampled(checker):
    r = subprocess.run([checker, "524288", "2", "11"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
