"""The correctly rounded float square root of the weight, proximity and NCC
kernels (rtstereo_amd/csrc/sqrt_rn.h: the hardware square root moved to
RN(sqrt(x)) by exact integer midpoint tests, no double precision), compiled
here from the same header and checked against the C library's IEEE sqrtf on
every float of [1, 4), every denormal, every integer up to 2^18, specials
and a sample across all exponents, from approximations up to two ulp off
(seq/AdaptiveWeightAggregation.cpp:15-24 and seq/NCCDisparity.cpp:66 take
float square roots)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "rtstereo_amd", "csrc")


def test_sqrt_rn_matches_ieee(tmp_path):
    exe = str(tmp_path / "sqrt_rn_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", f"-I{CSRC}", "-o", exe, os.path.join(HERE, "sqrt_rn_check.c"),
                    "-lm"], check=True)
    r = subprocess.run([exe, "1000000", "5"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
