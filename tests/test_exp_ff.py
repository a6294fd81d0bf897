"""The device's single-precision exponential (indicatrix_amd/csrc/exp_ff.h).

The film's band sums and Planck illuminant need e^x rounded once to f32 (what
glibc's expf gives the oracle for these arguments), and the device computes
it without double precision: Tang's table method in float-float.  Here the
same source, compiled for the host, is held to the double-precision
exponential rounded to f32 over 6 M arguments, its table to 2^(j/64) in
exact decimal arithmetic, and its special values.  CPU only.
"""
import json
import os
import subprocess
from decimal import Decimal, getcontext

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exp_ff_is_the_correctly_rounded_exponential(tmp_path):
    exe = str(tmp_path / "exp_ff_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(ROOT, "tests", "native", "exp_ff_check.cpp")], check=True)
    out = json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)
    assert out["n"] > 6_000_000
    # a tie within 2^-45 of a rounding boundary is the only way to differ:
    # measured 1 in 30 M normal results
    assert out["bad_normal"] <= 3, out["bad_normal"]
    assert out["bad_edge"] == 0 and out["inf_ok"] and out["nan_ok"] and out["ninf_ok"]
    # subnormal results round twice (significand, then scaling): bounded, never above 1 ULP
    assert out["bad_subnormal"] < 0.01 * out["n"]
    getcontext().prec = 60
    ln2 = Decimal(2).ln()
    for j, (h, l) in enumerate(zip(out["hi"], out["lo"])):
        exact = (Decimal(j) / 64 * ln2).exp()
        pair = Decimal(float.fromhex(h)) + Decimal(float.fromhex(l))
        assert abs(pair - exact) / exact < Decimal(2) ** -47, j
        # hi is the f32 nearest 2^(j/64): the remainder is under half its ULP
        assert abs(exact - Decimal(float.fromhex(h))) <= Decimal(2) ** -24 * exact, j
