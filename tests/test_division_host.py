"""The quotient identity behind div_rcp (uwshcu_math.h), checked on the host.

For y = RN(1/b), q0 = a*y and q = fma(fma(-b, q0, a), y, q0) is the correctly rounded
a/b (Markstein, IBM J. Res. Dev. 34(1), 1990).  The saturation adjustment on the GPU
takes its divisions by recurring divisors this way; the GPU side of the same identity,
and of the unscaled division sequence, is checked in test_gpu_parity.py.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_reciprocal_quotient_is_the_ieee_quotient(tmp_path):
    exe = tmp_path / "divcheck"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", os.path.join(HERE, "hostemu", "divcheck.c"),
                    "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe), "1500000"], check=True, capture_output=True, text=True).stdout.split()
    n, bad = int(out[-2]), int(out[-1])
    assert n == 64 * 1500000
    assert bad == 0, " ".join(out)
