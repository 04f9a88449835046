"""The correctly rounded division of the pipelined Gauss-Seidel walk.

The LDS sweep kernel replaces x/ap/sor on its critical path by
div_rcp2(x, ap, 1/ap, sor, 1/sor): with y = 1/a correctly rounded, q0 = x*y
and one fused remainder correction give the correctly rounded quotient
(Markstein's theorem), twice, with one sign fix for zero results, so the
fields stay bit-identical to the reference's divisions.
tests/native/div_rcp_check.c applies the same operation sequence on the host
to random and adversarial operands (signed zeros included) and compares
every result with the plain divisions bit for bit; the GPU parity tests
cover the device code itself.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fma_corrected_quotient_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "div_rcp_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "native", "div_rcp_check.c"), "-lm"], check=True)
    for seed in ("1", "2"):
        p = subprocess.run([exe, seed, "5000000"], capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stdout + p.stderr
        assert "mismatches 0" in p.stdout
