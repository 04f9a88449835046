"""The mixed-precision cavity at Re = 1000, up to the full 256^3 workload.

Small meshes: the FP32 pc solves (colour-order "mixed" and multigrid
"mgcg-mixed") against the oracle's restatement, bit for bit.  Full size
(256^3, 16.8M cells, steady, relaxation 0.7 / 0.3), after the same outer
iterations from rest:
  - "mixed" (FP32 colour sweeps under FP64 defect correction at every
    residual check) against the FP64 colour order: per field relative L2
    <= 1e-6 (measured about 1e-9);
  - "mgcg-mixed" (FP32 V-cycle inside FP64 CG) against "mgcg": <= 1e-5.
    Each CG solve stops at a tenfold residual reduction, so the two
    preconditioners leave different unconverged remainders in pc, which the
    outer iterations carry (a difference of solver path, not of rounding:
    at 128^3 Re=400 it measures 1.8e-6 in u and 5.7e-6 in p);
velocity errors are relative to the norm of the whole velocity field (u, v,
w together: v alone is small in this cavity, and its own norm would inflate
its relative error), p to its own norm;
the pc right-hand side sums to zero (a closed cavity: inflow = outflow); and a
second run on a fresh context repeats the first bit for bit.
"""
import numpy as np
import pytest

import cfdl
import meshgen
import oracle_lib

pytestmark = pytest.mark.gpu

RE = 1000.0


@pytest.mark.parametrize("n,jit,order,gs", [(24, 0.0, "mixed", "reference"),
                                            (24, 0.1, "mgcg-mixed", "colored"),
                                            (16, 0.2, "mixed", "colored")])
def test_re1000_mixed_bitwise_against_the_oracle(n, jit, order, gs):
    mesh = cfdl.Mesh.box(n, jitter=jit).partition(4)
    eq = cfdl.UVWP.cavity(mesh, rho=1.0, mu=1.0 / RE, dt=float("inf"))
    g = meshgen.box_mesh(n, jitter=jit)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=4, rho=1.0, mu=1.0 / RE,
                               dt=float("inf"))
    for x in (eq, oc):
        x.set_relaxation(0.7, 0.3)
        x.set_pc_order(order)
        x.set_gs_order(gs)
    eq.run(1, 4)
    oc.run(1, 4)
    for name in ("u", "v", "w", "p", "gp", "mip", "phic"):
        assert np.array_equal(eq.get(name), oc.arr(name)), name
    assert [l[:2] for l in eq.log()] == [l[:2] for l in oc.log()]


def _errors(got, ref):
    vel = np.linalg.norm(np.concatenate([ref[k] for k in ("u", "v", "w")]))
    return {k: float(np.linalg.norm(got[k] - ref[k]) /
                     max(vel if k != "p" else np.linalg.norm(ref[k]), 1e-300))
            for k in ("u", "v", "w", "p")}


@pytest.fixture(scope="module")
def mesh256():
    return cfdl.Mesh.box(256).partition(4)


def _run(mesh, order, gs, nouter):
    eq = cfdl.UVWP.cavity(mesh, rho=1.0, mu=1.0 / RE, dt=float("inf"))
    eq.set_relaxation(0.7, 0.3)
    eq.set_pc_order(order)
    eq.set_gs_order(gs)
    eq.run(1, nouter)
    out = {k: eq.get(k) for k in ("u", "v", "w", "p")}
    out["log"] = eq.log()
    # the pc system of the last iteration: its right-hand side is the net
    # inflow of every cell, which sums to zero in a closed cavity
    eq.calc_coef_p()
    b = eq.get("b")
    out["bsum"] = (float(np.sum(b)), float(np.sum(np.abs(b))))
    eq.close()
    return out


def test_256_mixed_precision_against_fp64(mesh256):
    K = 4
    fp64_col = _run(mesh256, "colored", "colored", K)
    mixed = _run(mesh256, "mixed", "colored", K)
    fp64_mg = _run(mesh256, "mgcg", "colored", K)
    mixed_mg = _run(mesh256, "mgcg-mixed", "colored", K)
    for ref, got, bound in ((fp64_col, mixed, 1e-6), (fp64_mg, mixed_mg, 1e-5)):
        assert [l[:2] for l in ref["log"]] == [l[:2] for l in got["log"]]
        err = _errors(got, ref)
        assert max(err.values()) <= bound, err
    for r in (fp64_col, mixed, fp64_mg, mixed_mg):
        assert abs(r["bsum"][0]) <= 1e-12 * r["bsum"][1]
    again = _run(mesh256, "mgcg-mixed", "colored", K)
    for k in ("u", "v", "w", "p"):
        assert np.array_equal(again[k], mixed_mg[k]), k
