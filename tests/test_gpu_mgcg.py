"""GPU parity of the multigrid-preconditioned CG pressure-correction solve
(pc orders "mgcg" and "mgcg-mixed", an extension) and of the steady mode
(dt = infinity), against the oracle, bit for bit.

The oracle restates the solve operation by operation, including the blocked
order of every dot product and norm, so the fields and the logged residuals
(not only the iteration counts) are equal.  The meshes cover every shape of
the hierarchy: one level (4^3, the V-cycle is the coarse sweeps on the mesh),
small levels in the one-workgroup kernel only (8^3 .. 32^3), and grid-wide
coarse levels above it (40^3).
"""
import numpy as np
import pytest

import cfdl
import meshgen
import oracle_lib

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "w", "p", "gp", "mip", "phic", "gpc", "ap", "anb", "b", "bu", "d", "dc"]


def make_pair(n, jitter, nsubd, mu, dt):
    mesh = cfdl.Mesh.box(n, jitter=jitter).partition(nsubd)
    eq = cfdl.UVWP.cavity(mesh, rho=1.0, mu=mu, dt=dt)
    g = meshgen.box_mesh(n, jitter=jitter)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=nsubd, rho=1.0, mu=mu, dt=dt)
    return eq, oc


def check_equal(eq, oc, exact_pc=True):
    for name in FIELDS:
        a, b = eq.get(name), oc.arr(name)
        if name == "gpc":
            a, b = a[:3 * oc.c.ne], b[:3 * oc.c.ne]
        assert np.array_equal(a, b), "%s: max |diff| %g" % (name, np.max(np.abs(a - b)))
    gl, ol = eq.log(), oc.log()
    assert len(gl) == len(ol)
    for a, b in zip(gl, ol):
        assert a[:2] == b[:2], (a, b)
        if a[0] == "pc" and exact_pc:  # blocked reductions in the oracle's order: equal
            assert a[2:] == b[2:], (a, b)
        else:
            for x, y in zip(a[2:], b[2:]):
                assert x == pytest.approx(y, rel=1e-12, abs=1e-300), (a, b)


@pytest.mark.parametrize("n,jit,nsubd,order,steady,nouter", [
    (4, 0.0, 4, "mgcg", False, 3),
    (8, 0.0, 4, "mgcg", False, 6),
    (12, 0.15, 4, "mgcg", True, 6),
    (12, 0.15, 1, "mgcg-mixed", True, 6),
    (9, 0.1, 2, "mgcg-mixed", False, 6),
    (20, 0.2, 4, "mgcg", True, 4),
    (32, 0.0, 4, "mgcg-mixed", True, 4),
    (40, 0.0, 4, "mgcg", True, 3),
    (40, 0.1, 4, "mgcg-mixed", True, 3),
])
def test_mgcg_bitwise(n, jit, nsubd, order, steady, nouter):
    """Relaxed config mode (0.7 / 0.3), pseudo-transient or steady; every
    field and every solver line equal to the oracle's."""
    dt = float("inf") if steady else 0.01
    eq, oc = make_pair(n, jit, nsubd, 0.0025, dt)
    for x in (eq, oc):
        x.set_relaxation(0.7, 0.3)
        x.set_pc_order(order)
    for k in range(nouter):
        eq.run(1, 1)
        oc.run(1, 1)
    check_equal(eq, oc)
    assert any(l[0] == "pc" and l[1] > 0 for l in eq.log())


def test_mgcg_with_colored_gs_and_reset():
    """With the colour order of solve_gs, and across a second run from the
    same state on a context that switched orders in between (the colour-order
    "mixed" solve in the middle logs tree-reduced norms that agree with the
    oracle's sequential ones to rounding, so its lines are compared to
    1e-12)."""
    eq, oc = make_pair(16, 0.1, 4, 0.001, float("inf"))
    for x in (eq, oc):
        x.set_relaxation(0.7, 0.3)
        x.set_gs_order("colored")
        x.set_pc_order("mgcg-mixed")
    eq.run(1, 3)
    oc.run(1, 3)
    for x in (eq, oc):
        x.set_pc_order("mixed")
    eq.run(1, 2)
    oc.run(1, 2)
    for x in (eq, oc):
        x.set_pc_order("mgcg")
    eq.run(1, 3)
    oc.run(1, 3)
    check_equal(eq, oc, exact_pc=False)
    mg_lines = [l for l in eq.log()[-12:] if l[0] == "pc"]
    assert mg_lines == [l for l in oc.log()[-12:] if l[0] == "pc"]


@pytest.mark.parametrize("n,jit,nsubd,order", [(8, 0.1, 4, "reference"), (10, 0.0, 1, "reference"),
                                               (12, 0.1, 4, "mixed")])
def test_steady_mode_bitwise(n, jit, nsubd, order):
    """dt = infinity drops the transient terms (rho V / dt and the Rhie-Chow
    transient correction are exactly zero); with relaxation the loop is
    steady SIMPLE.  Same fields and solver lines as the oracle."""
    eq, oc = make_pair(n, jit, nsubd, 0.01, float("inf"))
    for x in (eq, oc):
        x.set_relaxation(0.7, 0.3)
        x.set_pc_order(order)
    eq.run(2, 3)
    oc.run(2, 3)
    for name in ["u", "v", "w", "p", "mip", "gp", "ap", "dc"]:
        assert np.array_equal(eq.get(name), oc.arr(name)), name
    assert [l[:2] for l in eq.log()] == [l[:2] for l in oc.log()]


def test_steady_mode_needs_relaxation():
    """Without momentum relaxation a steady cell has ap - sum(anb) = 0, so
    dc = V / 0: the context refuses to run instead of producing NaN."""
    mesh = cfdl.Mesh.box(4).partition(1)
    eq = cfdl.UVWP.cavity(mesh, rho=1.0, mu=0.01, dt=float("inf"))
    with pytest.raises(cfdl.CfdlError):
        eq.run(1, 1)


@pytest.mark.parametrize("n,jit,nsubd,order,gs", [(12, 0.15, 4, "mixed", "colored"),
                                                  (16, 0.2, 4, "mgcg-mixed", "colored"),
                                                  (9, 0.1, 2, "reference", "reference")])
def test_fp32_assembly_bitwise(n, jit, nsubd, order, gs):
    """FP32 coefficient assembly (an extension): the FP32 operations of
    k_coef_uvw6<float> / k_coef_p6<float> against the oracle's float
    restatement, bit for bit, within the steady relaxed loop."""
    eq, oc = make_pair(n, jit, nsubd, 0.001, float("inf"))
    for x in (eq, oc):
        x.set_relaxation(0.7, 0.3)
        x.set_pc_order(order)
        x.set_gs_order(gs)
        x.set_assembly_precision(32)
    eq.run(1, 4)
    oc.run(1, 4)
    for name in ["u", "v", "w", "p", "gp", "mip", "ap", "anb", "b", "bu", "bv", "bw", "d", "dc"]:
        a, b = eq.get(name), oc.arr(name)
        assert np.array_equal(a, b), "%s: max |diff| %g" % (name, np.max(np.abs(a - b)))
    assert [l[:2] for l in eq.log()] == [l[:2] for l in oc.log()]
