"""GPU parity: the HIP path against the oracle, bit for bit.

Every field the reference's solve_uvwp touches is compared with the oracle
(the CPU restatement pinned to the reference build) after whole outer
iterations and routine by routine.  The residual logs must have the same
solver iteration counts; residual values come from fixed-order parallel
reductions and agree to 1e-12 relative.
"""
import numpy as np
import pytest

import cfdl
import meshgen
import oracle_lib

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "w", "p", "u0", "v0", "w0", "gu", "gv", "gw", "gp", "mip", "mip0",
          "bu", "bv", "bw", "d", "dc", "ap", "b", "anb", "phic"]


def make_pair(n, jitter=0.0, nsubd=4, rho=1.0, mu=0.01, dt=0.01, nit=100):
    mesh = cfdl.Mesh.box(n, jitter=jitter).partition(nsubd)
    eq = cfdl.UVWP.cavity(mesh, rho=rho, mu=mu, dt=dt, nit=nit)
    g = meshgen.box_mesh(n, jitter=jitter)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=nsubd, rho=rho, mu=mu, dt=dt, nit=nit)
    return eq, oc


def assert_fields_equal(eq, oc, names=FIELDS):
    for name in names:
        a, b = eq.get(name), oc.arr(name)
        if name == "gpc":
            a, b = a[:3 * oc.c.ne], b[:3 * oc.c.ne]
        assert np.array_equal(a, b), "%s: max |diff| %g" % (name, np.max(np.abs(a - b)))


def assert_logs_match(gl, ol):
    assert len(gl) == len(ol)
    for a, b in zip(gl, ol):
        assert a[0] == b[0] and a[1] == b[1], (a, b)
        for x, y in zip(a[2:], b[2:]):
            assert x == pytest.approx(y, rel=1e-12, abs=1e-300), (a, b)


@pytest.mark.parametrize("n,jitter,nsubd,ntstep,ncoef", [
    (4, 0.0, 4, 1, 2),
    (8, 0.0, 4, 2, 3),
    (6, 0.2, 4, 2, 3),
    (8, 0.2, 1, 2, 2),
    (10, 0.15, 4, 1, 3),
    (9, 0.0, 4, 1, 2),
])
def test_run_bitwise(n, jitter, nsubd, ntstep, ncoef):
    eq, oc = make_pair(n, jitter, nsubd)
    eq.run(ntstep, ncoef)
    oc.run(ntstep, ncoef)
    assert_fields_equal(eq, oc)
    assert_logs_match(eq.log(), oc.log())
