"""GPU parity of the distorted-channel input (an extension): parabolic
inlet, zero-gradient outlet balanced to the inflow, no-slip walls, on
jittered meshes whose cells are randomly renumbered then RCM-reordered.
Bitwise against the oracle on small meshes in every pc order; on the full
128^3 workload, properties that do not depend on size (outflow equals
inflow, finite fields, the pc solve converging)."""
import numpy as np
import pytest

import cfdl
import meshgen
import oracle_lib

pytestmark = pytest.mark.gpu


def make_pair(nx, ny, lx, jit, renum, nsubd, mu, dt):
    m = cfdl.Mesh.box(nx, ny, ny, lx=lx, jitter=jit)
    g = meshgen.box_mesh(nx, ny, ny, lx=lx, jitter=jit)
    if renum:
        m.renumber(renum)
        g = meshgen.renumber(g, renum)
    m.partition(nsubd)
    eq = cfdl.UVWP.channel(m, rho=1.0, mu=mu, dt=dt)
    oc = oracle_lib.OracleCase(g, meshgen.channel_bc(g), nsubd=nsubd, rho=1.0, mu=mu, dt=dt)
    return m, eq, oc


@pytest.mark.parametrize("nx,ny,jit,renum,nsubd,order,gs,steady", [
    (12, 6, 0.0, 0, 4, "reference", "reference", True),
    (10, 8, 0.2, 5, 4, "reference", "reference", False),
    (10, 8, 0.2, 5, 4, "mixed", "colored", True),
    (16, 8, 0.1, 9, 4, "mgcg", "reference", True),
    (16, 8, 0.2, 9, 2, "mgcg-mixed", "colored", True),
    (20, 10, 0.2, 3, 1, "reference", "colored", True),
])
def test_channel_bitwise(nx, ny, jit, renum, nsubd, order, gs, steady):
    m, eq, oc = make_pair(nx, ny, 2.0, jit, renum, nsubd, 1.0 / 200.0,
                          float("inf") if steady else 0.01)
    for x in (eq, oc):
        x.set_relaxation(0.7, 0.3)
        x.set_pc_order(order)
        x.set_gs_order(gs)
    for _ in range(2):
        eq.run(1, 3)
        oc.run(1, 3)
    for name in ["u", "v", "w", "p", "gp", "mip", "phic", "ap", "anb", "b", "bu", "bv", "bw",
                 "d", "dc"]:
        a, b = eq.get(name), oc.arr(name)
        assert np.array_equal(a, b), "%s: max |diff| %g" % (name, np.max(np.abs(a - b)))
    gl, ol = eq.log(), oc.log()
    assert [l[:2] for l in gl] == [l[:2] for l in ol]


def boundary_flux(m, mip, kinds):
    """Net mass flux through the sections of each kind (3 inlet, 4 outlet)."""
    i = m.info()
    ne = i["ne"]
    idx = m.get("ef2nb_idx")
    L = idx[-1] - 1
    fsg = m.get("ef2nb")[L:]
    bs = np.abs(m.get("bs"))
    esec = m.get("esec").reshape(-1, 2)
    out = {}
    for s, name in enumerate(m.names):
        if s == 0 or kinds.get(name) is None:
            continue
        el = np.arange(esec[s][0], esec[s][1] + 1)
        p = bs[el - ne - 1]
        e, lf = p >> 5, p & 31
        fg = fsg[idx[e - 1] + lf - 2]
        out[kinds[name]] = out.get(kinds[name], 0.0) + float(np.sum(mip[fg - 1]))
    return out


def test_channel_128_distorted_renumbered_properties():
    """The distorted-channel workload at full size: 128^3 cells jittered by
    +-0.2 h, randomly renumbered then RCM-ordered, Re = 200, steady, the
    mixed-precision multigrid pc solve and colour-order momentum solves.
    After a few outer iterations the outflow equals the inflow (the
    parabolic inlet of peak 1 carries 4/9 of the section), every field is
    finite, and every pc solve reached its tenfold reduction."""
    n = 128
    m = cfdl.Mesh.box(n, jitter=0.2, seed=20261202).renumber(20261202).partition(4)
    eq = cfdl.UVWP.channel(m, rho=1.0, mu=1.0 / 200.0, dt=float("inf"))
    eq.set_relaxation(0.7, 0.3)
    eq.set_pc_order("mgcg-mixed")
    eq.set_gs_order("colored")
    eq.run(1, 4)
    q = boundary_flux(m, eq.get("mip"), {"west": 3, "east": 4})
    assert q[3] == pytest.approx(-4.0 / 9.0, rel=1e-3)
    assert abs(q[3] + q[4]) <= 1e-12
    for name in ("u", "v", "w", "p", "mip"):
        assert np.all(np.isfinite(eq.get(name))), name
    pc = [l for l in eq.log() if l[0] == "pc"]
    assert len(pc) == 4 and all(l[1] < 100 and l[3] <= l[2] / 10 for l in pc)


def test_channel_fp32_assembly_bitwise():
    """The channel boundaries in the FP32 assembly, against the oracle."""
    m, eq, oc = make_pair(12, 8, 2.0, 0.2, 7, 4, 1.0 / 200.0, float("inf"))
    for x in (eq, oc):
        x.set_relaxation(0.7, 0.3)
        x.set_pc_order("mgcg-mixed")
        x.set_gs_order("colored")
        x.set_assembly_precision(32)
    eq.run(1, 4)
    oc.run(1, 4)
    for name in ["u", "v", "w", "p", "mip", "ap", "anb", "b", "bu", "dc"]:
        assert np.array_equal(eq.get(name), oc.arr(name)), name
