"""Batches (continuous batching of independent cases, an extension): every
case run inside a batch -- grouped launches over cases of different sizes
and types, attached and detached between runs -- equals the same case run
alone through cfdl_run, bit for bit (fields and solver logs).  The cases
alone are pinned to the oracle by the channel, multigrid and FP32-assembly
tests; one small batched case is also compared with the oracle directly."""
import numpy as np
import pytest

import cfdl
import meshgen
import oracle_lib

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "w", "p", "u0", "mip", "mip0", "gp", "gu", "gpc", "phic", "ap", "anb", "b",
          "bu", "bv", "bw", "d", "dc"]


def make_case(kind, n, re, jitter=0.1, renum=0, order="mgcg-mixed", asm=32):
    m = cfdl.Mesh.box(n, jitter=jitter)
    if renum:
        m.renumber(renum)
    m.partition(1)
    mk = cfdl.UVWP.channel if kind == "channel" else cfdl.UVWP.cavity
    eq = mk(m, rho=1.0, mu=1.0 / re, dt=float("inf"))
    eq.set_relaxation(0.7, 0.3)
    eq.set_gs_order("colored")
    eq.set_pc_order(order)
    eq.set_assembly_precision(asm)
    return eq


def same(a, b):
    for name in FIELDS:
        x, y = a.get(name), b.get(name)
        assert np.array_equal(x, y), "%s: max |diff| %g" % (name, np.nanmax(np.abs(x - y)))
    assert a.log() == b.log()


SPECS = [("cavity", 16, 100.0, 0.0, 0), ("channel", 24, 200.0, 0.2, 7),
         ("cavity", 40, 1000.0, 0.1, 0), ("channel", 16, 150.0, 0.1, 3),
         ("cavity", 24, 400.0, 0.2, 11), ("channel", 32, 1000.0, 0.0, 0)]


def test_batch_equals_alone_with_cases_joining_and_leaving():
    """Six cases (cavity and channel, 16^3 .. 40^3, Re 100 .. 1000,
    jittered, some renumbered): two start, the others join after one and two
    runs, two leave early; every case equals its run alone with the same
    sequence of run(1, 2) calls."""
    alone = [make_case(*s) for s in SPECS]
    mixed = [make_case(*s) for s in SPECS]
    b = cfdl.Batch()
    joins = {0: [0, 1], 1: [2, 3], 2: [4, 5]}
    leaves = {3: [1, 3]}
    runs = [0] * len(SPECS)
    inside = set()
    for r in range(5):
        for k in joins.get(r, []):
            b.attach(mixed[k])
            inside.add(k)
        for k in leaves.get(r, []):
            b.detach(mixed[k])
            inside.discard(k)
        b.run(1, 2)
        for k in inside:
            runs[k] += 1
    b.close()
    for k, eq in enumerate(alone):
        for _ in range(runs[k]):
            eq.run(1, 2)
    for k in range(len(SPECS)):
        assert runs[k] >= 2
        same(mixed[k], alone[k])
        assert any(l[0] == "pc" and l[1] > 0 for l in mixed[k].log())


def test_batch_serving_sizes_fp64():
    """The serving sizes' end points (16^3 and 96^3, cavity and channel) with
    the FP64 assembly and the FP64 multigrid preconditioner."""
    specs = [("cavity", 96, 300.0, 0.0, 0), ("channel", 16, 700.0, 0.2, 5),
             ("channel", 96, 120.0, 0.1, 0), ("cavity", 16, 1000.0, 0.0, 0)]
    alone = [make_case(*s, order="mgcg", asm=64) for s in specs]
    mixed = [make_case(*s, order="mgcg", asm=64) for s in specs]
    b = cfdl.Batch()
    for eq in mixed:
        b.attach(eq)
    b.run(1, 3)
    b.close()
    for eq in alone:
        eq.run(1, 3)
    for x, y in zip(mixed, alone):
        same(x, y)


def test_batch_block_size_does_not_change_results():
    """Lane counts 64 and 1024 and the tuned one give the same bits."""
    res = []
    for blk in (64, 1024, 0):
        eqs = [make_case(*s) for s in SPECS[:3]]
        b = cfdl.Batch(block=blk)
        for eq in eqs:
            b.attach(eq)
        b.run(1, 3)
        assert b.block() in (64, 128, 256, 512, 1024)
        b.close()
        res.append(eqs)
    for k in range(3):
        same(res[0][k], res[1][k])
        same(res[0][k], res[2][k])


def test_batch_non_finite_case_leaves_the_others_unchanged():
    """A NaN planted in one case's velocity (from the host) makes that
    case's residuals non-finite; the other cases' bits do not change."""
    alone = [make_case(*s) for s in SPECS[:4]]
    mixed = [make_case(*s) for s in SPECS[:4]]
    bad = mixed[2]
    u = bad.get("u")
    u[5] = np.nan
    bad.set("u", u)
    b = cfdl.Batch()
    for eq in mixed:
        b.attach(eq)
    b.run(1, 2)
    b.close()
    for k in (0, 1, 3):
        alone[k].run(1, 2)
        same(mixed[k], alone[k])
    assert not all(np.isfinite(l[2]) for l in bad.log())


def test_batched_case_equals_the_oracle():
    """A 10^3 jittered, renumbered channel with FP32 assembly and the FP32
    multigrid preconditioner, in a batch with a cavity, against the
    oracle's restatement."""
    g = meshgen.box_mesh(10, jitter=0.15)
    g = meshgen.renumber(g, 4)
    oc = oracle_lib.OracleCase(g, meshgen.channel_bc(g), nsubd=1, rho=1.0, mu=1.0 / 200.0,
                               dt=float("inf"))
    oc.set_relaxation(0.7, 0.3)
    oc.set_gs_order("colored")
    oc.set_pc_order("mgcg-mixed")
    oc.set_assembly_precision(32)
    eq = make_case("channel", 10, 200.0, 0.15, 4)
    other = make_case("cavity", 12, 500.0)
    b = cfdl.Batch()
    b.attach(other)
    b.attach(eq)
    for _ in range(2):
        b.run(1, 2)
        oc.run(1, 2)
    b.close()
    for name in ["u", "v", "w", "p", "mip", "gp", "phic", "ap", "anb", "b", "d", "dc"]:
        assert np.array_equal(eq.get(name), oc.arr(name)), name
    assert [l[:2] for l in eq.log()] == [l[:2] for l in oc.log()]


def test_batch_rejects_cases_it_cannot_run():
    b = cfdl.Batch()
    eq = make_case("cavity", 8, 100.0)
    eq.set_gs_order("reference")
    with pytest.raises(cfdl.CfdlError):
        b.attach(eq)
    eq.set_gs_order("colored")
    b.attach(eq)
    with pytest.raises(cfdl.CfdlError):
        b.attach(eq)  # twice
    other = make_case("cavity", 8, 100.0, asm=64)
    with pytest.raises(cfdl.CfdlError):
        b.attach(other)  # another assembly precision
    eq.set_pc_order("mixed")
    with pytest.raises(cfdl.CfdlError):
        b.run(1, 1)  # the case's pc order changed while attached
    eq.set_pc_order("mgcg-mixed")
    b.run(1, 1)
    b.close()
    eq.run(1, 1)  # detached by close: runs on its own stream again


def test_batch_graphs_equal_direct_launches_after_every_run():
    """The captured graphs of the iteration's segments (replayed many times
    per run, recaptured when a case joins) against the same batch launched
    directly and against every case alone, compared after every run."""
    specs = SPECS[:4]
    alone = [make_case(*s) for s in specs]
    with_graphs = [make_case(*s) for s in specs]
    direct = [make_case(*s) for s in specs]
    bg, bd = cfdl.Batch(), cfdl.Batch(graphs=False)
    for k in range(3):
        bg.attach(with_graphs[k])
        bd.attach(direct[k])
    for r in range(4):
        if r == 2:
            bg.attach(with_graphs[3])
            bd.attach(direct[3])
        bg.run(1, 2)
        bd.run(1, 2)
        for k in range(len(specs)):
            if k < 3 or r >= 2:
                alone[k].run(1, 2)
        for k in range(len(specs)):
            same(with_graphs[k], direct[k])
            same(with_graphs[k], alone[k])
    captures, launches = bg.graph_stats()
    assert captures >= 8 and launches > 2 * captures
    assert bd.graph_stats() == (0, 0)
    bg.close()
    bd.close()
