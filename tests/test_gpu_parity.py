"""GPU parity: the HIP path against the oracle, bit for bit.

Every field the reference's solve_uvwp touches is compared with the oracle
(the CPU restatement pinned to the reference build) after whole outer
iterations and routine by routine.  The residual logs must have the same
solver iteration counts; residual values come from fixed-order parallel
reductions and agree to 1e-12 relative.
"""
import hashlib
import json
import os

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
        if name == "h0":  # construct_energy leaves the halo part of phi0 unset
            a, b = a[:oc.c.ne], b[:oc.c.ne]
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
    (12, 0.1, 2, 1, 3),
    (10, 0.1, 8, 2, 2),
    (16, 0.0, 16, 1, 2),
])
def test_run_bitwise(n, jitter, nsubd, ntstep, ncoef):
    """Whole outer iterations against the oracle on uniform and jittered
    meshes, with 1, 2, 4, 8 and 16 subdomains (the oracle's own runs with 2
    and 8 subdomains are checked against the reference)."""
    eq, oc = make_pair(n, jitter, nsubd)
    eq.run(ntstep, ncoef)
    oc.run(ntstep, ncoef)
    assert_fields_equal(eq, oc)
    assert_logs_match(eq.log(), oc.log())


def test_relaxation_config_mode_bitwise():
    """Config mode (alpha_u=0.7, alpha_p=0.3) against the oracle's same
    extension; the reference has no relaxation, so this pins the GPU path to
    the restatement, and alpha=1 pins the restatement to the reference."""
    eq, oc = make_pair(8, 0.1, 4, mu=0.0025)
    eq.set_relaxation(0.7, 0.3)
    oc.set_relaxation(0.7, 0.3)
    eq.run(2, 3)
    oc.run(2, 3)
    assert_fields_equal(eq, oc)
    assert_logs_match(eq.log(), oc.log())


@pytest.mark.parametrize("mode", ["tiled", "level"])
def test_32cube_one_step(monkeypatch, mode):
    monkeypatch.setenv("CFDL_GS_MODE", mode)
    eq, oc = make_pair(32, 0.0, 4, mu=0.0025)
    eq.run(1, 2)
    oc.run(1, 2)
    assert_fields_equal(eq, oc, ["u", "v", "w", "p", "mip", "gp", "phic"])
    assert_logs_match(eq.log(), oc.log())


def test_nonuniform_properties_bitwise():
    """Per-cell density and viscosity, random in [0.8, 1.2] and [0.005, 0.015],
    on a jittered mesh.  The face values of rho and mu, the Rhie-Chow and
    pressure-correction coefficients and the arithmetic-mean density of
    update_uvwp weight neighbouring cells differently
    (src/equations/mod_uvwp.f90:183-226, 334, 411, 475-485), which uniform
    properties cannot show.  Fields and solver lines match the oracle."""
    eq, oc = make_pair(9, 0.15, 4)
    rng = np.random.default_rng(11)
    ne = oc.c.ne
    rho = 0.8 + 0.4 * rng.random(ne)
    mu = 0.005 + 0.01 * rng.random(ne)
    eq.set_properties(rho, mu)
    oc.arr("rho")[:] = rho
    oc.arr("mu")[:] = mu
    eq.run(2, 3)
    oc.run(2, 3)
    assert_fields_equal(eq, oc)
    assert_logs_match(eq.log(), oc.log())


@pytest.mark.parametrize("n,jit,nsubd,sym", [(9, 0.15, 4, ("north", "bottom")), (8, 0.2, 1, ("east",)),
                                             (12, 0.1, 8, ("south", "west", "north"))])
def test_symmetry_sections_bitwise(n, jit, nsubd, sym):
    """Symmetry sections (mod_uvwp.f90:517-546): the halo velocity is the
    inner cell's reflected about the face normal, its pressure the inner
    cell's, and the momentum assembly drops the wall-shear term of
    'zero_flux' faces (:266).  Jittered meshes put the normals off the axes.
    Fields and solver lines match the oracle, which matches the reference
    build bit for bit on the same binding."""
    mesh = cfdl.Mesh.box(n, jitter=jit).partition(nsubd)
    eq = cfdl.UVWP.cavity(mesh, rho=1.0, mu=0.01, dt=0.02)
    for name in sym:
        eq.set_section_bc(mesh, name, "symmetry")
    g = meshgen.box_mesh(n, jitter=jit)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g, sym), nsubd=nsubd, rho=1.0, mu=0.01, dt=0.02)
    eq.run(2, 3)
    oc.run(2, 3)
    assert_fields_equal(eq, oc)
    assert_logs_match(eq.log(), oc.log())


def test_set_section_bc_rejects_bad_input():
    mesh = cfdl.Mesh.box(4).partition(1)
    eq = cfdl.UVWP.cavity(mesh)
    with pytest.raises(cfdl.CfdlError):
        eq.set_section_bc(mesh, "cells", "symmetry")
    with pytest.raises(cfdl.CfdlError):
        eq.set_section_bc(mesh, "north", "outflow")  # not a BC kind
    with pytest.raises(cfdl.CfdlError):
        eq.set_bc_kind(1, 5, "symmetry")  # cells, not halos


def test_64cube_jittered_relaxed_bitwise():
    """A 64^3 box with interior vertices jittered by up to 0.2 h, Re = 400 and
    the 0.7/0.3 relaxation of the 128^3 benchmark, two outer iterations.  At
    this size the pressure-correction sweeps run full 16 x 8 x 4 tiles with
    32-level batches, and the momentum sweeps their own tiles; every field
    equals the oracle's bits."""
    eq, oc = make_pair(64, 0.2, 4, mu=0.0025)
    eq.set_relaxation(0.7, 0.3)
    oc.set_relaxation(0.7, 0.3)
    eq.run(1, 2)
    oc.run(1, 2)
    assert_fields_equal(eq, oc, ["u", "v", "w", "p", "mip", "gp", "phic", "ap", "anb", "b"])
    assert_logs_match(eq.log(), oc.log())


@pytest.mark.parametrize("n,jit,relax", [(8, 0.1, False), (12, 0.15, True), (32, 0.0, True)])
def test_colored_pc_order_bitwise(n, jit, relax):
    """The colour order of the pressure-correction smoother (an extension):
    one launch per colour on the GPU, against the oracle's sequential
    Gauss-Seidel that visits the cells in the same order.  Both colour every
    subdomain greedily in its local order.  Fields and solver lines match
    bit for bit."""
    eq, oc = make_pair(n, jit, 4, mu=0.0025)
    if relax:
        eq.set_relaxation(0.7, 0.3)
        oc.set_relaxation(0.7, 0.3)
    eq.set_pc_order("colored")
    oc.set_pc_order("colored")
    eq.run(2, 3)
    oc.run(2, 3)
    assert_fields_equal(eq, oc, ["u", "v", "w", "p", "mip", "gp", "phic"])
    assert_logs_match(eq.log(), oc.log())


def _ocall(oc, fn, *args):
    getattr(oracle_lib.lib(), fn)(oc.p, *args)


@pytest.mark.parametrize("scale", [0.0, 1e10])
@pytest.mark.parametrize("order", ["reference", "colored", "mixed"])
def test_pc_solve_stopping_early_bitwise(order, scale):
    """The 4-subdomain pc solve stopping before its iteration cap (quirk Q1
    keeps every ordinary solve at the cap).  With the face fluxes scaled by
    1e10 the accumulated residual passes the tenfold test early (iteration
    30 in the reference order, 50 in the colour orders); with zero fluxes
    the right-hand side is zero and the solve stops on entry.  The blocks
    queued after the stop (halo exchanges that do not read the active flag)
    are covered by the comparison of phic."""
    eq, oc = make_pair(8, 0.1, 4)
    eq.set_pc_order(order)
    oc.set_pc_order(order)
    eq.run(1, 1)
    oc.run(1, 1)
    eq.set("mip", eq.get("mip") * scale)
    om = oc.arr("mip")
    om *= scale
    assert_fields_equal(eq, oc, ["mip"])
    eq.calc_coef_p()
    _ocall(oc, "or_calc_coef_p")
    eq.set_a_0("phic")
    oc.arr("phic")[:] = 0.0
    eq.clear_log()
    n0 = len(oc.log())
    eq.solve("pc")
    oracle_lib.lib().or_solve(oc.p, b"pc", oc.c.phic)
    assert_fields_equal(eq, oc, ["phic"])
    gl, ol = eq.log(), oc.log()[n0:]
    assert len(gl) == len(ol) == 1
    assert_logs_match(gl, ol)
    assert gl[0][1] == (0 if scale == 0.0 else 30 if order == "reference" else 50)


def test_routines_one_by_one():
    """Each C-ABI routine against the matching oracle routine, from a state
    reached after two outer iterations (non-trivial fields)."""
    eq, oc = make_pair(7, 0.2, 4)
    eq.run(1, 2)
    oc.run(1, 2)
    C = __import__("ctypes")
    eq.update_boundaries()
    _ocall(oc, "or_update_boundaries")
    assert_fields_equal(eq, oc, ["u", "v", "w", "p", "mip"])
    eq.calc_coef_uvw()
    _ocall(oc, "or_calc_coef_uvw")
    assert_fields_equal(eq, oc, ["ap", "anb", "bu", "bv", "bw", "d", "dc"])
    eq.solve_gs("u", "u", "bu")
    L = oracle_lib.lib()
    c = oc.c
    L.or_solve_gs.argtypes = [C.POINTER(oracle_lib.OrCase), C.c_char_p, oracle_lib.PD, oracle_lib.PD,
                              oracle_lib.PD, oracle_lib.PD, oracle_lib.PI, oracle_lib.PI,
                              C.c_int, C.c_int, C.c_int, C.c_int]
    L.or_solve_gs(oc.p, b"u", c.u, c.ap, c.anb, c.bu, c.ef2nb_idx, c.ef2nb, c.ne, c.nf, c.nbf, c.nit)
    assert_fields_equal(eq, oc, ["u"])
    eq.calc_grad("u", "gu")
    L.or_calc_grad(c.u, c.gu, c.xc, c.yc, c.zc, c.ef2nb_idx, c.ef2nb, c.ne, c.nf, c.nbf)
    assert_fields_equal(eq, oc, ["gu"])
    eq.calc_mip(True)
    _ocall(oc, "or_calc_mip", 1)
    assert_fields_equal(eq, oc, ["mip"])
    eq.calc_coef_p()
    _ocall(oc, "or_calc_coef_p")
    assert_fields_equal(eq, oc, ["ap", "anb", "b"])
    eq.set_a_0("phic")
    oc.arr("phic")[:] = 0.0
    eq.solve("pc")
    L.or_solve(oc.p, b"pc", c.phic)
    assert_fields_equal(eq, oc, ["phic"])
    eq.adjust_pc()
    L.or_adjust_pc(oc.p, c.phic, oc.arr("phic")[0])
    assert_fields_equal(eq, oc, ["phic"])
    eq.calc_grad("phic", "gpc")
    L.or_calc_grad(c.phic, c.gpc, c.xc, c.yc, c.zc, c.ef2nb_idx, c.ef2nb, c.ne, c.nf, c.nbf)
    assert_fields_equal(eq, oc, ["gpc"])
    eq.update_uvwp()
    L.or_update_uvwp(oc.p, c.phic)
    assert_fields_equal(eq, oc, ["p", "gp", "mip"])
    eq.update_time()
    _ocall(oc, "or_update_time")
    assert_fields_equal(eq, oc, ["u0", "v0", "w0", "mip0"])
    assert_logs_match(eq.log(), oc.log())


def test_repeatable():
    """Two runs of the same case give identical bits (fixed-order reductions,
    no atomics)."""
    a, _ = make_pair(12, 0.1, 4)
    b, _ = make_pair(12, 0.1, 4)
    a.run(1, 2)
    b.run(1, 2)
    for name in ("u", "p", "mip"):
        assert np.array_equal(a.get(name), b.get(name))
    assert a.log() == b.log()


def test_set_get_roundtrip_and_errors():
    eq, _ = make_pair(4)
    x = np.arange(eq.sizes["u"], dtype=np.float64)
    eq.set("u", x)
    assert np.array_equal(eq.get("u"), x)
    with pytest.raises(cfdl.CfdlError):
        eq.set("u", x[:-1])
    with pytest.raises(cfdl.CfdlError):
        eq.set_relaxation(0.0, 1.0)


@pytest.mark.parametrize("tile,mode,lds,span,pipe,waves", [
    ("2", "tiled", "1", "16", "1", "4"), ("3", "tiled", "1", "3", "1", "4"),
    ("3", "tiled", "0", "16", "1", "4"), ("8", "tiled", "1", "16", "1", "4"),
    ("8", "tiled", "1", "16", "0", "4"), ("8", "level", "1", "16", "1", "4"),
    ("12", "tiled", "1", "32", "1", "4"), ("4x2x3", "tiled", "1", "5", "1", "4"),
    ("16x8x4", "tiled", "1", "28", "1", "4"), ("8", "tiled", "1", "16", "1", "1"),
    ("8", "tiled", "1", "16", "0", "1"), ("12", "tiled", "1", "32", "1", "1"),
    ("16x8x4", "tiled", "1", "28", "1", "1"), ("12", "tiled", "1", "63", "1", "4")])
def test_sweep_schedules_bitwise(monkeypatch, tile, mode, lds, span, pipe, waves):
    """The pressure-correction sweeps run as one launch per dependency level
    or as space-time tiles launched batch by batch (in LDS, or from global
    memory with CFDL_TILE_LDS=0).  In LDS, level segments of at most 64 cells
    take the register-pipelined walk unless CFDL_TILE_PIPE=0, and a batch is
    gathered and written back by four wavefronts (whose idle three prefetch
    the next launch's batches during the walk), or by one with
    CFDL_TILE_WAVES=1.  Small tiles and spans put many tile boundaries into
    small meshes; tile 12 with span 32 needs more than 64 KB of LDS per batch
    and has wider segments, and with span 63 a batch holds a subdomain's
    whole sweep (the walk's most segments); tiles may have a different edge
    per axis.  Every schedule gives the oracle's bits."""
    monkeypatch.setenv("CFDL_TILE", tile)
    monkeypatch.setenv("CFDL_GS_MODE", mode)
    monkeypatch.setenv("CFDL_TILE_LDS", lds)
    monkeypatch.setenv("CFDL_TILE_SPAN", span)
    monkeypatch.setenv("CFDL_TILE_PIPE", pipe)
    monkeypatch.setenv("CFDL_TILE_WAVES", waves)
    eq, oc = make_pair(12, 0.15, 4)
    eq.run(1, 2)
    oc.run(1, 2)
    assert_fields_equal(eq, oc, ["u", "v", "w", "p", "mip", "gp", "phic"])
    assert_logs_match(eq.log(), oc.log())


def _scale_rows(oc, rows, s, rhs):
    """Multiplies rows of the shared system by the power of two s: ap, the
    row's anb entries and the right-hand sides (names in rhs).  The solution
    of every row is unchanged, bit for bit."""
    idx = oc.arr("ef2nb_idx")
    ap, anb = oc.arr("ap"), oc.arr("anb")
    for r in rows:
        ap[r] *= s
        anb[idx[r] - 1:idx[r + 1] - 1] *= s
        for name in rhs:
            oc.arr(name)[r] *= s


def _tiny_rows(oc, rows, names, seed):
    """Decouples rows from their neighbours (anb = 0) and gives them values
    of 2^-1000 to 2^-1060 in the arrays named: their Gauss-Seidel updates
    then divide tiny numbers in every sweep, with quotients near and below
    the least normal double, and their last bits survive to the end."""
    rng = np.random.default_rng(seed)
    idx, anb = oc.arr("ef2nb_idx"), oc.arr("anb")
    for r in rows:
        anb[idx[r] - 1:idx[r + 1] - 1] = 0.0
        for name in names:
            oc.arr(name)[r] = np.ldexp(1.0 + rng.random(), -1000 - int(rng.integers(61)))


@pytest.mark.parametrize("case", ["tiny_rows", "small_rows", "large_rows"])
def test_pc_sweep_division_fallback(monkeypatch, case):
    """The LDS sweep divides by precomputed reciprocals of ap and sor only
    where that is exact: dividends of frexp exponent within +-750 and
    divisors within [2^-64, 2^64].  A batch that meets anything else is
    recomputed with the hardware division: here decoupled rows with
    right-hand sides near 2^-1030 (where about one chained reciprocal
    quotient in ten is not correctly rounded) and rows whose ap is scaled
    out of the divisor range.  (Dividends beyond 2^750 cannot reach a sweep:
    the initial residual, a root mean square, overflows and the solver
    stops before it.)  The
    pressure-correction solve stays bitwise equal to the oracle's.  Tile 3
    puts such rows and ordinary ones into different batches."""
    monkeypatch.setenv("CFDL_TILE", "3")
    eq, oc = make_pair(10, 0.1, 4)
    eq.run(1, 2)
    oc.run(1, 2)
    eq.calc_mip(True)
    _ocall(oc, "or_calc_mip", 1)
    eq.calc_coef_p()
    _ocall(oc, "or_calc_coef_p")
    ne = oc.c.ne
    if case == "tiny_rows":
        _tiny_rows(oc, range(3, ne, 7), ["b"], 5)
    else:
        _scale_rows(oc, range(3, ne, 7), 2.0 ** (-100 if case == "small_rows" else 100), ["b"])
    for name in ("ap", "anb", "b"):
        eq.set(name, oc.arr(name))
    eq.set_a_0("phic")
    oc.arr("phic")[:] = 0.0
    eq.solve("pc")
    oracle_lib.lib().or_solve(oc.p, b"pc", oc.c.phic)
    assert_fields_equal(eq, oc, ["phic"])
    assert np.count_nonzero(oc.arr("phic")[:ne]) > ne // 2


@pytest.mark.parametrize("case", ["tiny_rows", "small_rows"])
def test_momentum_sweep_division_fallback(monkeypatch, case):
    """The three momentum solves share one LDS walk; its reciprocal division
    falls back to the hardware division batch by batch as the pressure-
    correction sweep does (decoupled rows with values near 2^-1030, or rows
    scaled by 2^-100).  u, v and w equal the oracle's solve_gs results."""
    import ctypes as C
    monkeypatch.setenv("CFDL_TILE", "3")
    eq, oc = make_pair(10, 0.1, 4)
    eq.run(1, 2)
    oc.run(1, 2)
    eq.update_boundaries()
    _ocall(oc, "or_update_boundaries")
    eq.calc_coef_uvw()
    _ocall(oc, "or_calc_coef_uvw")
    if case == "tiny_rows":
        _tiny_rows(oc, range(2, oc.c.ne, 5), ["bu", "bv", "bw", "u", "v", "w"], 6)
    else:
        _scale_rows(oc, range(2, oc.c.ne, 5), 2.0 ** -100, ["bu", "bv", "bw"])
    for name in ("ap", "anb", "bu", "bv", "bw", "u", "v", "w"):
        eq.set(name, oc.arr(name))
    eq.solve_gs_uvw()
    L, c = oracle_lib.lib(), oc.c
    L.or_solve_gs.argtypes = [C.POINTER(oracle_lib.OrCase), C.c_char_p, oracle_lib.PD, oracle_lib.PD,
                              oracle_lib.PD, oracle_lib.PD, oracle_lib.PI, oracle_lib.PI,
                              C.c_int, C.c_int, C.c_int, C.c_int]
    for phi, b in (("u", "bu"), ("v", "bv"), ("w", "bw")):
        L.or_solve_gs(oc.p, phi.encode(), getattr(c, phi), c.ap, c.anb, getattr(c, b),
                      c.ef2nb_idx, c.ef2nb, c.ne, c.nf, c.nbf, c.nit)
    assert_fields_equal(eq, oc, ["u", "v", "w"])


@pytest.mark.parametrize("n,jit,nsubd,nt,nc", [(6, 0.0, 4, 2, 2), (8, 0.2, 4, 1, 3), (7, 0.1, 1, 2, 2)])
def test_energy_equation_bitwise(n, jit, nsubd, nt, nc):
    """The energy equation solved after every solve_uvwp (src/equations/
    mod_energy.f90:58-80) against the oracle's restatement, which is pinned to
    the reference's own solve_energy: temperature, enthalpy, gradients, the
    coefficient arrays it leaves and the 'e' solver lines."""
    eq, oc = make_pair(n, jit, nsubd)
    eq.run(nt, nc, energy=True)
    oc.run(nt, nc, energy=True)
    assert_fields_equal(eq, oc, ["t", "gt", "h", "h0", "gh", "u", "v", "w", "p", "ap", "anb", "b"])
    assert_logs_match(eq.log(), oc.log())


def test_energy_routines_one_by_one():
    """BC setters, coefficients with non-uniform thermal properties, and the
    temperature update, each against the oracle's routine on the same state."""
    eq, oc = make_pair(6, 0.15, 4)
    eq.run(1, 2)
    oc.run(1, 2)
    rng = np.random.default_rng(3)
    ne = oc.c.ne
    cp, tc = 900.0 + 200.0 * rng.random(ne), 2.0 + 6.0 * rng.random(ne)
    eq.set_thermal(cp, tc)
    oc.arr("cp")[:] = cp
    oc.arr("tc")[:] = tc
    eq.construct_energy()
    oracle_lib.lib().or_construct_energy(oc.p)
    eq.update_boundaries_energy()
    oracle_lib.lib().or_update_boundaries_energy(oc.p)
    assert_fields_equal(eq, oc, ["t", "h"])
    eq.solve_energy()
    oracle_lib.lib().or_solve_energy(oc.p)
    assert_fields_equal(eq, oc, ["t", "gt", "h", "gh", "ap", "anb", "b"])


@pytest.mark.parametrize("n,jit,relax,nsubd,nit", [(8, 0.1, False, 4, 100), (12, 0.15, True, 4, 100),
                                                    (32, 0.0, True, 4, 100), (10, 0.1, False, 2, 25),
                                                    (12, 0.0, True, 8, 7)])
def test_mixed_pc_order_bitwise(n, jit, relax, nsubd, nit):
    """The mixed-precision colour order (an extension): FP32 colour sweeps
    on the correction system, FP64 defect correction at every residual check.
    The oracle restates the same single-precision operations in C, so fields
    and solver lines match bit for bit.  An iteration cap that is not a
    multiple of ten (25, 7) leaves a correction after the last check (none at
    all with 7), which the final update adds."""
    eq, oc = make_pair(n, jit, nsubd, mu=0.0025, nit=nit)
    if relax:
        eq.set_relaxation(0.7, 0.3)
        oc.set_relaxation(0.7, 0.3)
    eq.set_pc_order("mixed")
    oc.set_pc_order("mixed")
    eq.run(2, 3)
    oc.run(2, 3)
    assert_fields_equal(eq, oc, ["u", "v", "w", "p", "mip", "gp", "phic"])
    assert_logs_match(eq.log(), oc.log())


def test_serial_assembly_kernels_bitwise(monkeypatch):
    """The one-thread-per-cell assembly kernels (k_coef_uvw, k_coef_p), which
    meshes with other row lengths would use, selected with CFDL_COEF_SERIAL=1
    when the context is created: fields and solver lines match the oracle."""
    monkeypatch.setenv("CFDL_COEF_SERIAL", "1")
    eq, oc = make_pair(9, 0.15, 4)
    eq.set_relaxation(0.7, 0.3)
    oc.set_relaxation(0.7, 0.3)
    eq.run(2, 3)
    oc.run(2, 3)
    assert_fields_equal(eq, oc)
    assert_logs_match(eq.log(), oc.log())


@pytest.mark.parametrize("n,jit,nsubd,pc", [(8, 0.1, 4, "reference"), (12, 0.15, 4, "mixed"),
                                            (10, 0.1, 1, "reference"), (32, 0.0, 4, "colored")])
def test_colored_gs_order_bitwise(n, jit, nsubd, pc):
    """The colour order of solve_gs (an extension): one launch per colour of
    the global system for the momentum solves (and, with one subdomain, the
    pc solve), against the oracle's sequential solve_gs in the same order;
    alone and with the pc smoother's orders.  Fields and solver lines match."""
    eq, oc = make_pair(n, jit, nsubd, mu=0.0025)
    eq.set_relaxation(0.7, 0.3)
    oc.set_relaxation(0.7, 0.3)
    eq.set_gs_order("colored")
    oc.set_gs_order("colored")
    eq.set_pc_order(pc)
    oc.set_pc_order(pc)
    eq.run(2, 3)
    oc.run(2, 3)
    assert_fields_equal(eq, oc, ["u", "v", "w", "p", "mip", "gp", "phic"])
    assert_logs_match(eq.log(), oc.log())


def test_serving_alone_vs_mix():
    """Continuous batching: cavity and channel requests of different sizes
    arrive while others run, join the batch, and leave when their scaled
    residual reaches tol or after max_outer iterations.  Each result (and
    its iteration count) equals the same request solved alone on a fresh
    context in the same chunks, with room to spare and with a budget of two
    contexts (admissions wait, idle contexts are closed)."""
    from cfdl import serving
    R = serving.Request
    # four slots for six requests: the last two join when others leave
    reqs = [R(0, 0.0, "cavity", 16, 150.0), R(1, 0.0, "channel", 24, 400.0),
            R(2, 0.0, "cavity", 16, 900.0), R(3, 0.01, "channel", 16, 250.0),
            R(4, 0.02, "cavity", 24, 120.0), R(5, 0.03, "channel", 16, 600.0)]
    kw = dict(tol=3e-2, max_outer=60, chunk=5, keep_fields=True)
    out = serving.serve(reqs, max_cases=4, **kw)
    assert out["solves_per_s"] > 0 and out["p99_s"] >= out["p50_s"] > 0 and not out["failed"]
    tight = serving.footprint(24) + serving.footprint(16)
    out2 = serving.serve(reqs, max_cases=4, hbm_budget=tight, **kw)
    assert out2["contexts_created"] >= out["contexts_created"]
    for r in reqs:
        pool = serving.ContextPool()
        try:
            alone, outer, worst = serving.solve_alone(pool, r, tol=3e-2, max_outer=60, chunk=5)
        finally:
            pool.close()
        for o in (out, out2):
            assert o["outer"][r.idx] == outer and o["scaled_residual"][r.idx] == worst, r.idx
            for k in ("u", "v", "w", "p"):
                assert np.array_equal(o["fields"][r.idx][k], alone[k]), (r.idx, k)


def test_energy_equation_colored_gs_bitwise():
    """The energy solve goes through solve_gs too, so the colour order of
    solve_gs applies to it: momentum and energy in colour order, against the
    oracle in the same order, energy fields and solver lines bit for bit."""
    eq, oc = make_pair(9, 0.15, 4)
    eq.set_gs_order("colored")
    oc.set_gs_order("colored")
    eq.run(2, 2, energy=True)
    oc.run(2, 2, energy=True)
    assert_fields_equal(eq, oc, ["t", "gt", "h", "h0", "gh", "u", "v", "w", "p"])
    assert_logs_match(eq.log(), oc.log())


@pytest.mark.parametrize("dims,lens,jit,nsubd,order", [
    ((7, 5, 9), (1.4, 1.0, 2.0), 0.1, 4, "reference"), ((10, 4, 6), (2.0, 0.5, 1.0), 0.0, 2, "reference"),
    ((12, 6, 10), (1.2, 0.6, 1.0), 0.1, 4, "mixed")])
def test_non_cubic_box_bitwise(dims, lens, jit, nsubd, order):
    """Boxes with unequal cell counts and side lengths (the oracle matches the
    reference build on such boxes): another partition, other tile shapes and
    level structures.  Fields and solver lines match the oracle."""
    mesh = cfdl.Mesh.box(*dims, lx=lens[0], ly=lens[1], lz=lens[2], jitter=jit).partition(nsubd)
    eq = cfdl.UVWP.cavity(mesh, rho=1.0, mu=0.01, dt=0.02)
    g = meshgen.box_mesh(*dims, lx=lens[0], ly=lens[1], lz=lens[2], jitter=jit)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=nsubd, rho=1.0, mu=0.01, dt=0.02)
    eq.set_pc_order(order)
    oc.set_pc_order(order)
    eq.run(2, 3)
    oc.run(2, 3)
    assert_fields_equal(eq, oc)
    assert_logs_match(eq.log(), oc.log())


def test_cavity128_bench_configuration_digests():
    """The benchmark's own configuration (128^3 cavity, Re 400, 4 subdomains,
    under-relaxation 0.7/0.3), two outer iterations: every field's SHA-256
    digest equals the oracle's (tests/golden/make_digest128.py), so the
    headline workload itself is bit-identical, and the solver lines match."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "cavity128_relaxed_digests.json")
    ref = json.load(open(path))
    c = ref["case"]
    mesh = cfdl.Mesh.box(c["n"]).partition(c["nsubd"])
    eq = cfdl.UVWP.cavity(mesh, rho=c["rho"], mu=c["mu"], dt=c["dt"], nit=c["nit"])
    eq.set_relaxation(c["alpha_u"], c["alpha_p"])
    eq.run(c["ntstep"], c["ncoef"])
    for k, h in ref["digests"].items():
        got = hashlib.sha256(np.ascontiguousarray(eq.get(k), dtype="<f8").tobytes()).hexdigest()
        assert got == h, k
    assert_logs_match(eq.log(), [tuple(l) for l in ref["log"]])


def test_cavity128_reference_mode_matches_the_reference_build():
    """The 128^3 cavity in the reference's own algorithm (Re 400, 4
    subdomains, no under-relaxation), two outer iterations: every field's
    SHA-256 digest equals that of the reference build's fields
    (tests/golden/make_digest128_ref.py runs oracle/_ref/refsim_O2), and the
    solver lines agree (iteration counts exactly, residuals to the four
    digits the reference prints)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "cavity128_reference_digests.json")
    ref = json.load(open(path))
    c = ref["case"]
    mesh = cfdl.Mesh.box(c["n"]).partition(c["nsubd"])
    eq = cfdl.UVWP.cavity(mesh, rho=c["rho"], mu=c["mu"], dt=c["dt"], nit=c["nit"])
    eq.run(c["ntstep"], c["ncoef"])
    for k, h in ref["digests"].items():
        got = hashlib.sha256(np.ascontiguousarray(eq.get(k), dtype="<f8").tobytes()).hexdigest()
        assert got == h, k
    log = eq.log()
    assert [(l[0], l[1]) for l in log] == [(l[0], l[1]) for l in ref["log"]]
    for a, b in zip(log, ref["log"]):
        assert np.allclose(a[2:], b[2:], rtol=6e-4, atol=0), (a, b)
