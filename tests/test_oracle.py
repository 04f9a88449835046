"""CPU tests of the oracle: pinned to the reference and to closed forms.

The oracle (oracle/*.c) restates the reference operation by operation.  Here
it is checked (1) bit for bit against golden fixtures written by the
reference itself (tests/golden/make_golden.py runs the reference's Fortran
built from its own sources), (2) live against that build when it is present,
and (3) against what the mathematics fixes: closed cells, exact gradients of
linear fields, the textbook matrix inverse, the published quicksort, the
SIMPLEC identity dc = V/(ap - sum anb) = dt/rho, the Neumann row sums, and a
dense numpy Gauss-Seidel of the same splitting.
"""
import ctypes as C
import glob
import os
import random
import tempfile

import numpy as np
import pytest

import meshgen
import oracle_lib
import refsim_io

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
FIELDS = ["u", "v", "w", "p", "gu", "gv", "gw", "gp", "mip", "mip0", "bu", "bv", "bw", "d", "dc",
          "ap", "b", "anb", "phic"]
SETUP = ["ef2nb_idx", "ef2nb", "s2g", "bs", "xc", "yc", "zc", "vol", "aip", "rip"]


def case_from_params(prm):
    n, jit, nsubd, rho, mu, dt, nt, nc = prm
    n, nsubd, nt, nc = int(n), int(nsubd), int(nt), int(nc)
    g = meshgen.box_mesh(n, jitter=jit)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=nsubd, rho=rho, mu=mu, dt=dt)
    return oc, nt, nc


def es93(x):
    """The reference's residual print format (es9.3e2)."""
    return float("%.3e" % x)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_reference_goldens(path):
    g = np.load(path)
    oc, nt, nc = case_from_params(g["params"])
    for k in SETUP:
        assert np.array_equal(oc.arr(k), g[k]), k
    if "g2gf" in g:
        for k in ("gf2g", "g2gf", "g2gf_idx"):
            assert np.array_equal(oc.arr(k), g[k]), k
    energy = "t" in g
    oc.run(nt, nc, energy=energy)
    for k in FIELDS:
        assert np.array_equal(oc.arr(k), g[k]), k
    assert np.array_equal(oc.arr("gpc")[:len(g["gpc"])], g["gpc"])
    if energy:  # the reference's solve_energy after every solve_uvwp
        for k in ("t", "gt", "h", "h0", "gh"):
            assert np.array_equal(oc.arr(k)[:len(g[k])], g[k]), k
    log = oc.log()
    assert [l[0] for l in log] == list(g["log_name"])
    assert [l[1] for l in log] == list(g["log_it"])
    # values to the 4 digits the reference prints; res_max is left out for
    # solves with zero iterations, where the reference prints an unset local
    for l, r in zip(log, g["log_res"]):
        assert es93(l[2]) == r[0] and es93(l[3]) == r[1]
        if l[1] > 0:
            assert es93(l[4]) == r[2]


@pytest.mark.skipif(not refsim_io.have_refsim("O2"), reason="reference build absent")
@pytest.mark.parametrize("n,jit,nsubd", [(5, 0.0, 4), (7, 0.2, 4), (12, 0.1, 2), (6, 0.1, 8)])
def test_oracle_matches_reference_live(n, jit, nsubd):
    g = meshgen.box_mesh(n, jitter=jit)
    with tempfile.TemporaryDirectory() as td:
        mf, of = os.path.join(td, "m.bin"), os.path.join(td, "o.bin")
        meshgen.write_refsim_mesh(g, mf)
        log, _ = refsim_io.run_refsim(mf, of, nsubd, 2.0, 0.005, 0.02, 1, 3)
        d = refsim_io.read_dump(of)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=nsubd, rho=2.0, mu=0.005, dt=0.02)
    oc.run(1, 3)
    for k in FIELDS + SETUP:
        assert np.array_equal(oc.arr(k), d[k]), k
    assert [(l[0], l[1]) for l in oc.log()] == [(l[0], l[1]) for l in log]


# ---------------------------------------------------------------- closed forms
@pytest.mark.parametrize("n,jit", [(4, 0.0), (6, 0.25), (5, 0.3)])
def test_geometry_closed_cells_and_volume(n, jit):
    g = meshgen.box_mesh(n, jitter=jit)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=1)
    c = oc.c
    L = c.L
    idx, nb2 = oc.arr("ef2nb_idx"), oc.arr("ef2nb")[L:]
    aip = oc.arr("aip").reshape(-1, 3)
    for e in range(c.ne):
        s = np.zeros(3)
        for k in range(idx[e] - 1, idx[e + 1] - 1):
            f = nb2[k]
            s += np.sign(f) * aip[abs(f) - 1]
        assert np.allclose(s, 0.0, atol=1e-15)
    assert abs(oc.arr("vol").sum() - 1.0) < 1e-13
    assert c.nf == (6 * c.ne + c.nbf) // 2
    if jit == 0.0:
        h = 1.0 / n
        assert np.allclose(np.linalg.norm(aip, axis=1), h * h, rtol=1e-14)
        assert np.allclose(oc.arr("vol"), h ** 3, rtol=1e-13)


@pytest.mark.parametrize("jit", [0.0, 0.25])
def test_gradient_exact_for_linear_fields(jit):
    g = meshgen.box_mesh(5, jitter=jit)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=1)
    c = oc.c
    n = c.ne + c.nbf
    x, y, z = oc.arr("xc"), oc.arr("yc"), oc.arr("zc")
    phi = np.ascontiguousarray(0.3 + 1.5 * x - 2.0 * y + 0.25 * z)
    grad = np.zeros(3 * n)
    oracle_lib.lib().or_calc_grad(phi.ctypes.data_as(oracle_lib.PD), grad.ctypes.data_as(oracle_lib.PD),
                                  c.xc, c.yc, c.zc, c.ef2nb_idx, c.ef2nb, c.ne, c.nf, c.nbf)
    gr = grad[:3 * c.ne].reshape(-1, 3)
    assert np.allclose(gr, [1.5, -2.0, 0.25], atol=1e-11)


def test_matinv3_against_numpy_and_singular_fallback():
    rng = np.random.default_rng(3)
    for _ in range(50):
        A = rng.normal(size=(3, 3))
        A = A @ A.T + 0.1 * np.eye(3)
        B = np.zeros(9)
        Af = np.asfortranarray(A).reshape(-1, order="F").copy()
        oracle_lib.lib().or_matinv3(Af.ctypes.data_as(oracle_lib.PD), B.ctypes.data_as(oracle_lib.PD))
        assert np.allclose(B.reshape(3, 3, order="F"), np.linalg.inv(A), rtol=1e-10)
    A = np.array([1.0, 2.0, 3.0, 2.0, 4.0, 6.0, 3.0, 6.0, 9.0])  # rank one: det = 0
    B = np.zeros(9)
    oracle_lib.lib().or_matinv3(A.ctypes.data_as(oracle_lib.PD), B.ctypes.data_as(oracle_lib.PD))
    assert np.allclose(B.reshape(3, 3), np.eye(3) / 14.0)


def test_vec_weight():
    f = oracle_lib.lib().or_vec_weight
    P = lambda v: np.array(v, np.float64).ctypes.data_as(oracle_lib.PD)  # noqa: E731
    assert f(P([0, 0, 0]), P([-1, 0, 0]), P([1, 0, 0])) == 0.5
    assert f(P([0, 0, 0]), P([-1, 0, 0]), P([3, 0, 0])) == 0.25
    assert f(P([1, 1, 1]), P([1, 1, 1]), P([1, 1, 1])) == 0.5


def finley_reference(key, b):
    """Darel Rex Finley's published non-recursive quicksort, pure Python."""
    key, b = list(key), list(b)
    beg, end, i = [0] * 1000, [0] * 1000, 0
    end[0] = len(key)
    while i >= 0:
        L, R = beg[i], end[i] - 1
        if L < R:
            piv, b0 = key[L], b[L]
            while L < R:
                while key[R] >= piv and L < R:
                    R -= 1
                if L < R:
                    key[L], b[L] = key[R], b[R]
                    L += 1
                while key[L] <= piv and L < R:
                    L += 1
                if L < R:
                    key[R], b[R] = key[L], b[L]
                    R -= 1
            key[L], b[L] = piv, b0
            beg[i + 1], end[i + 1], end[i] = L + 1, end[i], L
            i += 1
            if end[i] - beg[i] > end[i - 1] - beg[i - 1]:
                beg[i], beg[i - 1] = beg[i - 1], beg[i]
                end[i], end[i - 1] = end[i - 1], end[i]
        else:
            i -= 1
    return key, b


def test_qsort_matches_published_algorithm():
    rnd = random.Random(7)
    for trial in range(300):
        n = rnd.randint(1, 60)
        nk = rnd.randint(1, 5)
        key = [rnd.randint(1, nk) for _ in range(n)]
        if trial % 3 == 0:  # runs, as the partition produces
            key.sort(key=lambda v: (v % 2, v))
        b = list(range(1, n + 1))
        ek, eb = finley_reference(key, b)
        k = np.array(key, np.int32)
        bb = np.array(b, np.int32)
        oracle_lib.lib().or_qsort_key_nrec(k.ctypes.data_as(oracle_lib.PI), bb.ctypes.data_as(oracle_lib.PI), n)
        assert list(k) == ek and list(bb) == eb
        assert ek == sorted(key)


# -------------------------------------------------------- equation invariants
def assembled(n=6, jit=0.2, steps=(1, 2)):
    g = meshgen.box_mesh(n, jitter=jit)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=4, rho=1.0, mu=0.01, dt=0.01)
    oc.run(*steps)
    return oc


def test_simplec_identity_and_positive_coefficients():
    oc = assembled()
    lib = oracle_lib.lib()
    lib.or_update_boundaries(oc.p)
    lib.or_calc_coef_uvw(oc.p)
    c = oc.c
    idx, ap, anb = oc.arr("ef2nb_idx"), oc.arr("ap"), oc.arr("anb")
    vol, dc, d = oc.arr("vol"), oc.arr("dc"), oc.arr("d")
    rows = np.array([anb[idx[e] - 1:idx[e + 1] - 1].sum() for e in range(c.ne)])
    assert np.all(anb >= 0.0)
    assert np.allclose(ap - rows, vol / c.dt, rtol=1e-10)   # ap0 = rho V / dt
    assert np.allclose(dc, c.dt / 1.0, rtol=1e-10)          # dc = V/ap0 = dt/rho
    assert np.allclose(d, vol / ap, rtol=0)


def test_pressure_matrix_neumann_symmetric_and_mass_balance():
    oc = assembled()
    lib = oracle_lib.lib()
    lib.or_update_boundaries(oc.p)
    lib.or_calc_coef_uvw(oc.p)
    lib.or_calc_mip(oc.p, 1)
    lib.or_calc_coef_p(oc.p)
    c = oc.c
    L = c.L
    idx, ap, anb, b = oc.arr("ef2nb_idx"), oc.arr("ap"), oc.arr("anb"), oc.arr("b")
    nb1 = oc.arr("ef2nb")[:L]
    for e in range(c.ne):
        row = anb[idx[e] - 1:idx[e + 1] - 1]
        s = 0.0
        for v in row:
            s = s + v
        assert ap[e] == s  # zero row sum, bit for bit (ap accumulates the same terms)
    # symmetry: a_PN == a_NP up to the last bits (wt vs 1-wt)
    for e in range(c.ne):
        for k in range(idx[e] - 1, idx[e + 1] - 1):
            nb, lf = nb1[k] >> 5, nb1[k] & 31
            if lf == 0:
                continue
            back = anb[idx[nb - 1] - 1 + lf - 1]
            assert back == pytest.approx(anb[k], rel=1e-12)
    assert abs(b.sum()) < 1e-14 * max(1e-30, np.abs(b).max()) * c.ne


def test_gauss_seidel_equals_dense_splitting():
    """One symmetric GS sweep of the oracle equals forward/backward
    substitution with the dense matrix of the same splitting (numpy)."""
    oc = assembled(n=4, jit=0.1)
    lib = oracle_lib.lib()
    lib.or_update_boundaries(oc.p)
    lib.or_calc_coef_uvw(oc.p)
    c = oc.c
    L, ne, n = c.L, c.ne, c.ne + c.nbf
    idx, ap, anb, bu = oc.arr("ef2nb_idx"), oc.arr("ap"), oc.arr("anb"), oc.arr("bu")
    nb1 = oc.arr("ef2nb")[:L]
    A = np.zeros((ne, ne))
    rhs = bu.copy()
    u = oc.arr("u").copy()
    for e in range(ne):
        A[e, e] = ap[e]
        for k in range(idx[e] - 1, idx[e + 1] - 1):
            nb = nb1[k] >> 5
            if nb <= ne:
                A[e, nb - 1] -= anb[k]
            else:
                rhs[e] += anb[k] * u[nb - 1]
    x = u[:ne].copy()
    Lw, Up, D = np.tril(A, -1), np.triu(A, 1), np.diag(np.diag(A))
    x = np.linalg.solve(D + Lw, rhs - Up @ x)
    x = np.linalg.solve(D + Up, rhs - Lw @ x)
    phi = oc.arr("u").copy()
    lib.or_smoother_gs.argtypes = [C.c_char_p, oracle_lib.PD, oracle_lib.PD, oracle_lib.PD, oracle_lib.PD,
                                   oracle_lib.PI, oracle_lib.PI, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.or_smoother_gs(b"u", phi.ctypes.data_as(oracle_lib.PD), c.ap, c.anb, c.bu, c.ef2nb_idx, c.ef2nb,
                       ne, c.nf, c.nbf, 1)
    assert np.allclose(phi[:ne], x, rtol=1e-12, atol=1e-15)
    assert np.array_equal(phi[ne:n], u[ne:n])  # halos are read, never written


def test_pc_residual_history_fixed_point():
    """Quirk Q1 of the reference: the subdomain solver's res_f accumulates
    onto its previous value, so it tends to ~1/4 and the solve always uses
    the full nit iterations."""
    oc = assembled(n=8, jit=0.0, steps=(1, 2))
    pcs = [l for l in oc.log() if l[0] == "pc"]
    assert all(l[1] == 100 for l in pcs)
    assert all(0.2 < l[3] < 0.3 for l in pcs)


def test_relaxation_one_is_reference_mode():
    ref = assembled(n=6, jit=0.1, steps=(1, 2))
    g = meshgen.box_mesh(6, jitter=0.1)
    b = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=4, rho=1.0, mu=0.01, dt=0.01)
    b.set_relaxation(1.0, 1.0)
    b.run(1, 2)
    for k in ("u", "p", "mip"):
        assert np.array_equal(ref.arr(k), b.arr(k))
    c = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=4, rho=1.0, mu=0.01, dt=0.01)
    c.set_relaxation(0.7, 0.3)
    c.run(1, 2)
    assert not np.array_equal(ref.arr("u"), c.arr("u"))


def _residuals(oc, names=("u", "v", "w"), rhs=("bu", "bv", "bw")):
    """b - (ap*phi - sum anb*phi_nb) per cell, for each momentum component."""
    c = oc.c
    ne, L = c.ne, c.L
    idx, ap, anb = oc.arr("ef2nb_idx"), oc.arr("ap"), oc.arr("anb")
    nb = (oc.arr("ef2nb")[:L] >> 5) - 1
    row = np.repeat(np.arange(ne), np.diff(idx))
    out = []
    for name, bn in zip(names, rhs):
        phi = oc.arr(name)
        s = np.zeros(ne)
        np.add.at(s, row, anb * phi[nb])
        out.append(oc.arr(bn) - (ap * phi[:ne] - s))
    return out


def test_momentum_relaxation_keeps_the_residual_of_the_current_iterate():
    """Config mode, momentum: Patankar's implicit under-relaxation divides ap
    by alpha_u and adds (ap/alpha_u - ap)*phi to b, so the residual of the
    current iterate is unchanged (a converged field stays converged) while the
    diagonal grows.  The neighbour coefficients are untouched, and d and dc
    follow the relaxed ap (d = V/ap, dc = V/(ap - sum anb))."""
    cases = []
    for alpha in (1.0, 0.7):
        oc = assembled(n=6, jit=0.15, steps=(1, 2))
        lib = oracle_lib.lib()
        lib.or_update_boundaries(oc.p)
        oc.set_relaxation(alpha, 1.0)
        lib.or_calc_coef_uvw(oc.p)
        cases.append(oc)
    a, b = cases
    assert np.array_equal(a.arr("anb"), b.arr("anb"))
    assert np.array_equal(b.arr("ap"), a.arr("ap") / 0.7)
    ra, rb = _residuals(a), _residuals(b)
    for x, y, bn in zip(ra, rb, ("bu", "bv", "bw")):
        scale = np.abs(a.arr(bn)).max() + np.abs(a.arr("ap")).max() * 1e-3
        assert np.allclose(x, y, rtol=0, atol=1e-12 * scale)
    c = b.c
    idx, anb, vol = b.arr("ef2nb_idx"), b.arr("anb"), b.arr("vol")
    rows = np.add.reduceat(anb, idx[:-1] - 1)
    assert np.allclose(b.arr("d") * b.arr("ap"), vol, rtol=1e-14)
    assert np.allclose(b.arr("dc") * (b.arr("ap") - rows), vol, rtol=1e-12)
    assert c.ne == len(vol)


def test_pressure_relaxation_scales_the_correction():
    """Config mode, pressure: update_uvwp adds alpha_p*pc to p and
    alpha_p*gpc to gp, and corrects the face mass fluxes with the whole pc
    (continuity is enforced in full; only the pressure is relaxed)."""
    rng = np.random.default_rng(7)
    cases = []
    pc = None
    for alpha in (1.0, 0.3):
        oc = assembled(n=5, jit=0.1, steps=(1, 2))
        c = oc.c
        n = c.ne + c.nbf
        if pc is None:
            pc = rng.standard_normal(n)
            gpc = rng.standard_normal(3 * n)
        oc.arr("gpc")[:] = gpc
        p0, gp0 = oc.arr("p").copy(), oc.arr("gp").copy()
        oc.set_relaxation(1.0, alpha)
        oracle_lib.lib().or_update_uvwp(oc.p, pc.ctypes.data_as(oracle_lib.PD))
        cases.append((oc, p0, gp0))
    (a, pa, ga), (b, pb, gb) = cases
    ne = a.c.ne
    assert np.array_equal(pa, pb) and np.array_equal(ga, gb)
    assert np.allclose(b.arr("p")[:ne] - pb[:ne], 0.3 * (a.arr("p")[:ne] - pa[:ne]),
                       rtol=1e-12, atol=1e-14)
    assert np.allclose(b.arr("gp")[:3 * ne] - gb[:3 * ne],
                       0.3 * (a.arr("gp")[:3 * ne] - ga[:3 * ne]), rtol=1e-12, atol=1e-14)
    assert np.array_equal(a.arr("mip"), b.arr("mip"))
    assert np.array_equal(a.arr("p")[ne:], pa[ne:])  # halos are not touched


@pytest.mark.skipif(not refsim_io.have_refsim("O2"), reason="reference build absent")
@pytest.mark.parametrize("n,jit,nsubd,nt,nc", [(5, 0.0, 4, 2, 2), (8, 0.2, 4, 1, 3), (6, 0.1, 1, 2, 2)])
def test_oracle_energy_matches_reference_live(n, jit, nsubd, nt, nc):
    """The energy equation (never solved by the reference's main loop) solved
    after every solve_uvwp by the reference's own solve_energy and by the
    oracle: temperature, enthalpy, both gradients, the shared coefficient
    arrays it leaves and the 'e' solver lines."""
    g = meshgen.box_mesh(n, jitter=jit)
    with tempfile.TemporaryDirectory() as td:
        mf = os.path.join(td, "m.bin")
        meshgen.write_refsim_mesh(g, mf)
        log, _ = refsim_io.run_refsim(mf, os.path.join(td, "r.bin"), nsubd, 1.0, 0.01, 0.01, nt, nc,
                                      energy=True)
        d = refsim_io.read_dump(os.path.join(td, "r.bin"))
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=nsubd, rho=1.0, mu=0.01, dt=0.01)
    oc.run(nt, nc, energy=True)
    for k in ("t", "gt", "h", "h0", "gh", "u", "p", "ap", "anb", "b"):
        assert np.array_equal(oc.arr(k)[:len(d[k])], d[k]), k
    el = [l for l in oc.log() if l[0] == "e"]
    rl = [l for l in log if l[0] == "e"]
    assert [l[1] for l in el] == [l[1] for l in rl]
    for a, b in zip(el, rl):
        assert es93(a[2]) == b[2] and es93(a[3]) == b[3]


def test_energy_closed_forms():
    """Energy assembly on a uniform box at rest (mip = 0): no convection, so
    ap - sum(anb) = rho*V/dt exactly where a cell has no boundary face; the
    interior diffusion coefficient is tc/cp*h (wt = 1/2, |dr| = h, A = h^2);
    the boundary one tc*A/ds/cp with ds = h/2; and b = ap0*h0 with zero
    deferred correction while the gradients are zero."""
    n = 5
    g = meshgen.box_mesh(n)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=1, rho=1.0, mu=0.01, dt=0.01)
    lib = oracle_lib.lib()
    lib.or_update_boundaries_energy(oc.p)
    lib.or_calc_coef_energy(oc.p)
    c = oc.c
    h = 1.0 / n
    idx, ap, anb, b = oc.arr("ef2nb_idx"), oc.arr("ap"), oc.arr("anb"), oc.arr("b")
    nb = oc.arr("ef2nb")[:c.L]
    ap0 = 1.0 * h ** 3 / 0.01
    for e in range(c.ne):
        row = slice(idx[e] - 1, idx[e + 1] - 1)
        halo = (nb[row] & 31) == 0
        assert np.allclose(anb[row][~halo], 5.0 / 1000.0 * h, rtol=1e-12)
        assert np.allclose(anb[row][halo], 5.0 * h * h / (h / 2) / 1000.0, rtol=1e-12)
        assert ap[e] == pytest.approx(anb[row].sum() + ap0, rel=1e-12)
        assert b[e] == pytest.approx(ap0 * 273.0 * 1000.0, rel=1e-14)


def _greedy_colours(sd, key=None):
    """Greedy colouring of a system, restated: visiting the cells in order of
    key (1-based cell -> rank; the cell index itself by default), a cell takes
    the least colour none of its interior neighbours earlier in that order has
    (halo entries, packed index > ne, are not cells of the system)."""
    ne, idx, ef = sd["ne"], sd["ef2nb_idx"], sd["ef2nb"]
    key = key if key is not None else (lambda e: e)
    col = np.zeros(ne + 1, dtype=np.int64)
    for e in sorted(range(1, ne + 1), key=key):
        used = {int(col[d]) for d in (int(ef[i - 1]) >> 5 for i in range(idx[e - 1], idx[e]))
                if d <= ne and key(d) < key(e)}
        c = 0
        while c in used:
            c += 1
        col[e] = c
    return col


def test_pc_colour_order_is_a_proper_greedy_colouring():
    """The colour order of the subdomain pressure-correction smoother (an
    extension, no reference counterpart) colours each subdomain greedily in
    global cell order, which gives two colours on a box, and visits every
    cell once, in the order (colour, local index); no two neighbouring cells
    share a colour, so each colour's updates are independent and the GPU may
    run them in parallel.  With it the solve takes another Gauss-Seidel path to
    the same linear systems: after six outer iterations the fields stay
    within 1e-3 of the reference order's and the solver lines agree in
    their iteration counts."""
    g = meshgen.box_mesh(10, jitter=0.15)
    ref = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=4, rho=1.0, mu=0.0025, dt=0.01)
    col = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=4, rho=1.0, mu=0.0025, dt=0.01)
    col.set_pc_order("colored")
    g2gf, g2gf_idx = col.arr("g2gf"), col.arr("g2gf_idx")
    for k in range(4):
        sd = col.subdomain(k)
        ne, idx, ef = sd["ne"], sd["ef2nb_idx"], sd["ef2nb"]
        gl = g2gf[g2gf_idx[k] - 1:g2gf_idx[k + 1] - 1]
        colours = _greedy_colours(sd, key=lambda e: gl[e - 1])
        assert list(col.pc_perm(k)) == sorted(range(1, ne + 1), key=lambda e: (colours[e], e))
        for e in range(1, ne + 1):
            for i in range(idx[e - 1], idx[e]):
                d = int(ef[i - 1]) >> 5
                if d <= ne:
                    assert colours[d] != colours[e]
        assert colours.max() + 1 == 2
    ref.run(2, 3)
    col.run(2, 3)
    assert [(l[0], l[1]) for l in ref.log()] == [(l[0], l[1]) for l in col.log()]
    for name in ("u", "v", "w"):
        assert np.max(np.abs(ref.arr(name) - col.arr(name))) < 1e-3
    assert not np.array_equal(ref.arr("p"), col.arr("p"))


@pytest.mark.skipif(not refsim_io.have_refsim("O2"), reason="reference build absent")
@pytest.mark.parametrize("n,jit,nsubd", [(6, 0.2, 4), (5, 0.1, 1)])
def test_oracle_matches_reference_nonuniform_properties(n, jit, nsubd):
    """Per-cell density and viscosity (random in [0.8, 1.2] and [0.005,
    0.015]) fed to the reference's own properties_t arrays: the face
    interpolations of rho and mu weigh the two cells differently, which the
    uniform cases cannot check.  Fields and solver lines equal the
    reference's bit for bit."""
    g = meshgen.box_mesh(n, jitter=jit)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=nsubd, rho=1.0, mu=0.01, dt=0.02)
    rng = np.random.default_rng(7)
    ne = oc.c.ne
    rho, mu = 0.8 + 0.4 * rng.random(ne), 0.005 + 0.01 * rng.random(ne)
    with tempfile.TemporaryDirectory() as td:
        mf, of = os.path.join(td, "m.bin"), os.path.join(td, "o.bin")
        meshgen.write_refsim_mesh(g, mf)
        log, _ = refsim_io.run_refsim(mf, of, nsubd, 1.0, 0.01, 0.02, 2, 3, props=(rho, mu))
        d = refsim_io.read_dump(of)
    oc.arr("rho")[:] = rho
    oc.arr("mu")[:] = mu
    oc.run(2, 3)
    for k in FIELDS:
        assert np.array_equal(oc.arr(k), d[k]), k
    assert [(l[0], l[1]) for l in oc.log()] == [(l[0], l[1]) for l in log]


@pytest.mark.skipif(not refsim_io.have_refsim("O2"), reason="reference build absent")
@pytest.mark.parametrize("n,jit,nsubd,sym", [(6, 0.2, 4, ("north", "bottom")), (5, 0.15, 1, ("east",))])
def test_oracle_matches_reference_symmetry(n, jit, nsubd, sym):
    """Symmetry sections: the driver rebinds their setters to the reference's
    symmetry (mod_uvwp.f90:517-546), which reflects the inner cell's velocity
    about the face normal and sets 'zero_flux', so the momentum assembly
    skips their wall-shear term (:266).  On a jittered mesh the normals are
    off the axes and every velocity component enters the reflection.  Fields
    and solver lines equal the reference's bit for bit."""
    g = meshgen.box_mesh(n, jitter=jit)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g, sym), nsubd=nsubd, rho=1.0, mu=0.01, dt=0.02)
    with tempfile.TemporaryDirectory() as td:
        mf, of = os.path.join(td, "m.bin"), os.path.join(td, "o.bin")
        meshgen.write_refsim_mesh(g, mf)
        log, _ = refsim_io.run_refsim(mf, of, nsubd, 1.0, 0.01, 0.02, 2, 3, symmetry=sym)
        d = refsim_io.read_dump(of)
    oc.run(2, 3)
    for k in FIELDS:
        assert np.array_equal(oc.arr(k), d[k]), k
    assert [l[:2] for l in oc.log()] == [l[:2] for l in log]
    for a, b in zip(oc.log(), log):  # the reference prints 4 significant digits
        assert np.allclose(a[2:], b[2:], rtol=6e-4, atol=0), (a, b)
    # the symmetry halos differ from a wall's: their velocity is not zero
    sec = [i for i, nm in enumerate(g["names"]) if nm in sym]
    esec = np.asarray(g["esec"]).reshape(-1, 2)
    halo = np.concatenate([np.arange(esec[i][0], esec[i][1] + 1) - 1 for i in sec])
    assert np.abs(d["u"][halo]).max() + np.abs(d["v"][halo]).max() > 0


@pytest.mark.parametrize("n,jit,relax,nit", [(10, 0.15, False, 100), (9, 0.0, True, 100),
                                             (8, 0.1, False, 25)])
def test_mixed_pc_solve_keeps_fp64_accuracy(n, jit, relax, nit):
    """The mixed-precision colour order (an extension): FP32 sweeps on the
    correction system ap*d - sum(anb*d_nb) = r, with the FP64 values updated
    and the FP64 residual recomputed at every check.  In exact arithmetic the
    sweeps of d from zero are the sweeps of pc itself, so the fields must
    stay close to the FP64 colour order's, with the same iteration counts,
    and still differ in the last bits.  The bounds are set between measured
    values: a solve that ends on a check leaves p within 3e-10 (relative L2),
    while a variant that never refreshes r (FP32 corrections accumulated over
    the whole solve) is off by 1.5e-8 or more, and a missed final update or a
    wrong correction by far more.  With a cap of 25 the last correction comes
    after the last check and keeps FP32 accuracy (p 1.2e-8); there mip still
    separates the two variants (5e-10 against 5e-9)."""
    g = meshgen.box_mesh(n, jitter=jit)
    runs = {}
    for mode in ("colored", "mixed"):
        oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=4, rho=1.0, mu=0.0025, dt=0.01,
                                   nit=nit)
        if relax:
            oc.set_relaxation(0.7, 0.3)
        oc.set_pc_order(mode)
        oc.run(2, 3)
        runs[mode] = oc
    a, b = runs["colored"], runs["mixed"]
    # velocity components against the velocity field's size (v and w of a
    # symmetric cavity are close to zero), p and mip against their own
    vel = np.linalg.norm(np.concatenate([a.arr(k) for k in ("u", "v", "w")]))
    if nit % 10 == 0:
        tol = {"u": 1e-10, "v": 1e-10, "w": 1e-10, "p": 2e-9, "mip": 1e-9}
    else:
        tol = {"u": 1e-10, "v": 1e-10, "w": 1e-10, "p": 5e-8, "mip": 2e-9}
    for k in ("u", "v", "w", "p", "mip"):
        x, y = a.arr(k), b.arr(k)
        scale = vel if k in ("u", "v", "w") else np.linalg.norm(x)
        assert np.linalg.norm(x - y) <= tol[k] * scale, k
    assert not np.array_equal(a.arr("p"), b.arr("p"))
    la, lb = a.log(), b.log()
    assert [l[:2] for l in la] == [l[:2] for l in lb]
    for x, y in zip(la, lb):
        assert np.allclose(x[2:], y[2:], rtol=1e-6, atol=0), (x, y)


def test_gs_colour_order_is_a_proper_greedy_colouring():
    """The colour order of solve_gs (an extension, no reference counterpart)
    visits every cell of the global system once, in the order (greedy
    colour, global index); no two neighbouring cells share a colour, so the
    GPU may update a colour at once.  The momentum solves take another
    Gauss-Seidel path to the same systems: after six outer iterations the
    fields stay within 1e-3 of the reference order's and the iteration
    counts agree."""
    g = meshgen.box_mesh(10, jitter=0.15)
    ref = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=4, rho=1.0, mu=0.0025, dt=0.01)
    col = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=4, rho=1.0, mu=0.0025, dt=0.01)
    col.set_gs_order("colored")
    c = col.c
    glob = dict(ne=c.ne, ef2nb_idx=np.ctypeslib.as_array(c.ef2nb_idx, shape=(c.ne + 1,)).copy(),
                ef2nb=np.ctypeslib.as_array(c.ef2nb, shape=(2 * c.L,))[:c.L].copy())
    colours = _greedy_colours(glob)
    assert list(col.gs_perm()) == sorted(range(1, c.ne + 1), key=lambda e: (colours[e], e))
    idx, ef = glob["ef2nb_idx"], glob["ef2nb"]
    for e in range(1, c.ne + 1):
        for i in range(idx[e - 1], idx[e]):
            d = int(ef[i - 1]) >> 5
            if d <= c.ne:
                assert colours[d] != colours[e]
    assert 2 <= colours.max() + 1 <= 4
    ref.run(2, 3)
    col.run(2, 3)
    assert [(l[0], l[1]) for l in ref.log()] == [(l[0], l[1]) for l in col.log()]
    for name in ("u", "v", "w"):
        assert np.max(np.abs(ref.arr(name) - col.arr(name))) < 1e-3
    assert not np.array_equal(ref.arr("u"), col.arr("u"))


def test_serving_request_stream_is_seeded_and_bounded():
    """The serving benchmark's stream: Poisson arrivals (mean gap 1/rate),
    edges from {16, 24, ..., 96}, Re log-uniform in [100, 1000], cavity and
    channel a fair draw, the same stream for the same seed."""
    from cfdl import serving
    a = serving.make_requests(512, 20.0, 7)
    assert a == serving.make_requests(512, 20.0, 7)
    assert [r.idx for r in a] == list(range(512))
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) > 0) and abs(np.mean(np.diff(t)) - 0.05) < 0.01
    assert {r.n for r in a} <= set(range(16, 97, 8)) and len({r.n for r in a}) == 11
    kinds = [r.kind for r in a]
    assert set(kinds) == {"cavity", "channel"} and abs(kinds.count("cavity") - 256) < 40
    re = np.array([r.re for r in a])
    assert re.min() >= 100.0 and re.max() <= 1000.0
    assert abs(np.median(np.log(re)) - 0.5 * (np.log(100.0) + np.log(1000.0))) < 0.2


@pytest.mark.skipif(not refsim_io.have_refsim("O2"), reason="reference build absent")
@pytest.mark.parametrize("dims,lens,jit,nsubd", [((7, 5, 9), (1.4, 1.0, 2.0), 0.1, 4),
                                                 ((10, 4, 6), (2.0, 0.5, 1.0), 0.0, 2)])
def test_oracle_matches_reference_non_cubic_box(dims, lens, jit, nsubd):
    """Boxes with unequal cell counts and side lengths: the bisection splits
    the longest axis of each group's bounding box (mod_agglomeration.f90:
    380-561), so the subdomains, their local order and the interface pairs
    differ from a cube's.  Set-up arrays, fields and solver lines equal the
    reference's bit for bit."""
    g = meshgen.box_mesh(*dims, lx=lens[0], ly=lens[1], lz=lens[2], jitter=jit)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=nsubd, rho=1.0, mu=0.01, dt=0.02)
    with tempfile.TemporaryDirectory() as td:
        mf, of = os.path.join(td, "m.bin"), os.path.join(td, "o.bin")
        meshgen.write_refsim_mesh(g, mf)
        log, _ = refsim_io.run_refsim(mf, of, nsubd, 1.0, 0.01, 0.02, 2, 3)
        d = refsim_io.read_dump(of)
    for k in ("gf2g", "g2gf", "g2gf_idx"):
        assert np.array_equal(oc.arr(k), d[k]), k
    oc.run(2, 3)
    for k in FIELDS + SETUP:
        assert np.array_equal(oc.arr(k), d[k]), k
    assert [l[:2] for l in oc.log()] == [l[:2] for l in log]
