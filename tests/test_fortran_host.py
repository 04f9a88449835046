"""GPU tests of the Fortran host side of the boundary.

dropsim is the reference's own program (its setup, mod_physics,
mod_subdomains, mod_energy ... and the same driver as the reference build)
with the GPU build's drop-in mod_uvwp / mod_solver linked in place of the
reference's two hot-path modules; it must dump exactly the reference's bits
and print the reference's solver lines.  dropsim_res also links the drop-in
mod_physics, which keeps the uvwp state on the device between output steps
(update_time on the device, the host copy refreshed every 10th step and at
the last); the same bits are required of it.  cfdlite_gpu is the standalone
Fortran host program of the package (iso_c_binding into libcfdl_hip.so).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import meshgen
import oracle_lib
import refsim_io

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPSIM = os.path.join(REPO, "oracle", "_ref", "dropsim")
DROPSIM_RES = os.path.join(REPO, "oracle", "_ref", "dropsim_res")
HOST = os.path.join(REPO, "cfd-lite_amd", "bin", "cfdlite_gpu")
KEYS = ["ef2nb_idx", "ef2nb", "s2g", "bs", "xc", "vol", "aip", "u", "v", "w", "p", "u0", "gu",
        "gv", "gw", "gp", "gpc", "mip", "mip0", "bu", "bv", "bw", "d", "dc", "ap", "b", "anb",
        "phic"]


def solver_lines(stdout):
    return [l for l in stdout.splitlines() if l.strip().split(" ")[0] in ("u", "v", "w", "pc", "e")]


@pytest.mark.skipif(not refsim_io.have_refsim("O2"), reason="reference build absent")
@pytest.mark.parametrize("binary", [DROPSIM, DROPSIM_RES], ids=["dropsim", "dropsim_res"])
@pytest.mark.parametrize("n,jit,nsubd,nt,nc,energy,sym", [
    (8, 0.0, 4, 2, 3, False, ()), (7, 0.2, 4, 1, 3, False, ()), (6, 0.1, 1, 2, 2, False, ()),
    (6, 0.15, 4, 2, 2, True, ()), (5, 0.0, 1, 1, 3, True, ()),
    (7, 0.2, 4, 2, 2, False, ("north", "bottom")), (5, 0.1, 4, 12, 1, False, ())])
def test_dropin_modules_reproduce_the_reference(binary, n, jit, nsubd, nt, nc, energy, sym):
    """Both programs also end with the reference's own write_vtubin, which
    reads the drop-in's uvwp_t: the VTU files must be identical.  With
    energy, every outer iteration also calls solve_energy (the drop-in
    mod_energy on the GPU in dropsim).  With sym, both drivers bind those
    sections' setters to their module's symmetry."""
    if not os.path.exists(binary):
        pytest.skip("%s not built" % os.path.basename(binary))
    g = meshgen.box_mesh(n, jitter=jit)
    vtu = os.path.join("VTK", "output-%d" % (nt + 1), "uvwp.vtu")
    with tempfile.TemporaryDirectory() as td:
        mf = os.path.join(td, "m.bin")
        meshgen.write_refsim_mesh(g, mf)
        refsim_io.run_refsim(mf, os.path.join(td, "r.bin"), nsubd, 1.0, 0.01, 0.01, nt, nc,
                             project=os.path.join(td, "r") + "/", energy=energy, symmetry=sym)
        ref_out = refsim_io.run_refsim.last_stdout
        refsim_io.run_refsim(mf, os.path.join(td, "d.bin"), nsubd, 1.0, 0.01, 0.01, nt, nc,
                             binary=binary, project=os.path.join(td, "d") + "/", energy=energy,
                             symmetry=sym)
        drop_out = refsim_io.run_refsim.last_stdout
        r = refsim_io.read_dump(os.path.join(td, "r.bin"))
        d = refsim_io.read_dump(os.path.join(td, "d.bin"))
        rv = open(os.path.join(td, "r", vtu), "rb").read()
        dv = open(os.path.join(td, "d", vtu), "rb").read()
    for k in KEYS + (["t", "gt", "h", "h0", "gh"] if energy else []):
        assert np.array_equal(r[k], d[k]), k
    assert solver_lines(ref_out) == solver_lines(drop_out)
    assert rv == dv


@pytest.mark.skipif(not os.path.exists(HOST), reason="cfdlite_gpu not built")
@pytest.mark.parametrize("au,ap", [(1.0, 1.0), (0.7, 0.3)])
def test_fortran_host_program_matches_oracle(au, ap):
    n, nt, nc = 8, 2, 2
    with tempfile.TemporaryDirectory() as td:
        dump = os.path.join(td, "f.bin")
        p = subprocess.run([HOST, str(n), "4", "1.0", "0.0025", "0.01", str(nt), str(nc), repr(au),
                            repr(ap), dump], capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stdout + p.stderr
        raw = np.fromfile(dump, dtype="<f8")
    g = meshgen.box_mesh(n)
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=4, rho=1.0, mu=0.0025, dt=0.01)
    oc.set_relaxation(au, ap)
    oc.run(nt, nc)
    off = 0
    for name in ("u", "v", "w", "p", "mip", "gp", "phic"):
        ref = oc.arr(name)
        assert np.array_equal(raw[off:off + ref.size], ref), name
        off += ref.size
    lines = solver_lines(p.stdout)
    log = oc.log()
    assert len(lines) == len(log)
    for line, (name, it, ri, rf, rm) in zip(lines, log):
        f = line.split()
        assert f[0] == name and int(f[1]) == it
        assert float(f[2]) == float("%.3e" % ri) and float(f[3]) == float("%.3e" % rf)
