"""VTU output in the reference's layout (src/modules/mod_vtu_output.f90).

The reference build (oracle/_ref/refsim_O2) ends its run with the
reference's own write_vtubin when given a project path, like src/main.f90:89.
The CPU test feeds the fields it dumped (and the per-cell mass imbalance,
summed in CSR order from its mip) to cfdl_mesh_write_vtu and requires the
same file, byte for byte; the GPU test runs the same case on the GPU and
writes the file from device fields and the device mass-imbalance kernel.
"""
import os
import tempfile

import numpy as np
import pytest

import cfdl
import meshgen
import refsim_io

needs_ref = pytest.mark.skipif(not refsim_io.have_refsim("O2"),
                               reason="reference build oracle/_ref/refsim_O2 not built")


def _reference_case(td, n, jit, nsubd, nt, nc):
    g = meshgen.box_mesh(n, jitter=jit)
    mf = os.path.join(td, "mesh.bin")
    meshgen.write_refsim_mesh(g, mf)
    proj = os.path.join(td, "ref") + "/"
    refsim_io.run_refsim(mf, os.path.join(td, "r.bin"), nsubd, 1.0, 0.01, 0.01, nt, nc, project=proj)
    vtu = os.path.join(proj, "VTK", "output-%d" % (nt + 1), "uvwp.vtu")
    mesh = cfdl.Mesh.from_elements(g["x"], g["y"], g["z"], g["etype"], g["esec"], g["e2vx"],
                                   names=g["names"])
    assert mesh.names == g["names"]
    return g, mesh, refsim_io.read_dump(os.path.join(td, "r.bin")), open(vtu, "rb").read()


def _imbalance(d):
    """Sum over each cell's CSR entries of mip(face)*sign, entry by entry."""
    ne, L = d["ne"], 2 * d["nf"] - d["nbf"]
    idx, fsg = d["ef2nb_idx"], d["ef2nb"][L:]
    s = np.zeros(ne)
    rows = np.diff(idx)
    assert np.all(rows == rows[0])
    k0 = idx[:-1] - 1
    for k in range(rows[0]):
        f = fsg[k0 + k]
        s = s + d["mip"][np.abs(f) - 1] * np.where(f > 0, 1.0, -1.0)
    return s


@needs_ref
@pytest.mark.parametrize("n,jit,nsubd", [(5, 0.0, 4), (6, 0.15, 1)])
def test_writer_matches_reference_writer(n, jit, nsubd):
    with tempfile.TemporaryDirectory() as td:
        g, mesh, d, ref = _reference_case(td, n, jit, nsubd, 1, 2)
        ne = d["ne"]
        out = os.path.join(td, "mine.vtu")
        mesh.write_vtu(out, [("u", d["u"][:ne]), ("v", d["v"][:ne]), ("w", d["w"][:ne]),
                             ("p", d["p"][:ne]), ("gpc", d["gpc"][:3 * ne], 3),
                             ("mip", _imbalance(d))])
        mine = open(out, "rb").read()
    assert len(mine) == len(ref)
    assert mine == ref


@pytest.mark.gpu
@needs_ref
def test_gpu_run_writes_the_reference_file():
    n, nsubd, nt, nc = 6, 4, 1, 2
    with tempfile.TemporaryDirectory() as td:
        g, mesh, d, ref = _reference_case(td, n, 0.1, nsubd, nt, nc)
        mesh.partition(nsubd)
        eq = cfdl.UVWP.cavity(mesh, rho=1.0, mu=0.01, dt=0.01)
        eq.run(nt, nc)
        path = eq.write_vtu(mesh, os.path.join(td, "gpu"), nt + 1)
        mine = open(path, "rb").read()
        assert np.array_equal(eq.mass_imbalance(), _imbalance(d))
    assert mine == ref
