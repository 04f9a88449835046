"""CGNS mesh files (libcfdl_cgns.so, include/cfdl_cgns.h), CPU.

The reference reads its mesh from CGNS with the CGNS library
(src/setup/cell_input.f90:36-93, src/modules/mod_cgns.f90:17-148).  Neither
that library nor a CGNS file is in the reference tree or this image, so the
file format here is pinned only by its own writer (a round trip must rebuild
the same mesh bit for bit, and the file must be HDF5), not by a file the CGNS
library wrote: parity unpinned for the format itself.  The mesh the reader
builds goes through the same element-list ingest (cfdl_mesh_from_elements)
as the box generator, whose set-up arrays equal the oracle's and the
reference build's (test_setup_abi.py, test_oracle.py)."""
import os

import numpy as np
import pytest

import cfdl
from cfdl import _lib

pytestmark = pytest.mark.skipif(not os.path.exists(_lib.cgns_library_path()),
                                reason="libcfdl_cgns.so not built (no HDF5 headers)")

ARRAYS = ["x", "y", "z", "etype", "esec", "e2vx", "ef2nb_idx", "ef2nb", "s2g", "bs", "xc", "yc",
          "zc", "vol", "aip", "rip"]


def _same(a, b):
    assert a.info() == b.info()
    for k in ARRAYS:
        x, y = a.get(k), b.get(k)
        assert x.dtype == y.dtype and np.array_equal(x, y), k


@pytest.mark.parametrize("dims,jit,renum", [((6, 5, 7), 0.2, 0), ((8, 8, 8), 0.0, 0),
                                            ((7, 6, 5), 0.1, 11)])
def test_round_trip_rebuilds_the_same_mesh(tmp_path, dims, jit, renum):
    m = cfdl.Mesh.box(*dims, jitter=jit)
    if renum:
        m.renumber(renum)
    f = tmp_path / "m.cgns"
    m.write_cgns(f)
    assert open(f, "rb").read(8) == b"\x89HDF\r\n\x1a\n"  # HDF5 storage
    r = cfdl.Mesh.read_cgns(f)
    assert r.names == m.names
    _same(r, m)
    # it partitions and binds the cavity BCs like the generated mesh
    r.partition(4)
    m.partition(4)
    for k in ("gf2g", "g2gf", "g2gf_idx"):
        assert np.array_equal(r.get(k), m.get(k)), k


def test_read_rejects_what_is_not_a_cgns_mesh(tmp_path):
    with pytest.raises(cfdl.CfdlError):
        cfdl.Mesh.read_cgns(tmp_path / "missing.cgns")
    bad = tmp_path / "text.cgns"
    bad.write_text("not an HDF5 file\n")
    with pytest.raises(cfdl.CfdlError):
        cfdl.Mesh.read_cgns(bad)


def test_library_exports_the_cgns_header():
    import ctypes
    import re
    lib = ctypes.CDLL(_lib.cgns_library_path())
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "include", "cfdl_cgns.h")).read()
    names = re.findall(r"\b(cfdl_\w+)\s*\(", hdr)
    assert set(names) == set(_lib.cgns_exported_names())
    for n in names:
        assert hasattr(lib, n), n
