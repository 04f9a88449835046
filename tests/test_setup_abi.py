"""CPU tests of the product's host side: the C ABI library loads, exports
what include/cfdl.h declares, carries gfx950 code only, and its host-side
setup (connectivity, geometry, partition, subdomain order) equals the
oracle's bit for bit.  No GPU compute is called here."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import cfdl
import meshgen
import oracle_lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "cfdl.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cfdl_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = cfdl.load_library()
    names = declared_functions()
    assert len(names) > 30
    for n in names:
        assert hasattr(lib, n), n
    # and the Python binding declares a signature for each of them
    from cfdl._lib import exported_names
    assert sorted(exported_names()) == names


def test_library_targets_gfx950_only():
    data = open(cfdl.library_path(), "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx\w+)", data))
    assert targets == {b"gfx950"}


def test_product_does_not_link_or_import_the_oracle():
    out = subprocess.run(["ldd", cfdl.library_path()], capture_output=True, text=True).stdout
    assert "oracle" not in out
    pkg = os.path.join(REPO, "cfd-lite_amd")
    for root, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h", ".f90")):
                src = open(os.path.join(root, f), errors="replace").read()
                assert "oracle_lib" not in src and "liboracle" not in src, f


def test_no_device_gives_a_clean_error():
    import ctypes
    lib = cfdl.load_library()
    if os.environ.get("HIP_VISIBLE_DEVICES", None) is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be visible")
    m = cfdl.Mesh.box(3).partition(1)
    with pytest.raises(cfdl.CfdlError):
        cfdl.UVWP.cavity(m)


@pytest.mark.parametrize("dims,jit,nsubd", [((4, 4, 4), 0.0, 4), ((8, 8, 8), 0.2, 4),
                                              ((9, 9, 9), 0.0, 4), ((12, 8, 6), 0.15, 4),
                                              ((10, 10, 10), 0.1, 2), ((8, 8, 8), 0.1, 8),
                                              ((7, 9, 5), 0.0, 3), ((6, 6, 6), 0.0, 1)])
def test_setup_equals_oracle(dims, jit, nsubd):
    nx, ny, nz = dims
    m = cfdl.Mesh.box(nx, ny, nz, jitter=jit).partition(nsubd)
    g = meshgen.box_mesh(nx, ny, nz, jitter=jit)
    for k in ("x", "y", "z"):
        assert np.array_equal(m.get(k), g[k]), k
    assert np.array_equal(m.get("e2vx"), g["e2vx"])
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=nsubd)
    for k in ("ef2nb_idx", "ef2nb", "s2g", "bs", "xc", "yc", "zc", "vol", "aip", "rip", "gf2g",
              "g2gf", "g2gf_idx"):
        assert np.array_equal(m.get(k), oc.arr(k)), k


def test_setup_from_elements_matches_box():
    g = meshgen.box_mesh(5, jitter=0.1)
    a = cfdl.Mesh.from_elements(g["x"], g["y"], g["z"], g["etype"], g["esec"], g["e2vx"]).partition(4)
    b = cfdl.Mesh.box(5, jitter=0.1).partition(4)
    for k in ("ef2nb", "s2g", "aip", "xc", "g2gf"):
        assert np.array_equal(a.get(k), b.get(k)), k


def test_setup_rejects_bad_input():
    g = meshgen.box_mesh(3)
    bad = g["e2vx"].copy()
    bad[0] = g["nvx"] + 5
    with pytest.raises(cfdl.CfdlError):
        cfdl.Mesh.from_elements(g["x"], g["y"], g["z"], g["etype"], g["esec"], bad)
    with pytest.raises(cfdl.CfdlError):
        cfdl.Mesh.box(0)
    m = cfdl.Mesh.box(2)
    with pytest.raises(cfdl.CfdlError):
        m.partition(0)


def test_partition_large_is_fast():
    import time
    t = time.time()
    m = cfdl.Mesh.box(64).partition(4)
    assert time.time() - t < 30.0
    idx = m.get("g2gf_idx")
    assert list(np.diff(idx)) == [64 ** 3 // 4] * 4


def test_bench_refuses_several_gpus():
    p = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0 and "one GPU" in p.stderr


def test_cavity_binding_needs_a_top_section():
    """BCs bind by section name (mod_eqn_setup.f90:46-67): a mesh whose
    sections carry no names has no 'top', and the cavity binding refuses it
    (before it looks for a GPU) instead of binding every section as a wall."""
    g = meshgen.box_mesh(4)
    mesh = cfdl.Mesh.from_elements(g["x"], g["y"], g["z"], g["etype"], g["esec"], g["e2vx"]).partition(4)
    with pytest.raises(cfdl.CfdlError, match="top"):
        cfdl.UVWP.cavity(mesh)
