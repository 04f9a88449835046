"""cfdl_calc_residual_noalloc (the entry point that allocates nothing)
against cfdl_calc_residual, which allocates its scratch: equal results, bit
for bit, on two calls in a row, the second with a smaller system.  Device
memory comes from the HIP runtime the library itself uses (ctypes), and the
values are also checked against a numpy restatement of calc_residual
(src/modules/mod_solver.f90:230-253) to rounding."""
import ctypes as C

import numpy as np
import pytest

import cfdl

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libcfdl_hip.so uses."""

    def __init__(self):
        cfdl.load_library()
        self.h = C.CDLL("libamdhip64.so.7")  # the soname the library loaded
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
        self.live.append(p)
        return p

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def system(n, seed):
    """A hexahedral box's CSR (halos after the cells) with random values."""
    m = cfdl.Mesh.box(n, jitter=0.1)
    i = m.info()
    ne, nbf = i["ne"], i["nbf"]
    idx = m.get("ef2nb_idx")
    L = idx[-1] - 1
    nbr = ((m.get("ef2nb")[:L].view(np.uint32) >> 5).astype(np.int32) - 1)
    row = (idx - 1).astype(np.int32)
    rng = np.random.default_rng(seed)
    anb = rng.uniform(0.0, 1.0, L)
    ap = rng.uniform(3.0, 7.0, ne)
    b = rng.uniform(-1.0, 1.0, ne)
    phi = rng.uniform(-1.0, 1.0, ne + nbf)
    return ne, row, nbr, ap, anb, b, phi


def numpy_residual(ne, row, nbr, ap, anb, b, phi):
    """calc_residual's per-cell residual in CSR order (every row of a box
    has six entries), so r and its signed max are exact; the rms to
    rounding."""
    r = b.copy()
    for k in range(6):
        r = r + anb[row[:-1] + k] * phi[nbr[row[:-1] + k]]
    r = r - ap * phi[:ne]
    return np.sqrt(np.sum(r * r) / ne), max(r.max(), 0.0)


def call(lib, fn, hip, sysarrs):
    ne, row, nbr, ap, anb, b, phi = sysarrs
    ptrs = [hip.put(a) for a in (row, nbr, ap, anb, b, phi)]
    res, rmax = C.c_double(), C.c_double()
    rc = getattr(lib, fn)(ne, *ptrs, C.byref(res), C.byref(rmax), None)
    assert rc == 0, lib.cfdl_last_error()
    return res.value, rmax.value


def test_noalloc_residual_equals_allocating_entry_point():
    lib = cfdl.load_library()
    hip = Hip()
    try:
        for n, seed in ((48, 1), (20, 2)):  # the second call has a smaller input
            s = system(n, seed)
            a = call(lib, "cfdl_calc_residual", hip, s)
            z = call(lib, "cfdl_calc_residual_noalloc", hip, s)
            assert a == z, (n, a, z)
            ref = numpy_residual(*s)
            assert a[0] == pytest.approx(ref[0], rel=1e-12)
            assert a[1] == ref[1]
    finally:
        hip.free()


def test_noalloc_residual_rejects_sizes():
    lib = cfdl.load_library()
    res, rmax = C.c_double(), C.c_double()
    nul = C.c_void_p(8)  # never dereferenced: the size check comes first
    assert lib.cfdl_calc_residual_noalloc((1 << 30) + 1, nul, nul, nul, nul, nul, nul,
                                          C.byref(res), C.byref(rmax), None) == 1
    assert lib.cfdl_calc_residual_noalloc(0, nul, nul, nul, nul, nul, nul,
                                          C.byref(res), C.byref(rmax), None) == 1
