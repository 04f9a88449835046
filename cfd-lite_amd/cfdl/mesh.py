"""Host-side setup: element lists -> connectivity, geometry, partition.

Mirrors what the reference's pre-processing hands to the hot path
(geometry_t, src/setup/mod_mg_lvl_uns.f90:46-58; cell_input,
src/setup/cell_input.f90:10-165) using the C++ setup in libcfdl_hip.so.
"""
import ctypes as C

import numpy as np

from ._lib import PD, PI, VP, check, check_cgns, load_cgns_library, load_library

HEXA_8, QUAD_4 = 17, 7

_I32 = ("ef2nb_idx", "ef2nb", "s2g", "bs", "esec", "etype", "e2vx", "gf2g", "g2gf", "g2gf_idx")
_F64 = ("xc", "yc", "zc", "vol", "aip", "rip", "x", "y", "z")


class Mesh:
    """A hexahedral mesh with the reference's connectivity and geometry."""

    def __init__(self, handle, names=None):
        self._lib = load_library()
        self._h = handle
        self.names = names

    @classmethod
    def box(cls, nx, ny=None, nz=None, lx=1.0, ly=1.0, lz=1.0, jitter=0.0, seed=20261202):
        """Unit box of nx*ny*nz hexahedra (i fastest), QUAD_4 sections
        bottom, south, east, north, west, top after the cells."""
        lib = load_library()
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        h = VP()
        check(lib.cfdl_mesh_box(nx, ny, nz, lx, ly, lz, jitter, seed, C.byref(h)))
        return cls(h, ["cells", "bottom", "south", "east", "north", "west", "top"])

    @classmethod
    def from_elements(cls, x, y, z, etype, esec, e2vx, names=None):
        lib = load_library()
        x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
        et = np.ascontiguousarray(etype, np.int32)
        es = np.ascontiguousarray(esec, np.int32)
        ev = np.ascontiguousarray(e2vx, np.int32)
        h = VP()
        check(lib.cfdl_mesh_from_elements(len(x), x.ctypes.data_as(PD), y.ctypes.data_as(PD),
                                          z.ctypes.data_as(PD), len(et), et.ctypes.data_as(PI),
                                          es.ctypes.data_as(PI), ev.ctypes.data_as(PI), C.byref(h)))
        if names is not None:
            arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
            check(lib.cfdl_mesh_set_section_names(h, len(names), arr))
        return cls(h, names)

    @classmethod
    def read_cgns(cls, path):
        """The mesh of a CGNS file (first base, first zone; HDF5 storage),
        as the reference's cell_input reads it (cfdl_mesh_read_cgns)."""
        lib = load_cgns_library()
        h = VP()
        check_cgns(lib.cfdl_mesh_read_cgns(str(path).encode(), C.byref(h)))
        ints = [C.c_int() for _ in range(6)]  # ne, nf, nbf, nsec, nsubd, nvx
        check(load_library().cfdl_mesh_info(h, *[C.byref(i) for i in ints]))
        names = []
        buf = C.create_string_buffer(64)
        for k in range(ints[3].value):
            check_cgns(lib.cfdl_cgns_section_name(h, k, buf, 64))
            names.append(buf.value.decode())
        return cls(h, names)

    def write_cgns(self, path):
        """Vertices, named sections and element lists as a CGNS file
        (cfdl_mesh_write_cgns)."""
        lib = load_cgns_library()
        check_cgns(lib.cfdl_mesh_write_cgns(self._h, str(path).encode()))

    def renumber(self, seed):
        """Seeded random cell permutation, then reverse Cuthill-McKee
        (cfdl_mesh_renumber); before partition()."""
        check(self._lib.cfdl_mesh_renumber(self._h, seed))
        return self

    def partition(self, nsubd):
        """Oct-tree bisection + subdomain-local order (cell_input.f90:131-151)."""
        check(self._lib.cfdl_mesh_partition(self._h, nsubd))
        return self

    def info(self):
        v = [C.c_int() for _ in range(6)]
        check(self._lib.cfdl_mesh_info(self._h, *[C.byref(a) for a in v]))
        return dict(zip(("ne", "nf", "nbf", "nsec", "nsubd", "nvx"), (a.value for a in v)))

    def _size(self, name):
        i = self.info()
        ne, nf, nbf, nsec, nsubd, nvx = (i[k] for k in ("ne", "nf", "nbf", "nsec", "nsubd", "nvx"))
        L = 2 * nf - nbf
        return {"ef2nb_idx": ne + 1, "ef2nb": 2 * L, "s2g": nf, "bs": nbf, "esec": 2 * nsec,
                "etype": nsec, "e2vx": 8 * (ne + nbf), "gf2g": ne if nsubd else 0,
                "g2gf": ne if nsubd else 0, "g2gf_idx": nsubd + 1 if nsubd else 0,
                "xc": ne + nbf, "yc": ne + nbf, "zc": ne + nbf, "vol": ne, "aip": 3 * nf,
                "rip": 3 * nf, "x": nvx, "y": nvx, "z": nvx}[name]

    def write_vtu(self, path, arrays):
        """VTU file of cell arrays in the reference's write_vtubin layout
        (src/modules/mod_vtu_output.f90:6-326).  arrays: [(name, values)] or
        [(name, values, ncomp)], values in cell order with at least ne*ncomp
        entries (ncomp 1 by default; extra halo entries are not written)."""
        ne = self.info()["ne"]
        keep, names, ncomp = [], [], []
        for item in arrays:
            name, a = item[0], np.ascontiguousarray(item[1], dtype=np.float64)
            k = item[2] if len(item) > 2 else 1
            if a.size < k * ne:
                raise ValueError("%s: %d values for %d cells" % (name, a.size, ne))
            keep.append(a)
            names.append(name.encode())
            ncomp.append(k)
        nv = len(keep)
        cn = (C.c_char_p * max(nv, 1))(*names)
        nc = (C.c_int * max(nv, 1))(*ncomp)
        dp = (PD * max(nv, 1))(*[a.ctypes.data_as(PD) for a in keep])
        check(self._lib.cfdl_mesh_write_vtu(self._h, str(path).encode(), nv, cn, nc, dp))

    def get(self, name):
        n = self._size(name)
        if name in _I32:
            a = np.empty(n, np.int32)
            check(self._lib.cfdl_mesh_get_i32(self._h, name.encode(), a.ctypes.data_as(PI), n))
        elif name in _F64:
            a = np.empty(n, np.float64)
            check(self._lib.cfdl_mesh_get_f64(self._h, name.encode(), a.ctypes.data_as(PD), n))
        else:
            raise KeyError(name)
        return a

    def mg_get(self, level, name):
        """One array of the agglomeration multigrid hierarchy that the "mgcg"
        pc solve builds for this mesh (cfdl_mesh_mg_get); None when the
        level or array does not exist."""
        cnt = C.c_int64()
        check(self._lib.cfdl_mesh_mg_get(self._h, level, name.encode(), None, 0, C.byref(cnt)))
        if cnt.value < 0:
            return None
        a = np.empty(max(cnt.value, 1), np.int32)
        check(self._lib.cfdl_mesh_mg_get(self._h, level, name.encode(), a.ctypes.data_as(PI),
                                         cnt.value, C.byref(cnt)))
        return a[:cnt.value]

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h is not None:
            self._lib.cfdl_mesh_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
