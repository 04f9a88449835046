"""ctypes view of the oracle (oracle/liboracle.so) for the tests.

Test infrastructure: the oracle is the checker.  Arrays are exposed as numpy
views on the C memory of an or_case (layout mirrors oracle/oracle.h).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB_PATH = os.path.join(REPO, "oracle", "liboracle.so")

PI = C.POINTER(C.c_int)
PD = C.POINTER(C.c_double)


class OrSubd(C.Structure):
    _fields_ = [("ne", C.c_int), ("nf", C.c_int), ("nbf", C.c_int),
                ("ef2nb_idx", PI), ("ef2nb", PI),
                ("ap", PD), ("b", PD), ("anb", PD), ("phic", PD)]


class OrIntf(C.Structure):
    _fields_ = [("c1", C.c_int), ("c2", C.c_int), ("ncs", C.c_int), ("index1", PI), ("index2", PI)]


class OrLog(C.Structure):
    _fields_ = [("it", C.c_int), ("res_i", C.c_double), ("res_f", C.c_double),
                ("res_max", C.c_double), ("name", C.c_char * 8)]


class OrCase(C.Structure):
    _fields_ = [("nvx", C.c_int), ("nsec", C.c_int), ("nelem", C.c_int), ("ne", C.c_int),
                ("nbf", C.c_int), ("nf", C.c_int), ("L", C.c_int),
                ("x", PD), ("y", PD), ("z", PD), ("e2vx", PI), ("etype", PI), ("esec", PI),
                ("bckind", PI),
                ("ef2nb_idx", PI), ("ef2nb", PI), ("s2g", PI), ("bs", PI),
                ("xc", PD), ("yc", PD), ("zc", PD), ("vol", PD), ("aip", PD), ("rip", PD),
                ("nsubd", C.c_int), ("gf2g", PI), ("g2gf", PI), ("g2gf_idx", PI),
                ("sd", C.POINTER(OrSubd)), ("intf", C.POINTER(OrIntf)),
                ("dt", C.c_double), ("nit", C.c_int), ("alpha_u", C.c_double),
                ("alpha_p", C.c_double), ("rho", PD), ("mu", PD),
                ("u", PD), ("v", PD), ("w", PD), ("p", PD), ("u0", PD), ("v0", PD), ("w0", PD),
                ("gu", PD), ("gv", PD), ("gw", PD), ("gp", PD), ("gpc", PD),
                ("mip", PD), ("mip0", PD),
                ("bu", PD), ("bv", PD), ("bw", PD), ("d", PD), ("dc", PD),
                ("ap", PD), ("anb", PD), ("b", PD), ("phic", PD),
                ("nlog", C.c_int), ("caplog", C.c_int), ("log", C.POINTER(OrLog)),
                ("cp", PD), ("tc", PD), ("t", PD), ("gt", PD), ("h", PD), ("h0", PD), ("gh", PD),
                ("pc_order", C.c_int), ("pc_perm", C.POINTER(PI)),
                ("gs_order", C.c_int), ("gs_perm", PI), ("mg", C.c_void_p),
                ("inlet", C.c_double * 5), ("asm32", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(OrCase)
        L.or_case_new.restype = P
        L.or_case_new.argtypes = [C.c_int, PD, PD, PD, C.c_int, PI, PI, PI, PI]
        for fn in ("or_connectivity", "or_partition"):
            getattr(L, fn).restype = C.c_int
        L.or_partition.argtypes = [P, C.c_int]
        L.or_connectivity.argtypes = [P]
        L.or_geometry.argtypes = [P]
        L.or_construct_subdomains.argtypes = [P]
        L.or_init_physics.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_int]
        L.or_case_free.argtypes = [P]
        L.or_qsort_key_nrec.argtypes = [PI, PI, C.c_int]
        L.or_vec_weight.restype = C.c_double
        L.or_vec_weight.argtypes = [PD, PD, PD]
        L.or_matinv3.argtypes = [PD, PD]
        L.or_calc_grad.argtypes = [PD, PD, PD, PD, PD, PI, PI, C.c_int, C.c_int, C.c_int]
        for fn in ("or_update_boundaries", "or_calc_coef_uvw", "or_calc_coef_p", "or_solve_uvwp",
                   "or_update_time", "or_construct_uvwp", "or_construct_energy",
                   "or_update_boundaries_energy", "or_calc_coef_energy", "or_solve_energy",
                   "or_update_time_energy"):
            getattr(L, fn).argtypes = [P]
        L.or_run_energy.argtypes = [P, C.c_int, C.c_int]
        L.or_calc_mip.argtypes = [P, C.c_int]
        L.or_adjust_pc.argtypes = [P, PD, C.c_double]
        L.or_update_uvwp.argtypes = [P, PD]
        L.or_run.argtypes = [P, C.c_int, C.c_int]
        L.or_solve.argtypes = [P, C.c_char_p, PD]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


class OracleCase:
    """A mesh + physics case run through the oracle."""

    def __init__(self, mesh, bckind, nsubd=4, rho=1.0, mu=0.01, dt=0.01, nit=100):
        L = lib()
        self._keep = [np.ascontiguousarray(mesh[k]) for k in ("x", "y", "z")]
        et = np.ascontiguousarray(mesh["etype"], np.int32)
        es = np.ascontiguousarray(mesh["esec"], np.int32)
        e2 = np.ascontiguousarray(mesh["e2vx"], np.int32)
        bk = np.ascontiguousarray(bckind, np.int32)
        x, y, z = self._keep
        self.p = L.or_case_new(int(mesh["nvx"]), _ptr(x, PD), _ptr(y, PD), _ptr(z, PD), len(et),
                               _ptr(et, PI), _ptr(es, PI), _ptr(e2, PI), _ptr(bk, PI))
        if L.or_connectivity(self.p) != 0:
            raise RuntimeError("oracle connectivity failed")
        L.or_geometry(self.p)
        if L.or_partition(self.p, nsubd) != 0:
            raise RuntimeError("oracle partition failed")
        if nsubd > 1:
            L.or_construct_subdomains(self.p)
        L.or_init_physics(self.p, rho, mu, dt, nit)
        L.or_construct_uvwp(self.p)
        L.or_construct_energy(self.p)

    def __del__(self):
        if getattr(self, "p", None) is not None and _lib is not None:
            _lib.or_case_free(self.p)
            self.p = None

    @property
    def c(self):
        return self.p.contents

    def arr(self, name, n=None, dtype=None):
        c = self.c
        sizes = self.sizes()
        n = sizes[name] if n is None else n
        ptr = getattr(c, name)
        return np.ctypeslib.as_array(ptr, shape=(n,))

    def sizes(self):
        c = self.c
        ne, nf, nbf, L = c.ne, c.nf, c.nbf, c.L
        n = ne + nbf
        s = dict(ef2nb_idx=ne + 1, ef2nb=2 * L, s2g=nf, bs=nbf, xc=n, yc=n, zc=n, vol=ne,
                 aip=3 * nf, rip=3 * nf, gf2g=ne, g2gf=ne, g2gf_idx=c.nsubd + 1, rho=ne, mu=ne,
                 mip=nf, mip0=nf, ap=ne, b=ne, anb=L, phic=n, cp=ne, tc=ne, t=n, h=n, h0=n,
                 gt=3 * n, gh=3 * n)
        for k in ("u", "v", "w", "p", "u0", "v0", "w0"):
            s[k] = n
        for k in ("gu", "gv", "gw", "gp", "gpc"):
            s[k] = 3 * n
        for k in ("bu", "bv", "bw", "d", "dc"):
            s[k] = ne
        return s

    def log(self):
        c = self.c
        return [(c.log[k].name.decode(), c.log[k].it, c.log[k].res_i, c.log[k].res_f,
                 c.log[k].res_max) for k in range(c.nlog)]

    def set_pc_order(self, order="reference"):
        """Colour order of the subdomain pc smoother (or_set_pc_order)."""
        L = lib()
        L.or_set_pc_order.argtypes = [C.c_void_p, C.c_int]
        L.or_set_pc_order(C.cast(self.p, C.c_void_p),
                          {"reference": 0, "colored": 1, "mixed": 2, "mgcg": 3, "mgcg-mixed": 4}[order])

    def set_gs_order(self, order="reference"):
        """Colour order of solve_gs (or_set_gs_order)."""
        L = lib()
        L.or_set_gs_order.argtypes = [C.c_void_p, C.c_int]
        L.or_set_gs_order(C.cast(self.p, C.c_void_p), {"reference": 0, "colored": 1}[order])

    def gs_perm(self):
        """Visiting order (1-based cells) of solve_gs under the colour order."""
        return np.ctypeslib.as_array(self.c.gs_perm, shape=(self.c.ne,)).copy()

    def pc_perm(self, k):
        """Visiting order (1-based local cells) of subdomain k's pc smoother
        under the colour order."""
        return np.ctypeslib.as_array(self.c.pc_perm[k], shape=(self.c.sd[k].ne,)).copy()

    def set_inlet_profile(self, umax=1.0, y0=0.0, y1=1.0, z0=0.0, z1=1.0):
        for k, v in enumerate((umax, y0, y1, z0, z1)):
            self.c.inlet[k] = v

    def set_assembly_precision(self, bits=64):
        self.c.asm32 = 1 if bits == 32 else 0

    def set_relaxation(self, alpha_u=1.0, alpha_p=1.0):
        self.c.alpha_u = alpha_u
        self.c.alpha_p = alpha_p

    def run(self, ntstep, ncoef, energy=False):
        (lib().or_run_energy if energy else lib().or_run)(self.p, ntstep, ncoef)

    def subdomain(self, k):
        s = self.c.sd[k]
        L = 2 * s.nf - s.nbf
        return dict(ne=s.ne, nf=s.nf, nbf=s.nbf,
                    ef2nb_idx=np.ctypeslib.as_array(s.ef2nb_idx, shape=(s.ne + 1,)).copy(),
                    ef2nb=np.ctypeslib.as_array(s.ef2nb, shape=(L,)).copy())

    def intf(self, c1, c2):
        t = self.c.intf[c1 * self.c.nsubd + c2]
        return np.ctypeslib.as_array(t.index1, shape=(max(t.ncs, 1),))[:t.ncs].copy()
