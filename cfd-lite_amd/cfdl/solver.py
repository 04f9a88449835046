"""Device-resident uvwp equation: the hot path of the reference.

Method names follow src/equations/mod_uvwp.f90 and src/modules/mod_solver.f90
(solve_uvwp, calc_coef_uvw, solve_gs, calc_grad, calc_mip, calc_coef_p,
solve, adjust_pc, update_uvwp) and src/modules/mod_physics.f90
(update_boundaries, update_time).  Every call runs HIP kernels in
libcfdl_hip.so; nothing here computes on the CPU.
"""
import ctypes as C

import numpy as np

from ._lib import PD, PI, VP, CfdlError, check, load_library

BC_WALL, BC_LID, BC_SYMMETRY, BC_INLET, BC_OUTLET = 0, 1, 2, 3, 4

FIELDS = {name: k for k, name in enumerate(
    ["u", "v", "w", "p", "u0", "v0", "w0", "gu", "gv", "gw", "gp", "gpc", "mip", "mip0",
     "bu", "bv", "bw", "d", "dc", "ap", "b", "anb", "phic", "work", "gwork",
     "t", "gt", "h", "h0", "gh"])}


def _fid(name):
    if isinstance(name, int):
        return name
    try:
        return FIELDS[name]
    except KeyError:
        raise CfdlError("unknown field %r" % (name,))


class UVWP:
    """uvwp_t + the shared work arrays, resident on the GPU."""

    def __init__(self, handle, sizes):
        self._lib = load_library()
        self._h = handle
        self.sizes = sizes

    @staticmethod
    def _sizes(ne, nf, nbf):
        n, L = ne + nbf, 2 * nf - nbf
        s = {}
        for k in ("u", "v", "w", "p", "u0", "v0", "w0", "phic", "work", "t", "h", "h0"):
            s[k] = n
        for k in ("gu", "gv", "gw", "gp", "gpc", "gwork", "gt", "gh"):
            s[k] = 3 * n
        for k in ("mip", "mip0"):
            s[k] = nf
        for k in ("bu", "bv", "bw", "d", "dc", "ap", "b"):
            s[k] = ne
        s["anb"] = L
        return s

    @classmethod
    def cavity(cls, mesh, rho=1.0, mu=0.01, dt=0.01, nit=100):
        """Lid-driven cavity binding ('top' -> lid, others -> wall,
        mod_uvwp.f90:73-78) on a partitioned Mesh."""
        lib = load_library()
        h = VP()
        check(lib.cfdl_ctx_create_cavity(mesh.handle, rho, mu, dt, nit, C.byref(h)))
        i = mesh.info()
        return cls(h, cls._sizes(i["ne"], i["nf"], i["nbf"]))

    @classmethod
    def channel(cls, mesh, rho=1.0, mu=0.005, dt=0.01, nit=100, umax=1.0):
        """Channel along x on a partitioned box Mesh (an extension): 'west'
        is a parabolic inlet of peak umax, 'east' a zero-gradient outlet
        scaled to the inflow, the other sections no-slip walls."""
        lib = load_library()
        h = VP()
        check(lib.cfdl_ctx_create_channel(mesh.handle, rho, mu, dt, nit, umax, C.byref(h)))
        i = mesh.info()
        return cls(h, cls._sizes(i["ne"], i["nf"], i["nbf"]))

    @classmethod
    def from_arrays(cls, ne, nf, nbf, ef2nb_idx, ef2nb, s2g, bs, xc, yc, zc, vol, aip, rip,
                    esec, bckind, nsubd, gf2g, g2gf, g2gf_idx, rho, mu, dt, nit=100):
        """Context from the reference's own arrays (geometry_t, meshds_t,
        mg_lvl_t%g2gf/gf2g, properties_t)."""
        lib = load_library()
        i32 = lambda a: np.ascontiguousarray(a, np.int32)  # noqa: E731
        f64 = lambda a: np.ascontiguousarray(a, np.float64)  # noqa: E731
        keep = [i32(ef2nb_idx), i32(ef2nb), i32(s2g), i32(bs), f64(xc), f64(yc), f64(zc),
                f64(vol), f64(aip), f64(rip), i32(esec), i32(bckind),
                i32(gf2g if gf2g is not None else np.zeros(1)),
                i32(g2gf if g2gf is not None else np.zeros(1)),
                i32(g2gf_idx if g2gf_idx is not None else np.zeros(2)), f64(rho), f64(mu)]
        ptr = [a.ctypes.data_as(PI if a.dtype == np.int32 else PD) for a in keep]
        h = VP()
        check(lib.cfdl_ctx_create(ne, nf, nbf, ptr[0], ptr[1], ptr[2], ptr[3], ptr[4], ptr[5],
                                  ptr[6], ptr[7], ptr[8], ptr[9], len(keep[11]), ptr[10], ptr[11],
                                  nsubd, ptr[12], ptr[13], ptr[14], ptr[15], ptr[16], dt, nit,
                                  C.byref(h)))
        return cls(h, cls._sizes(ne, nf, nbf))

    PC_ORDERS = {"reference": 0, "colored": 1, "mixed": 2, "mgcg": 3, "mgcg-mixed": 4}

    def set_pc_order(self, order="reference"):
        """Visiting order of the multi-subdomain pressure-correction smoother:
        "reference" (the reference's order, the default), "colored" (greedy
        colouring of each subdomain in global cell order, swept colour by
        colour; an extension) or
        "mixed" (the colour order with FP32 sweeps on the correction system
        under FP64 defect correction at every residual check; an extension),
        "mgcg" (conjugate gradients on the global system preconditioned by an
        agglomeration-multigrid V-cycle, the reference's stop rule; an
        extension) or "mgcg-mixed" (the same with the V-cycle in FP32)."""
        if order not in self.PC_ORDERS:
            raise CfdlError("pc order must be one of %s" % sorted(self.PC_ORDERS))
        check(self._lib.cfdl_set_pc_order(self._h, self.PC_ORDERS[order]))

    def set_assembly_precision(self, bits=64):
        """64 (the reference's FP64 assembly) or 32 (FP32 coefficients under
        the FP64 outer loop; an extension)."""
        check(self._lib.cfdl_set_assembly_precision(self._h, bits))

    GS_ORDERS = {"reference": 0, "colored": 1}

    def set_gs_order(self, order="reference"):
        """Visiting order of solve_gs (the momentum and energy solves, and the
        pc solve with one subdomain): "reference" (the reference's global
        order, the default) or "colored" (greedy colouring of the global
        system, swept colour by colour; an extension)."""
        if order not in self.GS_ORDERS:
            raise CfdlError("gs order must be one of %s" % sorted(self.GS_ORDERS))
        check(self._lib.cfdl_set_gs_order(self._h, self.GS_ORDERS[order]))

    BC_KINDS = {"wall": 0, "lid": 1, "symmetry": 2, "inlet": 3, "outlet": 4}

    def set_bc_kind(self, first, last, kind):
        """Boundary type of the halos first..last (1-based element numbers), as
        the setters dirichlet0 / lid / symmetry assign it (mod_uvwp.f90:493-570):
        "wall", "lid" or "symmetry"."""
        if kind not in self.BC_KINDS:
            raise CfdlError("bc kind must be one of %s" % sorted(self.BC_KINDS))
        check(self._lib.cfdl_set_bc_kind(self._h, first, last, self.BC_KINDS[kind]))

    def set_section_bc(self, mesh, name, kind):
        """set_bc_kind over the halos of the mesh's boundary section name."""
        if not mesh.names or name not in mesh.names[1:]:
            raise CfdlError("no boundary section %r" % (name,))
        esec = mesh.get("esec").reshape(-1, 2)
        s = mesh.names.index(name)
        self.set_bc_kind(int(esec[s][0]), int(esec[s][1]), kind)

    def set_dt(self, dt):
        """Pseudo-time step (phys_t%dt, mod_physics.f90:15); float("inf") is
        the steady mode (an extension: transient terms exactly zero, needs
        alpha_u < 1)."""
        check(self._lib.cfdl_set_dt(self._h, dt))

    def set_inlet_profile(self, umax=1.0, y0=0.0, y1=1.0, z0=0.0, z1=1.0):
        """Parabolic inlet of the "inlet" halos (channel extension)."""
        check(self._lib.cfdl_set_inlet_profile(self._h, umax, y0, y1, z0, z1))

    def set_relaxation(self, alpha_u=1.0, alpha_p=1.0):
        """Config-mode under-relaxation (1, 1 = the reference algorithm)."""
        check(self._lib.cfdl_set_relaxation(self._h, alpha_u, alpha_p))

    # ---- fields ----
    def get(self, name):
        fid = _fid(name)
        key = [k for k, v in FIELDS.items() if v == fid][0]
        a = np.empty(self.sizes[key], np.float64)
        check(self._lib.cfdl_get_field(self._h, fid, a.ctypes.data_as(PD), a.size))
        return a

    def set(self, name, values):
        fid = _fid(name)
        key = [k for k, v in FIELDS.items() if v == fid][0]
        a = np.ascontiguousarray(values, np.float64)
        if a.size != self.sizes[key]:
            raise CfdlError("field %s needs %d values, got %d" % (key, self.sizes[key], a.size))
        check(self._lib.cfdl_set_field(self._h, fid, a.ctypes.data_as(PD), a.size))

    # ---- routines (one per reference procedure) ----
    def update_boundaries(self):
        check(self._lib.cfdl_update_boundaries(self._h))

    def calc_coef_uvw(self):
        check(self._lib.cfdl_calc_coef_uvw(self._h))

    def solve_gs(self, cname, phi, b):
        check(self._lib.cfdl_solve_gs(self._h, cname.encode(), _fid(phi), _fid(b)))

    def solve_gs_uvw(self):
        check(self._lib.cfdl_solve_gs_uvw(self._h))

    def calc_grad(self, phi, grad):
        check(self._lib.cfdl_calc_grad(self._h, _fid(phi), _fid(grad)))

    def calc_mip(self, rhie_chow):
        check(self._lib.cfdl_calc_mip(self._h, 1 if rhie_chow else 0))

    def calc_coef_p(self):
        check(self._lib.cfdl_calc_coef_p(self._h))

    def set_a_0(self, name):
        check(self._lib.cfdl_set_a_0(self._h, _fid(name)))

    def solve(self, cname="pc"):
        check(self._lib.cfdl_solve(self._h, cname.encode()))

    def adjust_pc(self):
        check(self._lib.cfdl_adjust_pc(self._h))

    def update_uvwp(self):
        check(self._lib.cfdl_update_uvwp(self._h))

    def solve_uvwp(self):
        check(self._lib.cfdl_solve_uvwp(self._h))

    def update_time(self):
        check(self._lib.cfdl_update_time(self._h))

    def mass_imbalance(self):
        """Net outward mass flux per cell (ne values), computed on the GPU."""
        check(self._lib.cfdl_calc_mass_imbalance(self._h, FIELDS["work"]))
        return self.get("work")[:self.sizes["ap"]]

    def write_vtu(self, mesh, project, step):
        """The reference's end-of-run output of the uvwp equation
        (src/main.f90:89 -> write_vtubin): project/VTK/output-<step>/uvwp.vtu
        with u, v, w, p, gpc and the cell mass imbalance as 'mip'.  Returns
        the file path."""
        import os
        d = os.path.join(project, "VTK", "output-%d" % step)
        os.makedirs(d, exist_ok=True)
        path = os.path.join(d, "uvwp.vtu")
        ne = self.sizes["ap"]
        mesh.write_vtu(path, [("u", self.get("u")[:ne]), ("v", self.get("v")[:ne]),
                              ("w", self.get("w")[:ne]), ("p", self.get("p")[:ne]),
                              ("gpc", self.get("gpc")[:3 * ne], 3),
                              ("mip", self.mass_imbalance())])
        return path

    def run(self, ntstep, ncoef, energy=False):
        """ntstep x ncoef outer iterations (src/main.f90:50-63); energy=True
        also solves the energy equation after every solve_uvwp."""
        if energy:
            check(self._lib.cfdl_run_energy(self._h, ntstep, ncoef))
        else:
            check(self._lib.cfdl_run(self._h, ntstep, ncoef))

    # energy equation, src/equations/mod_energy.f90
    def set_properties(self, rho, mu):
        """Per-cell density and viscosity (properties_t%rho, %mu)."""
        rho = np.ascontiguousarray(rho, np.float64)
        mu = np.ascontiguousarray(mu, np.float64)
        if rho.size != self.sizes["ap"] or mu.size != self.sizes["ap"]:
            raise CfdlError("rho and mu need one value per cell")
        check(self._lib.cfdl_set_properties(self._h, rho.ctypes.data_as(PD), mu.ctypes.data_as(PD)))

    def set_thermal(self, cp, tc):
        cp = np.ascontiguousarray(cp, np.float64)
        tc = np.ascontiguousarray(tc, np.float64)
        if cp.size != self.sizes["ap"] or tc.size != self.sizes["ap"]:
            raise CfdlError("cp and tc need one value per cell")
        check(self._lib.cfdl_set_thermal(self._h, cp.ctypes.data_as(PD), tc.ctypes.data_as(PD)))

    def construct_energy(self):
        check(self._lib.cfdl_construct_energy(self._h))

    def update_boundaries_energy(self):
        check(self._lib.cfdl_update_boundaries_energy(self._h))

    def calc_coef_energy(self):
        check(self._lib.cfdl_calc_coef_energy(self._h))

    def solve_energy(self):
        check(self._lib.cfdl_solve_energy(self._h))

    def synchronize(self):
        check(self._lib.cfdl_synchronize(self._h))

    def reset(self):
        """Back to the state at creation (fields cleared, log cleared); the
        mesh, properties, BCs and settings stay."""
        check(self._lib.cfdl_ctx_reset(self._h))

    def log(self):
        """[(name, it, res_i, res_f, res_max)] one per linear solve."""
        n = C.c_int()
        check(self._lib.cfdl_log_count(self._h, C.byref(n)))
        out = []
        name = C.create_string_buffer(8)
        it = C.c_int()
        ri, rf, rm = C.c_double(), C.c_double(), C.c_double()
        for k in range(n.value):
            check(self._lib.cfdl_log_get(self._h, k, name, C.byref(it), C.byref(ri), C.byref(rf),
                                         C.byref(rm)))
            out.append((name.value.decode(), it.value, ri.value, rf.value, rm.value))
        return out

    def clear_log(self):
        check(self._lib.cfdl_log_clear(self._h))

    def set_profile(self, sample_every):
        """Time every sample_every-th pressure-correction sweep launch."""
        check(self._lib.cfdl_set_profile(self._h, sample_every))

    def stats(self):
        s = (C.c_double * 8)()
        check(self._lib.cfdl_stats(self._h, s, 8))
        keys = ("gs_launches", "pc_launches", "pc_cells", "outer_iterations", "sampled_launches",
                "launch_ms", "launch_cells", "pc_levels_per_sym_sweep")
        return dict(zip(keys, list(s)))

    def reset_stats(self):
        check(self._lib.cfdl_stats_reset(self._h))

    def close(self):
        if self._h is not None:
            self._lib.cfdl_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Continuous batching of independent cases (include/cfdl.h,
    cfdl_batch_*): run(ntstep, ncoef) does for every attached UVWP what
    UVWP.run does for one, with one grouped launch per step for all of them;
    each case's results equal those of the case run alone, bit for bit.
    Cases attach and detach between runs (up to 32 at a time)."""

    MAX_CASES = 32

    def __init__(self, block=0, graphs=True):
        self._lib = load_library()
        h = VP()
        check(self._lib.cfdl_batch_create(C.byref(h)))
        self._h = h
        self.cases = []
        if block:
            self.set_block(block)
        if not graphs:
            self.set_graphs(False)

    def set_graphs(self, on):
        """Outer iterations as captured HIP graphs (default) or direct launches."""
        check(self._lib.cfdl_batch_set_graphs(self._h, 1 if on else 0))

    def graph_stats(self):
        """(graph captures, graph launches) so far."""
        a, b = C.c_int64(0), C.c_int64(0)
        check(self._lib.cfdl_batch_graph_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def attach(self, eq):
        check(self._lib.cfdl_batch_attach(self._h, eq._h))
        self.cases.append(eq)

    def detach(self, eq):
        check(self._lib.cfdl_batch_detach(self._h, eq._h))
        self.cases.remove(eq)

    def __len__(self):
        n = C.c_int(0)
        check(self._lib.cfdl_batch_size(self._h, C.byref(n)))
        return n.value

    def set_block(self, block):
        """Lane count of the element-wise launches (0: tuned)."""
        check(self._lib.cfdl_batch_set_block(self._h, int(block)))

    def block(self):
        b = C.c_int(0)
        check(self._lib.cfdl_batch_block(self._h, C.byref(b)))
        return b.value

    def run(self, ntstep, ncoef):
        check(self._lib.cfdl_batch_run(self._h, int(ntstep), int(ncoef)))

    def synchronize(self):
        check(self._lib.cfdl_batch_synchronize(self._h))

    def close(self):
        if self._h is not None:
            self._lib.cfdl_batch_destroy(self._h)
            self._h = None
            self.cases = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
