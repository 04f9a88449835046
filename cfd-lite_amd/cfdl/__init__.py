"""cfdl -- MI355X (gfx950) SIMPLE pressure-velocity loop of CFD-Lite.

Host-side Python mirror of the reference's hot-path interface
(src/equations/mod_uvwp.f90, src/modules/mod_solver.f90): the same routine
names and argument meaning, backed by libcfdl_hip.so through its C ABI
(include/cfdl.h).  There is no CPU fallback: if the HIP library is missing
or no GPU is visible, the calls raise.

    import sys; sys.path.insert(0, "<repo>/cfd-lite_amd")
    import cfdl
    mesh = cfdl.Mesh.box(32)            # hex box + connectivity + geometry
    mesh.partition(4)                   # 4 subdomains, reference order
    eq = cfdl.UVWP.cavity(mesh, rho=1.0, mu=0.01, dt=0.01)
    eq.run(ntstep=10, ncoef=3)          # src/main.f90:50-63
    u = eq.get("u")
"""
from ._lib import CfdlError, load_library, library_path
from .mesh import Mesh
from .solver import UVWP, Batch, FIELDS, BC_WALL, BC_LID, BC_SYMMETRY, BC_INLET, BC_OUTLET

__all__ = ["CfdlError", "load_library", "library_path", "Mesh", "UVWP", "Batch", "FIELDS",
           "BC_WALL", "BC_LID", "BC_SYMMETRY", "BC_INLET", "BC_OUTLET"]
