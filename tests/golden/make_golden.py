"""Write the golden fixtures: the reference itself, run on small meshes.

Runs oracle/_ref/refsim_O2 (the reference's Fortran compiled from its own
sources with the driver in oracle/refdrv) on the seeded meshes of
meshgen.box_mesh and stores inputs (mesh parameters), setup arrays, final
fields and the residual lines the reference prints (4 significant digits,
its es9.3 format) in tests/golden/<case>.npz.

    python tests/golden/make_golden.py [case ...]

Needs the reference tree to build oracle/_ref; the fixtures it writes are
data and travel with the repository.
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))
import meshgen  # noqa: E402
import refsim_io  # noqa: E402

# name: (n, jitter, nsubd, rho, mu, dt, ntstep, ncoef); energy_* cases also
# call the reference's solve_energy after every solve_uvwp (REFSIM_ENERGY=1)
CASES = {
    "energy6_jitter": (6, 0.15, 4, 1.0, 0.01, 0.01, 2, 2),
    "energy7_1sd": (7, 0.1, 1, 1.0, 0.01, 0.01, 1, 3),
    "cavity8": (8, 0.0, 4, 1.0, 0.01, 0.01, 2, 3),
    "cavity6_jitter": (6, 0.2, 4, 1.0, 0.01, 0.01, 2, 3),
    "cavity8_jitter_1sd": (8, 0.2, 1, 1.0, 0.01, 0.01, 2, 2),
    "cavity9_odd": (9, 0.0, 4, 1.0, 0.01, 0.01, 1, 2),
    "cavity10_re400": (10, 0.15, 4, 1.0, 0.0025, 0.01, 1, 3),
    "cavity4_rho5": (4, 0.0, 4, 5.0, 0.01, 0.01, 2, 3),
}

KEEP = ["ef2nb_idx", "ef2nb", "s2g", "bs", "xc", "yc", "zc", "vol", "aip", "rip", "gf2g", "g2gf",
        "g2gf_idx", "u", "v", "w", "p", "gu", "gv", "gw", "gp", "gpc", "mip", "mip0", "bu", "bv",
        "bw", "d", "dc", "ap", "b", "anb", "phic", "t", "gt", "h", "h0", "gh"]


def main(names):
    for name, (n, jit, nsubd, rho, mu, dt, nt, nc) in CASES.items():
        if names and name not in names:
            continue
        mesh = meshgen.box_mesh(n, jitter=jit)
        with tempfile.TemporaryDirectory() as td:
            mf, of = os.path.join(td, "m.bin"), os.path.join(td, "o.bin")
            meshgen.write_refsim_mesh(mesh, mf)
            log, _ = refsim_io.run_refsim(mf, of, nsubd, rho, mu, dt, nt, nc, opt="O2",
                                          energy=name.startswith("energy"))
            d = refsim_io.read_dump(of)
        out = {k: d[k] for k in KEEP if k in d}
        out["params"] = np.array([n, jit, nsubd, rho, mu, dt, nt, nc], np.float64)
        out["log_name"] = np.array([l[0] for l in log])
        out["log_it"] = np.array([l[1] for l in log], np.int32)
        out["log_res"] = np.array([l[2:] for l in log], np.float64)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, "written,", len(log), "solver lines")


if __name__ == "__main__":
    main(sys.argv[1:])
