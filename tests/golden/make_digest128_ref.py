"""Write the digests of the 128^3 cavity in reference mode, from the reference.

Runs oracle/_ref/refsim_O2 (the reference's Fortran compiled from its own
sources) on the 128^3 lid-driven cavity (Re 400: rho 1, mu 0.0025; 4
subdomains; dt 0.01; no under-relaxation, the reference's algorithm) for one
pseudo time step of two outer iterations, and stores the SHA-256 digest of
each field's little-endian FP64 bytes and the solver lines the reference
prints (4 significant digits) in tests/golden/cavity128_reference_digests.json.
The GPU test compares the digests of its own fields with them.

    python tests/golden/make_digest128_ref.py

Needs the reference tree to build oracle/_ref; takes several minutes (the
reference's set-up).
"""
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))
import meshgen  # noqa: E402
import refsim_io  # noqa: E402

CASE = {"n": 128, "nsubd": 4, "rho": 1.0, "mu": 0.0025, "dt": 0.01, "nit": 100,
        "ntstep": 1, "ncoef": 2}
FIELDS = ["u", "v", "w", "p", "mip", "gp", "phic", "ap", "b", "anb", "bu", "d", "dc"]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def main():
    c = CASE
    g = meshgen.box_mesh(c["n"])
    with tempfile.TemporaryDirectory() as td:
        mf, of = os.path.join(td, "m.bin"), os.path.join(td, "o.bin")
        meshgen.write_refsim_mesh(g, mf)
        log, _ = refsim_io.run_refsim(mf, of, c["nsubd"], c["rho"], c["mu"], c["dt"],
                                      c["ntstep"], c["ncoef"], nit=c["nit"], timeout=7200)
        d = refsim_io.read_dump(of)
    out = {"case": c, "source": "oracle/_ref/refsim_O2 (the reference build)",
           "digests": {k: digest(d[k]) for k in FIELDS},
           "log": [[l[0], int(l[1]), float(l[2]), float(l[3]), float(l[4])] for l in log]}
    with open(os.path.join(HERE, "cavity128_reference_digests.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
