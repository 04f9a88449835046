"""Write the digests of the 128^3 benchmark cavity after two outer iterations.

The benchmark's configuration (128^3 lid-driven cavity, Re = 400, 4
subdomains, dt = 0.01, under-relaxation 0.7/0.3) is too large for a test to
run the oracle on (its set-up alone takes minutes), and its fields are too
large to keep.  This script runs the oracle (oracle/liboracle.so) once on it
and stores the SHA-256 digest of each field's little-endian FP64 bytes and
the solver lines in tests/golden/cavity128_relaxed_digests.json; the GPU test
compares the digests of its own fields with them.

    python tests/golden/make_digest128.py

The under-relaxation has no counterpart in the reference, so the reference
build cannot write these; the oracle's relaxed mode is pinned by its
reduction to the reference at 1/1 and by GPU = oracle at small sizes.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))
import meshgen  # noqa: E402
import oracle_lib  # noqa: E402

CASE = {"n": 128, "nsubd": 4, "rho": 1.0, "mu": 0.0025, "dt": 0.01, "nit": 100,
        "alpha_u": 0.7, "alpha_p": 0.3, "ntstep": 1, "ncoef": 2}
FIELDS = ["u", "v", "w", "p", "mip", "gp", "phic", "ap", "b", "anb", "bu", "d", "dc"]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def main():
    c = CASE
    g = meshgen.box_mesh(c["n"])
    oc = oracle_lib.OracleCase(g, meshgen.cavity_bc(g), nsubd=c["nsubd"], rho=c["rho"],
                               mu=c["mu"], dt=c["dt"], nit=c["nit"])
    oc.set_relaxation(c["alpha_u"], c["alpha_p"])
    oc.run(c["ntstep"], c["ncoef"])
    out = {"case": c, "digests": {k: digest(oc.arr(k)) for k in FIELDS},
           "log": [[l[0], int(l[1]), float(l[2]), float(l[3]), float(l[4])] for l in oc.log()]}
    with open(os.path.join(HERE, "cavity128_relaxed_digests.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
