"""Run the reference build (oracle/_ref/refsim_*) and read what it dumps.

Test infrastructure.  The dump layout is written by oracle/refdrv/drv_main.f90.
"""
import os
import re
import resource
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF_DIR = os.path.join(REPO, "oracle", "_ref")

# one residual line per linear solve, src/modules/mod_solver.f90:6
_LINE = re.compile(r"^\s+(\S+)\s+(\d+)\s+(\S+)\s+(\S+)\s+(\S+)\s*$")


def refsim_path(opt="O2"):
    return os.path.join(REF_DIR, "refsim_" + opt)


def have_refsim(opt="O2"):
    return os.path.exists(refsim_path(opt))


def _unlimited_stack():
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_refsim(meshfile, outfile, nsubd, rho, mu, dt, ntstep, ncoef, nit=100, opt="O2", timeout=600,
               binary=None, project=None, energy=False, props=None, symmetry=()):
    """project: a directory path ending in '/'; the run then ends with the
    reference's write_vtubin into project/VTK/output-<ntstep+1>/uvwp.vtu.
    props: (rho, mu) per-cell arrays, written next to outfile and read by the
    driver instead of the constants rho, mu.
    symmetry: boundary section names whose uvwp setter the driver rebinds to
    the reference's symmetry (mod_uvwp.f90:517-546)."""
    args = [binary or refsim_path(opt), meshfile, outfile, str(nsubd), repr(float(rho)), repr(float(mu)),
            repr(float(dt)), str(ntstep), str(ncoef), str(nit)]
    if project is not None:
        args.append(project)
    env = dict(os.environ, REFSIM_ENERGY="1" if energy else "0")
    env.pop("REFSIM_PROPS", None)
    env["REFSIM_SYMMETRY"] = " ".join(symmetry)
    if props is not None:
        pf = outfile + ".props"
        with open(pf, "wb") as f:
            f.write(np.ascontiguousarray(props[0], np.float64).tobytes())
            f.write(np.ascontiguousarray(props[1], np.float64).tobytes())
        env["REFSIM_PROPS"] = pf
    p = subprocess.run(args, capture_output=True, text=True, timeout=timeout,
                       preexec_fn=_unlimited_stack, env=env)
    if p.returncode != 0:
        raise RuntimeError("refsim failed: %s\n%s" % (p.returncode, p.stderr[-2000:]))
    log, secs = [], None
    run_refsim.last_stdout = p.stdout
    for line in p.stdout.splitlines():
        m = _LINE.match(line)
        if m and m.group(1) in ("u", "v", "w", "pc", "e"):
            log.append((m.group(1), int(m.group(2)), float(m.group(3)), float(m.group(4)),
                        float(m.group(5))))
        if line.startswith("refsim: seconds"):
            secs = float(line.split()[-1])
    return log, secs


class _Reader:
    def __init__(self, path):
        self.buf = open(path, "rb").read()
        self.off = 0

    def take(self, dtype, n):
        a = np.frombuffer(self.buf, dtype=dtype, count=n, offset=self.off).copy()
        self.off += a.nbytes
        return a


def read_dump(path):
    r = _Reader(path)
    ne, nf, nbf, nsubd = (int(v) for v in r.take("<i4", 4))
    L = 2 * nf - nbf
    n = ne + nbf
    d = dict(ne=ne, nf=nf, nbf=nbf, nsubd=nsubd)
    d["ef2nb_idx"] = r.take("<i4", ne + 1)
    d["ef2nb"] = np.concatenate([r.take("<i4", L), r.take("<i4", L)])
    d["s2g"] = r.take("<i4", nf)
    d["bs"] = r.take("<i4", nbf)
    for k in ("xc", "yc", "zc"):
        d[k] = r.take("<f8", n)
    d["vol"] = r.take("<f8", ne)
    d["aip"] = r.take("<f8", 3 * nf)
    d["rip"] = r.take("<f8", 3 * nf)
    if nsubd > 1:
        d["gf2g"] = r.take("<i4", ne)
        d["g2gf"] = r.take("<i4", ne)
        d["g2gf_idx"] = r.take("<i4", nsubd + 1)
        d["subd"] = []
        d["intf"] = {}
        for c in range(nsubd):
            sne, snf, snbf = (int(v) for v in r.take("<i4", 3))
            sL = 2 * snf - snbf
            sd = dict(ne=sne, nf=snf, nbf=snbf, ef2nb_idx=r.take("<i4", sne + 1), ef2nb=r.take("<i4", sL))
            d["subd"].append(sd)
            for cnb in range(nsubd):
                ncs = int(r.take("<i4", 1)[0])
                d["intf"][(c, cnb)] = r.take("<i4", ncs)
    for k in ("u", "v", "w", "p", "u0", "v0", "w0"):
        d[k] = r.take("<f8", n)
    for k in ("gu", "gv", "gw", "gp"):
        d[k] = r.take("<f8", 3 * n)
    d["gpc"] = r.take("<f8", 3 * ne)
    d["mip"] = r.take("<f8", nf)
    d["mip0"] = r.take("<f8", nf)
    for k in ("bu", "bv", "bw", "d", "dc"):
        d[k] = r.take("<f8", ne)
    d["ap"] = r.take("<f8", ne)
    d["b"] = r.take("<f8", ne)
    d["anb"] = r.take("<f8", L)
    d["phic"] = r.take("<f8", n)
    if r.off < len(r.buf):  # energy fields (REFSIM_ENERGY=1)
        d["t"] = r.take("<f8", n)
        d["gt"] = r.take("<f8", 3 * ne)
        d["h"] = r.take("<f8", n)
        d["h0"] = r.take("<f8", ne)
        d["gh"] = r.take("<f8", 3 * ne)
    assert r.off == len(r.buf), (r.off, len(r.buf))
    return d
