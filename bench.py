"""SIMPLE throughput on one MI355X: cell-iterations per second.

Headline workload (the GPU FP64 configuration): lid-driven cavity on a
uniform 128x128x128 hex box (2,097,152 cells, unstructured face lists),
Re=400 (rho=1, mu=0.0025), lid u=1, fields start at zero; steady SIMPLE(C)
(dt = infinity: no pseudo-time term, the loop that converges to the 1e-6
scaled residual) with under-relaxation 0.7 (u) and 0.3 (p), the reference's
exact-order Gauss-Seidel solvers and 4-subdomain pressure-correction solver,
FP64.  One step = one outer iteration (update_boundaries + solve_uvwp).
Inputs are resident in HBM when the timed region starts.  The timed steps
are a prefix of the run (residual_target_met false); the "converge" block
holds runs to the 1e-6 scaled residual.

Prints one JSON line.  value = 2,097,152 * K / seconds for K timed steps.
roofline: the pressure-correction Gauss-Seidel sweep kernel k_gs_tiles_lds (the dominant
kernel), 108 algorithmic bytes per cell update (FP64 CSR sweep: b, ap, 6
anb, 6 neighbour indices, own phi read and write) times the cells of one
launch, over the mean launch duration: HIP events on the context stream
around every 7th sweep of the timed region (all of its launches, back to
back), elapsed time divided by launches.  traffic: the HBM bytes per launch
that rocprofv3 counted for this kernel on the 128^3 workload with the
default schedule (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors a 256 MiB
device copy gives), reported for that workload only; --traffic-kib gives
another measured figure.
cpu_baseline: the reference Fortran (flang -O2 build in oracle/_ref) on one
host core, 64^3 sample (see --cpu-*), or the oracle's C restatement when the
reference build is absent; "port" beside it: the oracle's restatement of the
headline algorithm itself (steady, relaxed) on one core, 64^3 sample.
Extras in the same line (--no-extras skips them): "converge" (runs to the
1e-6 scaled residual: the 32^3 Re=100 cavity in the reference order and in
the multigrid mode, the 128^3 workload in the multigrid mode), "modes_128"
and "modes_256" (the extension solvers on the 128^3 workload and on the
256^3 Re=1000 cavity: cell-iterations/s, whole-iteration HBM fraction from
the algorithmic bytes of every kernel, field error against the FP64 runs),
"channel_128" (the distorted channel: 128^3 jittered, renumbered and RCM
ordered, Re=200).
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cfd-lite_amd"))

METRIC = "SIMPLE cell-iters/s at 1e-6 residual; batched solves/s + p50/p99 latency"
BYTES_PER_CELL_SWEEP = 108.0
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per cell of one outer iteration (each array element a
# kernel touches counted once; FP64 8 B, FP32 and int32 4 B; hexahedral
# interior cell): 2108 B for assembly, gradients, face fluxes and the update,
# 100 + 316 n per momentum solve of n Gauss-Seidel iterations, and for the
# pc solve of n iterations:
#   reference / colour order: 100 + 226 n  (two FP64 sweeps of 108 B and a
#     residual every tenth iteration)
#   "mixed" (FP32 colour sweeps): 184 + 146 n  (FP32 copies of ap and anb,
#     two 68 B sweeps, an FP64 residual every tenth iteration)
#   "mgcg" (FP64 V-cycle): 190 + 580 n  (initial residual and the Galerkin
#     sums; per CG iteration a forward and a backward colour sweep of 108 B,
#     the restriction 104 B and prolongation 20 B on the mesh, 8/7 of that
#     over the coarse levels, plus 196 B of CG vectors and the 100 B product)
#   "mgcg-mixed" (FP32 V-cycle): 274 + 442 n
BYTES_FIXED = 2108.0
PC_BYTES = {"reference": (100.0, 226.0), "colored": (100.0, 226.0), "mixed": (184.0, 146.0),
            "mgcg": (190.0, 580.0), "mgcg-mixed": (274.0, 442.0)}


def iteration_bytes(log, pc_order):
    """Algorithmic bytes per cell of the outer iterations in log (one line
    per linear solve, u v w pc per iteration): the sum over them."""
    b0, b1 = PC_BYTES[pc_order]
    total = 0.0
    for name, it, ri, rf, rm in log:
        if name == "pc":
            total += BYTES_FIXED + b0 + b1 * it
        else:
            total += 100.0 + 316.0 * it
    return total


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def progress(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def scaled_residual(log):
    """Per-equation initial residual divided by its max over the first five
    outer iterations; returns the largest over u, v, w, pc at the last one."""
    hist = {}
    for name, it, ri, rf, rm in log:
        hist.setdefault(name, []).append(ri)
    worst = 0.0
    for name, h in hist.items():
        ref = max(h[:5]) if h else 0.0
        if ref > 0:
            worst = max(worst, h[-1] / ref)
    return worst


def run_to_convergence(eq, ne, tstep_every, max_outer, tol, check_every=10):
    """Outer iterations until every equation's scaled residual (its initial
    residual over the maximum of its first five) is at most tol, or max_outer.
    Returns (outer iterations, seconds, final worst scaled residual)."""
    hist = {}
    k = 0
    worst = float("inf")
    eq.clear_log()
    eq.synchronize()
    t = time.perf_counter()
    while k < max_outer:
        for _ in range(min(check_every, max_outer - k)):
            eq.update_boundaries()
            eq.solve_uvwp()
            k += 1
            if k % tstep_every == 0:
                eq.update_time()
        for name, it, ri, rf, rm in eq.log():
            hist.setdefault(name, []).append(ri)
        eq.clear_log()
        worst = 0.0
        for h in hist.values():
            ref = max(h[:5])
            worst = max(worst, h[-1] / ref if ref > 0 else 0.0)
        print("bench: outer %d, worst scaled residual %.3e" % (k, worst), file=sys.stderr, flush=True)
        if len(next(iter(hist.values()))) >= 5 and worst <= tol:
            break
    eq.synchronize()
    return k, time.perf_counter() - t, worst


# HBM KiB per launch of the pc sweep kernel on the 128^3 cavity, default
# schedule, from rocprofv3 counter runs (profiles/k_gs_tiles_lds_traffic_
# cavity128_fp64.md): mean FETCH_SIZE 6246.0 KiB x 1.99975 (the factor a copy
# of known size calibrates) + mean WRITE_SIZE 463.2 KiB x 1.0 over the 7400
# launches of one outer iteration.  About half the fetch is the prefetch of
# the next launch's batches, which that launch then fetches again.
TRAFFIC_KIB_128 = 12953.6

# host <-> device traffic of one drop-in solve_uvwp call
# (cfd-lite_amd/fortran/dropin/mod_uvwp.f90: put_state, get_state, get_work)
PUT_STATE = ("u", "v", "w", "p", "u0", "v0", "w0", "gu", "gv", "gw", "gp", "mip", "mip0")
GET_STATE = ("u", "v", "w", "p", "gu", "gv", "gw", "gp", "gpc", "mip", "bu", "bv", "bw", "d", "dc")
GET_WORK = ("ap", "anb", "b", "phic")
PULL_EXTRA = ("u0", "v0", "w0", "mip0")


def pcie_rate(eq, ne, tstep_every, steps=30, rho=None, mu=None):
    """Outer iterations as the drop-in Fortran modules run them, with host
    arrays (pageable, as the Fortran host's):
      * per_call -- with the reference's mod_physics: the state goes to the
        device before every solve_uvwp and the state and work arrays come
        back after it;
      * resident -- with the drop-in mod_physics: the properties go up and
        the new solver lines come back every call, update_time runs on the
        device, and the state and work arrays come back every 10th time step.
    Returns both rates; the headline value stays the device-resident rate."""
    host = {k: eq.get(k) for k in set(PUT_STATE + GET_STATE + GET_WORK + PULL_EXTRA)}
    rho = np.ones(ne) if rho is None else rho
    mu = np.full(ne, 0.0025) if mu is None else mu
    out = {}
    for mode in ("per_call", "resident"):
        eq.synchronize()
        moved = 0
        t = time.perf_counter()
        for k in range(steps):
            if mode == "per_call":
                for name in PUT_STATE:
                    eq.set(name, host[name])
                    moved += host[name].nbytes
            eq.set_properties(rho, mu)
            moved += rho.nbytes + mu.nbytes
            eq.update_boundaries()
            eq.solve_uvwp()
            eq.log()
            eq.clear_log()
            if mode == "per_call":
                for name in GET_STATE + GET_WORK:
                    host[name] = eq.get(name)
                    moved += host[name].nbytes
            if (k + 1) % tstep_every == 0:
                eq.update_time()
                if mode == "resident" and ((k + 1) // tstep_every) % 10 == 0:
                    for name in GET_STATE + GET_WORK + PULL_EXTRA:
                        host[name] = eq.get(name)
                        moved += host[name].nbytes
        eq.synchronize()
        secs = time.perf_counter() - t
        out[mode] = {"value": round(ne * steps / secs, 1), "unit": "cell-iterations/s",
                     "ms_per_step": round(1e3 * secs / steps, 3),
                     "host_device_bytes_per_step": int(moved / steps), "steps": steps}
    return out


def write_ref_mesh(mesh, path):
    """Mesh file read by the reference driver (oracle/refdrv/drv_cell.f90)."""
    info = mesh.info()
    names = mesh.names
    etype, esec = mesh.get("etype"), mesh.get("esec")
    with open(path, "wb") as f:
        f.write(np.array([info["nvx"], info["nsec"]], np.int32).tobytes())
        for s in range(info["nsec"]):
            f.write(names[s].encode().ljust(32, b" "))
            f.write(np.array([etype[s], esec[2 * s], esec[2 * s + 1]], np.int32).tobytes())
        for k in ("x", "y", "z"):
            f.write(mesh.get(k).astype("<f8").tobytes())
        f.write(mesh.get("e2vx").astype("<i4").tobytes())


def cpu_baseline(n, nouter, mu):
    import cfdl
    ref = os.path.join(REPO, "oracle", "_ref", "refsim_O2")
    ncells = n ** 3
    if os.path.exists(ref):
        mesh = cfdl.Mesh.box(n)
        with tempfile.TemporaryDirectory() as td:
            mf, of = os.path.join(td, "mesh.bin"), os.path.join(td, "out.bin")
            write_ref_mesh(mesh, mf)

            def unlimited():
                import resource
                resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY,) * 2)
                os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})

            p = subprocess.run([ref, mf, of, "4", "1.0", repr(mu), "0.01", "1", str(nouter), "100"],
                               capture_output=True, text=True, timeout=600, preexec_fn=unlimited)
            if p.returncode != 0:
                raise RuntimeError("reference run failed: " + p.stderr[-500:])
            secs = [float(l.split()[-1]) for l in p.stdout.splitlines() if l.startswith("refsim: seconds")][0]
        return {"value": ncells * nouter / secs, "unit": "cell-iterations/s", "cores": 1,
                "kind": "reference",
                "sample": "%d^3 lid-driven cavity, Re=%g, %d outer iterations of the reference "
                          "Fortran (flang -O2, 4 subdomains, dt=0.01, no relaxation) on one host "
                          "core, setup excluded; %.1f s" % (n, 1.0 / mu, nouter, secs)}
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib  # test infrastructure: the oracle as checker/baseline only
    mesh = cfdl.Mesh.box(n)
    g = {"nvx": mesh.info()["nvx"], "x": mesh.get("x"), "y": mesh.get("y"), "z": mesh.get("z"),
         "etype": mesh.get("etype"), "esec": mesh.get("esec"), "e2vx": mesh.get("e2vx")}
    kind = np.array([-1, 0, 0, 0, 0, 0, 1], np.int32)
    oc = oracle_lib.OracleCase(g, kind, nsubd=4, rho=1.0, mu=mu, dt=0.01)
    t = time.perf_counter()
    oc.run(1, nouter)
    secs = time.perf_counter() - t
    return {"value": ncells * nouter / secs, "unit": "cell-iterations/s", "cores": 1,
            "kind": "port",
            "sample": "%d^3 lid-driven cavity, Re=%g, %d outer iterations of the oracle's C "
                      "restatement on one host core; %.1f s" % (n, 1.0 / mu, nouter, secs)}


def port_baseline(n, nouter, mu):
    """The oracle's C restatement of the headline algorithm itself (steady,
    relaxation 0.7 / 0.3, reference-order solvers) on one host core, on an
    n^3 sample: the like-for-like CPU rate beside the reference's own."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib  # test infrastructure: the oracle as checker/baseline only
    import cfdl
    mesh = cfdl.Mesh.box(n)
    g = {"nvx": mesh.info()["nvx"], "x": mesh.get("x"), "y": mesh.get("y"), "z": mesh.get("z"),
         "etype": mesh.get("etype"), "esec": mesh.get("esec"), "e2vx": mesh.get("e2vx")}
    kind = np.array([-1, 0, 0, 0, 0, 0, 1], np.int32)
    oc = oracle_lib.OracleCase(g, kind, nsubd=4, rho=1.0, mu=mu, dt=float("inf"))
    oc.set_relaxation(0.7, 0.3)
    oc.run(1, 1)
    t = time.perf_counter()
    oc.run(1, nouter)
    secs = time.perf_counter() - t
    return {"value": n ** 3 * nouter / secs, "unit": "cell-iterations/s", "cores": 1,
            "kind": "port", "cpu": cpu_model(),
            "sample": "%d^3 lid-driven cavity, Re=%g, steady, relaxation 0.7/0.3, %d outer "
                      "iterations of the oracle's C restatement (the headline algorithm, "
                      "reference-order solvers) on one host core; %.1f s" % (n, 1.0 / mu, nouter, secs)}


def run_mode(cfdl, mesh, mu, order, gs, steps, warmup, channel=False, keep=False):
    """Steady outer iterations of one solver mode from rest on a fresh
    context: warmup untimed, then steps timed.  Returns the rate, the whole
    iteration's HBM fraction (algorithmic bytes of every kernel over the wall
    time) and, with keep, the fields."""
    ne = mesh.info()["ne"]
    mk = cfdl.UVWP.channel if channel else cfdl.UVWP.cavity
    eq = mk(mesh, rho=1.0, mu=mu, dt=float("inf"))
    eq.set_relaxation(0.7, 0.3)
    eq.set_pc_order(order)
    eq.set_gs_order(gs)
    for _ in range(warmup):
        eq.update_boundaries()
        eq.solve_uvwp()
    eq.clear_log()
    eq.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        eq.update_boundaries()
        eq.solve_uvwp()
    eq.synchronize()
    secs = time.perf_counter() - t
    log = eq.log()
    nbytes = ne * iteration_bytes(log, order)
    out = {"pc_order": order, "gs_order": gs, "value": round(ne * steps / secs, 1),
           "unit": "cell-iterations/s", "ms_per_step": round(1e3 * secs / steps, 3),
           "steps": steps, "warmup": warmup,
           "hbm_frac_whole_iteration": round(nbytes / secs / (HBM_PEAK_GBS * 1e9), 4),
           "pc_iterations": [l[1] for l in log if l[0] == "pc"][-3:],
           "dtype": "f32 pc preconditioner / sweeps, f64 otherwise" if "mixed" in order else "f64"}
    if keep:
        out["fields"] = {k: eq.get(k)[:ne] for k in ("u", "v", "w", "p")}
    eq.close()
    return out


def field_error(a, b):
    """Relative L2 difference per field (velocities against the velocity
    norm, p against its own)."""
    vel = np.linalg.norm(np.concatenate([b[k] for k in ("u", "v", "w")]))
    return {k: float(np.linalg.norm(a[k] - b[k]) / (vel if k != "p" else np.linalg.norm(b[k])))
            for k in ("u", "v", "w", "p")}


def converge(cfdl, mesh, mu, order, gs, max_outer, tol, channel=False, keep=False):
    """Steady outer iterations from rest to the scaled residual tol or
    max_outer iterations (the configurations' criterion: 1e-6 or 500)."""
    ne = mesh.info()["ne"]
    mk = cfdl.UVWP.channel if channel else cfdl.UVWP.cavity
    eq = mk(mesh, rho=1.0, mu=mu, dt=float("inf"))
    eq.set_relaxation(0.7, 0.3)
    eq.set_pc_order(order)
    eq.set_gs_order(gs)
    k, secs, worst = run_to_convergence(eq, ne, 1 << 30, max_outer, tol)
    out = {"pc_order": order, "gs_order": gs, "outer_iterations": k, "seconds": round(secs, 3),
           "converged": bool(worst <= tol), "scaled_residual": worst, "tol": tol,
           "cell_iterations_per_s": round(ne * k / secs, 1)}
    if keep:
        out["fields"] = {f: eq.get(f)[:ne] for f in ("u", "v", "w", "p")}
    eq.close()
    return out


def same_residual(cfdl, mesh, mu, plain, fast_modes, max_outer):
    """The fast modes from rest to the scaled residual the plain FP64 run
    reached (within max_outer): their time to that residual against the
    plain run's."""
    r = {}
    for order, gs in fast_modes:
        f = converge(cfdl, mesh, mu, order, gs, max_outer, max(plain["scaled_residual"], 1e-12))
        if f["converged"] and f["seconds"] > 0:
            f["time_ratio_vs_plain_fp64"] = round(plain["seconds"] / f["seconds"], 2)
        r[order + "/" + gs] = f
    return r


def extras(cfdl, args, base_mesh, headline):
    """The extension modes and the runs to the 1e-6 residual (see the module
    docstring); each section records its error instead of failing the line."""
    out = {}

    def section(name, fn):
        progress("section %s" % name)
        t = time.perf_counter()
        try:
            out[name] = fn()
        except Exception as e:  # noqa: BLE001 - recorded in the line
            out[name] = {"error": "%s: %s" % (type(e).__name__, e)}
        progress("section %s done in %.1f s" % (name, time.perf_counter() - t))

    def conv32():
        m = cfdl.Mesh.box(32).partition(4)
        # to 1e-6 (past the configuration's cap of 500, which stops it at
        # about 5e-6) for the time to solution
        ref = converge(cfdl, m, 0.01, "reference", "reference", 1000, args.tol, keep=True)
        fast = converge(cfdl, m, 0.01, "mgcg-mixed", "colored", 1000, args.tol, keep=True)
        fast["rel_l2_vs_fp64_reference_order_converged"] = field_error(fast.pop("fields"),
                                                                       ref.pop("fields"))
        r = {"case": "32^3 cavity, Re=100, steady, relaxation 0.7/0.3 (the CPU configuration), "
                     "to the 1e-6 scaled residual",
             "fp64_reference_order": ref, "mgcg_mixed": fast}
        if ref["converged"] and fast["converged"]:
            r["time_to_solution_ratio"] = round(ref["seconds"] / fast["seconds"], 2)
        # the oracle is bitwise the reference-order GPU path, so it needs the
        # same outer iterations: its time to solution from a timed sample
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import oracle_lib  # test infrastructure: the CPU rate only
        g = {"nvx": m.info()["nvx"], "x": m.get("x"), "y": m.get("y"), "z": m.get("z"),
             "etype": m.get("etype"), "esec": m.get("esec"), "e2vx": m.get("e2vx")}
        oc = oracle_lib.OracleCase(g, np.array([-1, 0, 0, 0, 0, 0, 1], np.int32), nsubd=4,
                                   rho=1.0, mu=0.01, dt=float("inf"))
        oc.set_relaxation(0.7, 0.3)
        t = time.perf_counter()
        oc.run(1, 20)
        per = (time.perf_counter() - t) / 20
        r["cpu_port_seconds_to_solution"] = round(per * ref["outer_iterations"], 2)
        r["cpu_port_seconds_per_iteration"] = round(per, 4)
        return r

    def modes128():
        r = {"case": "128^3 cavity, Re=%g, steady, relaxation 0.7/0.3, 10 outer iterations "
                     "after 2" % args.re,
             "reference": {"value": headline["value"], "hbm_frac_whole_iteration":
                           headline["hbm_frac_whole_iteration"]}}
        for order in ("colored", "mixed", "mgcg", "mgcg-mixed"):
            r[order] = run_mode(cfdl, base_mesh, 1.0 / args.re, order, "colored", 10, 2)
        return r

    def conv128():
        mu = 1.0 / args.re
        plain = converge(cfdl, base_mesh, mu, "reference", "reference", 500, args.tol)
        progress("128^3 plain FP64: %d outer iterations, scaled residual %.3e, %.1f s"
                 % (plain["outer_iterations"], plain["scaled_residual"], plain["seconds"]))
        return {"case": "128^3 cavity, Re=%g, steady, relaxation 0.7/0.3, 1e-6 or 500 outer "
                        "iterations (the configuration's criterion); the fast modes run to the "
                        "residual the plain FP64 path reached" % args.re,
                "plain_fp64_reference_order": plain,
                "same_residual": same_residual(cfdl, base_mesh, mu, plain,
                                               (("mixed", "colored"), ("mgcg-mixed", "colored")),
                                               2000)}

    def modes256():
        m = cfdl.Mesh.box(256).partition(4)
        K = 4
        runs = {}
        for order, gs in (("reference", "reference"), ("colored", "colored"),
                          ("mixed", "colored"), ("mgcg", "colored"), ("mgcg-mixed", "colored")):
            runs[order] = run_mode(cfdl, m, 1e-3, order, gs, K, 0, keep=True)
            progress("256^3 %s: %.1f M cell-iterations/s" % (order, runs[order]["value"] / 1e6))
        for order in ("colored", "mixed", "mgcg", "mgcg-mixed"):
            runs[order]["rel_l2_vs_fp64_reference_order"] = field_error(
                runs[order]["fields"], runs["reference"]["fields"])
        for order, ref in (("mixed", "colored"), ("mgcg-mixed", "mgcg")):
            runs[order]["rel_l2_vs_fp64_same_solver"] = field_error(
                runs[order]["fields"], runs[ref]["fields"])
        for v in runs.values():
            v.pop("fields")
        base = runs["reference"]["value"]
        for order in runs:
            runs[order]["x_fp64_reference_order"] = round(runs[order]["value"] / base, 2)
        # the mixed-precision configuration's run: 1e-6 or 500 outer iterations
        runs["mixed_500"] = converge(cfdl, m, 1e-3, "mixed", "colored", 500, args.tol)
        return dict(case="256^3 cavity, Re=1000, steady, relaxation 0.7/0.3, %d outer "
                         "iterations from rest (mixed_500: to 1e-6 or 500)" % K, **runs)

    def channel128():
        m = cfdl.Mesh.box(128, jitter=0.2, seed=20261202).renumber(20261202).partition(4)
        r = run_mode(cfdl, m, 1.0 / 200.0, "mgcg-mixed", "colored", 10, 2, channel=True)
        return dict(case="128^3 channel, interior vertices jittered +-0.2h, cells shuffled then "
                         "RCM ordered, parabolic inlet u_max=1, zero-gradient outlet, Re=200, "
                         "steady", **r)

    def transfers128():
        mu = 1.0 / args.re
        eq = cfdl.UVWP.cavity(base_mesh, rho=1.0, mu=mu, dt=float("inf"))
        eq.set_relaxation(args.alpha_u, args.alpha_p)
        ne = base_mesh.info()["ne"]
        r = pcie_rate(eq, ne, args.tstep_every, 30, np.ones(ne), np.full(ne, mu))
        eq.close()
        return dict(case="the headline configuration (128^3, reference order) driven the way "
                         "the Fortran drop-in modules drive it, host arrays pageable; "
                         "per_call: reference mod_physics, resident: drop-in mod_physics",
                    device_resident=headline["value"], **r)

    def serving():
        sat = serving_run(cfdl, 64, 1000.0, args.tol, 100, 10, 32, warm=2)
        stream = serving_run(cfdl, 48, 4.0, args.tol, 100, 10, 32, warm=2)
        return {"saturated": sat, "stream_4_per_s": stream,
                "note": "the configuration's 512-request stream to 1e-6 or 500 outer "
                        "iterations: bench.py --serving"}

    section("converge_cavity32", conv32)
    section("modes_128", modes128)
    section("converge_cavity128", conv128)
    section("modes_256", modes256)
    section("channel_128", channel128)
    section("dropin_transfers_128", transfers128)
    section("serving", serving)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128, help="cells per box edge")
    ap.add_argument("--re", type=float, default=400.0)
    ap.add_argument("--alpha-u", type=float, default=0.7)
    ap.add_argument("--alpha-p", type=float, default=0.3)
    ap.add_argument("--cpu-n", type=int, default=64)
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--port-iters", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline line (no converge / modes / channel blocks)")
    ap.add_argument("--sample-every", type=int, default=7, help="time every k-th pc sweep")
    ap.add_argument("--pseudo-transient", action="store_true",
                    help="the reference's pseudo-time loop (dt=0.01, update_time after every "
                         "--tstep-every outer iterations) instead of steady SIMPLE")
    ap.add_argument("--tstep-every", type=int, default=3)
    ap.add_argument("--converge", action="store_true",
                    help="only run from rest to a 1e-6 scaled residual (or --max-outer)")
    ap.add_argument("--max-outer", type=int, default=500)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--steady", action="store_true",
                    help="with --converge: dt = infinity (the default elsewhere)")
    ap.add_argument("--pc-order", choices=("reference", "colored", "mixed", "mgcg", "mgcg-mixed"),
                    default="reference",
                    help="pc solver of --converge runs")
    ap.add_argument("--gs-order", choices=("reference", "colored"), default="reference",
                    help="order of solve_gs of --converge runs")
    ap.add_argument("--assembly-bits", type=int, choices=(32, 64), default=64,
                    help="coefficient assembly precision of --converge runs")
    ap.add_argument("--serving", action="store_true",
                    help="only run the serving stream and report solves/s with p50/p99 latency")
    ap.add_argument("--serving-requests", type=int, default=512)
    ap.add_argument("--serving-rate", type=float, default=10.0, help="arrivals per second")
    ap.add_argument("--serving-cases", type=int, default=32, help="cases in one batch")
    ap.add_argument("--serving-max-outer", type=int, default=500)
    ap.add_argument("--serving-chunk", type=int, default=10,
                    help="outer iterations per batch run (convergence checked between runs)")
    ap.add_argument("--serving-warm", type=int, default=2,
                    help="contexts per (kind, N) created before the stream starts")
    ap.add_argument("--pcie", action="store_true",
                    help="also time outer iterations with the host transfers of the drop-in "
                         "Fortran solve_uvwp (state up, state and work arrays down)")
    ap.add_argument("--traffic-kib", type=float, default=None,
                    help="measured HBM KiB per sweep launch (rocprofv3 FETCH_SIZE + WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; --gpus %d is not supported "
              "(the solver does not shard the mesh over GPUs)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import cfdl
    if args.serving:
        serving_line(cfdl, args)
        return
    mu = 1.0 / args.re
    t0 = time.perf_counter()
    mesh = cfdl.Mesh.box(args.n).partition(4)
    steady = not args.pseudo_transient
    dt = float("inf") if steady else 0.01
    eq = cfdl.UVWP.cavity(mesh, rho=1.0, mu=mu, dt=dt, nit=100)
    eq.set_relaxation(args.alpha_u, args.alpha_p)
    ne = mesh.info()["ne"]
    progress("setup %.1f s, %d cells" % (time.perf_counter() - t0, ne))

    if args.converge:
        if args.steady:
            eq.set_dt(float("inf"))
        eq.set_pc_order(args.pc_order)
        eq.set_gs_order(args.gs_order)
        eq.set_assembly_precision(args.assembly_bits)
        k, secs, worst = run_to_convergence(eq, ne, args.tstep_every, args.max_outer, args.tol)
        print(json.dumps({
            "metric": METRIC, "mode": "converge", "pc_order": args.pc_order,
            "assembly_bits": args.assembly_bits,
            "gs_order": args.gs_order, "steady": bool(args.steady or steady),
            "value": round(ne * k / secs, 1),
            "unit": "cell-iterations/s", "outer_iterations": k, "seconds": round(secs, 3),
            "converged": worst <= args.tol, "scaled_residual": worst, "tol": args.tol,
            "n_gpus": 1, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "lid-driven cavity %d^3 hex (%d cells), Re=%g, under-relaxation "
                                   "%g/%g" % (args.n, ne, args.re, args.alpha_u, args.alpha_p)}}),
              flush=True)
        return

    step_no = [0]

    def step():
        eq.update_boundaries()
        eq.solve_uvwp()
        step_no[0] += 1
        if not steady and step_no[0] % args.tstep_every == 0:
            eq.update_time()

    if args.pcie:
        r = pcie_rate(eq, ne, args.tstep_every)
        print(json.dumps({"mode": "pcie", "metric": METRIC, **r}), flush=True)

    for _ in range(args.warmup):
        step()
    eq.synchronize()
    progress("warmup done")
    eq.reset_stats()
    eq.set_profile(args.sample_every)
    eq.synchronize()
    t = time.perf_counter()
    for k in range(args.steps):
        step()
        progress("step %d queued" % (k + 1))
    eq.synchronize()
    secs = time.perf_counter() - t
    eq.set_profile(0)
    st = eq.stats()
    log = eq.log()
    value = ne * args.steps / secs
    kernel_ms = st["launch_ms"]
    cells = st["launch_cells"]
    achieved = BYTES_PER_CELL_SWEEP * cells / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    traffic_kib = args.traffic_kib
    schedule_env = any(os.environ.get(k) for k in ("CFDL_TILE", "CFDL_TILE_SPAN", "CFDL_GS_MODE",
                                                  "CFDL_TILE_LDS", "CFDL_TILE_PIPE", "CFDL_TILE_WAVES",
                                                  "CFDL_TILE_PREFETCH"))
    if traffic_kib is None and args.n == 128 and not schedule_env:
        traffic_kib = TRAFFIC_KIB_128
    traffic = None if traffic_kib is None else round(traffic_kib * 1024.0)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "algorithmic_bytes_per_launch": round(BYTES_PER_CELL_SWEEP * cells),
                "kernel": "k_gs_tiles_lds<true, true, 1, 4> (pressure-correction Gauss-Seidel, space-time tiles in LDS)",
                "launch_ms": round(kernel_ms, 5), "cells_per_launch": round(cells, 1),
                "sampled_launches": int(st["sampled_launches"])}
    hbm_whole = ne * iteration_bytes(log[-4 * args.steps:], "reference") / secs / (HBM_PEAK_GBS * 1e9)
    progress("pc level launches %d (%.0f cells each), %d levels per symmetric sweep; "
             "outer iteration %.1f ms" % (st["pc_launches"], st["pc_cells"] / max(st["pc_launches"], 1),
                                          st["pc_levels_per_sym_sweep"], 1e3 * secs / args.steps))
    base = None
    if not args.no_cpu:
        base = cpu_baseline(args.cpu_n, args.cpu_iters, mu)
        base["cpu"] = cpu_model()
        progress("cpu baseline %.3g cell-iterations/s" % base["value"])
        try:
            base["port_like_for_like"] = port_baseline(args.cpu_n, args.port_iters, mu)
        except Exception as e:  # noqa: BLE001 - recorded in the line
            base["port_like_for_like"] = {"error": str(e)}
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-iterations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * secs / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "lid-driven cavity %d^3 hex (%d cells), Re=%g, %s FP64 SIMPLE(C) "
                               "with exact-order Gauss-Seidel and 4-subdomain pc solver, "
                               "under-relaxation %g/%g" % (
                                   args.n, ne, args.re, "steady" if steady else "pseudo-transient",
                                   args.alpha_u, args.alpha_p),
                   "mesh": "%dx%dx%d" % (args.n, args.n, args.n), "subdomains": 4,
                   "dt": "inf" if steady else 0.01,
                   "outer_iterations_timed": args.steps,
                   "scaled_residual_reached": scaled_residual(log),
                   "residual_target_met": False,
                   "hbm_frac_whole_iteration": round(hbm_whole, 4)},
        "roofline": roofline,
        "cpu_baseline": base,
    }
    if not args.no_extras:
        eq.close()
        out["extras"] = extras(cfdl, args, mesh, {"value": out["value"],
                                                  "hbm_frac_whole_iteration": round(hbm_whole, 4)})
    print(json.dumps(out), flush=True)


def serving_run(cfdl, nreq, rate, tol, max_outer, chunk, max_cases, seed=20261202, warm=0):
    """The serving stream (an extension; the reference solves one case per
    run): seeded Poisson arrivals of steady cavity and channel solves from
    rest (N^3, N in {16, ..., 96}, Re log-uniform in [100, 1000], kind a fair
    draw), continuously batched: admitted into one batch as they arrive and
    memory frees up, all cases in the same grouped launches, each run to its
    scaled residual tol or max_outer outer iterations."""
    from cfdl import serving
    reqs = serving.make_requests(nreq, rate, seed)
    progress("serving %d requests, %g per s, tol %g or %d outer iterations"
             % (nreq, rate, tol, max_outer))
    t0 = time.perf_counter()
    last = [t0]
    seen = [0]

    def on_run(ncases):
        seen[0] += 1
        now = time.perf_counter()
        if now - last[0] > 20.0:
            last[0] = now
            progress("serving: %.0f s, batch run %d, %d cases in the batch"
                     % (now - t0, seen[0], ncases))

    out = serving.serve(reqs, max_cases=max_cases, tol=tol, max_outer=max_outer, chunk=chunk,
                        warm=warm, on_run=on_run)
    cell_its = sum(r.n ** 3 * out["outer"][r.idx] for r in reqs)
    return {"metric": "batched solves/s with p50/p99 latency",
            "solves_per_s": round(out["solves_per_s"], 3),
            "p50_ms": round(1e3 * out["p50_s"], 1), "p99_ms": round(1e3 * out["p99_s"], 1),
            "wall_s": round(out["wall_s"], 2), "requests": nreq, "arrival_rate_per_s": rate,
            "max_cases_in_batch": max_cases,
            "mean_cases_per_batch_run": round(out["case_runs"] / max(out["runs"], 1), 2),
            "outer_iterations_per_batch_run": chunk,
            "converged": int(sum(out["converged"].values())), "failed": len(out["failed"]),
            "tol": tol, "max_outer": max_outer,
            "cell_iterations_per_s": round(cell_its / out["wall_s"], 1),
            "contexts_created": out["contexts_created"],
            "contexts_warmed_before_start": out["contexts_warmed"],
            "warm_up_s": round(out["warm_s"], 2),
            "host_seconds": {"admission": round(out["admit_s"], 2),
                             "batch_runs": round(out["run_s"], 2),
                             "retirement": round(out["retire_s"], 2)},
            "case": "steady lid-driven cavity or channel (50/50) on N^3 hex, N in {16,...,96}, "
                    "Re log-uniform in [100, 1000], relaxation 0.7/0.3, from rest",
            "pc_order": "mgcg-mixed", "gs_order": "colored", "assembly": "f32",
            "dtype": "f32 assembly and pc preconditioner, f64 otherwise", "data": "synthetic"}


def serving_line(cfdl, args):
    out = serving_run(cfdl, args.serving_requests, args.serving_rate, args.tol,
                      args.serving_max_outer, args.serving_chunk, args.serving_cases,
                      warm=args.serving_warm)
    print(json.dumps(dict(out, mode="serving")), flush=True)


if __name__ == "__main__":
    main()
