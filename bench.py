"""SIMPLE throughput on one MI355X: cell-iterations per second.

Workload (the GPU FP64 configuration): lid-driven cavity on a uniform
128x128x128 hex box (2,097,152 cells, unstructured face lists), Re=400
(rho=1, mu=0.0025), lid u=1, fields start at zero; the reference SIMPLE(C)
outer iteration (solve_uvwp) with its exact-order Gauss-Seidel solvers and
4-subdomain pressure-correction solver, FP64, pseudo-time step dt=0.01
with 3 outer iterations per step, plus config-mode under-relaxation 0.7 (u)
and 0.3 (p).  One step = one outer iteration (update_boundaries +
solve_uvwp; update_time after every third).  Inputs are resident in HBM when
the timed region starts.

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
reference build is absent.
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
# cavity128_fp64.md): mean FETCH_SIZE 3630.8 KiB x 1.99974 + mean WRITE_SIZE
# 445.7 KiB x 1.0 over the 7400 launches of one outer iteration
TRAFFIC_KIB_128 = 7706.4

# host <-> device traffic of one drop-in solve_uvwp call
# (cfd-lite_amd/fortran/dropin/mod_uvwp.f90: put_state, get_state, get_work)
PUT_STATE = ("u", "v", "w", "p", "u0", "v0", "w0", "gu", "gv", "gw", "gp", "mip", "mip0")
GET_STATE = ("u", "v", "w", "p", "gu", "gv", "gw", "gp", "gpc", "mip", "bu", "bv", "bw", "d", "dc")
GET_WORK = ("ap", "anb", "b", "phic")


def pcie_rate(eq, ne, args, steps=4):
    """Outer iterations as the drop-in Fortran modules run them: the host
    arrays (pageable, as the Fortran host's) go to the device before every
    solve_uvwp and come back after it.  Printed as its own JSON line; the
    headline value stays the device-resident rate."""
    host = {k: eq.get(k) for k in set(PUT_STATE + GET_STATE + GET_WORK)}
    nbytes = 8 * (sum(host[k].size for k in PUT_STATE) +
                  sum(host[k].size for k in GET_STATE + GET_WORK))
    eq.synchronize()
    t = time.perf_counter()
    for k in range(steps):
        for name in PUT_STATE:
            eq.set(name, host[name])
        eq.update_boundaries()
        eq.solve_uvwp()
        for name in GET_STATE + GET_WORK:
            host[name] = eq.get(name)
        if (k + 1) % args.tstep_every == 0:
            eq.update_time()
    eq.synchronize()
    secs = time.perf_counter() - t
    print(json.dumps({"mode": "pcie", "metric": METRIC, "value": round(ne * steps / secs, 1),
                      "unit": "cell-iterations/s", "ms_per_step": round(1e3 * secs / steps, 3),
                      "host_device_bytes_per_step": nbytes, "steps": steps}), flush=True)


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
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sample-every", type=int, default=7, help="time every k-th pc sweep")
    ap.add_argument("--tstep-every", type=int, default=3,
                    help="outer iterations per pseudo-time step (update_time after each group)")
    ap.add_argument("--converge", action="store_true",
                    help="run from rest to a 1e-6 scaled residual (or --max-outer) instead")
    ap.add_argument("--max-outer", type=int, default=500)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--steady", action="store_true",
                    help="dt = infinity: steady SIMPLE (no pseudo-time term), the mode that "
                         "converges to the 1e-6 scaled residual")
    ap.add_argument("--pc-order", choices=("reference", "colored", "mixed", "mgcg", "mgcg-mixed"),
                    default="reference",
                    help="visiting order of the pressure-correction smoother; 'colored' is an "
                         "extension (greedy colouring, one launch per colour) reported on its "
                         "own line, not as the headline value; 'mixed' also runs the colour "
                         "order with FP32 sweeps under FP64 defect correction and reports its "
                         "field error against the FP64 colour order")
    ap.add_argument("--gs-order", choices=("reference", "colored"), default="reference",
                    help="visiting order of solve_gs (the momentum solves) in the --pc-order "
                         "extension lines; the headline value always keeps the reference order")
    ap.add_argument("--serving", action="store_true",
                    help="also run a serving stream (Poisson arrivals of independent cavity "
                         "solves of edge 16..96 and Re 100..1000, each the reference's 10 x 3 "
                         "outer iterations) and report solves/s with p50/p99 latency")
    ap.add_argument("--serving-requests", type=int, default=512)
    ap.add_argument("--serving-rate", type=float, default=10.0, help="arrivals per second")
    ap.add_argument("--serving-slots", type=int, default=8, help="requests in flight")
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
    mu = 1.0 / args.re
    t0 = time.perf_counter()
    mesh = cfdl.Mesh.box(args.n).partition(4)
    eq = cfdl.UVWP.cavity(mesh, rho=1.0, mu=mu, dt=0.01, nit=100)
    eq.set_relaxation(args.alpha_u, args.alpha_p)
    ne = mesh.info()["ne"]
    print("bench: setup %.1f s, %d cells" % (time.perf_counter() - t0, ne), file=sys.stderr, flush=True)

    if args.converge:
        if args.steady:
            eq.set_dt(float("inf"))
        eq.set_pc_order(args.pc_order)
        eq.set_gs_order(args.gs_order)
        k, secs, worst = run_to_convergence(eq, ne, args.tstep_every, args.max_outer, args.tol)
        print(json.dumps({
            "metric": METRIC, "mode": "converge", "pc_order": args.pc_order,
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
        if step_no[0] % args.tstep_every == 0:
            eq.update_time()

    if args.pcie:
        pcie_rate(eq, ne, args)

    for _ in range(args.warmup):
        step()
    eq.synchronize()
    print("bench: warmup done", file=sys.stderr, flush=True)
    eq.reset_stats()
    eq.set_profile(args.sample_every)
    eq.synchronize()
    t = time.perf_counter()
    for k in range(args.steps):
        step()
        print("bench: step %d queued" % (k + 1), file=sys.stderr, flush=True)
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
                                                  "CFDL_TILE_LDS", "CFDL_TILE_PIPE"))
    if traffic_kib is None and args.n == 128 and not schedule_env:
        traffic_kib = TRAFFIC_KIB_128
    traffic = None if traffic_kib is None else round(traffic_kib * 1024.0)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "algorithmic_bytes_per_launch": round(BYTES_PER_CELL_SWEEP * cells),
                "kernel": "k_gs_tiles_lds<true, true, 1> (pressure-correction Gauss-Seidel, space-time tiles in LDS)",
                "launch_ms": round(kernel_ms, 5), "cells_per_launch": round(cells, 1),
                "sampled_launches": int(st["sampled_launches"])}
    base = None if args.no_cpu else cpu_baseline(args.cpu_n, args.cpu_iters, mu)
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-iterations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * secs / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "lid-driven cavity %d^3 hex (%d cells), Re=%g, FP64 SIMPLE(C) with "
                               "exact-order Gauss-Seidel and 4-subdomain pc solver, "
                               "under-relaxation %g/%g" % (args.n, ne, args.re, args.alpha_u,
                                                          args.alpha_p),
                   "mesh": "%dx%dx%d" % (args.n, args.n, args.n), "subdomains": 4, "dt": 0.01,
                   "outer_iterations_timed": args.steps,
                   "scaled_residual_reached": scaled_residual(log)},
        "roofline": roofline,
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)
    print("bench: pc level launches %d (%.0f cells each), %d levels per symmetric sweep; "
          "outer iteration %.1f ms" % (st["pc_launches"], st["pc_cells"] / max(st["pc_launches"], 1),
                                       st["pc_levels_per_sym_sweep"], 1e3 * secs / args.steps),
          file=sys.stderr)
    if args.serving:
        serving_line(args)
    if args.pc_order in ("colored", "mixed"):
        ref = colored_rate(cfdl, mesh, mu, ne, args, "colored")
    if args.pc_order == "mixed":
        colored_rate(cfdl, mesh, mu, ne, args, "mixed", ref)


def serving_line(args):
    """The serving stream (an extension; the reference solves one case per
    run): seeded Poisson arrivals, each request a cavity solved from rest on
    a context kept per edge, in the fastest mode (mixed pc solve, colour
    order of solve_gs).  Printed as its own JSON line."""
    from cfdl import serving
    reqs = serving.make_requests(args.serving_requests, args.serving_rate, 20261202)
    print("bench: serving %d requests" % len(reqs), file=sys.stderr, flush=True)
    out = serving.serve(reqs, slots=args.serving_slots, pc_order="mixed", gs_order="colored")
    cells = sum(n ** 3 for _, n, _ in reqs)
    print(json.dumps({"mode": "serving", "metric": "solves/s with p50/p99 latency",
                      "solves_per_s": round(out["solves_per_s"], 3),
                      "p50_ms": round(1e3 * out["p50_s"], 1), "p99_ms": round(1e3 * out["p99_s"], 1),
                      "wall_s": round(out["wall_s"], 2), "requests": len(reqs),
                      "arrival_rate_per_s": args.serving_rate, "slots": args.serving_slots,
                      "cell_iterations_per_s": round(30 * cells / out["wall_s"], 1),
                      "case": "lid-driven cavity N^3, N in {16,...,96}, Re log-uniform in "
                              "[100, 1000], 10 x 3 outer iterations from rest, 4 subdomains",
                      "pc_order": "mixed", "gs_order": "colored",
                      "dtype": "f32 pc sweeps, f64 otherwise", "data": "synthetic"}), flush=True)


def colored_rate(cfdl, mesh, mu, ne, args, order, fp64=None):
    """The same outer iterations with a colour-order pressure-correction
    smoother (extensions; the headline value keeps the reference order),
    from rest on a fresh context.  Printed as its own JSON line.  With fp64
    (the FP64 colour order's fields after the same iterations) the line also
    gives each field's relative L2 difference from them.  Returns the
    fields."""
    eq = cfdl.UVWP.cavity(mesh, rho=1.0, mu=mu, dt=0.01, nit=100)
    eq.set_relaxation(args.alpha_u, args.alpha_p)
    eq.set_pc_order(order)
    eq.set_gs_order(args.gs_order)
    n = 0
    for k in range(args.warmup + args.steps):
        if k == args.warmup:
            eq.synchronize()
            t = time.perf_counter()
        eq.update_boundaries()
        eq.solve_uvwp()
        n += 1
        if n % args.tstep_every == 0:
            eq.update_time()
    eq.synchronize()
    secs = time.perf_counter() - t
    fields = {k: eq.get(k)[:ne] for k in ("u", "v", "w", "p")}
    mode = "pc-" + order + ("/gs-colored" if args.gs_order == "colored" else "")
    line = {"mode": mode, "metric": METRIC, "value": round(ne * args.steps / secs, 1),
            "unit": "cell-iterations/s", "ms_per_step": round(1e3 * secs / args.steps, 3),
            "steps": args.steps, "outer_iterations_from_rest": args.warmup + args.steps,
            "dtype": "f32 pc sweeps, f64 otherwise" if order == "mixed" else "f64",
            "scaled_residual_reached": scaled_residual(eq.log())}
    if fp64 is not None:
        vel = np.linalg.norm(np.concatenate([fp64[k] for k in ("u", "v", "w")]))
        line["rel_l2_vs_fp64_colored"] = {
            k: float(np.linalg.norm(fields[k] - fp64[k]) /
                     (vel if k != "p" else np.linalg.norm(fp64[k]))) for k in fields}
    print(json.dumps(line), flush=True)
    return fields


if __name__ == "__main__":
    main()
