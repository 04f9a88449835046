"""Serving: continuous batching of independent cavity and channel solves on
one GPU (an extension; the reference solves one case per run,
src/main.f90:25-38).

A request is a steady solve from rest of a lid-driven cavity or of a channel
(parabolic inlet u_max = 1, zero-gradient outlet, no-slip walls) on a cube of
edge N (N^3 hex cells) at Reynolds number Re (rho = 1, mu = 1/Re), with
under-relaxation 0.7 / 0.3, run until its scaled residual (each equation's
initial residual over the largest of its first five, the worst equation)
is at most tol, or max_outer outer iterations.

The engine is one host loop around one cfdl.Batch:
  * admission: requests that have arrived are admitted in arrival order
    while the batch has a free slot and the case's device footprint fits the
    HBM budget (idle contexts of other shapes are closed to make room);
  * execution: every admitted case advances `chunk` outer iterations per
    batch run, all cases in the same grouped launches, so a request that
    arrives while others are on the GPU joins them at the next run;
  * retirement: after each run every case's new residuals are read; a case
    that has converged or reached max_outer leaves the batch (its latency is
    completion minus arrival), and so does a case whose residuals are not
    finite (evicted, reported as failed) -- the other cases are not touched;
  * contexts are kept per (kind, N) and reset between requests (every field
    cleared on the device, the viscosity set from Re), so a case gives the
    same bits alone and in any mix.
"""
import collections
import time

import numpy as np

from .mesh import Mesh
from .solver import UVWP, Batch

# Device footprint of one context: mesh, fields, the momentum system in
# colour numbering, the multigrid hierarchy and the FP32 copies; about
# 64 MiB plus 1.3 KB per cell.
BYTES_PER_CONTEXT = 64 << 20
BYTES_PER_CELL = 1344
KINDS = ("cavity", "channel")


def footprint(n):
    """Device bytes of one context for a cube of edge n."""
    return BYTES_PER_CONTEXT + BYTES_PER_CELL * n ** 3


def scaled_residual(hist):
    """hist: {equation: [initial residuals, one per outer iteration]}; the
    worst equation's last initial residual over the largest of its first
    five (inf if none is finite)."""
    worst = 0.0
    for h in hist.values():
        if not h:
            continue
        ref = max(h[:5])
        last = h[-1]
        if not (np.isfinite(ref) and np.isfinite(last)):
            return float("inf")
        if ref > 0:
            worst = max(worst, last / ref)
    return worst


# one request: its index in the stream, arrival (s from the start), "cavity"
# or "channel", the cube edge n and the Reynolds number
Request = collections.namedtuple("Request", ["idx", "arrival", "kind", "n", "re"])


def make_requests(count, rate, seed, sizes=tuple(range(16, 97, 8)), re_range=(100.0, 1000.0),
                  kinds=KINDS):
    """Poisson arrivals at rate per second; N uniform over sizes; Re
    log-uniform over re_range; the kind a fair draw over kinds."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    ns = rng.choice(np.asarray(sizes), count)
    res = np.exp(rng.uniform(np.log(re_range[0]), np.log(re_range[1]), count))
    ks = rng.integers(0, len(kinds), count)
    return [Request(i, float(a), kinds[int(k)], int(b), float(c))
            for i, (a, b, c, k) in enumerate(zip(t, ns, res, ks))]


class ContextPool:
    """Contexts kept per (kind, n); one host mesh per n, shared by the
    contexts built from it; bytes counts the device footprint of every
    context held, idle or in use."""

    def __init__(self, pc_order="mgcg-mixed", assembly_bits=32, jitter=0.0):
        self.pc_order, self.assembly_bits, self.jitter = pc_order, assembly_bits, jitter
        self.bytes = 0
        self.created = 0
        self._idle = {}
        self._all = []
        self._meshes = {}

    def mesh(self, n):
        if n not in self._meshes:
            self._meshes[n] = Mesh.box(n, jitter=self.jitter).partition(1)
        return self._meshes[n]

    def _close(self, key, item):
        item[1].close()
        self._all.remove(item)
        self.bytes -= footprint(key[1])

    def warm(self, shapes, count, budget):
        """Creates up to count idle contexts per (kind, n) in shapes, as far
        as the budget allows; returns how many it made."""
        made = []
        for kind, n in shapes:
            for _ in range(count):
                if self.bytes + footprint(n) > budget:
                    break
                made.append(((kind, n), self._create(kind, n)))
        for (kind, n), item in made:
            self.release(kind, n, item)
        return len(made)

    def acquire(self, kind, n, budget):
        """A context for (kind, n): an idle one, else a new one if it fits the
        budget after closing idle contexts of other shapes; None if not."""
        key = (kind, n)
        if self._idle.get(key):
            return self._idle[key].pop()
        need = footprint(n)
        while self.bytes + need > budget:
            victim = next((k for k, v in self._idle.items() if v), None)
            if victim is None:
                return None
            self._close(victim, self._idle[victim].pop())
        return self._create(kind, n)

    def _create(self, kind, n):
        mesh = self.mesh(n)
        mk = UVWP.channel if kind == "channel" else UVWP.cavity
        eq = mk(mesh, rho=1.0, mu=0.01, dt=float("inf"))
        try:
            eq.set_relaxation(0.7, 0.3)
            eq.set_gs_order("colored")
            eq.set_pc_order(self.pc_order)
            eq.set_assembly_precision(self.assembly_bits)
        except Exception:
            eq.close()
            raise
        item = (mesh, eq)
        self._all.append(item)
        self.bytes += footprint(n)
        self.created += 1
        return item

    def release(self, kind, n, item, ok=True):
        """Back to the idle set; a context whose case failed is closed."""
        if ok:
            self._idle.setdefault((kind, n), []).append(item)
        else:
            self._close((kind, n), item)

    def close(self):
        for _, eq in self._all:
            eq.close()
        for m in self._meshes.values():
            m.close()
        self._all, self._idle, self._meshes, self.bytes = [], {}, {}, 0


def prepare(eq, n, re):
    """A context to the request's initial state: rest, mu = 1/re."""
    eq.reset()
    ne = n ** 3
    eq.set_properties(np.ones(ne), np.full(ne, 1.0 / re))


def solve_alone(pool, req, tol=1e-6, max_outer=500, chunk=10):
    """One request on its own context, in the same chunks as serve() runs
    it; returns (fields, outer iterations, scaled residual)."""
    item = pool.acquire(req.kind, req.n, float("inf"))
    mesh, eq = item
    try:
        prepare(eq, req.n, req.re)
        hist = collections.defaultdict(list)
        outer, worst = 0, float("inf")
        while outer < max_outer:
            eq.run(1, chunk)
            outer += chunk
            for name, _, ri, _, _ in eq.log():
                hist[name].append(ri)
            eq.clear_log()
            worst = scaled_residual(hist)
            if worst <= tol or not np.isfinite(worst):
                break
        ne = req.n ** 3
        return {k: eq.get(k)[:ne] for k in ("u", "v", "w", "p")}, outer, worst
    finally:
        pool.release(req.kind, req.n, item)


def serve(requests, hbm_budget=200e9, max_cases=32, tol=1e-6, max_outer=500, chunk=10,
          pc_order="mgcg-mixed", assembly_bits=32, block=0, warm=0, keep_fields=False,
          on_run=None):
    """Serves the requests (arrival times in seconds from the start) and
    returns a dict: wall seconds, solves per second, p50 / p99 latency
    (completion minus arrival), per-request latency, outer iterations,
    final scaled residual, whether it converged, the failed (evicted)
    requests and, with keep_fields, each request's u, v, w, p.  warm > 0
    creates that many contexts per (kind, N) of the stream before the clock
    starts (a server's start-up), as far as the budget allows.  on_run, if
    given, is called after every batch run with the number of cases in it."""
    if chunk < 1 or max_outer < chunk:
        raise ValueError("chunk must be in 1 .. max_outer")
    for r in requests:
        if footprint(r.n) > hbm_budget:
            raise ValueError("request %d (N=%d) needs more than the HBM budget" % (r.idx, r.n))
    pending = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.idx)))
    pool = ContextPool(pc_order, assembly_bits)
    tw = time.perf_counter()
    shapes = sorted({(r.kind, r.n) for r in requests}, key=lambda s: (s[1], s[0]))
    warmed = pool.warm(shapes, warm, hbm_budget) if warm > 0 else 0
    warm_s = time.perf_counter() - tw
    batch = Batch(block)
    active = {}  # id(eq) -> [req, item, hist, outer]
    out = {"latency_s": {}, "outer": {}, "scaled_residual": {}, "converged": {}, "failed": [],
           "fields": {} if keep_fields else None, "runs": 0, "case_runs": 0,
           "admit_s": 0.0, "run_s": 0.0, "retire_s": 0.0}
    try:
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            ta = time.perf_counter()
            while pending and pending[0].arrival <= now and len(active) < max_cases:
                req = pending[0]
                item = pool.acquire(req.kind, req.n, hbm_budget)
                if item is None:
                    break  # waits for a running case to free its memory
                pending.popleft()
                prepare(item[1], req.n, req.re)
                batch.attach(item[1])
                active[id(item[1])] = [req, item, collections.defaultdict(list), 0]
            out["admit_s"] += time.perf_counter() - ta
            if not active:
                wait = pending[0].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
                continue
            tr = time.perf_counter()
            batch.run(1, chunk)
            batch.synchronize()
            out["run_s"] += time.perf_counter() - tr
            tr = time.perf_counter()
            out["runs"] += 1
            out["case_runs"] += len(active)
            if on_run:
                on_run(len(active))
            done = []
            for key, st in active.items():
                req, item, hist, _ = st
                eq = item[1]
                st[3] += chunk
                for name, _, ri, _, _ in eq.log():
                    hist[name].append(ri)
                eq.clear_log()
                worst = scaled_residual(hist)
                finite = np.isfinite(worst)
                if worst <= tol or st[3] >= max_outer or not finite:
                    done.append((key, finite, worst))
            for key, finite, worst in done:
                req, item, hist, outer = active.pop(key)
                eq = item[1]
                if keep_fields and finite:
                    ne = req.n ** 3
                    out["fields"][req.idx] = {k: eq.get(k)[:ne] for k in ("u", "v", "w", "p")}
                batch.detach(eq)
                out["latency_s"][req.idx] = time.perf_counter() - t0 - req.arrival
                out["outer"][req.idx] = outer
                out["scaled_residual"][req.idx] = worst
                out["converged"][req.idx] = bool(worst <= tol)
                if not finite:
                    out["failed"].append(req.idx)
                pool.release(req.kind, req.n, item, ok=finite)
            out["retire_s"] += time.perf_counter() - tr
        wall = time.perf_counter() - t0
    finally:
        batch.close()
        pool.close()
    ok = [i for i in out["latency_s"] if i not in out["failed"]]
    lat = np.asarray([out["latency_s"][i] for i in ok]) if ok else np.zeros(1)
    out.update({"wall_s": wall, "solves_per_s": len(ok) / wall if wall > 0 else 0.0,
                "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "contexts_created": pool.created, "contexts_warmed": warmed,
                "warm_s": warm_s})
    return out
