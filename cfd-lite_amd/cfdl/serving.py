"""Serving: a stream of independent lid-driven cavity solves on one GPU.

Each request is a cavity of edge N (N^3 hex cells, 4 subdomains) at Reynolds
number Re, solved by the reference's loop (src/main.f90:50-63: ntstep pseudo
time steps of ncoef outer iterations, from rest).  Requests are admitted in
arrival order when a slot is free and their device footprint fits the HBM
budget; each admitted request runs on its own context, whose HIP stream runs
concurrently with the other slots'.  Contexts are kept per N and reset
between requests (every field cleared, the viscosity set from Re), so a case
gives the same bits whether it runs alone or in any mix.

There is no counterpart in the reference, which solves one case per run.
"""
import concurrent.futures as cf
import threading
import time

import numpy as np

from .mesh import Mesh
from .solver import FIELDS, UVWP

# the fields of solve_uvwp's state and work arrays (the energy fields are not
# used by these solves)
STATE_FIELDS = [f for f in FIELDS if f not in ("t", "gt", "h", "h0", "gh")]
# device footprint of one context (fields, mesh, the pc system in two
# numberings, sweep schedules, the FP32 copies of the mixed mode): measured
# on MI355X from the free device memory around context creation, 375 MB at
# 64^3 and 1.11 GB at 96^3, i.e. about 64 MiB plus 1.2 KB per cell
BYTES_PER_CONTEXT = 64 << 20
BYTES_PER_CELL = 1280


def footprint(n):
    """Device bytes of one context for edge n."""
    return BYTES_PER_CONTEXT + BYTES_PER_CELL * n ** 3


class ContextPool:
    """Contexts cached per box edge, each with its mesh; bytes counts the
    footprint of every context held, idle or in use."""

    def __init__(self, pc_order="reference", gs_order="reference", nsubd=4):
        self.pc_order, self.gs_order, self.nsubd = pc_order, gs_order, nsubd
        self.bytes = 0
        self._free = {}
        self._all = []
        self._lock = threading.Lock()

    def take(self, n):
        """An idle context for edge n, or None."""
        with self._lock:
            return self._free[n].pop() if self._free.get(n) else None

    def reserve(self, n, budget):
        """Counts a new context for edge n if it fits the budget, closing idle
        contexts of other edges to make room; False (nothing counted) if it
        does not fit even then."""
        need = footprint(n)
        with self._lock:
            while self.bytes + need > budget:
                victim = next(((k, v) for k, v in self._free.items() if v), None)
                if victim is None:
                    return False
                k, items = victim
                mesh, eq = items.pop()
                eq.close()
                mesh.close()
                self._all.remove((mesh, eq))
                self.bytes -= footprint(k)
            self.bytes += need
            return True

    def create(self, n):
        """A new context for edge n (its footprint already reserved)."""
        mesh = Mesh.box(n).partition(self.nsubd)
        eq = UVWP.cavity(mesh, rho=1.0, mu=0.01, dt=0.01, nit=100)
        eq.set_pc_order(self.pc_order)
        eq.set_gs_order(self.gs_order)
        with self._lock:
            self._all.append((mesh, eq))
        return mesh, eq

    def acquire(self, n, budget=float("inf")):
        item = self.take(n)
        if item is None:
            if not self.reserve(n, budget):
                raise ValueError("a context for N=%d does not fit the HBM budget" % n)
            item = self.create(n)
        return item

    def release(self, n, item):
        with self._lock:
            self._free.setdefault(n, []).append(item)

    def close(self):
        with self._lock:
            for mesh, eq in self._all:
                eq.close()
                mesh.close()
            self._all, self._free, self.bytes = [], {}, 0


def solve_case(pool, n, re, ntstep=10, ncoef=3, item=None):
    """One request: reset a context for edge n (item, or one from the pool)
    to rest with mu = 1/re (rho 1, lid speed 1), run ntstep x ncoef outer
    iterations, return u, v, w, p."""
    mesh, eq = item if item is not None else pool.acquire(n)
    try:
        ne = n ** 3
        for f in STATE_FIELDS:
            eq.set(f, np.zeros(eq.sizes[f]))
        eq.set_properties(np.ones(ne), np.full(ne, 1.0 / re))
        eq.clear_log()
        eq.run(ntstep, ncoef)
        return {k: eq.get(k)[:ne] for k in ("u", "v", "w", "p")}
    finally:
        pool.release(n, (mesh, eq))


def make_requests(count, rate, seed, sizes=tuple(range(16, 97, 8)), re_range=(100.0, 1000.0)):
    """Poisson arrivals at rate per second; N uniform over sizes; Re
    log-uniform over re_range.  Returns [(arrival_s, n, re)]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    ns = rng.choice(np.asarray(sizes), count)
    res = np.exp(rng.uniform(np.log(re_range[0]), np.log(re_range[1]), count))
    return [(float(a), int(b), float(c)) for a, b, c in zip(t, ns, res)]


def serve(requests, slots=4, hbm_budget=200e9, pc_order="reference", gs_order="reference",
          ntstep=10, ncoef=3, keep_fields=False):
    """Runs the requests ([(arrival_s, n, re)], arrival times from the start)
    and returns a dict: wall seconds, solves per second, the p50 and p99
    latencies (completion minus arrival, seconds), per-request latencies and,
    with keep_fields, each request's fields.  A request is admitted, in
    arrival order, when a slot is free and either an idle context of its edge
    exists or a new one fits hbm_budget with every context held."""
    pool = ContextPool(pc_order, gs_order)
    cond = threading.Condition()
    busy = [0]
    lat = [0.0] * len(requests)
    fields = [None] * len(requests)
    errors = []

    def run(i, arrival, n, re, item):
        try:
            if item is None:
                item = pool.create(n)
            out = solve_case(pool, n, re, ntstep, ncoef, item)
            if keep_fields:
                fields[i] = out
        except Exception as exc:  # noqa: BLE001 -- reported after the stream
            errors.append((i, exc))
        finally:
            lat[i] = time.perf_counter() - t0 - arrival
            with cond:
                busy[0] -= 1
                cond.notify_all()

    try:
        with cf.ThreadPoolExecutor(max_workers=slots) as ex:
            t0 = time.perf_counter()
            for i, (arrival, n, re) in enumerate(requests):
                if footprint(n) > hbm_budget:
                    raise ValueError("request %d (N=%d) needs more than the HBM budget" % (i, n))
                wait = t0 + arrival - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
                got = []

                def admit():
                    if busy[0] >= slots:
                        return False
                    item = pool.take(n)
                    if item is None and not pool.reserve(n, hbm_budget):
                        return False
                    got.append(item)  # None: a new context, its bytes reserved
                    return True

                with cond:
                    cond.wait_for(admit)
                    busy[0] += 1
                ex.submit(run, i, arrival, n, re, got[0])
        wall = time.perf_counter() - t0
    finally:
        pool.close()
    if errors:
        raise RuntimeError("request %d failed: %s" % errors[0])
    ms = np.asarray(lat)
    return {"wall_s": wall, "solves_per_s": len(requests) / wall,
            "p50_s": float(np.percentile(ms, 50)), "p99_s": float(np.percentile(ms, 99)),
            "latency_s": lat, "fields": fields if keep_fields else None}
