"""Serving loop: continuous batching of airfoil steady solves on one GPU.

Requests (an airfoil polygon and a Case) arrive over time.  Each is meshed on a host thread
pool (the C++ mesher releases the GIL), queued, and admitted first-come-first-served into a
free slot of one device arena whose slot count is set by an HBM budget
(batch.slots_fitting).  Every admitted slot runs its own steady check loop on the device
(batch.Driver); a slot whose run ends is reported and refilled at once while the others keep
running.  Latency of a request = completion time - arrival time.

A request's result does not depend on what shares the arena: slots never read each other's
data and each slot's check schedule is its own, so a request served in any mix gives bitwise
the result of the same request solved alone (same precision and launch tile).

No counterpart in the reference, which solves one airfoil per process
(src/quick_airfoil.cpp:46-95).
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    polygon: object          # (points, patches) as geometry.reorder_points returns
    case: object             # config.Case
    arrival: float = 0.0     # seconds after the start of serving


@dataclass
class Outcome:
    rid: int
    arrival: float
    meshed: float = float("nan")
    admitted: float = float("nan")
    completed: float = float("nan")
    slot: int = -1
    n_cells: int = 0
    result: dict = field(default_factory=dict)
    error: str = ""

    @property
    def latency(self) -> float:
        return self.completed - self.arrival


def poisson_arrivals(rate_per_second: float, n: int, seed: int = 0) -> np.ndarray:
    """Arrival times of a Poisson process (exponential gaps), starting with the first gap."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_second, size=n))


def percentile(values, q: float) -> float:
    """q-th percentile (nearest rank) of the values; nan when empty."""
    v = sorted(values)
    if not v:
        return float("nan")
    rank = max(1, int(np.ceil(q / 100.0 * len(v))))
    return float(v[rank - 1])


class MeshCache:
    """Least-recently-used cache of built meshes keyed by (polygon bytes, patches, mesh options,
    physics and freestream: the side boundaries carry the freestream state).  A repeated request
    reuses its mesh and viscous tables; only the initial state is rebuilt.  Thread-safe; a cached
    mesh is only read (admission copies it to the device)."""

    def __init__(self, capacity: int = 64):
        import threading
        from collections import OrderedDict
        self.capacity = capacity
        self._items = OrderedDict()
        self._lock = threading.Lock()
        self.hits = self.misses = 0

    @staticmethod
    def key(request: Request) -> str:
        import hashlib
        import json
        points, patches = request.polygon
        c = request.case
        physics = dict(model=c.model, sa=c.sa_model, law=c.law, mu=c.mu_star, s=c.sutherland_star,
                       nu=c.sa_nu_inf, vcfl=c.viscous_cfl, sacfl=c.sa_cfl, probe=c.wall_probe_min_h, cfl=c.cfl,
                       flux=c.flux, gamma=c.gamma, bc=list(c.bc_kinds), outlet=c.outlet_p_star, mach=c.mach,
                       aoa=c.aoa)
        h = hashlib.sha256(np.ascontiguousarray(points, dtype=np.float64).tobytes())
        h.update(np.ascontiguousarray(patches, dtype=np.int64).tobytes())
        h.update(json.dumps([c.mesh_options, physics], sort_keys=True, default=str).encode())
        return h.hexdigest()

    def get(self, key):
        with self._lock:
            if key in self._items:
                self._items.move_to_end(key)
                self.hits += 1
                return self._items[key]
            self.misses += 1
            return None

    def put(self, key, value):
        with self._lock:
            self._items[key] = value
            self._items.move_to_end(key)
            while len(self._items) > self.capacity:
                self._items.popitem(last=False)


_CACHE = MeshCache()


def build_request_mesh(request: Request, cache: MeshCache | None = _CACHE):
    """(mesh, params, u0) of a request: the case's mesh around its airfoil (from the cache when
    the same airfoil, mesh and physics came before), freestream start at the request's Mach
    number and angle of attack."""
    from . import mesh as M
    points, patches = request.polygon
    key = MeshCache.key(request) if cache is not None else None
    built = cache.get(key) if cache is not None else None
    if built is None:
        built = M.build_case_mesh(request.case, points, patches)
        if cache is not None:
            cache.put(key, built)
    mesh = built
    params = M.params_of(request.case)
    return mesh, params, mesh.initialize(params, request.case.freestream)


CAPACITY_FIELDS = ("cells", "faces", "walls", "csr", "terms", "groups", "members")


class DeviceBackend:
    """The slot arena and its steady driver (batch.Batch + batch.Driver).

    `caps` is one capacity per slot (slots may differ: e.g. classes sized for LF1, LF2 and
    level-7 meshes); or one capacity with `n_slots` copies."""

    def __init__(self, caps, features, n_slots: int | None = None, *, precision: int = 32, tile: int = 64,
                 batch_trials: int = 32):
        from . import batch as B
        caps = list(caps) if isinstance(caps, (list, tuple)) else [caps] * int(n_slots)
        self.batch = B.Batch.with_slots(caps, features, precision=precision, tile=tile, batch_trials=batch_trials)
        self.driver = B.Driver(self.batch)
        self.caps = caps
        self.n_slots = len(caps)

    def need(self, mesh):
        from . import batch as B
        return B.capacity(mesh, self.batch.tile)[0]

    def fits_slot(self, mesh, slot: int, need=None) -> bool:
        need = need if need is not None else self.need(mesh)
        return all(getattr(need, k) <= getattr(self.caps[slot], k) for k in CAPACITY_FIELDS)

    def slot_size(self, slot: int) -> int:
        return int(self.caps[slot].cells)

    def fits(self, mesh) -> bool:
        need = self.need(mesh)
        return any(self.fits_slot(mesh, s, need) for s in range(self.n_slots))

    def start(self, slot: int, mesh, params, u0, case, mode: str, stop_time, budget: float):
        self.batch.admit(slot, mesh, params, u0)
        self.driver.begin(slot, case, mode, stop_time, budget)

    def running(self) -> int:
        return self.driver.running()

    def poll(self, seconds: float):
        return self.driver.poll(seconds)

    def finish(self, slot: int):
        self.batch.release(slot)

    def close(self):
        self.driver.close()
        self.batch.close()


class Server:
    """Continuous batching over a backend with `n_slots` slots (see DeviceBackend): the backend
    provides fits(mesh), fits_slot(mesh, slot), slot_size(slot), start, running, poll, finish."""

    def __init__(self, backend, *, mode: str = "steady", stop_time: float | None = None, budget: float = 300.0,
                 workers: int = 8, poll_seconds: float = 0.05, mesher=build_request_mesh,
                 clock=time.perf_counter, log=None):
        self.backend = backend
        self.mode, self.stop_time, self.budget = mode, stop_time, budget
        self.workers = workers
        self.poll_seconds = poll_seconds
        self.mesher = mesher
        self.clock = clock
        self.log = log

    def serve(self, requests) -> dict:
        """Serve every request (arrival offsets in seconds); returns the outcomes and a summary."""
        order = sorted(requests, key=lambda r: (r.arrival, r.rid))
        outcomes = {r.rid: Outcome(r.rid, r.arrival) for r in order}
        free = list(range(self.backend.n_slots))
        busy: dict[int, Request] = {}
        ready: list[tuple[Request, tuple]] = []
        meshing = []
        nxt = 0
        busy_slot_seconds = 0.0
        t0 = self.clock()
        last = t0
        with ThreadPoolExecutor(max_workers=self.workers) as pool:
            while nxt < len(order) or meshing or ready or busy:
                now = self.clock() - t0
                busy_slot_seconds += len(busy) * (self.clock() - last)
                last = self.clock()
                while nxt < len(order) and order[nxt].arrival <= now:
                    meshing.append((order[nxt], pool.submit(self.mesher, order[nxt])))
                    nxt += 1
                still = []
                for req, fut in meshing:
                    if not fut.done():
                        still.append((req, fut))
                        continue
                    out = outcomes[req.rid]
                    out.meshed = self.clock() - t0
                    try:
                        built = fut.result()
                    except Exception as exc:  # a mesher error ends this request only
                        out.error = f"mesher: {exc}"
                        out.completed = out.meshed
                        continue
                    if not self.backend.fits(built[0]):
                        out.error = "mesh exceeds the slot capacity"
                        out.completed = out.meshed
                        continue
                    ready.append((req, built))
                meshing = still
                ready.sort(key=lambda item: (item[0].arrival, item[0].rid))  # first come, first served
                while free and ready:
                    # First come, first served; the head request takes the smallest free slot
                    # that holds it, or waits for one.
                    req, (mesh, params, u0) = ready[0]
                    fitting = [s for s in free if self.backend.fits_slot(mesh, s)]
                    if not fitting:
                        break
                    slot = min(fitting, key=lambda s: (self.backend.slot_size(s), s))
                    ready.pop(0)
                    free.remove(slot)
                    out = outcomes[req.rid]
                    out.slot, out.n_cells = slot, mesh.n_cells
                    out.admitted = self.clock() - t0
                    self.backend.start(slot, mesh, params, u0, req.case, self.mode, self.stop_time, self.budget)
                    busy[slot] = req
                if busy:
                    for slot, res in self.backend.poll(self.poll_seconds):
                        req = busy.pop(slot)
                        out = outcomes[req.rid]
                        out.completed = self.clock() - t0
                        out.result = res
                        self.backend.finish(slot)
                        free.append(slot)
                        if self.log:
                            self.log(f"request {req.rid} done in slot {slot}: latency {out.latency:.2f} s, "
                                     f"{res.get('steps')} steps, {res.get('stop_reason')}")
                else:
                    time.sleep(0.001)
        makespan = self.clock() - t0
        done = [o for o in outcomes.values() if not o.error]
        lat = [o.latency for o in done]
        return {
            "outcomes": [outcomes[r.rid] for r in requests],
            "completed": len(done),
            "failed": len(outcomes) - len(done),
            "converged": sum(1 for o in done if o.result.get("converged")),
            "makespan_seconds": makespan,
            "solves_per_hour": 3600.0 * len(done) / makespan if makespan > 0 else float("nan"),
            "p50_latency_seconds": percentile(lat, 50),
            "p99_latency_seconds": percentile(lat, 99),
            "mean_queue_seconds": float(np.mean([o.admitted - o.meshed for o in done])) if done else float("nan"),
            "slot_occupancy": busy_slot_seconds / (makespan * self.backend.n_slots) if makespan > 0 else 0.0,
            "slots": self.backend.n_slots,
        }
