"""Many meshes in one device arena (amr_batch_* / amr_driver_* in include/amr_hip.h).

A Batch is a fixed set of slots on one GPU.  Each slot holds one mesh with its own time,
target, step size, counters and failure code; one captured trial graph advances every slot,
and finished, failed or empty slots are no-ops inside it.  A Driver runs the steady check
loop of the reference solver (src/quick_airfoil.cpp:66-95) on every slot independently, so
slots end at different times and can be refilled (the serving loop in serve.py).

No counterpart in the reference, which advances one mesh per process.  Errors raise
BackendError; nothing falls back to the host.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .session import BackendError, result_dict, steady_options


def capacity(mesh, tile: int = 64):
    """(amr_slot_capacity, amr_batch_features) of the smallest slot that holds `mesh`."""
    cap, feat = N.amr_slot_capacity(), N.amr_batch_features()
    if N.load_library().amr_batch_capacity(C.byref(mesh.view), tile, C.byref(cap), C.byref(feat)):
        raise BackendError(1, "invalid mesh or tile for a batch slot")
    return cap, feat


def envelope(caps, margin: float = 0.0, tile: int = 64) -> N.amr_slot_capacity:
    """Capacity covering every capacity in `caps`, grown by `margin` (fraction) and re-aligned."""
    out = N.amr_slot_capacity()
    for name in ("cells", "faces", "walls", "csr", "terms", "groups", "members"):
        v = max(getattr(c, name) for c in caps)
        v = int(np.ceil(v * (1.0 + margin)))
        if name in ("cells", "faces", "walls"):
            v = -(-v // tile) * tile
        setattr(out, name, v)
    return out


def union(features) -> N.amr_batch_features:
    out = N.amr_batch_features()
    for name in ("laminar", "sa", "small", "visc_group"):
        setattr(out, name, int(any(getattr(f, name) for f in features)))
    return out


def arena_bytes(caps, features, precision: int = 64, tile: int = 64) -> int:
    """Device bytes of an arena with these slot capacities (its single allocation)."""
    arr = (N.amr_slot_capacity * len(caps))(*caps)
    return int(N.load_library().amr_batch_bytes(arr, len(caps), C.byref(features), precision, tile))


def slots_fitting(cap, features, budget_bytes: int, precision: int = 64, tile: int = 64, limit: int = 4096) -> int:
    """Largest number of identical slots whose arena fits `budget_bytes` (0 if none)."""
    lo, hi = 0, limit
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if arena_bytes([cap] * mid, features, precision, tile) <= budget_bytes:
            lo = mid
        else:
            hi = mid - 1
    return lo


def _states(arr) -> list[dict]:
    return [dict(time=s.time, target=s.target, steps=s.steps, rejected=s.rejected, n_cells=s.n_cells,
                 finished=bool(s.finished), fatal=s.fatal) for s in arr]


class Batch:
    """Slots of one device arena.  Build with Batch.of_meshes(...) or Batch.with_slots(...)."""

    def __init__(self, handle, n: int, nv: int, precision: int, tile: int):
        self._lib = N.load_library()
        self.handle = handle
        self.n = n
        self.nv = nv
        self.precision = precision
        self.tile = tile
        self.meshes = [None] * n  # keep the host arrays behind each admitted view alive
        self._state_buf = (N.amr_slot_state * n)()

    @staticmethod
    def _options(precision, tile, graph, batch_trials):
        return N.amr_batch_options(precision, tile, 1 if graph else 0, batch_trials)

    @classmethod
    def of_meshes(cls, meshes, params, states, *, metas=None, precision: int = 64, tile: int = 64,
                  graph: bool = True, batch_trials: int = 32) -> "Batch":
        """One slot per mesh, each sized to its mesh and admitted with its initial state."""
        lib = N.load_library()
        n = len(meshes)
        views = (N.amr_mesh_view * n)(*[m.view for m in meshes])
        prm = (N.amr_params * n)(*params)
        arrays = [np.ascontiguousarray(u, dtype=np.float64) for u in states]
        ptrs = (N.c_double_p * n)(*[N.dptr(a) for a in arrays])
        meta = None
        if metas is not None:
            meta = (N.amr_state_meta * n)(*[N.amr_state_meta(*m) for m in metas])
        handle = C.c_void_p()
        err = N.err_buffer()
        opts = cls._options(precision, tile, graph, batch_trials)
        if lib.amr_batch_create(views, prm, ptrs, meta, n, C.byref(opts), C.byref(handle), err, N.ERR_LEN):
            raise BackendError(1, err.value.decode())
        nv = 5 if meshes[0].view.transport.sa_model else 4
        b = cls(handle, n, nv, precision, tile)
        b.meshes = list(meshes)
        return b

    @classmethod
    def with_slots(cls, caps, features, *, precision: int = 64, tile: int = 64, graph: bool = True,
                   batch_trials: int = 32) -> "Batch":
        """Empty slots of the given capacities; the kernel set is fixed by `features`."""
        lib = N.load_library()
        n = len(caps)
        arr = (N.amr_slot_capacity * n)(*caps)
        handle = C.c_void_p()
        err = N.err_buffer()
        opts = cls._options(precision, tile, graph, batch_trials)
        if lib.amr_batch_create_slots(arr, n, C.byref(features), C.byref(opts), C.byref(handle), err, N.ERR_LEN):
            raise BackendError(1, err.value.decode())
        return cls(handle, n, 5 if features.sa else 4, precision, tile)

    def _check(self, status, err):
        if status == 1:
            raise BackendError(1, err.value.decode())
        return status

    def admit(self, slot: int, mesh, params, u, *, time: float = 0.0, steps: int = 0, rejected: int = 0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        meta = N.amr_state_meta(time, steps, rejected)
        err = N.err_buffer()
        self._check(self._lib.amr_batch_admit(self.handle, slot, C.byref(mesh.view), C.byref(params), N.dptr(u),
                                              u.size, C.byref(meta), err, N.ERR_LEN), err)
        self.meshes[slot] = mesh

    def release(self, slot: int):
        err = N.err_buffer()
        self._check(self._lib.amr_batch_release(self.handle, slot, err, N.ERR_LEN), err)
        self.meshes[slot] = None

    def set_targets(self, targets, maxsteps=None):
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(targets, dtype=np.float64), (self.n,)))
        ms = None
        if maxsteps is not None:
            ms_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(maxsteps, dtype=np.int64), (self.n,)))
            ms = ms_arr.ctypes.data_as(C.POINTER(C.c_int64))
        err = N.err_buffer()
        self._check(self._lib.amr_batch_set_targets(self.handle, N.dptr(t), ms, err, N.ERR_LEN), err)

    def advance(self, until: str = "all", budget: float = 0.0) -> int:
        """Replay trials until all slots finish ("all"), any running slot ends ("any"), or an eighth
        of the running slots have ended / min(8, running // 32) polls have passed since the first
        ended ("group"); 0 done, 2 budget expired.  Polls are pipelined: slots still running when
        the condition is seen take up to batch_trials more trials (never past their targets)."""
        err = N.err_buffer()
        mode = {"all": 0, "any": 1, "group": 2}[until]
        return self._check(self._lib.amr_batch_advance(self.handle, mode, budget, self._state_buf, err, N.ERR_LEN),
                           err)

    def run(self, trials: int) -> list[dict]:
        """Exactly `trials` trial replays, then one poll."""
        err = N.err_buffer()
        self._check(self._lib.amr_batch_run(self.handle, trials, self._state_buf, err, N.ERR_LEN), err)
        return _states(self._state_buf)

    def set_launch_shape(self, flux_spread: int = 1, few_slots: int = 0, few_flux_spread: int = 1) -> None:
        """Threads per cell of the flux launches, and the spread used while at most `few_slots`
        slots run (latency-bound trials); results are bitwise independent of the shape."""
        err = N.err_buffer()
        self._check(self._lib.amr_batch_set_launch_shape(self.handle, flux_spread, few_slots, few_flux_spread, err,
                                                          N.ERR_LEN), err)

    def states(self) -> list[dict]:
        self._lib.amr_batch_slots(self.handle, self._state_buf)
        return _states(self._state_buf)

    def download(self, slot: int) -> tuple[np.ndarray, dict]:
        n_cells = self.states()[slot]["n_cells"]
        u = np.zeros(self.nv * n_cells)
        meta = N.amr_state_meta()
        err = N.err_buffer()
        self._check(self._lib.amr_batch_download(self.handle, slot, N.dptr(u), u.size, C.byref(meta), err,
                                                 N.ERR_LEN), err)
        return u, dict(time=meta.time, steps=meta.steps, rejected=meta.rejected)

    def monitor(self, mask, frames, mach, aoa):
        """Masked steady monitor: (out[n, 11], status[n]) as amr_batch_monitor."""
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(mask, dtype=np.int32), (self.n,)))
        f = np.ascontiguousarray(np.broadcast_to(np.asarray(frames, dtype=np.float64), (self.n, 6)))
        ma = np.ascontiguousarray(np.broadcast_to(np.asarray(mach, dtype=np.float64), (self.n,)))
        ao = np.ascontiguousarray(np.broadcast_to(np.asarray(aoa, dtype=np.float64), (self.n,)))
        out = np.zeros((self.n, 11))
        status = np.zeros(self.n, dtype=np.int32)
        err = N.err_buffer()
        self._check(self._lib.amr_batch_monitor(self.handle, m.ctypes.data_as(N.c_i32_p), N.dptr(f), N.dptr(ma),
                                                N.dptr(ao), N.dptr(out), status.ctypes.data_as(N.c_i32_p), err,
                                                N.ERR_LEN), err)
        return out, status

    def solve(self, cases, mode: str = "steady", budget: float = 300.0, stop_time: float | None = None) -> list[dict]:
        """Steady run on every occupied slot (cases[m] per slot) to the end of the last one."""
        opts = (N.amr_steady_options * self.n)(*[steady_options(c, mode, stop_time, budget) if c is not None
                                                 else N.amr_steady_options() for c in cases])
        res = (N.amr_steady_result * self.n)()
        err = N.err_buffer()
        self._check(self._lib.amr_batch_solve(self.handle, opts, res, err, N.ERR_LEN), err)
        return [result_dict(r) for r in res]

    def profile(self, trials: int) -> dict:
        out = (N.amr_kernel_time * 64)()
        count = C.c_int32()
        err = N.err_buffer()
        self._check(self._lib.amr_batch_profile(self.handle, trials, out, 64, C.byref(count), err, N.ERR_LEN), err)
        return {out[i].name.decode(): (out[i].total_ms, out[i].launches) for i in range(count.value)}

    def stats(self) -> dict:
        s = N.amr_gpu_stats()
        self._lib.amr_batch_stats(self.handle, C.byref(s))
        return {name: getattr(s, name) for name, _ in N.amr_gpu_stats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            self._lib.amr_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Driver:
    """Per-slot steady runs over a Batch (amr_driver_*)."""

    def __init__(self, batch: Batch):
        self._lib = N.load_library()
        self.batch = batch
        handle = C.c_void_p()
        err = N.err_buffer()
        if self._lib.amr_driver_create(batch.handle, C.byref(handle), err, N.ERR_LEN):
            raise BackendError(1, err.value.decode())
        self.handle = handle
        self._results = (N.amr_steady_result * batch.n)()
        self._done = np.zeros(batch.n, dtype=np.int32)

    def begin(self, slot: int, case, mode: str = "steady", stop_time: float | None = None, budget: float = 300.0):
        o = steady_options(case, mode, stop_time, budget)
        err = N.err_buffer()
        if self._lib.amr_driver_begin(self.handle, slot, C.byref(o), err, N.ERR_LEN):
            raise BackendError(1, err.value.decode())

    def running(self) -> int:
        return int(self._lib.amr_driver_running(self.handle))

    def poll(self, budget: float = 0.0) -> list[tuple[int, dict]]:
        """Advance until at least one run ends (or budget seconds pass): [(slot, result), ...]."""
        n_done = C.c_int32()
        err = N.err_buffer()
        status = self._lib.amr_driver_poll(self.handle, budget, self._results, self._done.ctypes.data_as(N.c_i32_p),
                                           C.byref(n_done), err, N.ERR_LEN)
        if status == 1:
            raise BackendError(1, err.value.decode())
        return [(int(m), result_dict(self._results[m])) for m in self._done[:n_done.value]]

    def close(self):
        if getattr(self, "handle", None):
            self._lib.amr_driver_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
