"""Python handle of one device-resident solve (amr_hip_* in include/amr_hip.h).

Mirrors the reference backend API (src/static_amr2d_gpu.h:17-24): create, advance to a target
time, download the accepted state, stats, destroy; plus one residual evaluation at a given
state (the reference's static_amr2d_rhs, src/static_amr2d.h:80-82).  Errors raise
`BackendError` with the library's message; a missing library or GPU raises, never falls back.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class BackendError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def device_count() -> int:
    return int(N.load_library().amr_hip_device_count())


def steady_options(case, mode: str = "steady", stop_time: float | None = None,
                   budget: float = 118.0) -> N.amr_steady_options:
    """amr_steady_options of a Case: monitor frame, steady criteria, mode, stop time, wall budget."""
    conv = case.convergence
    o = N.amr_steady_options()
    o.mach, o.aoa_degrees = case.mach, case.aoa
    o.moment[0], o.moment[1] = case.moment_reference
    for i, v in enumerate(conv["residual_region_xy_c"]):
        o.region[i] = v
    o.raw_tol, o.load_tol = conv["raw_tolerance"], conv["coefficient_change"]
    o.min_time, o.period, o.hold = conv["min_time_star"], conv["check_period_star"], conv["hold"]
    o.mode = {"steady": 0, "time": 1, "smoke": 2}[mode]
    o.stop_time = case.stop_time if stop_time is None else stop_time
    o.budget_seconds = budget
    return o


def result_dict(res: N.amr_steady_result) -> dict:
    out = {name: getattr(res, name) for name, _ in N.amr_steady_result._fields_ if name != "pad"}
    out["residual"] = list(res.residual)
    out["stop_reason"] = N.STOP_REASONS.get(res.reason, str(res.reason))
    return out


class Session:
    """One mesh advanced on the GPU (FP64 by default, FP32 with precision=32)."""

    def __init__(self, mesh, params: N.amr_params, u: np.ndarray, *, time: float = 0.0, steps: int = 0,
                 rejected: int = 0, precision: int = 64, batch_trials: int = 32, graph: bool = True,
                 event_timing: bool = False):
        self._lib = N.load_library()
        self.mesh = mesh  # keeps the host arrays behind the view alive
        self.view = mesh.view
        self.params = params
        self.nv = 5 if self.view.transport.sa_model else 4
        self.n_cells = int(self.view.n_cells)
        u = np.ascontiguousarray(u, dtype=np.float64)
        opts = N.amr_session_options(batch_trials, 1 if graph else 0, 1 if event_timing else 0, precision)
        meta = N.amr_state_meta(time, steps, rejected)
        handle = C.c_void_p()
        err = N.err_buffer()
        status = self._lib.amr_hip_create(C.byref(self.view), C.byref(params), N.dptr(u), u.size, C.byref(meta),
                                          C.byref(opts), C.byref(handle), err, N.ERR_LEN)
        if status:
            raise BackendError(status, err.value.decode())
        self.handle = handle
        self.meta = dict(time=time, steps=steps, rejected=rejected)

    def advance(self, tend: float, maxsteps: int = 50_000_000, wall_budget: float = 0.0) -> int:
        """Advance to tend.  Returns 0 (reached) or 2 (budget expired); raises on status 1."""
        meta = N.amr_state_meta(self.meta["time"], self.meta["steps"], self.meta["rejected"])
        err = N.err_buffer()
        status = self._lib.amr_hip_advance(self.handle, tend, maxsteps, wall_budget, C.byref(meta), err, N.ERR_LEN)
        self.meta = dict(time=meta.time, steps=meta.steps, rejected=meta.rejected)
        if status == 1:
            raise BackendError(1, err.value.decode())
        return status

    def download(self) -> np.ndarray:
        u = np.zeros(self.nv * self.n_cells)
        meta = N.amr_state_meta()
        err = N.err_buffer()
        if self._lib.amr_hip_download(self.handle, N.dptr(u), u.size, C.byref(meta), err, N.ERR_LEN):
            raise BackendError(1, err.value.decode())
        self.meta = dict(time=meta.time, steps=meta.steps, rejected=meta.rejected)
        return u

    def rhs(self, u: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros_like(u)
        err = N.err_buffer()
        if self._lib.amr_hip_rhs(self.handle, N.dptr(u), u.size, N.dptr(r), err, N.ERR_LEN):
            raise BackendError(1, err.value.decode())
        return r

    def monitor(self, region=(-1.0, 3.0, -1.0, 1.0), moment=(0.25, 0.0), mach=1.65, aoa=1.5) -> dict:
        """Raw residual norms and CL/CD/CM_z of the accepted state (FP64 device reductions)."""
        reg = np.ascontiguousarray(region, dtype=np.float64)
        mom = np.ascontiguousarray(moment, dtype=np.float64)
        out = np.zeros(11)
        err = N.err_buffer()
        if self._lib.amr_hip_monitor(self.handle, N.dptr(reg), N.dptr(mom), mach, aoa, N.dptr(out), err, N.ERR_LEN):
            raise BackendError(1, err.value.decode())
        return dict(residual=out[:5].copy(), cl=out[5], cd=out[6], cm=out[7], fx=out[8], fy=out[9], mz=out[10])

    def solve(self, case, mode: str = "steady", stop_time: float | None = None, budget: float = 118.0,
              history: int = 512) -> dict:
        """Steady (or time / smoke) driver on the device: the check loop of the reference solver."""
        o = steady_options(case, mode, stop_time, budget)
        res = N.amr_steady_result()
        hist = (N.amr_steady_check * max(1, history))()
        err = N.err_buffer()
        status = self._lib.amr_hip_solve(self.handle, C.byref(o), C.byref(res), hist, history, err, N.ERR_LEN)
        out = result_dict(res)
        out["history"] = [dict(time=h.time, steps=h.steps, elapsed=h.elapsed, residual=list(h.residual), cl=h.cl,
                               cd=h.cd, cm=h.cm, hold=h.hold) for h in hist[:min(res.checks, history)]]
        out["error"] = err.value.decode() if status else ""
        self.meta = dict(time=res.time, steps=res.steps, rejected=res.rejected)
        return out

    def run(self, trials: int, tend: float = 1e30) -> dict:
        """Exactly `trials` step trials (graph replays) toward tend; raises on a fatal status."""
        meta = N.amr_state_meta()
        err = N.err_buffer()
        if self._lib.amr_hip_run(self.handle, tend, trials, C.byref(meta), err, N.ERR_LEN):
            raise BackendError(1, err.value.decode())
        self.meta = dict(time=meta.time, steps=meta.steps, rejected=meta.rejected)
        return self.meta

    def profile(self, trials: int, tend: float = 1e30) -> dict:
        """Per-kernel device milliseconds over `trials` directly launched trials {name: (ms, launches)}."""
        out = (N.amr_kernel_time * 64)()
        count = C.c_int32()
        err = N.err_buffer()
        if self._lib.amr_hip_profile(self.handle, tend, trials, out, 64, C.byref(count), err, N.ERR_LEN):
            raise BackendError(1, err.value.decode())
        return {out[i].name.decode(): (out[i].total_ms, out[i].launches) for i in range(count.value)}

    def stats(self) -> dict:
        s = N.amr_gpu_stats()
        self._lib.amr_hip_stats(self.handle, C.byref(s))
        return {name: getattr(s, name) for name, _ in N.amr_gpu_stats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            self._lib.amr_hip_destroy(self.handle)
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
