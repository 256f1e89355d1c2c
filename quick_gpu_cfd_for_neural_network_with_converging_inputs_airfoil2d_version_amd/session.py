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
