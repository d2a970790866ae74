"""Device context: owns the GPU state of one simulation through the C ABI."""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .boundary import SIDES, DomainBoundary
from .grid import Grid
from .pressure import PressureSolver


class DeviceContext:
    def __init__(self, grid: Grid, boundary: DomainBoundary, pressure: PressureSolver):
        L = _lib.lib()
        if _lib.device_count() < 1:
            raise RuntimeError("no HIP device: the cfd2d_amd solver runs on the GPU only")
        self.grid = grid
        nx, ny = grid.nx, grid.ny
        self._keep = []   # host arrays referenced by the descriptor during creation

        def keep(a):
            a = np.ascontiguousarray(a)
            self._keep.append(a)
            return a

        d = _lib.GridDesc()
        d.nx, d.ny, d.dx, d.dy = nx, ny, grid.dx, grid.dy
        d.periodic = 1 if boundary.periodic else 0
        d.fluid = _lib.ptr(keep(grid.fluid.astype(np.uint8)))
        for k, side in enumerate(SIDES):
            bc = boundary.bcs[side]
            d.bc_kind[k] = _lib.BC_KINDS[bc.kind]
            d.bc_velocity[k] = float(bc.velocity)
            d.open[k] = _lib.ptr(keep(boundary.open[side].astype(np.uint8)))
            if side in boundary.inflow_values:
                d.inflow[k] = _lib.ptr(keep(boundary.inflow_values[side].astype(np.float64)))
            else:
                d.inflow[k] = None
        d.pcode = _lib.ptr(keep(pressure.code))
        d.pcomp = _lib.ptr(keep(pressure.comp))
        d.ncomp = len(pressure.neumann_components)
        d.comp_count = _lib.ptr(keep(pressure.comp_count.astype(np.float64)))
        d.diag_table = _lib.ptr(keep(pressure.diag_table))
        d.inv_table = _lib.ptr(keep(pressure.inv_table))
        h = ctypes.c_void_p()
        _lib.check(L.cfd_ctx_create(ctypes.byref(d), ctypes.byref(h)), "context creation")
        self.handle = h
        self._keep = []
        _lib.track(self)

    def close(self) -> None:
        if getattr(self, "handle", None):
            _lib.lib().cfd_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        if _lib.closing_allowed():
            try:
                self.close()
            except Exception:
                pass

    # fields ------------------------------------------------------------------
    def shape(self, name: str) -> tuple[int, int]:
        nx, ny = self.grid.nx, self.grid.ny
        if name in ("u", "F", "hu", "hu_old"):
            return (nx + 1, ny + 2)
        if name in ("v", "G", "hv", "hv_old"):
            return (nx + 2, ny + 1)
        if name in ("p", "nut", "diff_k", "diff_omega", "alpha", "rho", "mu"):
            return (nx + 2, ny + 2)
        if name in ("x", "xprev"):
            return (nx, ny)
        raise KeyError(name)

    def download(self, name: str) -> np.ndarray:
        out = np.empty(self.shape(name))
        _lib.check(_lib.lib().cfd_download(self.handle, _lib.FIELDS[name], _lib.ptr(out)),
                   f"download {name}")
        return out

    def upload(self, name: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape(name):
            raise ValueError(f"{name} has shape {a.shape}, expected {self.shape(name)}")
        _lib.check(_lib.lib().cfd_upload(self.handle, _lib.FIELDS[name], _lib.ptr(a)),
                   f"upload {name}")

    def add_scalar(self, desc: _lib.ScalarDesc) -> int:
        i = ctypes.c_int32(-1)
        _lib.check(_lib.lib().cfd_add_scalar(self.handle, ctypes.byref(desc), ctypes.byref(i)),
                   "add scalar")
        return int(i.value)

    def download_scalar(self, sid: int) -> np.ndarray:
        out = np.empty((self.grid.nx + 2, self.grid.ny + 2))
        _lib.check(_lib.lib().cfd_download_scalar(self.handle, sid, _lib.ptr(out)),
                   "download scalar")
        return out

    def upload_scalar(self, sid: int, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != (self.grid.nx + 2, self.grid.ny + 2):
            raise ValueError(f"scalar field has shape {a.shape}")
        _lib.check(_lib.lib().cfd_upload_scalar(self.handle, sid, _lib.ptr(a)), "upload scalar")

    # state -------------------------------------------------------------------
    def get_state(self) -> _lib.State:
        s = _lib.State()
        _lib.check(_lib.lib().cfd_get_state(self.handle, ctypes.byref(s)), "get state")
        return s

    def set_state(self, s: _lib.State) -> None:
        _lib.check(_lib.lib().cfd_set_state(self.handle, ctypes.byref(s)), "set state")

    def initial_iterate(self) -> tuple[int, int]:
        """(mode, has_history) of the pressure solves' initial iterate: mode 1
        extrapolates in time, and has_history says whether ``xprev`` holds the
        solution before the last one."""
        m, h = ctypes.c_int32(0), ctypes.c_int32(0)
        self.call("cfd_initial_iterate", ctypes.byref(m), ctypes.byref(h))
        return int(m.value), int(h.value)

    def apply_bc(self) -> None:
        _lib.check(_lib.lib().cfd_apply_bc(self.handle), "apply bc")

    def project_initial(self, prm: _lib.StepParams) -> None:
        _lib.check(_lib.lib().cfd_project_initial(self.handle, ctypes.byref(prm)),
                   "initial projection")

    def step(self, prm: _lib.StepParams, count: int = 1) -> None:
        _lib.check(_lib.lib().cfd_step(self.handle, ctypes.byref(prm), int(count)), "step")

    def call(self, fn: str, *args) -> int:
        """``fn(handle, *args)`` of the C interface, status checked."""
        return _lib.check(getattr(_lib.lib(), fn)(self.handle, *args), fn)

    def sync(self) -> None:
        _lib.check(_lib.lib().cfd_sync(self.handle), "sync")

    def set_spin_limit(self, seconds: float) -> None:
        """Bound of the device's waits on other work-groups (default 2 s);
        clears an earlier give-up.  0 makes every wait give up at once."""
        _lib.check(_lib.lib().cfd_set_spin_limit(self.handle, float(seconds)), "spin limit")

    def elapsed_ms(self) -> float:
        ms = ctypes.c_double(0.0)
        _lib.check(_lib.lib().cfd_last_elapsed_ms(self.handle, ctypes.byref(ms)), "elapsed")
        return float(ms.value)

    def set_profiling(self, on: bool) -> None:
        _lib.check(_lib.lib().cfd_set_profiling(self.handle, 1 if on else 0), "profiling")

    def projection_times(self) -> tuple[float, int]:
        ms = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        _lib.check(_lib.lib().cfd_projection_times(self.handle, ctypes.byref(ms), ctypes.byref(n)),
                   "projection times")
        return float(ms.value), int(n.value)

    def records(self, first_n: int, count: int) -> list:
        if count <= 0:
            return []
        arr = (_lib.Record * count)()
        _lib.check(_lib.lib().cfd_read_records(self.handle, int(first_n), int(count), arr),
                   "read records")
        return list(arr)
