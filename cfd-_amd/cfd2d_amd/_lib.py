"""ctypes binding of libcfd2d_hip.so (declarations in include/cfd2d_hip.h).

The library is built in-tree by ``cfd-_amd/build.py``.  There is no CPU
fallback: importing the simulation classes works without a GPU, but creating a
device context raises ``RuntimeError`` when the library or a HIP device is
missing.
"""

from __future__ import annotations

import atexit
import ctypes
import itertools
import sys
import warnings
import weakref
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libcfd2d_hip.so"

CFD_OK = 0
CFD_ERR_ARG = -1
CFD_ERR_HIP = -2
CFD_ERR_STATE = -3
CFD_ERR_LIMIT = -4

BC_KINDS = {"noslip": 0, "slip": 1, "inflow": 2, "outflow": 3, "periodic": 4}
FIELDS = {"u": 0, "v": 1, "p": 2, "F": 3, "G": 4, "hu": 5, "hv": 6, "hu_old": 7, "hv_old": 8,
          "x": 9, "nut": 10, "diff_k": 11, "diff_omega": 12, "alpha": 13, "rho": 14, "mu": 15,
          "xprev": 16}
REC_NUT_MAX = 30     # record slot of the step's largest nu_t (10 + 5 * 4 scalar sums before it)
REC_FLOW_FORCE = 31  # record slot of the driving force after the step
REC_VOF = 32         # VOF: alpha sum over fluid cells, front cell, column-height cell
SOLVERS = {"pcg64": 0, "mgcg": 1, "learned": 2}
LEARNED_MODES = {"net": 0, "none": 1, "teacher": 2, "vcycle": 3}

_d = ctypes.c_double
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


class GridDesc(ctypes.Structure):
    _fields_ = [("nx", _i32), ("ny", _i32), ("dx", _d), ("dy", _d), ("periodic", _i32),
                ("fluid", _vp), ("bc_kind", _i32 * 4), ("bc_velocity", _d * 4),
                ("inflow", _vp * 4), ("open", _vp * 4), ("pcode", _vp), ("pcomp", _vp),
                ("ncomp", _i32), ("comp_count", _vp), ("diag_table", _vp), ("inv_table", _vp)]


class ScalarDesc(ctypes.Structure):
    _fields_ = [("diffusivity", _d), ("limiter", _i32), ("obstacle_fixed", _i32),
                ("obstacle_value", _d), ("bt", _vp * 4), ("bv", _vp * 4), ("wall_flux", _i32),
                ("boussinesq", _i32), ("gx", _d), ("gy", _d), ("t_ref", _d)]


class StepParams(ctypes.Structure):
    _fields_ = [("nu", _d), ("tau", _d), ("scheme_ab2", _i32), ("gamma_auto", _i32),
                ("gamma_fixed", _d), ("dt_fixed", _d), ("dt_model", _d), ("t_end", _d),
                ("clamp_t_end", _i32), ("solver", _i32), ("fx", _d), ("fy", _d), ("rtol", _d),
                ("maxiter", _i32), ("check_t_end", _i32), ("reliable_delta", _d), ("steady_tol", _d),
                ("max_steps", _i64), ("flow_control", _i32), ("pad_flow", _i32),
                ("flow_rate", _d)]


class MGLevelDesc(ctypes.Structure):
    _fields_ = [("nx", _i32), ("ny", _i32), ("diag", _vp), ("ce", _vp), ("cn", _vp)]


class MGDesc(ctypes.Structure):
    _fields_ = [("nlev", _i32), ("omega", ctypes.c_float), ("wc", ctypes.c_float),
                ("ncoarse", _i32), ("cinv", _vp), ("lev", MGLevelDesc * 16)]


class LearnedDesc(ctypes.Structure):
    _fields_ = [("net", _vp), ("level", _i32), ("mode", _i32), ("teacher_cycles", _i32),
                ("max_iter", _i32), ("fallback", _i32), ("use_graph", _i32),
                ("sigma", ctypes.c_float), ("one_a_one", _d), ("harvest_in", _vp),
                ("harvest_out", _vp), ("harvest_cap", _i64), ("coarse_fused", _i32),
                ("x0_extrapolate", _i32), ("sweeps", _i32)]


class ProbeDesc(ctypes.Structure):
    _fields_ = [("nsample", _i32), ("every", _i32), ("capacity", _i32), ("pad", _i32),
                ("src", _vp), ("corner", _vp), ("frac", _vp)]


class TurbDesc(ctypes.Structure):
    _fields_ = [("model", _i32), ("sid_k", _i32), ("sid_omega", _i32), ("pad", _i32),
                ("nu", _d), ("l2", _vp), ("wall_distance", _vp), ("wall_adjacent", _vp),
                ("omega_wall", _vp)]


class VofDesc(ctypes.Structure):
    _fields_ = [("rho1", _d), ("rho2", _d), ("mu1", _d), ("mu2", _d), ("sigma", _d),
                ("c_alpha", _d), ("cfl_alpha", _d), ("alpha_inflow", _d), ("alpha0", _vp),
                ("npinned", _i32), ("jacobi_only", _i32), ("pinned", _vp)]


class LearnedStats(ctypes.Structure):
    _fields_ = [("harvested", _i64), ("solves", _i64), ("iterations", _i64), ("fallbacks", _i64),
                ("last_iters", _i32), ("last_info", _i32), ("last_relres", _d),
                ("net_off", _i32), ("last_sweeps", _i32)]


ENS_MAX_CLASSES = 8
ENS_MAX_LEVELS = 12


class EnsClassDesc(ctypes.Structure):
    _fields_ = [("n", _i32), ("B", _i32), ("nlev", _i32), ("nu", _vp), ("lid", _vp),
                ("kappa", _vp), ("lev_n", _i32 * ENS_MAX_LEVELS), ("diag", _vp * ENS_MAX_LEVELS),
                ("ce", _vp * ENS_MAX_LEVELS), ("cn", _vp * ENS_MAX_LEVELS), ("cinv", _vp),
                ("one_a_one", _d)]


class EnsDesc(ctypes.Structure):
    _fields_ = [("nclass", _i32), ("cls", EnsClassDesc * ENS_MAX_CLASSES), ("tau", _d),
                ("rtol", _d), ("maxiter", _i32), ("omega", ctypes.c_float),
                ("wc", ctypes.c_float), ("autotune", _i32), ("use_graph", _i32),
                ("x0_extrapolate", _i32), ("pad", _i32)]


class EnsStats(ctypes.Structure):
    _fields_ = [("t", _d), ("dt", _d), ("gamma", _d), ("kinetic_energy", _d),
                ("divergence", _d), ("residual", _d), ("relres", _d), ("steps", _i64),
                ("sweeps_total", _i64), ("iters", _i32), ("info", _i32), ("pred_tx", _i32),
                ("pred_ty", _i32), ("pred_us", ctypes.c_float)]


class EnsTuning(ctypes.Structure):
    _fields_ = [("pred_tx", _i32), ("pred_ty", _i32), ("nlev", _i32), ("vec_batch", _i32),
                ("sweep_tx", _i32 * ENS_MAX_LEVELS), ("sweep_ty", _i32 * ENS_MAX_LEVELS),
                ("pred_us", ctypes.c_float), ("vec_us", ctypes.c_float),
                ("vec_default_us", ctypes.c_float), ("pad", ctypes.c_float),
                ("sweep_us", ctypes.c_float * ENS_MAX_LEVELS),
                ("sweep_default_us", ctypes.c_float * ENS_MAX_LEVELS)]


class State(ctypes.Structure):
    _fields_ = [("t", _d), ("dt", _d), ("dt_old", _d), ("gamma", _d), ("flow_force", _d),
                ("n", _i64), ("has_old", _i32), ("done", _i32), ("error", _i32), ("pad", _i32)]


class Record(ctypes.Structure):
    _fields_ = [("t", _d), ("dt", _d), ("gamma", _d), ("residual", _d), ("divergence", _d),
                ("kinetic_energy", _d), ("cg_relres", _d), ("mon", _d * 36), ("n", _i64),
                ("cg_iters", _i32), ("cg_info", _i32), ("cg_outer", _i32), ("nmon", _i32)]


_LIB = None
# objects holding device resources (anything with close()), by creation order
_LIVE: "weakref.WeakValueDictionary[int, object]" = weakref.WeakValueDictionary()
_SEQ = itertools.count()
_SHUT = False


def track(obj) -> None:
    """Register an object that owns device resources; whatever is still open
    at interpreter exit is closed, newest first, by :func:`_at_exit`."""
    _LIVE[next(_SEQ)] = obj


def closing_allowed() -> bool:
    """False once the library has shut down or Python is finalising: a
    ``__del__`` then makes no HIP call (the runtime may be half torn down)."""
    return not _SHUT and not sys.is_finalizing()


def _at_exit() -> None:
    """Close every tracked object, then ``cfd_shutdown`` (waits for the
    device).  Runs from ``atexit``, i.e. before the process's C exit handlers
    tear the HIP runtime and any profiler down."""
    global _SHUT
    if _LIB is None or _SHUT:
        return
    for key in sorted(_LIVE.keys(), reverse=True):
        obj = _LIVE.get(key)
        if obj is not None:
            try:
                obj.close()
            except Exception:
                pass
    try:
        _LIB.cfd_shutdown()
    finally:
        _SHUT = True


def lib() -> ctypes.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if _SHUT:
        raise RuntimeError("the cfd2d_amd library has been shut down (interpreter exit)")
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python cfd-_amd/build.py` "
                           "(there is no CPU fallback)")
    L = ctypes.CDLL(str(LIB_PATH))
    P = ctypes.POINTER
    sig = {
        "cfd_last_error": ([], ctypes.c_char_p),
        "cfd_shutdown": ([], _i32),
        "cfd_set_initial_iterate": ([_vp, _i32], _i32),
        "cfd_initial_iterate": ([_vp, P(_i32), P(_i32)], _i32),
        "cfd_device_count": ([P(_i32)], _i32),
        "cfd_ctx_create": ([P(GridDesc), P(_vp)], _i32),
        "cfd_ctx_destroy": ([_vp], _i32),
        "cfd_ctx_set_stream": ([_vp, _vp], _i32),
        "cfd_add_scalar": ([_vp, P(ScalarDesc), P(_i32)], _i32),
        "cfd_upload": ([_vp, _i32, _vp], _i32),
        "cfd_download": ([_vp, _i32, _vp], _i32),
        "cfd_upload_scalar": ([_vp, _i32, _vp], _i32),
        "cfd_download_scalar": ([_vp, _i32, _vp], _i32),
        "cfd_set_state": ([_vp, P(State)], _i32),
        "cfd_get_state": ([_vp, P(State)], _i32),
        "cfd_apply_bc": ([_vp], _i32),
        "cfd_set_multigrid": ([_vp, P(MGDesc)], _i32),
        "cfd_project_initial": ([_vp, P(StepParams)], _i32),
        "cfd_step": ([_vp, P(StepParams), _i32], _i32),
        "cfd_sync": ([_vp], _i32),
        "cfd_set_spin_limit": ([_vp, _d], _i32),
        "cfd_read_records": ([_vp, _i64, _i32, P(Record)], _i32),
        "cfd_last_elapsed_ms": ([_vp, P(_d)], _i32),
        "cfd_set_profiling": ([_vp, _i32], _i32),
        "cfd_projection_times": ([_vp, P(_d), P(_i64)], _i32),
        "cfd_pressure_solve": ([_vp, _vp, _vp, _i32, _d, _i32, P(_i32), P(_d)], _i32),
        "cfd_dev_alloc": ([_i64, P(_vp)], _i32),
        "cfd_dev_free": ([_vp], _i32),
        "cfd_memcpy": ([_vp, _vp, _i64, _i32], _i32),
        "cfd_momentum_tendencies": ([_vp, _vp, _vp, _i32, _i32, _d, _d, _d, _d, _i32, _vp, _vp,
                                     _vp], _i32),
        "cfd_obstacle_bc": ([_vp, _vp, _vp, _i32, _i32, _i32, _vp], _i32),
        "cfd_scalar_transport": ([_vp, _vp, _vp, _vp, _i32, _i32, _d, _d, _vp, _d, _i32, _i32]
                                 + [_vp] * 8 + [_i32, _d, _vp, _vp], _i32),
        "cfd_probes_set": ([_vp, P(ProbeDesc)], _i32),
        "cfd_probes_read": ([_vp, _i32, _vp, _vp, P(_i32), P(_i64)], _i32),
        "cfd_favg_set": ([_vp, _i32, _d], _i32),
        "cfd_favg_read": ([_vp, P(_d), _vp, _vp], _i32),
        "cfd_download_derived": ([_vp, _i32, _vp], _i32),
        "cfd_download_derived_noalloc": ([_vp, _i32, _vp], _i32),
        "cfd_set_turbulence": ([_vp, P(TurbDesc)], _i32),
        "cfd_set_vof": ([_vp, P(VofDesc)], _i32),
        "cfd_momentum_tendencies_var": ([_vp, _vp, _vp, _i32, _i32, _d, _d, _vp, _d, _i32, _vp,
                                         _vp, _vp], _i32),
        "cfd_set_learned": ([_vp, P(LearnedDesc)], _i32),
        "cfd_learned_stats": ([_vp, P(LearnedStats)], _i32),
        "cfd_learned_reset": ([_vp], _i32),
        "cfd_ens_create": ([P(EnsDesc), P(_vp)], _i32),
        "cfd_ens_destroy": ([_vp], _i32),
        "cfd_ens_set_net": ([_vp, _vp, ctypes.c_float], _i32),
        "cfd_ens_set_coarse": ([_vp, _i32], _i32),
        "cfd_ens_step": ([_vp, _i32], _i32),
        "cfd_ens_download": ([_vp, _i32, _i32, _i32, _vp], _i32),
        "cfd_ens_stats": ([_vp, _i32, _i32, P(EnsStats)], _i32),
        "cfd_ens_tuning": ([_vp, _i32, P(EnsTuning)], _i32),
        "cfd_unet_param_count": ([], _i64),
        "cfd_unet_create": ([_i32, _i32, P(_vp)], _i32),
        "cfd_unet_destroy": ([_vp], _i32),
        "cfd_unet_set_params": ([_vp, _vp], _i32),
        "cfd_unet_get_params": ([_vp, _vp], _i32),
        "cfd_unet_get_grad": ([_vp, _vp], _i32),
        "cfd_unet_zero_grad": ([_vp], _i32),
        "cfd_unet_forward": ([_vp, _vp, _vp], _i32),
        "cfd_unet_loss_grad": ([_vp, _vp, _vp, _vp, _d, P(_d)], _i32),
        "cfd_unet_adam": ([_vp, ctypes.c_float, ctypes.c_float], _i32),
        "cfd_unet_set_loss": ([_vp, _i32, _i32, _i32, ctypes.c_float], _i32),
        "cfd_unet_set_batch": ([_vp, _i32], _i32),
        "cfd_mfma_layout_check": ([_vp, _vp, _vp], _i32),
        "cfd_mfma_layout_check_k32": ([_vp, _vp, _vp], _i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _LIB = L
    atexit.register(_at_exit)
    return L


def exported_symbols() -> list[str]:
    return [n for n in ("cfd_last_error", "cfd_device_count", "cfd_shutdown",
                        "cfd_set_initial_iterate", "cfd_initial_iterate", "cfd_ctx_create",
                        "cfd_ctx_destroy", "cfd_ctx_set_stream", "cfd_add_scalar", "cfd_upload",
                        "cfd_download", "cfd_upload_scalar", "cfd_download_scalar",
                        "cfd_set_state", "cfd_get_state", "cfd_set_multigrid", "cfd_apply_bc",
                        "cfd_project_initial", "cfd_step", "cfd_sync", "cfd_set_spin_limit",
                        "cfd_read_records",
                        "cfd_last_elapsed_ms", "cfd_set_profiling", "cfd_projection_times",
                        "cfd_pressure_solve", "cfd_dev_alloc",
                        "cfd_dev_free", "cfd_memcpy", "cfd_momentum_tendencies",
                        "cfd_obstacle_bc", "cfd_scalar_transport", "cfd_probes_set",
                        "cfd_probes_read", "cfd_favg_set", "cfd_favg_read", "cfd_download_derived",
                        "cfd_download_derived_noalloc",
                        "cfd_set_turbulence", "cfd_momentum_tendencies_var", "cfd_set_vof",
                        "cfd_set_learned",
                        "cfd_learned_stats", "cfd_learned_reset", "cfd_ens_create",
                        "cfd_ens_destroy", "cfd_ens_set_net", "cfd_ens_set_coarse",
                        "cfd_ens_step", "cfd_ens_download", "cfd_ens_tuning",
                        "cfd_ens_stats", "cfd_unet_param_count",
                        "cfd_unet_create", "cfd_unet_destroy", "cfd_unet_set_params",
                        "cfd_unet_get_params", "cfd_unet_get_grad", "cfd_unet_zero_grad",
                        "cfd_unet_forward", "cfd_unet_loss_grad", "cfd_unet_adam",
                        "cfd_unet_set_loss", "cfd_unet_set_batch",
                        "cfd_mfma_layout_check", "cfd_mfma_layout_check_k32")]


def check(rc: int, what: str = "") -> int:
    """Map a status code to the reference's conventions: < 0 raises
    (ValueError for argument/limit errors, RuntimeError for HIP errors);
    > 0 (solver not converged) warns, as pressure.py:174-176 does."""
    if rc == CFD_OK:
        return rc
    msg = lib().cfd_last_error().decode(errors="replace")
    if rc > 0:
        warnings.warn(f"CG-Drucklöser nicht konvergiert ({rc} Iterationen)", stacklevel=3)
        return rc
    if rc in (CFD_ERR_ARG, CFD_ERR_LIMIT):
        raise ValueError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: {msg}")


def device_count() -> int:
    n = _i32(0)
    lib().cfd_device_count(ctypes.byref(n))
    return int(n.value)


class DeviceArray:
    """A device buffer holding a copy of a NumPy array (C-order), for the
    stateless kernel entry points."""

    def __init__(self, host: np.ndarray):
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = int(host.nbytes)
        p = _vp()
        check(lib().cfd_dev_alloc(self.nbytes, ctypes.byref(p)), "device alloc")
        self.ptr = p
        track(self)
        self.upload(host)

    def upload(self, host: np.ndarray) -> None:
        h = np.ascontiguousarray(host, dtype=self.dtype)
        check(lib().cfd_memcpy(self.ptr, h.ctypes.data_as(_vp), self.nbytes, 1), "copy to device")

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().cfd_memcpy(out.ctypes.data_as(_vp), self.ptr, self.nbytes, 2), "copy to host")
        return out

    def image(self, k: int) -> "DeviceImage":
        """View of image k of a (N, H, W) array (no copy)."""
        if len(self.shape) != 3 or not 0 <= k < self.shape[0]:
            raise IndexError(k)
        h, w = self.shape[1], self.shape[2]
        return DeviceImage(self, k * h * w * self.dtype.itemsize, (h, w))

    def free(self) -> None:
        if self.ptr:
            lib().cfd_dev_free(self.ptr)
            self.ptr = None

    close = free

    def __del__(self):  # pragma: no cover
        if closing_allowed():
            try:
                self.free()
            except Exception:
                pass


class DeviceImage:
    """A (H, W) window into a DeviceArray (keeps the owner alive)."""

    def __init__(self, owner: DeviceArray, offset: int, shape):
        self.owner, self.shape, self.dtype = owner, tuple(shape), owner.dtype
        self.ptr = _vp(owner.ptr.value + offset)

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().cfd_memcpy(out.ctypes.data_as(_vp), self.ptr, out.nbytes, 2), "copy to host")
        return out


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_vp)
