"""Learned flexible CG for the pressure system (``pressure_solver="cg_learned"``).

No counterpart in the reference (its solvers are SuperLU and Jacobi-PCG,
cfd2d/pressure.py:136-139).  The solve is an FP64 flexible CG ending on the
FP64 true residual; its preconditioner is an FP32 V-cycle over the aggregation
hierarchy of :mod:`cfd2d_amd.multigrid` that stops at a coarse level K (the
first level with at most ``TARGET_SIDE`` cells per side), where the U-Net of
:mod:`cfd2d_amd.unet` supplies the coarse correction (after an exact
correction in the constant vector), followed by two Jacobi sweeps on level K.
A solve that does not converge within ``max_iter`` iterations is continued by
the FP64 Jacobi-PCG.

Training data come from the solver's own solves: in teacher mode level K is
solved by V-cycles down to the dense coarsest level and each iteration stores
the pair (normalised coarse residual, normalised coarse correction) in device
memory.  :func:`train` fits the network to such pairs on the GPU with the
library's forward and backward kernels and Adam.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _lib
from .unet import UNet, init_params, param_count

TARGET_SIDE = 32
WEIGHTS = Path(__file__).resolve().parent / "weights" / "unet_pressure.npz"


@dataclass
class LearnedConfig:
    """How the learned solver runs.

    mode: "net" (the network), "none" (constant-vector correction only),
    "vcycle" (no network: the V-cycle continues below level K to the dense
    coarsest level, the complete multigrid preconditioner the network is
    measured against) or "teacher" (level K solved by
    ``teacher_cycles`` V-cycles; with ``harvest_cap`` > 0 the training pairs
    are stored).  ``params``/``sigma`` default to the weights shipped in
    ``cfd2d_amd/weights``; ``level`` None picks the level automatically.
    ``coarse_fused``: on a 32 x 32 level K, the levels of at most 64 x 64
    cells, the network and the level-K correction run in one work-group
    launch (k_lv_coarse) instead of ~25 launches; in mode "vcycle" the levels
    of at most 64 x 64 cells down to the dense solve run in one LDS-resident
    work-group (k_lv_vcoarse).  ``fused_legs`` (with ``coarse_fused``): the
    levels above that coarse end take one launch per level and leg (tiles
    staged in LDS) instead of two; ``fused_fine_up`` False keeps the fine
    level's up-leg as two launches, whose dot partials (row bands) make the
    solve bit-identical to the separate launches (tiles add them in another
    order).  ``fused_pair`` (mode "vcycle", with the fused legs): the two
    levels just above the LDS coarse end in one launch per leg, each
    work-group recomputing the cells it needs from the other level (closed
    domains whose level sizes halve exactly; the same fields bit for bit).
    ``sweeps``: damped-Jacobi sweeps per leg on those levels (2: the
    multigrid solver's V-cycle; 3 or 4: a stronger preconditioner, with the
    fused legs; "auto": 3 for a solve after one that took at least 4
    iterations, back to 2 after 8 consecutive 3-sweep solves of at most 2 --
    the extra sweep costs about 15 % of an iteration and pays where solves
    take many).
    ``x0``: "previous" starts each solve from the last solution (SciPy's
    warm start in the reference), "extrapolate" from the linear extrapolation
    of the last two solutions in time."""

    mode: str = "net"
    params: np.ndarray | None = None
    sigma: float | None = None
    level: int | None = None
    teacher_cycles: int = 8
    max_iter: int = 200
    fallback: bool = True
    use_graph: bool = True
    coarse_fused: bool = True
    fused_legs: bool = True
    fused_fine_up: bool = True
    fused_pair: bool = True
    sweeps: int | str = "auto"
    x0: str = "previous"
    harvest_cap: int = 0
    weights: str | None = None
    side: int = TARGET_SIDE
    extra: dict = field(default_factory=dict)

    @classmethod
    def from_spec(cls, spec) -> "LearnedConfig":
        if spec is None:
            return cls()
        if isinstance(spec, LearnedConfig):
            return spec
        if isinstance(spec, dict):
            return cls(**spec)
        raise ValueError(f"learned solver settings must be a dict, got {type(spec).__name__}")


def choose_level(hierarchy, side: int = TARGET_SIDE) -> int:
    """First level (>= 1) with at most `side` cells per side."""
    levels = hierarchy.levels
    for k in range(1, len(levels)):
        if max(levels[k]["nx"], levels[k]["ny"]) <= side:
            return k
    return len(levels) - 1


def image_shape(hierarchy, level: int) -> tuple[int, int]:
    L = hierarchy.levels[level]
    return ((L["nx"] + 7) // 8) * 8, ((L["ny"] + 7) // 8) * 8


def one_a_one(hierarchy, level: int) -> float:
    """1^T A_K 1: the Galerkin coefficient of the constant vector on level K."""
    return float(hierarchy.levels[level]["A"].sum())


def save_weights(path, params: np.ndarray, sigma: float, meta: dict | None = None) -> Path:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(path, params=np.asarray(params, dtype=np.float32),
                        sigma=np.float64(sigma), meta=np.array(repr(meta or {})))
    return path


def load_weights(path=None) -> tuple[np.ndarray, float]:
    path = Path(path) if path else WEIGHTS
    if not path.exists():
        raise ValueError(f"no trained network at {path}: train one with "
                         "cfd2d_amd.learned.train_default() or pass LearnedConfig(params=...)")
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].astype(np.float32)
        sigma = float(z["sigma"])
    if params.size != param_count():
        raise ValueError(f"{path}: {params.size} parameters, the network has {param_count()}")
    return params, sigma


class Attached:
    """A learned solver attached to a device context (keeps its buffers)."""

    def __init__(self, ctx, hierarchy, cfg: LearnedConfig, x0: str = "previous"):
        self.cfg = cfg
        self.level = choose_level(hierarchy, cfg.side) if cfg.level is None else int(cfg.level)
        self.region = (hierarchy.levels[self.level]["nx"], hierarchy.levels[self.level]["ny"])
        if not 0 <= self.level < len(hierarchy.levels):
            raise ValueError(f"level {self.level} outside the hierarchy")
        self.H, self.W = image_shape(hierarchy, self.level)
        self.net = None
        self.sigma = 1.0
        if cfg.mode not in _lib.LEARNED_MODES:
            raise ValueError(f"unknown learned-solver mode {cfg.mode!r}")
        if cfg.mode == "net":
            params, sigma = cfg.params, cfg.sigma
            if params is None:
                params, sigma = load_weights(cfg.weights)
            self.sigma = float(1.0 if sigma is None else sigma)
            self.net = UNet(self.H, self.W, params)
        self.hin = self.hout = None
        cap = int(cfg.harvest_cap) if cfg.mode == "teacher" else 0
        if cap > 0:
            z = np.zeros((cap, self.H, self.W), dtype=np.float32)
            self.hin, self.hout = _lib.DeviceArray(z), _lib.DeviceArray(z)
        d = _lib.LearnedDesc()
        d.net = self.net.handle if self.net else None
        d.level = self.level
        d.mode = _lib.LEARNED_MODES[cfg.mode]
        d.teacher_cycles = int(cfg.teacher_cycles)
        d.max_iter = int(cfg.max_iter)
        d.fallback = 1 if cfg.fallback else 0
        d.use_graph = 1 if cfg.use_graph else 0
        d.coarse_fused = ((1 | (2 if cfg.fused_legs else 0) | (0 if cfg.fused_fine_up else 4)
                           | (8 if cfg.fused_pair else 0))
                          if cfg.coarse_fused else 0)
        if cfg.x0 not in ("extrapolate", "previous"):
            raise ValueError(f"x0 must be 'extrapolate' or 'previous', got {cfg.x0!r}")
        # either setting asks for the extrapolated start (the solver's x0 or
        # the pressure solver's, which applies to every method)
        d.x0_extrapolate = 1 if "extrapolate" in (cfg.x0, x0) else 0
        if cfg.sweeps == "auto":
            d.sweeps = -1
        elif cfg.sweeps in (2, 3, 4):
            d.sweeps = int(cfg.sweeps)
        else:
            raise ValueError(f"sweeps must be 2, 3, 4 or 'auto', got {cfg.sweeps!r}")
        d.sigma = self.sigma
        d.one_a_one = one_a_one(hierarchy, self.level)
        d.harvest_in = self.hin.ptr if self.hin else None
        d.harvest_out = self.hout.ptr if self.hout else None
        d.harvest_cap = cap
        self._ctx = ctx
        _lib.check(_lib.lib().cfd_set_learned(ctx.handle, ctypes.byref(d)), "learned solver setup")

    def stats(self) -> dict:
        s = _lib.LearnedStats()
        _lib.check(_lib.lib().cfd_learned_stats(self._ctx.handle, ctypes.byref(s)), "learned stats")
        return {k: getattr(s, k) for k, _t in _lib.LearnedStats._fields_}

    def reset_stats(self) -> None:
        _lib.check(_lib.lib().cfd_learned_reset(self._ctx.handle), "learned reset")

    def harvest(self) -> tuple[np.ndarray, np.ndarray]:
        """The pairs stored so far, as host arrays (n, H, W)."""
        if self.hin is None:
            raise ValueError("no harvest buffers (teacher mode with harvest_cap > 0)")
        n = min(int(self.stats()["harvested"]), self.hin.shape[0])
        return self.hin.get()[:n], self.hout.get()[:n]


# ---------------------------------------------------------------------- #
# training
# ---------------------------------------------------------------------- #
def augment(x: np.ndarray, t: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """The symmetries of the 5-point Laplacian on a uniform grid: the two
    reflections, and for square images also the transpositions (8 maps)."""
    xs, ts = [x, x[:, ::-1, :], x[:, :, ::-1], x[:, ::-1, ::-1]], []
    for a in (t, t[:, ::-1, :], t[:, :, ::-1], t[:, ::-1, ::-1]):
        ts.append(a)
    if x.shape[1] == x.shape[2]:
        xs += [a.transpose(0, 2, 1) for a in xs]
        ts += [a.transpose(0, 2, 1) for a in ts]
    return (np.ascontiguousarray(np.concatenate(xs)), np.ascontiguousarray(np.concatenate(ts)))


def energy_norm2(t: np.ndarray, region: tuple[int, int], eps: float) -> np.ndarray:
    """t.A t per image, A = unit 5-point Laplacian with zero-flux sides on the
    region (rows, cols) + eps I (the operator of the energy loss)."""
    nx, ny = region
    e = t[:, :nx, :ny].astype(np.float64)
    val = eps * np.sum(e * e, axis=(1, 2))
    val += np.sum((e[:, 1:, :] - e[:, :-1, :]) ** 2, axis=(1, 2))
    val += np.sum((e[:, :, 1:] - e[:, :, :-1]) ** 2, axis=(1, 2))
    return val


def train(x: np.ndarray, t: np.ndarray, epochs: int = 10, batch: int = 8, lr: float = 1e-3,
          seed: int = 0, params: np.ndarray | None = None, log=None,
          time_limit: float | None = None, per_epoch: int | None = None,
          augment_data: bool = True, loss: str = "energy", region: tuple | None = None,
          eps: float = 0.01) -> tuple[np.ndarray, float, list]:
    """Fit the U-Net to pairs (x[k] -> t[k] / sigma) on the GPU.

    sigma = RMS of the targets.  Loss per pair, relative to the target: the
    energy norm of the error (loss="energy", the norm in which a coarse
    correction acts; A as in :func:`energy_norm2` on `region`, the level's
    cells, default the whole image) or its squared 2-norm (loss="mse").  Adam
    on mini-batches (gradients summed over the batch, scaled by 1/batch); the
    learning rate follows a cosine decay to lr/20.  With augment_data the
    symmetric images of every pair join the set (square regions only for the
    transpositions); per_epoch draws that many samples per epoch."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    sigma = float(np.sqrt(np.mean(t.astype(np.float64) ** 2)))
    if not sigma > 0:
        raise ValueError("training targets are all zero")
    ts = (t / sigma).astype(np.float32)
    tn = np.sum(ts.astype(np.float64) ** 2, axis=(1, 2))
    keep = tn > 0
    x, ts = x[keep], ts[keep]
    n, H, W = x.shape
    region = (H, W) if region is None else tuple(region)
    if augment_data:
        if region != (H, W):
            raise ValueError("augmentation needs the region to fill the image")
        x, ts = augment(x, ts)
    if loss == "energy":
        tn = energy_norm2(ts, region, eps)
    else:
        tn = np.sum(ts.astype(np.float64) ** 2, axis=(1, 2))
    n = x.shape[0]
    dx, dt = _lib.DeviceArray(x), _lib.DeviceArray(ts)
    net = UNet(H, W, init_params(seed) if params is None else params)
    if loss == "energy":
        net.set_loss("energy", region, eps)
    rng = np.random.default_rng(seed)
    per_epoch = n if per_epoch is None else min(int(per_epoch), n)
    steps_total = max(1, epochs * (per_epoch // batch))
    step = 0
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        order = rng.permutation(n)[:per_epoch]
        tot, cnt = 0.0, 0
        for b0 in range(0, per_epoch - batch + 1, batch):
            for k in order[b0:b0 + batch]:
                tot += net.loss_grad(dx.image(int(k)), dt.image(int(k)), None, float(tn[k]))
                cnt += 1
            rate = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * step / steps_total)))
            net.adam(float(rate), 1.0 / batch)
            step += 1
        hist.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {ep + 1}/{epochs}: mean relative loss {hist[-1]:.4f} "
                f"({time.time() - t0:.0f} s)")
        if time_limit is not None and time.time() - t0 > time_limit:
            break
    params = net.get_params()
    net.close()
    dx.free()
    dt.free()
    return params, sigma, hist


def harvest_from_cases(cases_spec, steps: int, cap_per_case: int, every: int = 1,
                       log=None) -> tuple[np.ndarray, np.ndarray]:
    """Run each case with the teacher-mode learned solver and collect the
    (input, correction) pairs of its pressure solves.  cases_spec: list of
    (factory, kwargs); every: keep every k-th pair."""
    xs, ts = [], []
    for factory, kw in cases_spec:
        cfg = LearnedConfig(mode="teacher", harvest_cap=cap_per_case)
        case = factory(pressure_solver="cg_learned", learned=cfg, **kw)
        sim = case.sim
        att = sim.pressure.learned
        for _ in range(steps):
            sim.step()
        sim._flush()
        x, t = att.harvest()
        xs.append(x[::every])
        ts.append(t[::every])
        if log:
            log(f"{kw}: {x.shape[0]} pairs on a {x.shape[1]}x{x.shape[2]} image, "
                f"{np.mean(sim.cg_iterations):.2f} iterations per solve")
        sim.close() if hasattr(sim, "close") else None
    shapes = {a.shape[1:] for a in xs}
    if len(shapes) != 1:
        raise ValueError(f"cases give different coarse image sizes {shapes}")
    return np.concatenate(xs), np.concatenate(ts)


# the training set and recipe of the shipped network
DEFAULT_CASES = ([(n, re) for n in (256, 512, 1024) for re in (100.0, 400.0, 1000.0, 3200.0, 5000.0)]
                 + [(2048, 1000.0), (2048, 5000.0)])


def train_default(out=None, steps: int = 650, seconds: float = 600.0, epochs: int = 250,
                  cases_list=None, side: int = TARGET_SIDE, seed: int = 0, log=None) -> Path:
    """Harvest pairs from lid-driven cavities (sizes 256-2048, Re 100-5000,
    pressure_tol 1e-6, `steps` steps each, teacher mode: about 50 000 pairs
    with the defaults) and train the network with the squared-error loss (lr
    1e-3 with cosine decay over `epochs` epochs of 2000 samples, batches of 8,
    augmentation by the grid symmetries) for at most `seconds`; writes the
    weights to `out`."""
    from . import cases as C

    xs, ts = [], []
    for n, re in (cases_list or DEFAULT_CASES):
        cfg = LearnedConfig(mode="teacher", harvest_cap=5 * steps + 64, side=side,
                            teacher_cycles=10)
        sim = C.lid_driven_cavity(re=re, n=n, pressure_solver="cg_learned", pressure_tol=1e-6,
                                  learned=cfg).sim
        for _ in range(steps):
            sim.step()
        sim._flush()
        x, t = sim.pressure.learned.harvest()
        xs.append(x)
        ts.append(t)
        if log:
            log(f"cavity n={n} Re={re:g}: {x.shape[0]} pairs")
        sim.close()
    x, t = np.concatenate(xs), np.concatenate(ts)
    params, sigma, hist = train(x, t, epochs=epochs, batch=8, lr=1e-3, per_epoch=2000, seed=seed,
                                loss="mse", log=log, time_limit=seconds)
    meta = {"pairs": int(x.shape[0]), "image": list(x.shape[1:]), "epochs": len(hist),
            "final_loss": float(hist[-1]), "side": side, "steps": steps}
    return save_weights(out or WEIGHTS, params, sigma, meta)
