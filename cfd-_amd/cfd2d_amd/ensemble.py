"""Batched ensembles of lid-driven and heated cavities on one GPU.

No counterpart in the reference (one case per process).  Each instance is a
unit-square cavity as built by ``cases.lid_driven_cavity`` (lid velocity 1,
nu = 1/Re) or ``cases.heated_cavity`` (nu = sqrt(Pr/Ra), temperature 0.5 /
-0.5 on the left / right walls, diffusivity 1/sqrt(Ra Pr), van Leer),
advanced with AB2, gamma "auto" and tau 0.5 from rest.  Instances of one side
length form a class; each kernel of a step serves a whole class.  The bulk is
FP32; the pressure solve is the learned flexible CG (FP64 vectors and dot
products, true residual <= rtol), whose coarse corrections for all instances
come from one batched pass of the shared U-Net per iteration; each solve
starts from the linear extrapolation in time of the instance's last two
solutions (``x0="extrapolate"``; "previous": from the last one).  With
``use_graph`` a step is one graph launch whose iteration count is controlled
on the device (a step that needs more iterations than the graph holds
continues in the next launch).  ``coarse="vcycle"`` replaces the network by the V-cycle carried on to the
dense coarsest solve (levels of at most 64 x 64 cells in LDS, one work-group
per instance).  With ``autotune`` the work-group shapes are chosen per class
by timing at setup: the predictor's, the pointwise V-cycle kernels' per
level, and the cells per thread of the FP64 vector kernels (``tuning``); no
choice changes a computed value.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .learned import load_weights
from .unet import UNet

SIDES = (64, 128, 256, 512)


@dataclass
class Instance:
    kind: str          # "lid" or "heated"
    n: int
    re: float = 100.0
    ra: float = 1e4
    pr: float = 0.71

    @property
    def nu(self) -> float:
        return 1.0 / self.re if self.kind == "lid" else float(np.sqrt(self.pr / self.ra))

    @property
    def kappa(self) -> float:
        return 0.0 if self.kind == "lid" else float(1.0 / np.sqrt(self.ra * self.pr))

    @property
    def lid(self) -> float:
        return 1.0 if self.kind == "lid" else 0.0

    def single_case(self, **kw):
        """The same cavity through the single-case path (cases.py)."""
        from . import cases

        if self.kind == "lid":
            return cases.lid_driven_cavity(re=self.re, n=self.n, **kw)
        return cases.heated_cavity(ra=self.ra, pr=self.pr, n=self.n, **kw)


def sample_instances(count: int = 256, seed: int = 0) -> list[Instance]:
    """The batched-ensemble workload: sides uniform over {64, 128, 256, 512};
    the first half lid cavities with Re log-uniform in [100, 5000], the second
    half heated cavities with Ra log-uniform in [1e4, 1e6] (Pr 0.71).  Draw
    order: all sides, then all Re, then all Ra (numpy default_rng(seed))."""
    rng = np.random.default_rng(seed)
    sides = rng.choice(SIDES, size=count)
    nl = count // 2
    res = 10.0 ** rng.uniform(2.0, np.log10(5000.0), nl)
    ras = 10.0 ** rng.uniform(4.0, 6.0, count - nl)
    out = [Instance("lid", int(sides[k]), re=float(res[k])) for k in range(nl)]
    out += [Instance("heated", int(sides[nl + k]), ra=float(ras[k])) for k in range(count - nl)]
    return out


def _cavity_hierarchy(n: int):
    """The multigrid hierarchy of the n x n pinned Neumann cavity (the
    pressure system of either cavity)."""
    from .boundary import BoundaryCondition as BC
    from .boundary import DomainBoundary
    from .grid import Grid
    from .multigrid import Hierarchy
    from .pressure import PressureSolver

    g = Grid(n, n, 1.0, 1.0)
    g.finalize()
    bcs = {s: BC("noslip") for s in ("left", "right", "bottom", "top")}
    ps = PressureSolver(g, DomainBoundary(g, bcs), method="cg_mixed", tol=1e-6)
    return Hierarchy(ps.fluid, ps.diag, n * n, n * n, False)


class Ensemble:
    """Instances advanced together; ``step(k)`` runs k steps of all of them."""

    def __init__(self, instances: list[Instance], net: bool | str = True, params=None,
                 sigma=None, rtol: float = 1e-6, maxiter: int = 200, autotune: bool = True,
                 use_graph: bool = True, x0: str = "extrapolate", coarse: str = "network"):
        self.instances = list(instances)
        sides = sorted({i.n for i in self.instances})
        if any(n not in SIDES for n in sides):
            raise ValueError(f"ensemble sides must be in {SIDES}")
        if len(sides) > 8:
            raise ValueError("at most 8 size classes")
        self.classes = [[k for k, i in enumerate(self.instances) if i.n == n] for n in sides]
        self.where = {}
        for c, members in enumerate(self.classes):
            for b, k in enumerate(members):
                self.where[k] = (c, b)
        d = _lib.EnsDesc()
        d.nclass = len(sides)
        d.tau, d.rtol, d.maxiter = 0.5, float(rtol), int(maxiter)
        d.omega, d.wc = 0.8, 1.0
        d.autotune = 1 if autotune else 0
        d.use_graph = 1 if use_graph else 0
        if x0 not in ("extrapolate", "previous"):
            raise ValueError(f"x0 must be 'extrapolate' or 'previous', got {x0!r}")
        d.x0_extrapolate = 1 if x0 == "extrapolate" else 0
        keep = []
        for c, n in enumerate(sides):
            members = [self.instances[k] for k in self.classes[c]]
            h = _cavity_hierarchy(n)
            cd = d.cls[c]
            cd.n, cd.B, cd.nlev = n, len(members), len(h.levels)
            arrs = [np.array([getattr(m, a) for m in members], dtype=np.float64)
                    for a in ("nu", "lid", "kappa")]
            keep += arrs
            cd.nu, cd.lid, cd.kappa = (_lib.ptr(a) for a in arrs)
            for l, L in enumerate(h.levels):
                cd.lev_n[l] = L["nx"]
                if l == 0:
                    continue
                ops = [np.ascontiguousarray(L[k], dtype=np.float32) for k in ("diag", "ce", "cn")]
                keep += ops
                cd.diag[l], cd.ce[l], cd.cn[l] = (_lib.ptr(a) for a in ops)
            cinv = np.ascontiguousarray(h.cinv, dtype=np.float32)
            keep.append(cinv)
            cd.cinv = _lib.ptr(cinv)
            K = int(np.log2(n // 32))
            cd.one_a_one = float(h.levels[K]["A"].sum())
        hdl = ctypes.c_void_p()
        _lib.check(_lib.lib().cfd_ens_create(ctypes.byref(d), ctypes.byref(hdl)), "ensemble create")
        self._h = hdl
        _lib.track(self)
        self.net = None
        if coarse not in ("network", "vcycle"):
            raise ValueError(f"coarse must be 'network' or 'vcycle', got {coarse!r}")
        self.coarse = coarse
        if coarse == "vcycle":
            _lib.check(_lib.lib().cfd_ens_set_coarse(self._h, 1), "ensemble coarse end")
        elif net:
            if params is None:
                params, sigma = load_weights(net if isinstance(net, str) else None)
            self.net = UNet(32 * len(self.instances), 32, params)
            _lib.check(_lib.lib().cfd_unet_set_batch(self.net.handle, 32), "network batch")
            _lib.check(_lib.lib().cfd_ens_set_net(self._h, self.net.handle,
                                                  float(1.0 if sigma is None else sigma)),
                       "ensemble network")
        self.steps = 0

    @property
    def cells(self) -> int:
        return sum(i.n * i.n for i in self.instances)

    def step(self, count: int = 1) -> None:
        _lib.check(_lib.lib().cfd_ens_step(self._h, int(count)), "ensemble step")
        self.steps += count

    def field(self, k: int, name: str) -> np.ndarray:
        c, b = self.where[k]
        n = self.instances[k].n
        fid, shape = {"u": (0, (n + 1, n + 2)), "v": (1, (n + 2, n + 1)),
                      "p": (2, (n + 2, n + 2)), "T": (3, (n + 2, n + 2))}[name]
        out = np.empty(shape, dtype=np.float32)
        _lib.check(_lib.lib().cfd_ens_download(self._h, c, b, fid, _lib.ptr(out)), "download")
        return out

    def stats(self, k: int) -> dict:
        c, b = self.where[k]
        s = _lib.EnsStats()
        _lib.check(_lib.lib().cfd_ens_stats(self._h, c, b, ctypes.byref(s)), "stats")
        return {name: getattr(s, name) for name, _t in _lib.EnsStats._fields_}

    def tuning(self, cls: int) -> dict:
        """What the autotuner chose for size class ``cls`` (index into
        ``classes``): the predictor's work-group shape, the work-group shape of
        the pointwise V-cycle kernels per level, the FP64 vector kernels' cells
        per thread and batch, and the time per launch of each choice and of the
        default shape."""
        t = _lib.EnsTuning()
        _lib.check(_lib.lib().cfd_ens_tuning(self._h, int(cls), ctypes.byref(t)), "tuning")
        nl = t.nlev
        return {"n": self.instances[self.classes[cls][0]].n,
                "predictor": [t.pred_tx, t.pred_ty], "predictor_us": round(t.pred_us, 2),
                "sweeps": [[t.sweep_tx[l], t.sweep_ty[l]] for l in range(nl)],
                "sweeps_us": [round(t.sweep_us[l], 2) for l in range(nl)],
                "sweeps_default_us": [round(t.sweep_default_us[l], 2) for l in range(nl)],
                "vector_batch": t.vec_batch, "vector_us": round(t.vec_us, 2),
                "vector_default_us": round(t.vec_default_us, 2)}

    def close(self) -> None:
        if self._h:
            _lib.lib().cfd_ens_destroy(self._h)
            self._h = None
        if self.net is not None:
            self.net.close()
            self.net = None

    def __del__(self):  # pragma: no cover
        if _lib.closing_allowed():
            try:
                self.close()
            except Exception:
                pass
