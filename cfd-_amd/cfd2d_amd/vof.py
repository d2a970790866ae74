"""Two-phase flow with the volume-of-fluid method (interface of
cfd2d/vof.py:136-306, interFoam's approach).

The phase fraction alpha lives on the device: its bounded van Leer transport
with interface compression, the mixture density and viscosity, the CSF
surface tension and the variable-density pressure solve run inside the step
(csrc/vof.hip).  The host sets the model up and reads alpha on demand.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .models import Model


class VOF(Model):
    """Phase 1 (alpha = 1, e.g. water) and phase 2 (e.g. air); ``alpha0``
    (nx, ny) the initial fraction; gravity becomes the simulation's body
    force; ``sigma`` > 0 adds surface tension; ``c_alpha`` the compression
    factor; ``cfl_alpha`` the largest Courant number of one alpha sub-step;
    ``alpha_inflow`` the fraction entering through open boundaries."""

    kind = "multiphase"

    def __init__(self, alpha0, rho1: float = 1000.0, rho2: float = 1.0, nu1: float = 1e-6,
                 nu2: float = 1.48e-5, sigma: float = 0.07, gravity=(0.0, -9.81),
                 c_alpha: float = 1.0, cfl_alpha: float = 0.25, alpha_inflow: float = 0.0,
                 multigrid: bool = True):
        self.alpha0 = np.asarray(alpha0, dtype=float)
        self.rho1, self.rho2 = float(rho1), float(rho2)
        self.mu1, self.mu2 = float(nu1 * rho1), float(nu2 * rho2)
        self.sigma = float(sigma)
        self.gravity = (float(gravity[0]), float(gravity[1]))
        self.c_alpha = float(c_alpha)
        self.cfl = float(cfl_alpha)
        self.alpha_inflow = float(alpha_inflow)
        self.multigrid = bool(multigrid)   # preconditioner of the variable-density CG
        self.title = (f"VOF zweiphasig: rho1 = {rho1:g}, rho2 = {rho2:g}, sigma = {sigma:g}, "
                      f"g = {self.gravity}, c_alpha = {c_alpha:g}")

    def attach(self, sim) -> None:
        super().attach(sim)
        g = sim.grid
        if self.alpha0.shape != (g.nx, g.ny):
            raise ValueError(f"alpha0 muss die Form {(g.nx, g.ny)} haben.")
        a0 = np.ascontiguousarray(np.where(g.fluid_interior, np.clip(self.alpha0, 0, 1), 0.0))
        pinned = np.ascontiguousarray(np.array(sim.pressure._pinned, dtype=np.int32).reshape(-1))
        d = _lib.VofDesc()
        d.rho1, d.rho2, d.mu1, d.mu2 = self.rho1, self.rho2, self.mu1, self.mu2
        d.sigma, d.c_alpha, d.cfl_alpha = self.sigma, self.c_alpha, self.cfl
        d.alpha_inflow = self.alpha_inflow
        d.alpha0 = _lib.ptr(a0)
        d.npinned = len(sim.pressure._pinned)
        d.jacobi_only = 0 if self.multigrid else 1
        d.pinned = _lib.ptr(pinned) if d.npinned else None
        sim._flush()
        sim._ctx.call("cfd_set_vof", ctypes.byref(d))
        sim.variable_density = True
        sim.variable_viscosity = True
        sim.body_force = (sim.body_force[0] + self.gravity[0], sim.body_force[1] + self.gravity[1])
        self.volume0 = self.volume()

    # device fields ---------------------------------------------------------------
    @property
    def alpha(self) -> np.ndarray:
        self.sim._flush()
        return self.sim._ctx.download("alpha")

    @alpha.setter
    def alpha(self, arr) -> None:
        self.sim._flush()
        self.sim._ctx.upload("alpha", np.asarray(arr, dtype=float))

    def volume(self) -> float:
        g = self.sim.grid
        return float(self.alpha[1:-1, 1:-1][g.fluid_interior].sum() * g.dx * g.dy)

    def dt_limit(self, sim):
        """Capillary (Brackbill) and gravity-wave limits (vof.py:256-268)."""
        g = sim.grid
        lim = None
        if self.sigma > 0:
            h = min(g.dx, g.dy)
            lim = np.sqrt(0.5 * (self.rho1 + self.rho2) * h ** 3 / (2 * np.pi * self.sigma))
        gmag = np.hypot(*self.gravity)
        if gmag > 0:
            lg = np.sqrt(min(g.dx, g.dy) / gmag)
            lim = lg if lim is None else min(lim, lg)
        return lim

    def record_monitors(self, sim, rec) -> dict:
        g = sim.grid
        s, front, height = (rec.mon[_lib.REC_VOF + k] for k in range(3))
        vol = float(s * g.dx * g.dy)
        return {"volumen_phase1": vol,
                "volumenfehler_rel": vol / max(self.volume0, 1e-300) - 1.0,
                "front_x": float(g.xc[int(front) - 1] + 0.5 * g.dx) if front >= 1 else 0.0,
                "hoehe_links": float(g.yc[int(height) - 1] + 0.5 * g.dy) if height >= 1 else 0.0}

    def monitors(self, sim) -> dict:
        g = sim.grid
        a = self.alpha[1:-1, 1:-1]
        vol = float(a[g.fluid_interior].sum() * g.dx * g.dy)
        row = np.where(a[:, 0] > 0.5)[0]
        col = np.where(a[0, :] > 0.5)[0]
        return {"volumen_phase1": vol,
                "volumenfehler_rel": vol / max(self.volume0, 1e-300) - 1.0,
                "front_x": float(g.xc[row.max()] + 0.5 * g.dx) if row.size else 0.0,
                "hoehe_links": float(g.yc[col.max()] + 0.5 * g.dy) if col.size else 0.0}

    def state(self) -> dict:
        return {"alpha": self.alpha, "volume0": np.array(self.volume0)}

    def set_state(self, sim, state: dict) -> None:
        self.alpha = state["alpha"]
        if "volume0" in state:
            self.volume0 = float(state["volume0"])

    def fields(self, sim) -> dict:
        solid = sim.grid.solid
        return {"alpha": ("Volumenanteil α (Phase 1)", np.where(solid, np.nan, self.alpha[1:-1, 1:-1])),
                "rho": ("Dichte ρ", np.where(solid, np.nan, sim.rho[1:-1, 1:-1]))}

    def field_keys(self) -> list[str]:
        return ["alpha", "rho"]
