"""Transported scalars and the model hook protocol (interface of
cfd2d/models.py:29-426, laminar part).

A :class:`ScalarField` lives on the device: its transport (TVD convection +
diffusion, explicit Euler), ghost fill and monitors run inside the fused step.
The host keeps the boundary description and reads the field on demand.
Boussinesq buoyancy is added to the momentum tendencies on the device.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .boundary import SIDES

LIMITERS = {"upwind": 0, "vanleer": 1, "minmod": 2, "superbee": 3, "linear": 4, "mc": 5}


@dataclass
class ScalarBC:
    kind: str = "zerogradient"
    value: object = 0.0

    _CODES = {"zerogradient": 0, "fixed": 1, "flux": 2}

    def __post_init__(self):
        if self.kind not in self._CODES:
            raise ValueError(f"Unbekannte Skalar-Randbedingung {self.kind!r}, "
                             f"erlaubt: {list(self._CODES)}")

    @property
    def code(self) -> int:
        return self._CODES[self.kind]

    def describe(self) -> str:
        names = {"zerogradient": "Gradient null", "fixed": "fester Wert", "flux": "fester Fluss"}
        v = self.value
        vs = f" = {v:g}" if isinstance(v, (int, float)) else ""
        return names[self.kind] + (vs if self.kind != "zerogradient" else "")


def default_scalar_bcs(sim, inflow_value: float = 0.0) -> dict:
    """Inflow sides get a fixed value, all others zero gradient (models.py:61-70)."""
    return {side: (ScalarBC("fixed", inflow_value) if sim.bcs[side].kind == "inflow"
                   else ScalarBC("zerogradient")) for side in SIDES}


class ScalarField:
    """Transport ``d(phi)/dt + div(u phi) = div(D grad phi)`` on the device."""

    def __init__(self, name: str, diffusivity: float, bcs: dict | None = None,
                 obstacle: ScalarBC | None = None, limiter: str = "vanleer",
                 initial: float | np.ndarray = 0.0, source=None, lower=None, upper=None,
                 label: str | None = None):
        if limiter not in LIMITERS:
            raise ValueError(f"Unbekannter Limiter {limiter!r}, erlaubt: {list(LIMITERS)}")
        if source is not None or lower is not None or upper is not None:
            raise ValueError("source terms and clipping (turbulence models) are not part of "
                             "the GPU step")
        self.name = name
        self.label = label or name
        self.D0 = float(diffusivity)
        self.bcs_in = bcs
        self.obstacle = obstacle or ScalarBC("zerogradient")
        self.limiter = limiter
        self.initial = initial
        self.sim = None
        self.sid = None
        self.boussinesq = None   # (gx, gy, T_ref) when the field drives buoyancy

    def attach(self, sim) -> None:
        self.sim = sim
        g = sim.grid
        nx, ny = g.nx, g.ny
        phi = np.zeros((nx + 2, ny + 2))
        if np.isscalar(self.initial):
            phi[:] = float(self.initial)
        else:
            init = np.asarray(self.initial, dtype=float)
            if init.shape == (nx, ny):
                phi[1:-1, 1:-1] = init
            elif init.shape == phi.shape:
                phi[:] = init
            else:
                raise ValueError(f"Anfangsfeld für {self.name} hat falsche Form {init.shape}")
        self.diff = np.full((nx + 2, ny + 2), self.D0)
        bcs = default_scalar_bcs(sim)
        bcs.update(self.bcs_in or {})
        self.bcs = bcs
        coords = {"left": g.yc, "right": g.yc, "bottom": g.xc, "top": g.xc}
        self.bt, self.bv = {}, {}
        for side in SIDES:
            bc, s = bcs[side], coords[side]
            self.bt[side] = np.full(len(s), bc.code, dtype=np.int64)
            val = bc.value
            if callable(val):
                self.bv[side] = np.asarray(val(s), dtype=float)
            else:
                self.bv[side] = np.broadcast_to(np.asarray(val, dtype=float), s.shape).copy()
        d = _lib.ScalarDesc()
        d.diffusivity = self.D0
        d.limiter = LIMITERS[self.limiter]
        d.obstacle_fixed = 1 if self.obstacle.kind == "fixed" else 0
        d.obstacle_value = float(self.obstacle.value) if self.obstacle.kind == "fixed" else 0.0
        self._bt_keep = [np.ascontiguousarray(self.bt[s]) for s in SIDES]
        self._bv_keep = [np.ascontiguousarray(self.bv[s]) for s in SIDES]
        for k in range(4):
            d.bt[k] = _lib.ptr(self._bt_keep[k])
            d.bv[k] = _lib.ptr(self._bv_keep[k])
        d.wall_flux = 1 if any((self.bt[s] != 0).any() for s in SIDES) else 0
        if self.boussinesq is not None:
            d.boussinesq = 1
            d.gx, d.gy, d.t_ref = self.boussinesq
        self.sid = sim._ctx.add_scalar(d)
        sim._ctx.upload_scalar(self.sid, phi)

    # device field ------------------------------------------------------------
    @property
    def phi(self) -> np.ndarray:
        """Current field with ghosts, shape (nx+2, ny+2) (a host copy)."""
        self.sim._flush()
        return self.sim._ctx.download_scalar(self.sid)

    @phi.setter
    def phi(self, arr) -> None:
        self.sim._flush()
        self.sim._ctx.upload_scalar(self.sid, np.asarray(arr, dtype=float))

    def fill_ghosts(self) -> None:
        """Ghost and solid cells are filled on the device after every update
        and on every upload (models.py:139-183)."""
        self.phi = self.phi

    def dt_limit(self) -> float:
        g = self.sim.grid
        if self.D0 <= 0:
            return np.inf
        return 0.5 / (self.D0 * (1.0 / g.dx ** 2 + 1.0 / g.dy ** 2))

    def cell(self) -> np.ndarray:
        return np.where(self.sim.grid.solid, np.nan, self.phi[1:-1, 1:-1])

    def total(self) -> float:
        g = self.sim.grid
        return float(np.nansum(self.cell()) * g.dx * g.dy)

    def wall_flux(self, side: str) -> np.ndarray:
        g = self.sim.grid
        phi = self.phi
        code, val = self.bt[side], self.bv[side]
        if side == "left":
            inner, h = phi[1, 1:-1], g.dx
        elif side == "right":
            inner, h = phi[g.nx, 1:-1], g.dx
        elif side == "bottom":
            inner, h = phi[1:-1, 1], g.dy
        else:
            inner, h = phi[1:-1, g.ny], g.dy
        q = np.zeros_like(inner)
        f = code == 1
        q[f] = -self.D0 * (val[f] - inner[f]) / (0.5 * h)
        q[code == 2] = -val[code == 2]
        return q

    # monitor values from the device sums of a step record
    def _sums(self, rec) -> tuple[float, list[float]]:
        base = 10 + 5 * self.sid
        return rec.mon[base], [rec.mon[base + 1 + k] for k in range(4)]


class Model:
    kind = "model"
    title = "Modell"

    def attach(self, sim) -> None:
        self.sim = sim

    def pre_momentum(self, sim) -> None:
        pass

    def momentum_source(self, sim, hu, hv) -> None:
        pass

    def post_step(self, sim, dt) -> None:
        pass

    def dt_limit(self, sim):
        return None

    def monitors(self, sim) -> dict:
        return {}

    def record_monitors(self, sim, rec) -> dict:
        return {}

    def fields(self, sim) -> dict:
        return {}

    def field_keys(self) -> list[str]:
        """The names fields() returns, without reading the data."""
        return []

    def state(self) -> dict:
        return {}

    def set_state(self, sim, state: dict) -> None:
        pass

    def describe(self) -> str:
        return self.title


class PassiveScalar(Model):
    """Passive scalar (models.py:293-333)."""

    kind = "scalar"

    def __init__(self, name: str = "c", diffusivity: float = 1e-3, bcs: dict | None = None,
                 obstacle: ScalarBC | None = None, limiter: str = "vanleer", initial=0.0,
                 source=None, turbulent_schmidt: float = 0.7, label: str | None = None):
        self.field = ScalarField(name, diffusivity, bcs, obstacle, limiter, initial, source,
                                 label=label or f"Skalar {name}")
        self.title = f"Passiver Skalar '{name}' (D = {diffusivity:g}, {limiter})"
        self.sct = turbulent_schmidt
        self.name = name

    def attach(self, sim) -> None:
        super().attach(sim)
        self.field.attach(sim)
        sim.scalars[self.name] = self.field

    def dt_limit(self, sim):
        lim = self.field.dt_limit()
        return lim if np.isfinite(lim) else None

    def monitors(self, sim) -> dict:
        return {f"{self.name}_integral": self.field.total()}

    def record_monitors(self, sim, rec) -> dict:
        s, _ = self.field._sums(rec)
        g = sim.grid
        return {f"{self.name}_integral": float(s * g.dx * g.dy)}

    def fields(self, sim) -> dict:
        return {self.name: (self.field.label, self.field.cell())}

    def field_keys(self) -> list[str]:
        return [self.name]


class BoussinesqTemperature(Model):
    """Temperature with Boussinesq buoyancy ``f = -beta (T - T_ref) g``
    (models.py:339-426)."""

    kind = "thermal"

    def __init__(self, alpha: float, bcs: dict | None = None, gravity=(0.0, -1.0),
                 T_ref: float = 0.0, initial=0.0, obstacle: ScalarBC | None = None,
                 limiter: str = "vanleer", prandtl_turbulent: float = 0.85,
                 buoyancy: bool = True):
        self.field = ScalarField("T", alpha, bcs, obstacle, limiter, initial, label="Temperatur T")
        self.alpha = float(alpha)
        self.gravity = (float(gravity[0]), float(gravity[1]))
        self.T_ref = float(T_ref)
        self.prt = prandtl_turbulent
        self.buoyancy = buoyancy
        if buoyancy:
            self.field.boussinesq = (self.gravity[0], self.gravity[1], self.T_ref)
        self.title = (f"Temperatur (alpha = {alpha:.4g}"
                      + (f", Auftrieb beta*g = {self.gravity}" if buoyancy else ", ohne Auftrieb")
                      + ")")

    @property
    def T(self) -> np.ndarray:
        return self.field.phi

    def attach(self, sim) -> None:
        super().attach(sim)
        self.field.attach(sim)
        sim.scalars["T"] = self.field

    def dt_limit(self, sim):
        lim = self.field.dt_limit()
        return lim if np.isfinite(lim) else None

    def nusselt(self, side: str, L: float = 1.0, dT: float = 1.0) -> float | None:
        if not np.any(self.field.bt[side] == 1):
            return None
        q = self.field.wall_flux(side)
        return abs(float(np.mean(q))) * L / (self.alpha * dT) if q.size else None

    def local_nusselt(self, side: str, L: float = 1.0, dT: float = 1.0) -> np.ndarray:
        return np.abs(self.field.wall_flux(side)) * L / (self.alpha * dT)

    def monitors(self, sim) -> dict:
        out = {"T_mittel": float(np.nanmean(self.field.cell()))}
        for side in SIDES:
            nu = self.nusselt(side)
            if nu is not None:
                out[f"Nu_{side}"] = nu
        return out

    def record_monitors(self, sim, rec) -> dict:
        s, q = self.field._sums(rec)
        out = {"T_mittel": float(s / sim._n_fluid_cells)}
        for k, side in enumerate(SIDES):
            if np.any(self.field.bt[side] == 1):
                n = len(self.field.bt[side])
                out[f"Nu_{side}"] = abs(float(q[k] / n)) * 1.0 / (self.alpha * 1.0)
        return out

    def fields(self, sim) -> dict:
        return {"T": ("Temperatur T", self.field.cell())}

    def field_keys(self) -> list[str]:
        return ["T"]
