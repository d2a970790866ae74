"""Domain boundary conditions (interface of cfd2d/boundary.py:28-245).

The host part computes what the device needs once: which boundary faces are
open (adjacent inner cell is fluid), the inflow profile per open run of faces,
and the outflow (pressure-Dirichlet) masks.  Applying the conditions to u and
v happens on the device (``cfd_apply_bc``), inside every step.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .grid import Grid

SIDES = ("left", "right", "bottom", "top")
SIDE_NAMES_DE = {"left": "links", "right": "rechts", "bottom": "unten", "top": "oben"}
KINDS = ("noslip", "slip", "inflow", "outflow", "periodic")
KIND_NAMES_DE = {"noslip": "Haftende Wand", "slip": "Reibungsfreie Wand",
                 "inflow": "Einströmung", "outflow": "Ausströmung", "periodic": "Periodisch"}


@dataclass
class BoundaryCondition:
    kind: str = "noslip"
    velocity: float = 0.0
    profile: str = "uniform"

    def __post_init__(self):
        if self.kind not in KINDS:
            raise ValueError(f"Unbekannte Randbedingung {self.kind!r}, erlaubt: {KINDS}")
        if self.profile not in ("uniform", "parabolic"):
            raise ValueError("Einströmprofil muss 'uniform' oder 'parabolic' sein.")

    def describe(self) -> str:
        s = KIND_NAMES_DE[self.kind]
        if self.kind == "noslip" and self.velocity != 0.0:
            s += f" (bewegt, U={self.velocity:g})"
        if self.kind == "inflow":
            prof = "parabolisch" if self.profile == "parabolic" else "Blockprofil"
            s += f" ({prof}, U_mittel={self.velocity:g})"
        return s


def inflow_profile(coords: np.ndarray, open_mask: np.ndarray, h: float, mean_velocity: float,
                   profile: str) -> np.ndarray:
    """Normal inflow velocity per boundary face.  Each maximal run of open faces
    gets its own profile: constant, or 6 U (s - a)(b - s)/(b - a)^2 between the
    run's outer face edges a, b (boundary.py:63-89)."""
    out = np.zeros_like(coords)
    idx = np.flatnonzero(np.asarray(open_mask, dtype=bool))
    if idx.size == 0:
        return out
    breaks = np.flatnonzero(np.diff(idx) > 1)
    starts = np.concatenate(([idx[0]], idx[breaks + 1]))
    ends = np.concatenate((idx[breaks], [idx[-1]]))
    for k0, k1 in zip(starts, ends):
        a = coords[k0] - 0.5 * h
        b = coords[k1] + 0.5 * h
        s = coords[k0:k1 + 1]
        if profile == "parabolic":
            out[k0:k1 + 1] = 6.0 * mean_velocity * (s - a) * (b - s) / (b - a) ** 2
        else:
            out[k0:k1 + 1] = mean_velocity
    return out


class DomainBoundary:
    def __init__(self, grid: Grid, bcs: dict):
        missing = [s for s in SIDES if s not in bcs]
        if missing:
            raise ValueError(f"Randbedingungen fehlen für: {missing}")
        self.grid, self.bcs = grid, bcs
        pl = bcs["left"].kind == "periodic"
        pr = bcs["right"].kind == "periodic"
        if pl != pr:
            raise ValueError("Periodische Ränder müssen links UND rechts gesetzt werden.")
        if bcs["bottom"].kind == "periodic" or bcs["top"].kind == "periodic":
            raise ValueError("Periodische Ränder werden nur in x-Richtung (links/rechts) "
                             "unterstützt.")
        self.periodic = pl
        if self.periodic:
            grid.set_periodic_x()
        f, nx, ny = grid.fluid, grid.nx, grid.ny
        self.open = {"left": f[1, 1:ny + 1].copy(), "right": f[nx, 1:ny + 1].copy(),
                     "bottom": f[1:nx + 1, 1].copy(), "top": f[1:nx + 1, ny].copy()}
        self.inflow_values = {}
        for side, bc in bcs.items():
            if bc.kind != "inflow":
                continue
            along_y = side in ("left", "right")
            vals = inflow_profile(grid.yc if along_y else grid.xc, self.open[side],
                                  grid.dy if along_y else grid.dx, bc.velocity, bc.profile)
            self.inflow_values[side] = (1.0 if side in ("left", "bottom") else -1.0) * vals

    @property
    def has_outflow(self) -> bool:
        return any(bc.kind == "outflow" for bc in self.bcs.values())

    def outflow_mask(self, side: str) -> np.ndarray:
        if self.bcs[side].kind != "outflow":
            return np.zeros_like(self.open[side])
        return self.open[side]

    def describe(self) -> str:
        return "\n".join(f"  {SIDE_NAMES_DE[s]:>6}: {self.bcs[s].describe()}" for s in SIDES)
