"""Staggered MAC grid and obstacle mask (same interface as cfd2d/grid.py:26-152).

Index convention (Griebel, Dornseifer & Neunhoeffer): cells (i, j) with a ghost
layer run over 0..nx+1 and 0..ny+1; p sits at cell centres, u[i, j] on the
face between cells i and i+1 (shape (nx+1, ny+2)), v[i, j] on the face between
cells j and j+1 (shape (nx+2, ny+1)).  ``fluid`` is True for fluid cells; the
ghost ring is False except for the mirrored columns of an x-periodic grid.

This is host-side setup: the mask is built here and uploaded once to the
device context of a :class:`~cfd2d_amd.solver.Simulation`.
"""

from __future__ import annotations

import numpy as np
from scipy import ndimage


class Grid:
    def __init__(self, nx: int, ny: int, lx: float, ly: float):
        if nx < 4 or ny < 4:
            raise ValueError("Das Gitter braucht mindestens 4 Zellen pro Richtung.")
        if lx <= 0 or ly <= 0:
            raise ValueError("Gebietslängen müssen positiv sein.")
        self.nx, self.ny = int(nx), int(ny)
        self.lx, self.ly = float(lx), float(ly)
        self.dx = self.lx / self.nx
        self.dy = self.ly / self.ny
        self.fluid = np.zeros((self.nx + 2, self.ny + 2), dtype=np.bool_)
        self.fluid[1:-1, 1:-1] = True
        self.periodic_x = False

    # coordinates -------------------------------------------------------------
    @property
    def xc(self) -> np.ndarray:
        return (np.arange(self.nx) + 0.5) * self.dx

    @property
    def yc(self) -> np.ndarray:
        return (np.arange(self.ny) + 0.5) * self.dy

    @property
    def xn(self) -> np.ndarray:
        return np.arange(self.nx + 1) * self.dx

    @property
    def yn(self) -> np.ndarray:
        return np.arange(self.ny + 1) * self.dy

    def cell_centers(self) -> tuple[np.ndarray, np.ndarray]:
        return np.meshgrid(self.xc, self.yc, indexing="ij")

    # obstacles ---------------------------------------------------------------
    def interior_solid_padded(self) -> np.ndarray:
        s = ~self.fluid.copy()
        s[0, :] = s[-1, :] = s[:, 0] = s[:, -1] = False
        return s

    @property
    def solid(self) -> np.ndarray:
        return ~self.fluid[1:-1, 1:-1]

    @property
    def fluid_interior(self) -> np.ndarray:
        return self.fluid[1:-1, 1:-1]

    def add_obstacle(self, mask: np.ndarray) -> None:
        mask = np.asarray(mask, dtype=bool)
        if mask.shape != (self.nx, self.ny):
            raise ValueError(
                f"Hindernismaske hat Form {mask.shape}, erwartet {(self.nx, self.ny)}")
        self.fluid[1:-1, 1:-1] &= ~mask

    def finalize(self) -> int:
        """Make fluid pockets solid: interior fluid cells with no path of
        face neighbours through fluid to a cell of the interior's outer ring.
        A single closed fluid region is left as it is.  Returns the number of
        cells made solid (behaviour of grid.py:100-125)."""
        inner = self.fluid[1:-1, 1:-1]
        ring = np.ones_like(inner)
        ring[1:-1, 1:-1] = False
        # flood fill from the ring's fluid cells over face neighbours
        reached = ndimage.binary_propagation(ring & inner, mask=inner)
        pocket = inner & ~reached
        if not pocket.any():
            return 0
        if not reached.any() and ndimage.label(inner)[1] <= 1:
            return 0
        self.fluid[1:-1, 1:-1] = reached
        return int(pocket.sum())

    def set_periodic_x(self) -> None:
        self.periodic_x = True
        self.fluid[0, :] = self.fluid[self.nx, :]
        self.fluid[self.nx + 1, :] = self.fluid[1, :]
        self.fluid[:, 0] = False
        self.fluid[:, -1] = False

    # face classification -----------------------------------------------------
    def u_fluid_faces(self) -> np.ndarray:
        return self.fluid[:-1, :] & self.fluid[1:, :]

    def v_fluid_faces(self) -> np.ndarray:
        return self.fluid[:, :-1] & self.fluid[:, 1:]

    @property
    def n_fluid(self) -> int:
        return int(self.fluid.sum())

    def __repr__(self) -> str:  # pragma: no cover
        return (f"Grid(nx={self.nx}, ny={self.ny}, lx={self.lx}, ly={self.ly}, "
                f"dx={self.dx:.4g}, dy={self.dy:.4g}, Fluidzellen={self.n_fluid})")
