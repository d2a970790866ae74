"""Staircase obstacle masks (the subset of cfd2d/geometry.py used by the
obstacle-channel workloads: circle, ellipse, rectangle; geometry.py:17-42,
166-195).  A mask is True where the cell centre lies inside the body."""

from __future__ import annotations

import numpy as np

from .grid import Grid


def circle(grid: Grid, xc: float, yc: float, radius: float) -> np.ndarray:
    if radius <= 0:
        raise ValueError("Radius muss positiv sein.")
    X, Y = grid.cell_centers()
    return (X - xc) ** 2 + (Y - yc) ** 2 <= radius ** 2


def ellipse(grid: Grid, xc: float, yc: float, a: float, b: float,
            angle_deg: float = 0.0) -> np.ndarray:
    X, Y = grid.cell_centers()
    t = np.deg2rad(angle_deg)
    xr = (X - xc) * np.cos(t) + (Y - yc) * np.sin(t)
    yr = -(X - xc) * np.sin(t) + (Y - yc) * np.cos(t)
    return (xr / a) ** 2 + (yr / b) ** 2 <= 1.0


def rectangle(grid: Grid, x0: float, y0: float, x1: float, y1: float) -> np.ndarray:
    xa, xb = sorted((x0, x1))
    ya, yb = sorted((y0, y1))
    X, Y = grid.cell_centers()
    return (X >= xa) & (X <= xb) & (Y >= ya) & (Y <= yb)


def build_obstacle(grid: Grid, spec: dict) -> np.ndarray:
    """Mask from a description dict with key ``typ``: ``kreis`` (x, y, r),
    ``ellipse`` (x, y, a, b, [winkel]) or ``rechteck`` (x0, y0, x1, y1)."""
    typ = spec.get("typ", "").lower()
    if typ == "kreis":
        return circle(grid, spec["x"], spec["y"], spec["r"])
    if typ == "ellipse":
        return ellipse(grid, spec["x"], spec["y"], spec["a"], spec["b"], spec.get("winkel", 0.0))
    if typ == "rechteck":
        return rectangle(grid, spec["x0"], spec["y0"], spec["x1"], spec["y1"])
    raise ValueError(f"Unbekannter oder nicht unterstützter Hindernistyp: {typ!r}")
