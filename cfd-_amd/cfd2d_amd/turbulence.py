"""Turbulence models (interface of cfd2d/turbulence.py): the Smagorinsky LES
model and Menter's k-omega SST (2003, low-Re form).

The per-step work runs on the device (csrc/turb.hip): nu_t before the
time-step control, the stress-form momentum diffusion of nu + nu_t, and for
SST the transport of k and omega with their blended diffusivities, sources
and clips.  The host does the one-time set-up the reference does in
``attach``: wall distances (SciPy's exact Euclidean distance transform), the
mixing-length limit and the viscous omega of the wall cells.
"""

from __future__ import annotations

import ctypes

import numpy as np
from scipy import ndimage

from . import _lib
from .boundary import SIDES
from .models import Model, ScalarBC, ScalarField

KAPPA = 0.41


# --------------------------------------------------------------------------- #
# wall geometry (turbulence.py:27-85)
# --------------------------------------------------------------------------- #
def wall_mask_padded(sim) -> np.ndarray:
    """Cells counting as wall, (nx+2, ny+2): inner solids and the ghost
    row/column of every no-slip side."""
    wall = sim.grid.interior_solid_padded()
    edges = {"left": (0, slice(None)), "right": (-1, slice(None)),
             "bottom": (slice(None), 0), "top": (slice(None), -1)}
    for side, idx in edges.items():
        if sim.bcs[side].kind == "noslip":
            wall[idx] = True
    return wall


def wall_distance(sim) -> np.ndarray:
    """Distance from every cell centre to the nearest wall face, (nx+2, ny+2);
    half a cell at least, zero off the fluid.  Periodic in x: distances are
    taken on three copies of the domain side by side."""
    g = sim.grid
    wall = wall_mask_padded(sim)
    if not wall.any():
        return np.full(wall.shape, 1e6)
    if sim.periodic:
        core = wall[1:-1, :]
        n = core.shape[0]
        far = ndimage.distance_transform_edt(~np.concatenate([core, core, core], axis=0),
                                             sampling=(g.dx, g.dy))[n:2 * n]
        dist = np.empty(wall.shape)
        dist[1:-1] = far
        dist[0], dist[-1] = far[-1], far[0]
    else:
        dist = ndimage.distance_transform_edt(~wall, sampling=(g.dx, g.dy))
    half = 0.5 * min(g.dx, g.dy)
    dist = np.maximum(dist - half, half)
    dist[~g.fluid] = 0.0
    return dist


def wall_adjacent_cells(sim):
    """Fluid cells next to a wall and their distance to it (half the cell
    width across the face; the smaller one where there are several)."""
    g = sim.grid
    wall, fl = wall_mask_padded(sim), g.fluid
    adj = np.zeros_like(fl)
    dist = np.full(fl.shape, np.inf)
    c = (slice(1, -1), slice(1, -1))
    for nb, h in (((slice(2, None), slice(1, -1)), g.dx), ((slice(0, -2), slice(1, -1)), g.dx),
                  ((slice(1, -1), slice(2, None)), g.dy), ((slice(1, -1), slice(0, -2)), g.dy)):
        hit = wall[nb] & fl[c]
        adj[c] |= hit
        dist[c] = np.where(hit, np.minimum(dist[c], 0.5 * h), dist[c])
    return adj, dist


def _set_model(sim, model: int, **arrays) -> None:
    d = _lib.TurbDesc()
    d.model = model
    d.nu = sim.nu
    keep = {}
    for name in ("l2", "wall_distance", "omega_wall"):
        if name in arrays:
            keep[name] = np.ascontiguousarray(arrays[name], dtype=np.float64)
            setattr(d, name, _lib.ptr(keep[name]))
    if "wall_adjacent" in arrays:
        keep["wall_adjacent"] = np.ascontiguousarray(arrays["wall_adjacent"], dtype=np.uint8)
        d.wall_adjacent = _lib.ptr(keep["wall_adjacent"])
    d.sid_k = arrays.get("sid_k", -1)
    d.sid_omega = arrays.get("sid_omega", -1)
    sim._flush()
    sim._ctx.call("cfd_set_turbulence", ctypes.byref(d))
    sim.variable_viscosity = True


# --------------------------------------------------------------------------- #
class Smagorinsky(Model):
    """LES: nu_t = (C_s Delta)^2 |S|, the length limited to kappa d near walls
    (turbulence.py:130-162)."""

    kind = "turbulence"

    def __init__(self, cs: float = 0.17, wall_damping: bool = True):
        self.cs = float(cs)
        self.wall_damping = wall_damping
        self.title = (f"LES Smagorinsky (C_s = {cs:g}"
                      + (", Wanddämpfung" if wall_damping else "") + ")")

    def attach(self, sim) -> None:
        super().attach(sim)
        g = sim.grid
        length = np.full((g.nx, g.ny), self.cs * np.sqrt(g.dx * g.dy))
        if self.wall_damping:
            length = np.minimum(length, KAPPA * wall_distance(sim)[1:-1, 1:-1])
        self.l2 = length ** 2
        _set_model(sim, 1, l2=self.l2)

    def record_monitors(self, sim, rec) -> dict:
        return {"nu_t_max": float(rec.mon[_lib.REC_NUT_MAX])}

    def monitors(self, sim) -> dict:
        return {"nu_t_max": float(sim.nu_t.max())}

    def fields(self, sim) -> dict:
        return {"nu_t": ("Turbulente Viskosität ν_t",
                         np.where(sim.grid.solid, np.nan, sim.nu_t[1:-1, 1:-1]))}

    def field_keys(self) -> list[str]:
        return ["nu_t"]


class _SSTField(ScalarField):
    """k or omega: a device scalar whose diffusivity the SST model sets each
    step (read back from the device)."""

    def __init__(self, *args, diff_field: str, **kw):
        super().__init__(*args, **kw)
        self._diff_field = diff_field

    @property
    def diff(self) -> np.ndarray:
        if self.sim is None or not self.sim.variable_viscosity:
            return np.full((self.sim.grid.nx + 2, self.sim.grid.ny + 2), self.D0)
        self.sim._flush()
        return self.sim._ctx.download(self._diff_field)

    @diff.setter
    def diff(self, value) -> None:   # the constant field the base class sets up
        pass

    def dt_limit(self) -> float:
        g = self.sim.grid
        dmax = float(self.diff[g.fluid].max()) if g.fluid.any() else self.D0
        if dmax <= 0:
            return np.inf
        return 0.5 / (dmax * (1.0 / g.dx ** 2 + 1.0 / g.dy ** 2))


class KOmegaSST(Model):
    """k-omega SST (Menter, Kuntz & Langtry 2003), low-Re form
    (turbulence.py:168-325).  k = 0 at walls; omega of the wall cells is the
    viscous value 6 nu / (beta1 y^2); inflow k and omega from the turbulence
    intensity and the viscosity ratio; zero gradient elsewhere."""

    kind = "turbulence"
    a1 = 0.31
    beta_star = 0.09
    sk1, so1, beta1, gamma1 = 0.85, 0.5, 0.075, 5.0 / 9.0
    sk2, so2, beta2, gamma2 = 1.0, 0.856, 0.0828, 0.44
    device_dt_limit = True     # the step evaluates dt_limit on the device

    def __init__(self, intensity: float = 0.05, viscosity_ratio: float = 10.0,
                 U_ref: float = 1.0, limiter: str = "upwind"):
        self.I = float(intensity)
        self.ratio = float(viscosity_ratio)
        self.U_ref = float(U_ref)
        self.limiter = limiter
        self.title = (f"RANS k-ω SST (Menter 2003), I = {intensity:g}, "
                      f"ν_t/ν = {viscosity_ratio:g}")

    def attach(self, sim) -> None:
        super().attach(sim)
        nu = sim.nu
        k0 = 1.5 * (self.I * self.U_ref) ** 2
        w0 = k0 / (nu * self.ratio)
        self.k_in, self.w_in = k0, w0
        kb, wb = {}, {}
        for side in SIDES:
            kind = sim.bcs[side].kind
            if kind == "inflow":
                kb[side], wb[side] = ScalarBC("fixed", k0), ScalarBC("fixed", w0)
            elif kind == "noslip":
                kb[side], wb[side] = ScalarBC("fixed", 0.0), ScalarBC("zerogradient")
            else:
                kb[side], wb[side] = ScalarBC("zerogradient"), ScalarBC("zerogradient")
        g = sim.grid
        d = np.maximum(wall_distance(sim)[1:-1, 1:-1], 1e-12)
        adj, dist = wall_adjacent_cells(sim)
        self.d = d
        self.wall_adj = adj[1:-1, 1:-1]
        di = dist[1:-1, 1:-1]
        self.w_wall = np.where(self.wall_adj,
                               6.0 * nu / (self.beta1 * np.where(np.isfinite(di), di, 1.0) ** 2),
                               0.0)
        # omega starts at w0 with the wall value in the wall cells
        # (turbulence.py:242-246 right after the field's attach)
        w_init = np.full((g.nx + 2, g.ny + 2), w0)
        inner = w_init[1:-1, 1:-1]
        inner[self.wall_adj] = self.w_wall[self.wall_adj]
        self.k = _SSTField("k", nu, kb, ScalarBC("fixed", 0.0), self.limiter, k0,
                           label="Turbulente kin. Energie k", diff_field="diff_k")
        self.w = _SSTField("omega", nu, wb, ScalarBC("zerogradient"), self.limiter, w_init,
                           label="Spez. Dissipationsrate ω", diff_field="diff_omega")
        self.k.attach(sim)
        self.w.attach(sim)
        sim.scalars["k"] = self.k
        sim.scalars["omega"] = self.w
        _set_model(sim, 2, wall_distance=self.d, wall_adjacent=self.wall_adj,
                   omega_wall=self.w_wall, sid_k=self.k.sid, sid_omega=self.w.sid)

    def dt_limit(self, sim):
        return min(self.k.dt_limit(), self.w.dt_limit())

    def record_monitors(self, sim, rec) -> dict:
        s, _ = self.k._sums(rec)
        return {"k_mittel": float(s / sim._n_fluid_cells),
                "nu_t_max": float(rec.mon[_lib.REC_NUT_MAX])}

    def monitors(self, sim) -> dict:
        return {"k_mittel": float(np.nanmean(self.k.cell())), "nu_t_max": float(sim.nu_t.max())}

    def fields(self, sim) -> dict:
        solid = sim.grid.solid
        return {"k": (self.k.label, self.k.cell()),
                "omega": ("log10 ω (spez. Dissipationsrate)",
                          np.log10(np.maximum(self.w.cell(), 1e-30))),
                "nu_t": ("Turbulente Viskosität ν_t",
                         np.where(solid, np.nan, sim.nu_t[1:-1, 1:-1]))}

    def field_keys(self) -> list[str]:
        return ["k", "omega", "nu_t"]


TURBULENCE_MODELS = {
    "laminar": None,
    "kOmegaSST": KOmegaSST,
    "Smagorinsky": Smagorinsky,
}
