"""CPU FP64 oracle of the turbulence models on the projection step.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker, never by the
product package.  Restates, in NumPy on host arrays, the reference's
variable-viscosity path (cfd2d/turbulence.py) on top of ``cfd_oracle.OSim``:

=============================  ==============================================
``wall_mask``                  turbulence.py:27-39
``wall_distance``              turbulence.py:42-69 (SciPy's exact EDT)
``wall_adjacent``              turbulence.py:72-85
``strain_rate_squared``        turbulence.py:88-99
``centred_gradient``           turbulence.py:102-106
``nut_ghosts``                 turbulence.py:109-124
``OSmagorinsky``               turbulence.py:130-162
``OSST``                       turbulence.py:168-325 (Menter 2003, low-Re)
``backward_facing_step``       cases.py:133-155
``turbulent_channel``          cases.py:258-288
``add_turbulence``             cases.py:326-337
=============================  ==============================================

Pinned by ``tests/golden/turbulence.npz`` (written by the reference):
the stress-form momentum kernel, the Smagorinsky cavity and the SST step.
"""

from __future__ import annotations

import numpy as np
from scipy import ndimage

from .cfd_oracle import OBC, OGrid, OScalar, OSim, SIDES

KAPPA = 0.41


def wall_mask(sim) -> np.ndarray:
    """Cells that count as wall, with the ghost ring: interior solids plus the
    ghost row/column of every no-slip side."""
    g = sim.grid
    wall = ~g.fluid.copy()
    wall[0, :] = wall[-1, :] = wall[:, 0] = wall[:, -1] = False
    for side, sl in (("left", (0, slice(None))), ("right", (-1, slice(None))),
                     ("bottom", (slice(None), 0)), ("top", (slice(None), -1))):
        if sim.bcs[side].kind == "noslip":
            wall[sl] = True
    return wall


def wall_distance(sim) -> np.ndarray:
    """Distance of the cell centres to the nearest wall face, (nx+2, ny+2):
    exact Euclidean distance transform to the wall cells' centres minus half
    a cell, at least half a cell; zero off the fluid.  Periodic in x: the
    field is tiled three times before the transform."""
    g = sim.grid
    wall = wall_mask(sim)
    if not wall.any():
        return np.full(wall.shape, 1e6)
    if sim.periodic:
        core = wall[1:-1, :]
        n = core.shape[0]
        d = ndimage.distance_transform_edt(~np.concatenate([core] * 3, axis=0),
                                           sampling=(g.dx, g.dy))[n:2 * n]
        dist = np.empty(wall.shape)
        dist[1:-1] = d
        dist[0] = d[-1]
        dist[-1] = d[0]
    else:
        dist = ndimage.distance_transform_edt(~wall, sampling=(g.dx, g.dy))
    h = 0.5 * min(g.dx, g.dy)
    dist = np.maximum(dist - h, h)
    dist[~g.fluid] = 0.0
    return dist


def wall_adjacent(sim):
    """Fluid cells with a wall neighbour and their distance to it (half the
    cell width across that face; the smaller one for several)."""
    g = sim.grid
    wall, fl = wall_mask(sim), g.fluid
    adj = np.zeros_like(fl)
    dist = np.full(fl.shape, np.inf)
    c = (slice(1, -1), slice(1, -1))
    for nb, h in (((slice(2, None), slice(1, -1)), g.dx), ((slice(0, -2), slice(1, -1)), g.dx),
                  ((slice(1, -1), slice(2, None)), g.dy), ((slice(1, -1), slice(0, -2)), g.dy)):
        hit = wall[nb] & fl[c]
        adj[c] |= hit
        dist[c] = np.where(hit, np.minimum(dist[c], 0.5 * h), dist[c])
    return adj, dist


def strain_rate_squared(sim) -> np.ndarray:
    """S^2 = 2 S_ij S_ij at the cell centres, (nx, ny); the shear part is
    averaged from the four nodes of the cell."""
    g = sim.grid
    u, v = sim.u, sim.v
    dudx = (u[1:, 1:-1] - u[:-1, 1:-1]) / g.dx
    dvdy = (v[1:-1, 1:] - v[1:-1, :-1]) / g.dy
    shear = (u[:, 1:] - u[:, :-1]) / g.dy + (v[1:, :] - v[:-1, :]) / g.dx
    sh2 = 0.25 * (shear[:-1, :-1] ** 2 + shear[1:, :-1] ** 2 + shear[:-1, 1:] ** 2
                  + shear[1:, 1:] ** 2)
    return 2.0 * (dudx ** 2 + dvdy ** 2) + sh2


def centred_gradient(phi, dx, dy):
    return ((phi[2:, 1:-1] - phi[:-2, 1:-1]) / (2 * dx),
            (phi[1:-1, 2:] - phi[1:-1, :-2]) / (2 * dy))


def nut_ghosts(sim, nut) -> None:
    """Ghost cells of nu_t: 0 next to no-slip walls and in solids, a copy of
    the inner cell otherwise (periodic: the wrapped cell)."""
    g = sim.grid
    nx, ny = g.nx, g.ny
    nut[~g.fluid] = 0.0
    if sim.periodic:
        nut[0, :] = nut[nx, :]
        nut[nx + 1, :] = nut[1, :]
    else:
        nut[0, :] = 0.0 if sim.bcs["left"].kind == "noslip" else nut[1, :]
        nut[nx + 1, :] = 0.0 if sim.bcs["right"].kind == "noslip" else nut[nx, :]
    nut[:, 0] = 0.0 if sim.bcs["bottom"].kind == "noslip" else nut[:, 1]
    nut[:, ny + 1] = 0.0 if sim.bcs["top"].kind == "noslip" else nut[:, ny]
    solid = ~g.fluid.copy()
    solid[0, :] = solid[-1, :] = solid[:, 0] = solid[:, -1] = False
    nut[solid] = 0.0


class OSmagorinsky:
    kind = "turbulence"

    def __init__(self, cs=0.17, wall_damping=True):
        self.cs, self.wall_damping = float(cs), wall_damping

    def attach(self, sim):
        sim.variable_viscosity = True
        g = sim.grid
        l = np.full((g.nx, g.ny), self.cs * np.sqrt(g.dx * g.dy))
        if self.wall_damping:
            l = np.minimum(l, KAPPA * wall_distance(sim)[1:-1, 1:-1])
        self.l2 = l ** 2
        self.pre_momentum(sim)

    def pre_momentum(self, sim):
        S = np.sqrt(strain_rate_squared(sim))
        sim.nu_t[1:-1, 1:-1] = np.where(sim.grid.interior, self.l2 * S, 0.0)
        nut_ghosts(sim, sim.nu_t)

    def momentum_source(self, sim, hu, hv):
        pass

    def post_step(self, sim, dt):
        pass

    def dt_limit(self, sim):
        return None

    def monitors(self, sim):
        return {"nu_t_max": float(sim.nu_t.max())}


class OSST:
    """k-omega SST (Menter, Kuntz & Langtry 2003), low-Re form."""

    kind = "turbulence"
    a1, beta_star = 0.31, 0.09
    sk1, so1, beta1, gamma1 = 0.85, 0.5, 0.075, 5.0 / 9.0
    sk2, so2, beta2, gamma2 = 1.0, 0.856, 0.0828, 0.44

    def __init__(self, intensity=0.05, viscosity_ratio=10.0, U_ref=1.0, limiter="upwind"):
        self.I, self.ratio, self.U_ref, self.limiter = (float(intensity), float(viscosity_ratio),
                                                        float(U_ref), limiter)

    def attach(self, sim):
        sim.variable_viscosity = True
        nu = sim.nu
        k0 = 1.5 * (self.I * self.U_ref) ** 2
        w0 = k0 / (nu * self.ratio)
        kb, wb = {}, {}
        for s in SIDES:
            kind = sim.bcs[s].kind
            kb[s] = (("fixed", k0) if kind == "inflow" else
                     ("fixed", 0.0) if kind == "noslip" else ("zerogradient", 0.0))
            wb[s] = ("fixed", w0) if kind == "inflow" else ("zerogradient", 0.0)
        self.k = OScalar("k", nu, kb, ("fixed", 0.0), self.limiter, k0,
                         source=self._k_source, lower=1e-14)
        self.w = OScalar("omega", nu, wb, ("zerogradient", 0.0), self.limiter, w0,
                         source=self._w_source, lower=1e-8)
        self.k.attach(sim)
        self.w.attach(sim)
        sim.scalars["k"], sim.scalars["omega"] = self.k, self.w
        self.d = np.maximum(wall_distance(sim)[1:-1, 1:-1], 1e-12)
        adj, dist = wall_adjacent(sim)
        self.wall_adj = adj[1:-1, 1:-1]
        di = dist[1:-1, 1:-1]
        self.w_wall = np.where(self.wall_adj,
                               6.0 * nu / (self.beta1 * np.where(np.isfinite(di), di, 1.0) ** 2),
                               0.0)
        self._wall_omega()
        self.pre_momentum(sim)

    def _wall_omega(self):
        inner = self.w.phi[1:-1, 1:-1]
        inner[self.wall_adj] = self.w_wall[self.wall_adj]
        self.w.fill_ghosts()

    def _kw(self):
        return (np.maximum(self.k.phi[1:-1, 1:-1], 0.0),
                np.maximum(self.w.phi[1:-1, 1:-1], 1e-12))

    def pre_momentum(self, sim):
        g, nu = sim.grid, sim.nu
        k, w = self._kw()
        self.S2 = strain_rate_squared(sim)
        d = self.d
        kx, ky = centred_gradient(self.k.phi, g.dx, g.dy)
        wx, wy = centred_gradient(self.w.phi, g.dx, g.dy)
        self.cross = kx * wx + ky * wy
        CD = np.maximum(2.0 * self.so2 / w * self.cross, 1e-10)
        sk = np.sqrt(k)
        visc = 500 * nu / (d ** 2 * w)
        arg1 = np.minimum(np.maximum(sk / (self.beta_star * w * d), visc),
                          4 * self.so2 * k / (CD * d ** 2))
        self.F1 = np.tanh(arg1 ** 4)
        arg2 = np.maximum(2 * sk / (self.beta_star * w * d), visc)
        F2 = np.tanh(arg2 ** 2)
        nut = self.a1 * k / np.maximum(self.a1 * w, np.sqrt(self.S2) * F2)
        sim.nu_t[1:-1, 1:-1] = np.where(g.interior, nut, 0.0)
        nut_ghosts(sim, sim.nu_t)
        F1 = self.F1
        for f, s1, s2 in ((self.k, self.sk1, self.sk2), (self.w, self.so1, self.so2)):
            f.diff[1:-1, 1:-1] = nu + (F1 * s1 + (1 - F1) * s2) * nut
            f.diff[0, :] = f.diff[1, :]
            f.diff[-1, :] = f.diff[-2, :]
            f.diff[:, 0] = f.diff[:, 1]
            f.diff[:, -1] = f.diff[:, -2]

    def _k_source(self, sim, field):
        k, w = self._kw()
        self.Pk = np.minimum(sim.nu_t[1:-1, 1:-1] * self.S2, 10.0 * self.beta_star * k * w)
        return self.Pk, self.beta_star * w

    def _w_source(self, sim, field):
        F1 = self.F1
        w = np.maximum(self.w.phi[1:-1, 1:-1], 1e-12)
        gam = F1 * self.gamma1 + (1 - F1) * self.gamma2
        beta = F1 * self.beta1 + (1 - F1) * self.beta2
        nut_in = sim.nu_t[1:-1, 1:-1]
        Pw = np.where(nut_in > 1e-14, gam * self.Pk / np.maximum(nut_in, 1e-14), gam * self.S2)
        CDw = 2.0 * (1 - F1) * self.so2 / w * self.cross
        return Pw + np.maximum(CDw, 0.0), beta * w + np.maximum(-CDw, 0.0) / w

    def momentum_source(self, sim, hu, hv):
        pass

    def post_step(self, sim, dt):
        self.S2 = strain_rate_squared(sim)
        self.k.step(dt)
        self.w.step(dt)
        self._wall_omega()

    def dt_limit(self, sim):
        return min(self.k.dt_limit(), self.w.dt_limit())

    def monitors(self, sim):
        return {"k_mittel": float(np.nanmean(self.k.cell())), "nu_t_max": float(sim.nu_t.max())}


def backward_facing_step(re=400.0, length=15.0, ny=40, nx=None, inlet_length=1.0,
                         **num) -> OSim:
    """Channel of height 1 behind a step of height 1/2, parabolic inflow above
    the step (cases.py:133-155)."""
    H, h = 1.0, 0.5
    if nx is None:
        nx = int(round(length / H * ny))
    g = OGrid(nx, ny, length, H)
    if inlet_length > 0:
        X, Y = g.cell_centers()
        g.add_obstacle((X >= 0.0) & (X <= inlet_length) & (Y >= 0.0) & (Y <= h))
    bcs = {"left": OBC("inflow", 1.0, "parabolic"), "right": OBC("outflow"),
           "bottom": OBC("noslip"), "top": OBC("noslip")}
    return OSim(g, bcs, nu=1.0 * H / re, **num)


def dean_cf(re_b):
    """Dean (1978): C_f = 0.073 Re_b^(-1/4) (cases.py:258-260)."""
    return 0.073 * re_b ** -0.25


def turbulent_channel(re_tau=180.0, ny=None, y_plus_1=1.0, model="kOmegaSST", **num):
    """Periodic channel of half height 1 at bulk velocity 1, Re_b from Dean's
    Re_tau = 0.09 Re_b^0.88, ny from the first cell's y+ (cases.py:263-288).
    Returns (sim, re_b)."""
    re_b = (re_tau / 0.09) ** (1 / 0.88)
    if ny is None:
        ny = int(np.ceil(re_tau / y_plus_1))
        ny += ny % 2
    g = OGrid(4, ny, 0.5, 2.0)
    bcs = {"left": OBC("periodic"), "right": OBC("periodic"), "bottom": OBC("noslip"),
           "top": OBC("noslip")}
    models = ([OSST(intensity=0.05, viscosity_ratio=50.0)] if model == "kOmegaSST" else
              [OSmagorinsky()] if model == "Smagorinsky" else [])
    sim = OSim(g, bcs, nu=2.0 / re_b, flow_rate=1.0, initial_velocity=(1.0, 0.0), models=models,
               **num)
    return sim, re_b


def add_turbulence(sim: OSim, model="kOmegaSST", U_ref=1.0, **kw) -> OSim:
    m = OSST(U_ref=U_ref, **kw) if model == "kOmegaSST" else OSmagorinsky(**kw)
    sim.add_model(m)
    return sim
