"""CPU FP64 oracle of the two-phase VOF model (cfd2d/vof.py).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker, never by the
product package.  ``cfd_oracle.vof_advect`` (C) restates the transport
kernel (vof.py:34-133); this module restates the model around it:

=============================  ==============================================
``OVOF.fill_ghosts``           vof.py:182-203
``OVOF.update_properties``     vof.py:205-210 (rho and the dynamic viscosity)
``OVOF.momentum_source``       vof.py:217-241 (CSF surface tension)
``OVOF.post_step``             vof.py:243-254 (sub-stepped transport)
``OVOF.dt_limit``              vof.py:256-268
``OVOF.monitors``              vof.py:270-281
``dam_break``                  cases.py:294-320
=============================  ==============================================

Pinned by ``tests/golden/vof.npz`` (written by the reference).
"""

from __future__ import annotations

import numpy as np

from .cfd_oracle import OBC, OGrid, OSim, vof_advect


class OVOF:
    kind = "multiphase"

    def __init__(self, alpha0, rho1=1000.0, rho2=1.0, nu1=1e-6, nu2=1.48e-5, sigma=0.07,
                 gravity=(0.0, -9.81), c_alpha=1.0, cfl_alpha=0.25, alpha_inflow=0.0):
        self.alpha0 = np.asarray(alpha0, dtype=float)
        self.rho1, self.rho2 = float(rho1), float(rho2)
        self.mu1, self.mu2 = float(nu1 * rho1), float(nu2 * rho2)
        self.sigma = float(sigma)
        self.gravity = (float(gravity[0]), float(gravity[1]))
        self.c_alpha, self.cfl, self.alpha_inflow = float(c_alpha), float(cfl_alpha), float(alpha_inflow)

    def attach(self, sim):
        self.sim = sim
        g = sim.grid
        self.alpha = np.zeros((g.nx + 2, g.ny + 2))
        self.alpha[1:-1, 1:-1] = np.where(g.interior, np.clip(self.alpha0, 0, 1), 0.0)
        self._tmp = np.zeros_like(self.alpha)
        sim.variable_density = True
        sim.variable_viscosity = True
        sim.body_force = (sim.body_force[0] + self.gravity[0], sim.body_force[1] + self.gravity[1])
        self.update_properties()
        self.volume0 = self.volume()

    def fill_ghosts(self):
        a, sim = self.alpha, self.sim
        g = sim.grid
        nx = g.nx
        if sim.periodic:
            a[0, :] = a[nx, :]
            a[nx + 1, :] = a[1, :]
        else:
            a[0, :] = a[1, :]
            a[-1, :] = a[-2, :]
        a[:, 0] = a[:, 1]
        a[:, -1] = a[:, -2]
        solid = ~g.fluid.copy()
        solid[0, :] = solid[-1, :] = solid[:, 0] = solid[:, -1] = False
        if solid.any():
            fl = g.fluid.astype(float)
            af = np.where(g.fluid, a, 0.0)
            s = np.zeros_like(a)
            c = np.zeros_like(a)
            s[1:-1, 1:-1] = af[2:, 1:-1] + af[:-2, 1:-1] + af[1:-1, 2:] + af[1:-1, :-2]
            c[1:-1, 1:-1] = fl[2:, 1:-1] + fl[:-2, 1:-1] + fl[1:-1, 2:] + fl[1:-1, :-2]
            a[solid] = np.where(c[solid] > 0, s[solid] / np.maximum(c[solid], 1), 0.0)

    def update_properties(self):
        self.fill_ghosts()
        a, sim = self.alpha, self.sim
        sim.rho[:] = a * self.rho1 + (1 - a) * self.rho2
        sim.nu_eff[:] = a * self.mu1 + (1 - a) * self.mu2

    def volume(self):
        g = self.sim.grid
        return float(self.alpha[1:-1, 1:-1][g.interior].sum() * g.dx * g.dy)

    def pre_momentum(self, sim):
        pass

    def momentum_source(self, sim, hu, hv):
        if self.sigma <= 0:
            return
        g, a = sim.grid, self.alpha
        As = a.copy()
        As[1:-1, 1:-1] = (4 * a[1:-1, 1:-1] + a[2:, 1:-1] + a[:-2, 1:-1] + a[1:-1, 2:]
                          + a[1:-1, :-2]) / 8.0
        gx = 0.5 * ((As[1:, 1:] - As[:-1, 1:]) + (As[1:, :-1] - As[:-1, :-1])) / g.dx
        gy = 0.5 * ((As[1:, 1:] - As[1:, :-1]) + (As[:-1, 1:] - As[:-1, :-1])) / g.dy
        gm = np.sqrt(gx ** 2 + gy ** 2) + 1e-8 / g.dx
        nxn, nyn = gx / gm, gy / gm
        kappa = np.zeros_like(a)
        kappa[1:-1, 1:-1] = -(0.5 * ((nxn[1:, 1:] - nxn[:-1, 1:]) + (nxn[1:, :-1] - nxn[:-1, :-1])) / g.dx
                              + 0.5 * ((nyn[1:, 1:] - nyn[1:, :-1]) + (nyn[:-1, 1:] - nyn[:-1, :-1])) / g.dy)
        rho = sim.rho
        fxs = self.sigma * 0.5 * (kappa[:-1, :] + kappa[1:, :]) * (a[1:, :] - a[:-1, :]) / g.dx
        fys = self.sigma * 0.5 * (kappa[:, :-1] + kappa[:, 1:]) * (a[:, 1:] - a[:, :-1]) / g.dy
        rx = 0.5 * (rho[:-1, :] + rho[1:, :])
        ry = 0.5 * (rho[:, :-1] + rho[:, 1:])
        hu[sim.u_fluid] += (fxs / rx)[sim.u_fluid]
        hv[sim.v_fluid] += (fys / ry)[sim.v_fluid]

    def post_step(self, sim, dt):
        g = sim.grid
        umax = max(np.abs(sim.u).max() / g.dx, np.abs(sim.v).max() / g.dy, 1e-12)
        nsub = max(1, int(np.ceil(dt * umax / self.cfl)))
        h = dt / nsub
        for _ in range(nsub):
            vof_advect(self.alpha, sim.u, sim.v, g.fluid, g.nx, g.ny, g.dx, g.dy, h,
                       self.c_alpha, sim.periodic, self.alpha_inflow, self._tmp)
            self.alpha[:] = self._tmp
            self.fill_ghosts()
        self.update_properties()

    def dt_limit(self, sim):
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

    def monitors(self, sim):
        g = sim.grid
        a = self.alpha[1:-1, 1:-1]
        out = {"volumen_phase1": self.volume(),
               "volumenfehler_rel": self.volume() / max(self.volume0, 1e-300) - 1.0}
        row = np.where(a[:, 0] > 0.5)[0]
        out["front_x"] = float(g.xc[row.max()] + 0.5 * g.dx) if row.size else 0.0
        col = np.where(a[0, :] > 0.5)[0]
        out["hoehe_links"] = float(g.yc[col.max()] + 0.5 * g.dy) if col.size else 0.0
        return out


def dam_break(n=96, size=0.584, column_width=0.1461, column_height=0.292, obstacle=True,
              sigma=0.07, open_top=True, **num) -> OSim:
    """Water column against the left wall of a square tank (cases.py:294-320)."""
    g = OGrid(n, n, size, size)
    X, Y = g.cell_centers()
    if obstacle:
        x0 = 0.5 * size - 0.012
        g.add_obstacle((X >= x0) & (X <= x0 + 0.024) & (Y >= 0.0) & (Y <= 0.048))
    alpha0 = ((X <= column_width) & (Y <= column_height)).astype(float)
    bcs = {"left": OBC("noslip"), "right": OBC("noslip"), "bottom": OBC("noslip"),
           "top": OBC("outflow") if open_top else OBC("noslip")}
    num.setdefault("gamma", 0.5)
    return OSim(g, bcs, nu=1e-6, models=[OVOF(alpha0, sigma=sigma)], **num)
