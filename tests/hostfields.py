"""The reference's derived fields restated with numpy on host copies of the
velocity (cfd2d/solver.py:526-537, 569-596), the checker for the device
kernels that compute them."""

import numpy as np


def node_vorticity(grid, u, v):
    g = grid
    return (v[1:, :] - v[:-1, :]) / g.dx - (u[:, 1:] - u[:, :-1]) / g.dy


def streamfunction(grid, u, v):
    g = grid
    fl = g.fluid
    u_open = fl[:-1, 1:-1] | fl[1:, 1:-1]
    v_open = fl[1:-1, 0] | fl[1:-1, 1]
    uf = np.where(u_open, u[:, 1:-1], 0.0)
    v0 = np.where(v_open, v[1:-1, 0], 0.0)
    psi = np.zeros((g.nx + 1, g.ny + 1))
    psi[1:, 0] = -np.cumsum(v0) * g.dx
    psi[:, 1:] = psi[:, :1] + np.cumsum(uf, axis=1) * g.dy
    return psi
