"""The reference's physics acceptance tests (tests/test_cfd2d.py:57-94) on the
GPU path, for the reference's solver ("direct") and the mixed-precision
flexible CG of the learned solver (V-cycle and network coarse ends): the lid
cavity against Ghia, Ghia & Shin's centre lines, the Poiseuille channel, the
cylinder at Re 20 and the backward-facing step.  Besides the short
trajectories that are checked bit for bit elsewhere, these pin the steady
states the mixed-precision solvers reach.

The centre-line data are Ghia, Ghia & Shin (1982), J. Comput. Phys. 48,
387-411, tables I and II (the reference holds the same table at
cfd2d/validation.py:15-34); the Poiseuille closed forms, the force
coefficients and the reattachment rule are restated from their definitions
(validation.py:69-80, 86-88, 131-152).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GHIA_Y = np.array([1.0000, 0.9766, 0.9688, 0.9609, 0.9531, 0.8516, 0.7344, 0.6172, 0.5000,
                   0.4531, 0.2813, 0.1719, 0.1016, 0.0703, 0.0625, 0.0547, 0.0000])
GHIA_X = np.array([1.0000, 0.9688, 0.9609, 0.9531, 0.9453, 0.9063, 0.8594, 0.8047, 0.5000,
                   0.2344, 0.2266, 0.1563, 0.0938, 0.0781, 0.0703, 0.0625, 0.0000])
GHIA_U = {
    100: [1.00000, 0.84123, 0.78871, 0.73722, 0.68717, 0.23151, 0.00332, -0.13641, -0.20581,
          -0.21090, -0.15662, -0.10150, -0.06434, -0.04775, -0.04192, -0.03717, 0.00000],
    1000: [1.00000, 0.65928, 0.57492, 0.51117, 0.46604, 0.33304, 0.18719, 0.05702, -0.06080,
           -0.10648, -0.27805, -0.38289, -0.29730, -0.22220, -0.20196, -0.18109, 0.00000],
}
GHIA_V = {
    100: [0.00000, -0.05906, -0.07391, -0.08864, -0.10313, -0.16914, -0.22445, -0.24533,
          0.05454, 0.17527, 0.17507, 0.16077, 0.12317, 0.10890, 0.10091, 0.09233, 0.00000],
    1000: [0.00000, -0.21388, -0.27669, -0.33714, -0.39188, -0.51550, -0.42665, -0.31966,
           0.02526, 0.32235, 0.33075, 0.37095, 0.32627, 0.30353, 0.29012, 0.27485, 0.00000],
}

SOLVERS = {
    "direct": dict(pressure_solver="direct"),
    "vcycle": dict(pressure_solver="cg_learned", pressure_tol=1e-6,
                   learned={"mode": "vcycle", "x0": "extrapolate"}),
    "net": dict(pressure_solver="cg_learned", pressure_tol=1e-6,
                learned={"mode": "net", "x0": "extrapolate"}),
}


def _ghia_errors(sim, re):
    """Largest deviation of u on the vertical and v on the horizontal centre
    line from Ghia et al. (the reference's cavity_centerline_errors)."""
    u = sim.sample_u(np.full_like(GHIA_Y, 0.5), GHIA_Y)
    v = sim.sample_v(GHIA_X, np.full_like(GHIA_X, 0.5))
    return (float(np.abs(u - np.array(GHIA_U[re])).max()),
            float(np.abs(v - np.array(GHIA_V[re])).max()))


def _cavity(re, n, solver):
    from cfd2d_amd import cases

    case = cases.lid_driven_cavity(re=re, n=n, **SOLVERS[solver])
    case.sim.sync_every = 200
    return case


@pytest.mark.parametrize("solver", ["direct", "vcycle", "net"])
def test_cavity_re100_matches_ghia(gpu, solver):
    """tests/test_cfd2d.py:57-62: 32^2, run to t = 40 or a residual of 1e-5,
    both centre-line errors below 0.03."""
    case = _cavity(100, 32, solver)
    reason = case.sim.run(40, steady_tol=1e-5, verbose=False)
    eu, ev = _ghia_errors(case.sim, 100)
    print(f"Re 100, 32^2, {solver}: {reason} after {case.sim.n} steps, "
          f"max error u {eu:.4f} v {ev:.4f}")
    assert eu < 0.03 and ev < 0.03
    case.sim.close()


@pytest.mark.parametrize("solver", ["direct", "vcycle"])
def test_cavity_re100_64_matches_readme_accuracy(gpu, solver):
    """README.md:126: at 64^2 the reference's steady state deviates from
    Ghia by 0.003 (u) and 0.008 (v); the GPU paths reach the same."""
    case = _cavity(100, 64, solver)
    case.sim.run(50, steady_tol=1e-5, verbose=False)
    eu, ev = _ghia_errors(case.sim, 100)
    print(f"Re 100, 64^2, {solver}: max error u {eu:.4f} v {ev:.4f}")
    assert eu < 0.0035 and ev < 0.0085
    case.sim.close()


@pytest.mark.parametrize("solver", ["vcycle", "net"])
def test_cavity_re1000_128_matches_ghia(gpu, solver):
    """README.md:127: at 128^2 and Re 1000 the reference deviates from Ghia by
    0.023 in u and v; the mixed-precision learned-solver family, every solve
    to a relative residual of 1e-6, reaches the same steady state."""
    case = _cavity(1000, 128, solver)
    reason = case.sim.run(50, steady_tol=1e-5, verbose=False)
    eu, ev = _ghia_errors(case.sim, 1000)
    print(f"Re 1000, 128^2, {solver}: {reason} after {case.sim.n} steps (t = {case.sim.t:.2f}), "
          f"max error u {eu:.4f} v {ev:.4f}")
    assert max(case.sim.cg_relres[-100:]) <= 1e-6
    assert eu < 0.026 and ev < 0.026
    case.sim.close()


@pytest.mark.parametrize("solver", ["direct", "vcycle"])
def test_poiseuille_profile_and_pressure_gradient(gpu, solver):
    """tests/test_cfd2d.py:65-77: plain inflow channel (uniform inflow, Re 20,
    4 x 1, 16 cells per H) to steady state: the profile at x = 3.6 within 0.02
    of 6y(1 - y), dp/dx within 3 % of -12 nu U / H^2."""
    from cfd2d_amd import cases

    case = cases.channel_flow(re=20, length=4, ny=16, profile="uniform", **SOLVERS[solver])
    case.sim.sync_every = 200
    case.sim.run(case.t_end, steady_tol=1e-6, verbose=False)
    s, g = case.sim, case.sim.grid
    y = g.yc
    u = s.sample_u(np.full_like(y, 3.6), y)
    ua = 6.0 * y * (1.0 - y)
    p = s.cell_pressure()[:, g.ny // 2]
    i1, i2 = int(0.6 * g.nx), int(0.9 * g.nx)
    dpdx = (p[i2] - p[i1]) / (g.xc[i2] - g.xc[i1])
    print(f"Poiseuille, {solver}: profile error {np.abs(u - ua).max():.4f}, "
          f"dp/dx {dpdx:.4f} vs {-12.0 * s.nu:.4f}")
    assert np.abs(u - ua).max() < 0.02
    assert dpdx == pytest.approx(-12.0 * s.nu, rel=0.03)
    s.close()


@pytest.mark.parametrize("solver", ["direct", "vcycle"])
def test_cylinder_re20_drag(gpu, solver):
    """tests/test_cfd2d.py:80-87: Schaefer-Turek cylinder at Re 20 on 220 x 41,
    run to t = 3: 5.0 < c_D < 6.1 and |c_L| < 0.05 (README.md:130: the
    reference gets c_D 5.47, c_L 0.0107)."""
    from cfd2d_amd import cases

    case = cases.cylinder_flow(re=20, nx=220, ny=41, t_end=3.0, **SOLVERS[solver])
    case.sim.sync_every = 200
    case.sim.run(case.t_end, verbose=False)
    fx, fy = case.sim.obstacle_forces()
    q = 0.5 * 1.0 * 1.0 ** 2 * 0.1          # 0.5 rho U^2 D
    cd, cl = fx / q, fy / q
    print(f"cylinder Re 20, {solver}: c_D {cd:.4f} c_L {cl:.5f}")
    assert 5.0 < cd < 6.1 and abs(cl) < 0.05
    case.sim.close()


def _reattachment(sim, x_step):
    """Distance from the step edge to the first change of the bottom wall's
    near-wall u from backflow (u < 0) to forward flow, interpolated linearly
    between faces (the reference's rule, validation.py:131-152)."""
    g = sim.grid
    x = np.arange(g.nx + 1) * g.dx
    u = sim.u[:, 1]
    m = x > x_step + 0.5 * g.dx
    xs, us = x[m], u[m]
    neg = np.flatnonzero(us < 0)
    if neg.size == 0:
        return None
    for k in range(neg[0], len(us) - 1):
        if us[k] < 0 <= us[k + 1]:
            return float(xs[k] - us[k] * (xs[k + 1] - xs[k]) / (us[k + 1] - us[k]) - x_step)
    return None


@pytest.mark.parametrize("solver", ["direct", "vcycle"])
def test_step_reattachment(gpu, solver):
    """tests/test_cfd2d.py:90-94: backward-facing step at Re 200 (8 x 1, 16
    cells per H) to t = 30 or a residual of 1e-4: a recirculation zone that
    reattaches between 1 and 5 step-heights-worth of length behind the edge."""
    from cfd2d_amd import cases

    case = cases.backward_facing_step(re=200, length=8, ny=16, t_end=30, **SOLVERS[solver])
    case.sim.sync_every = 200
    case.sim.run(case.t_end, steady_tol=1e-4, verbose=False)
    xr = _reattachment(case.sim, case.info["x_stufe"])
    print(f"step Re 200, {solver}: reattachment {xr}")
    assert xr is not None and 1.0 < xr < 5.0
    case.sim.close()
