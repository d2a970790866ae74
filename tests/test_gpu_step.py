"""The full GPU step (Simulation.step through the C ABI) against the CPU FP64
oracle and against the reference's own trajectories (tests/golden).

Tolerances: the explicit kernels are bit-exact; differences come only from
the pressure solve, whose iterates differ from SciPy's in the order of the
dot-product sums.  With a tight solver tolerance (1e-13) the fields agree to
~1e-12; with the reference's CG tolerance the difference is bounded by the
solve tolerance itself.
"""

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _cmp(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max())


def test_cavity_steps_match_oracle_tight(gpu):
    from cfd2d_amd import cases
    from oracle import cfd_oracle as O

    n = 48
    sim = cases.lid_driven_cavity(re=100, n=n, gamma=0.0, pressure_solver="cg",
                                  pressure_tol=1e-13, pressure_maxiter=20000).sim
    ref = O.cavity(re=100, n=n, gamma=0.0, pressure_solver="cg", pressure_tol=1e-13)
    ref.pressure.maxiter = 20000
    for k in range(30):
        sim.step()
        ref.step()
        assert sim.dt == ref.dt
        assert _cmp(sim.u, ref.u) < 1e-11, k
        assert _cmp(sim.v, ref.v) < 1e-11, k
    assert _cmp(sim.p, ref.p) < 1e-8
    assert np.allclose(sim.history.kinetic_energy, ref.hist["kinetic_energy"], rtol=1e-12,
                       atol=0)
    assert sim.history.divergence[-1] < 1e-10


def test_golden_cavity_c1_first_steps(gpu):
    """Config 1 (64^2, Re 100, gamma 0, AB2, CG rtol 1e-8) against the
    reference's own trajectory after 1 and 10 steps."""
    from cfd2d_amd import cases

    g = np.load(GOLDEN / "cavity64_c1.npz")
    sim = cases.lid_driven_cavity(re=100, n=64, gamma=0.0, scheme="ab2", tau=0.5,
                                  pressure_solver="cg", pressure_tol=1e-8).sim
    for k in range(1, 11):
        sim.step()
        if k in (1, 10):
            assert _cmp(sim.u, g[f"u_{k}"]) < 1e-8, k
            assert _cmp(sim.v, g[f"v_{k}"]) < 1e-8, k
            assert _cmp(sim.p, g[f"p_{k}"]) < 1e-6, k
    assert sim.history.dt == list(g["h_dt"][:10])
    assert sim.history.dt[0] == 1.52587890625e-3


def test_golden_channel_obstacles_scalar(gpu):
    """Obstacle channel with parabolic inflow, outflow, slip walls and a
    Superbee scalar on half the inlet; reference run with the LU solver."""
    from cfd2d_amd import cases

    g = np.load(GOLDEN / "channel64x16.npz")
    obstacles = [{"typ": "kreis", "x": 0.7, "y": 0.5, "r": 0.12},
                 {"typ": "kreis", "x": 1.5, "y": 0.3, "r": 0.08},
                 {"typ": "kreis", "x": 2.1, "y": 0.72, "r": 0.1},
                 {"typ": "kreis", "x": 2.8, "y": 0.45, "r": 0.06},
                 {"typ": "kreis", "x": 3.3, "y": 0.2, "r": 0.07}]
    case = cases.custom_obstacles(obstacles, lx=4.0, ly=1.0, nx=64, ny=16, re=200, L_ref=1.0,
                                  inflow_profile="parabolic", pressure_solver="direct")
    cases.add_scalar(case, "c", 1e-3, inflow_band=(0.25, 0.75), limiter="superbee")
    sim = case.sim
    assert np.array_equal(sim.grid.fluid, g["fluid"])
    for k in range(1, 31):
        sim.step()
        if k in (1, 30):
            assert _cmp(sim.u, g[f"u_{k}"]) < 1e-9, k
            assert _cmp(sim.v, g[f"v_{k}"]) < 1e-9, k
            assert _cmp(sim.p, g[f"p_{k}"]) < 1e-8, k
            assert _cmp(sim.scalars["c"].phi, g[f"c_{k}"]) < 1e-9, k
    assert np.allclose(sim.history.extra["c_integral"], g["hx_c_integral"], rtol=1e-9, atol=0)
    assert np.allclose(sim.history.divergence, g["h_divergence"], atol=1e-10)
    assert max(sim.history.divergence) < 1e-10


def test_golden_heated_cavity(gpu):
    from cfd2d_amd import cases

    g = np.load(GOLDEN / "heated24.npz")
    sim = cases.heated_cavity(ra=1e4, n=24).sim
    for _ in range(40):
        sim.step()
    assert _cmp(sim.u, g["u"]) < 1e-9
    assert _cmp(sim.scalars["T"].phi, g["T"]) < 1e-9
    for k in ("T_mittel", "Nu_left", "Nu_right"):
        assert np.allclose(sim.history.extra[k], g["hx_" + k], rtol=1e-9, atol=1e-12), k


def test_golden_periodic_scalar(gpu):
    from cfd2d_amd.boundary import BoundaryCondition as BC
    from cfd2d_amd.grid import Grid
    from cfd2d_amd.models import PassiveScalar
    from cfd2d_amd.solver import Simulation

    g = np.load(GOLDEN / "periodic64x8.npz")
    grid = Grid(64, 8, 1.0, 0.125)
    bcs = dict(left=BC("periodic"), right=BC("periodic"), bottom=BC("slip"), top=BC("slip"))
    X, _ = grid.cell_centers()
    c0 = ((X > 0.2) & (X < 0.4)).astype(float)
    sc = PassiveScalar("c", diffusivity=1e-3, limiter="mc", initial=c0)
    s = Simulation(grid, bcs, nu=1e-3, models=[sc], initial_velocity=(1.0, 0.0))
    assert _cmp(s.u, g["u_0"]) < 1e-12
    for _ in range(20):
        s.step()
    assert _cmp(s.u, g["u"]) < 1e-9
    assert _cmp(s.p, g["p"]) < 1e-8
    assert _cmp(sc.field.phi, g["c"]) < 1e-9


def test_golden_euler_cg(gpu):
    from cfd2d_amd import cases

    g = np.load(GOLDEN / "euler16.npz")
    sim = cases.lid_driven_cavity(re=100, n=16, scheme="euler", pressure_solver="cg",
                                  gamma=0.5).sim
    for _ in range(30):
        sim.step()
    assert _cmp(sim.u, g["u"]) < 1e-9
    assert max(sim.history.divergence) < 1e-8
    # same Jacobi-PCG iteration semantics as scipy's cg: counts agree within one
    assert np.abs(np.array(sim.cg_iterations) - g["cg_iters"]).max() <= 1


def test_golden_autogamma_run(gpu):
    """run(t_end) with adaptive dt, auto gamma and the shortened last step."""
    from cfd2d_amd import cases

    g = np.load(GOLDEN / "autogamma32.npz")
    sim = cases.lid_driven_cavity(re=1000, n=32).sim
    reason = sim.run(0.05, verbose=False)
    assert reason == str(g["reason"])
    assert sim.n == int(g["n"])
    assert sim.t == float(g["t"])
    assert sim.history.dt == list(g["h_dt"])
    assert sim.gamma == float(g["gamma"])
    assert _cmp(sim.u, g["u"]) < 1e-9


def test_instability_raises(gpu):
    from cfd2d_amd.boundary import BoundaryCondition as BC
    from cfd2d_amd.grid import Grid
    from cfd2d_amd.solver import Simulation

    grid = Grid(16, 16, 1.0, 1.0)
    bcs = {"left": BC("noslip"), "right": BC("noslip"), "bottom": BC("noslip"),
           "top": BC("noslip", 1.0)}
    sim = Simulation(grid, bcs, nu=1e-2, dt=1.0)
    with pytest.raises(FloatingPointError):
        for _ in range(200):
            sim.step()


# ---------------------------------------------------------------------------
# cg_mixed: FP64 flexible CG with the FP32 multigrid preconditioner
# ---------------------------------------------------------------------------
def test_mixed_solver_cavity_matches_fp64_oracle(gpu):
    """Velocity within 1e-5 of the FP64 oracle (LU) after 60 steps with the
    pressure solved to a true relative residual of 1e-6 on every step."""
    from cfd2d_amd import cases
    from oracle import cfd_oracle as O

    n = 64
    sim = cases.lid_driven_cavity(re=1000, n=n, pressure_solver="cg_mixed",
                                  pressure_tol=1e-6).sim
    ref = O.cavity(re=1000, n=n, pressure_solver="direct")
    for _ in range(60):
        sim.step()
        ref.step()
    assert max(sim.cg_relres) <= 1e-6
    assert max(sim.cg_iterations) < 30
    assert _cmp(sim.u, ref.u) < 1e-5
    assert _cmp(sim.v, ref.v) < 1e-5


def test_mixed_solver_channel_obstacles(gpu):
    from cfd2d_amd import cases

    g = np.load(GOLDEN / "channel64x16.npz")
    obstacles = [{"typ": "kreis", "x": 0.7, "y": 0.5, "r": 0.12},
                 {"typ": "kreis", "x": 1.5, "y": 0.3, "r": 0.08},
                 {"typ": "kreis", "x": 2.1, "y": 0.72, "r": 0.1},
                 {"typ": "kreis", "x": 2.8, "y": 0.45, "r": 0.06},
                 {"typ": "kreis", "x": 3.3, "y": 0.2, "r": 0.07}]
    case = cases.custom_obstacles(obstacles, lx=4.0, ly=1.0, nx=64, ny=16, re=200, L_ref=1.0,
                                  inflow_profile="parabolic", pressure_solver="cg_mixed",
                                  pressure_tol=1e-9)
    cases.add_scalar(case, "c", 1e-3, inflow_band=(0.25, 0.75), limiter="superbee")
    sim = case.sim
    for _ in range(30):
        sim.step()
    assert max(sim.cg_relres) <= 1e-9
    assert _cmp(sim.u, g["u_30"]) < 1e-7
    assert _cmp(sim.scalars["c"].phi, g["c_30"]) < 1e-7


def test_mixed_pressure_solve_equals_dct(gpu):
    """Standalone PressureSolver.solve (cg_mixed) against the exact DCT-II
    solution of the Neumann problem."""
    import scipy.fft

    from cfd2d_amd import cases

    n = 128
    sim = cases.lid_driven_cavity(re=100, n=n, pressure_solver="cg_mixed",
                                  pressure_tol=1e-10).sim
    rng = np.random.default_rng(5)
    rhs = rng.standard_normal((n, n))
    p = sim.pressure.solve(rhs)
    h = 1.0 / n
    lam = (2 - 2 * np.cos(np.pi * np.arange(n) / n)) / h ** 2
    L = lam[:, None] + lam[None, :]
    L[0, 0] = 1.0
    r = scipy.fft.dctn(-(rhs - rhs.mean()), type=2, norm="ortho") / L
    r[0, 0] = 0.0
    exact = scipy.fft.idctn(r, type=2, norm="ortho")
    assert sim.pressure.last_relres <= 1e-10
    assert np.abs(p - exact).max() < 1e-7 * np.abs(exact).max()


@pytest.mark.parametrize("solver", ["cg", "cg_learned"])
def test_device_wait_give_up_is_reported(gpu, solver):
    """The device's waits on other work-groups are bounded.  A zero limit,
    set from the host, makes every wait give up at once (nothing waits, so
    nothing can hang): the kernels still end, and the next read of the device
    state raises instead of returning meaningless fields.  Setting a limit
    again clears the error; a fresh context then steps normally."""
    from cfd2d_amd import cases

    print(f"give-up case: {solver}", flush=True)
    sim = cases.lid_driven_cavity(re=100, n=32, pressure_solver=solver, pressure_tol=1e-6).sim
    sim.step()
    sim._ctx.set_spin_limit(0.0)
    with pytest.raises(RuntimeError, match="gave up"):
        sim.step()
    sim._ctx.set_spin_limit(2.0)
    sim._ctx.sync()
    sim.close()
    fresh = cases.lid_driven_cavity(re=100, n=32, pressure_solver=solver, pressure_tol=1e-6).sim
    for _ in range(5):
        fresh.step()
    assert fresh.history.divergence[-1] < 1e-6
    fresh.close()


def test_scalar_dt_limit_binds(gpu):
    """ScalarField.dt_limit (models.py:210-215) as the binding time-step limit:
    a scalar whose diffusivity (0.05) is 50 times the viscosity sets
    dt = tau * 0.5 / (D (1/dx^2 + 1/dy^2)) exactly (solver.py:227-230), and the
    trajectory follows the oracle's step for step."""
    from cfd2d_amd import cases
    from oracle import cfd_oracle as O

    n, D = 32, 0.05
    case = cases.lid_driven_cavity(re=1000, n=n, pressure_solver="cg", pressure_tol=1e-12,
                                   pressure_maxiter=20000)
    rng = np.random.default_rng(4)
    phi0 = rng.uniform(0.0, 1.0, (n, n))
    cases.add_scalar(case, "c", diffusivity=D, limiter="superbee", initial=phi0)
    ref = O.cavity(re=1000, n=n, pressure_solver="cg", pressure_tol=1e-12)
    ref.pressure.maxiter = 20000
    O.add_scalar(ref, "c", D, limiter="superbee", initial=phi0)
    h = 1.0 / n
    limit = 0.5 * (0.5 / (D * (1.0 / h ** 2 + 1.0 / h ** 2)))
    for _ in range(10):
        case.sim.step()
        ref.step()
        assert case.sim.dt == ref.dt == limit
    case.sim._flush()
    assert np.abs(case.sim.u - ref.u).max() < 1e-11
    assert np.abs(case.sim.scalars["c"].phi - ref.scalars["c"].phi).max() < 1e-12
    case.sim.close()
