"""Pins of the CPU FP64 oracle (oracle/).

Each oracle function is checked against something it cannot agree with by
accident:
* golden vectors written by the reference itself (bit-exact);
* the DCT-II diagonalisation of the Neumann 5-point Laplacian (an exact
  Poisson solve independent of LU or CG);
* closed forms (time step of config 1, limiter formulas);
* invariants (divergence-free projection, TVD boundedness and conservation).
A planted mistake (a sign flip in the upwind blend, an off-by-one in the
Laplacian) is shown to fail these checks.
"""

import numpy as np
import pytest
import scipy.fft

from conftest import GOLDEN
from oracle import cfd_oracle as O


@pytest.fixture(scope="module")
def kg():
    return np.load(GOLDEN / "kernels_16x12.npz")


def _kern_args(kg, tag):
    return (kg[f"u_{tag}"], kg[f"v_{tag}"], kg[f"fluid_{tag}"], int(kg["nx"]), int(kg["ny"]),
            float(kg["dx"]), float(kg["dy"]))


@pytest.mark.parametrize("tag,periodic", [("np", False), ("p", True)])
def test_momentum_matches_reference_bitexact(kg, tag, periodic):
    u, v, fl, nx, ny, dx, dy = _kern_args(kg, tag)
    for gi, gamma in enumerate((0.0, 0.05, 1.0)):
        hu, hv = np.full_like(u, 7.0), np.full_like(v, 7.0)
        O.momentum_tendencies(u, v, fl, nx, ny, dx, dy, 0.0137, gamma, periodic, hu, hv)
        assert np.array_equal(hu, kg[f"mom_hu_{tag}_g{gi}"])
        assert np.array_equal(hv, kg[f"mom_hv_{tag}_g{gi}"])


def test_planted_upwind_sign_mistake_is_caught(kg):
    u, v, fl, nx, ny, dx, dy = _kern_args(kg, "np")
    hu, hv = np.zeros_like(u), np.zeros_like(v)
    O.momentum_tendencies(u, v, fl, nx, ny, dx, dy, 0.0137, -0.05, False, hu, hv)
    assert not np.allclose(hu, kg["mom_hu_np_g1"], atol=1e-6)


@pytest.mark.parametrize("tag,periodic", [("np", False), ("p", True)])
def test_obstacle_bc_matches_reference(kg, tag, periodic):
    u, v, fl, nx, ny, dx, dy = _kern_args(kg, tag)
    ub, vb = u.copy(), v.copy()
    O.obstacle_bc(ub, vb, fl, nx, ny, periodic)
    assert np.array_equal(ub, kg[f"obc_u_{tag}"])
    assert np.array_equal(vb, kg[f"obc_v_{tag}"])


@pytest.mark.parametrize("tag,periodic", [("np", False), ("p", True)])
@pytest.mark.parametrize("obs", [0, 1])
def test_scalar_transport_matches_reference(kg, tag, periodic, obs):
    u, v, fl, nx, ny, dx, dy = _kern_args(kg, tag)
    phi, diff = kg[f"phi_{tag}"], kg[f"diff_{tag}"]
    b = [kg[f"{k}_{s}_{tag}"] for s in "lrbt" for k in ("bt", "bv")]
    for lim in range(6):
        out = np.full_like(phi, 3.0)
        O.scalar_transport(phi, u, v, fl, nx, ny, dx, dy, diff, lim, periodic, *b, obs,
                           0.37 if obs else 0.0, out)
        assert np.array_equal(out, kg[f"tvd_{tag}_o{obs}_l{lim}"]), lim


def test_limiters_closed_form(kg):
    r = kg["lim_r"]
    ref = kg["lim_val"]
    vl = np.where(r <= 0, 0.0, 2 * r / (1 + r))
    mm = np.where(r <= 0, 0.0, np.minimum(r, 1.0))
    sb = np.where(r <= 0, 0.0, np.maximum(np.minimum(2 * r, 1.0), np.minimum(r, 2.0)))
    mc = np.where(r <= 0, 0.0, np.minimum(2 * r, np.minimum(0.5 * (1 + r), 2.0)))
    for k, expect in ((0, 0 * r), (1, vl), (2, mm), (3, sb), (4, 0 * r + 1), (5, mc)):
        got = np.array([O.limiter(x, k) for x in r])
        assert np.array_equal(got, expect)
        assert np.array_equal(got, ref[k])


def _neumann_dct_solve(rhs, dx, dy):
    """Exact solution of -Lap_h p = rhs with homogeneous Neumann walls, mean 0."""
    nx, ny = rhs.shape
    lx = (2 - 2 * np.cos(np.pi * np.arange(nx) / nx)) / dx ** 2
    ly = (2 - 2 * np.cos(np.pi * np.arange(ny) / ny)) / dy ** 2
    r = scipy.fft.dctn(rhs, type=2, norm="ortho")
    lam = lx[:, None] + ly[None, :]
    lam[0, 0] = 1.0
    p = r / lam
    p[0, 0] = 0.0
    return scipy.fft.idctn(p, type=2, norm="ortho")


@pytest.mark.parametrize("method", ["direct", "cg"])
def test_pressure_solve_equals_dct_exact(method):
    sim = O.cavity(re=100, n=24, pressure_solver=method, pressure_tol=1e-13)
    ps = sim.pressure
    ps.maxiter = 10000
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal((24, 24))
    p = ps.solve(rhs)                       # Lap p = rhs, p mean 0
    exact = _neumann_dct_solve(-(rhs - rhs.mean()), sim.grid.dx, sim.grid.dy)
    assert np.abs(p - exact).max() < 1e-10 * np.abs(exact).max()
    assert abs(p.mean()) < 1e-13


def test_oracle_dct_method_equals_lu_trajectory():
    """The oracle's "dct" pressure method (the FP64 reference of the 1024^2
    error figures) follows the reference's LU trajectory: a 32^2 Re 1000
    cavity with a van Leer scalar over 20 steps, fields to rounding."""
    a = O.cavity(re=1000, n=32, pressure_solver="direct")
    b = O.cavity(re=1000, n=32, pressure_solver="dct")
    for s in (a, b):
        O.add_scalar(s, "c", 1e-3, limiter="vanleer", initial=O.gaussian_blobs(s.grid))
        for _ in range(20):
            s.step()
    for f in ("u", "v", "p"):
        assert np.abs(getattr(a, f) - getattr(b, f)).max() < 1e-11, f
    assert np.abs(a.scalars["c"].phi - b.scalars["c"].phi).max() < 1e-12
    with pytest.raises(ValueError):
        O.obstacle_channel([(1.0, 0.5, 0.1)], nx=32, ny=8, pressure_solver="dct")


def test_oracle_openmp_cg_follows_scipy_cg():
    """The all-cores CPU baseline's C Jacobi-CG runs SciPy's loop: the same
    iteration counts (to one, the dot products are summed in another order)
    and the same fields to the solve tolerance as scipy.sparse.linalg.cg."""
    a = O.cavity(re=1000, n=48, pressure_solver="cg", pressure_tol=1e-8)
    b = O.cavity(re=1000, n=48, pressure_solver="cg_omp", pressure_tol=1e-8)
    for s in (a, b):
        for _ in range(10):
            s.step()
    assert all(abs(x - y) <= 1 for x, y in zip(a.cg_iters, b.cg_iters))
    assert np.abs(a.u - b.u).max() < 1e-9 and np.abs(a.p - b.p).max() < 1e-7


def test_planted_laplacian_mistake_is_caught():
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal((16, 16))
    rhs -= rhs.mean()
    good = _neumann_dct_solve(-rhs, 1 / 16, 1 / 16)
    bad = _neumann_dct_solve(-rhs, 1 / 16, 1 / 15)   # wrong spacing in y
    sim = O.cavity(re=100, n=16, pressure_solver="direct")
    p = sim.pressure.solve(rhs)
    assert np.abs(p - good).max() < 1e-10
    assert np.abs(p - bad).max() > 1e-4


def test_config1_time_step_closed_form():
    sim = O.cavity(re=100, n=64, gamma=0.0, scheme="ab2", tau=0.5, pressure_solver="cg",
                   pressure_tol=1e-8)
    assert sim.compute_dt() == 1.52587890625e-3
    sim.step()
    assert sim.dt == 1.52587890625e-3


def test_cavity_c1_trajectory_matches_reference():
    """Config 1 of the reference for 10 steps, compared bit for bit."""
    g = np.load(GOLDEN / "cavity64_c1.npz")
    sim = O.cavity(re=100, n=64, gamma=0.0, scheme="ab2", tau=0.5, pressure_solver="cg",
                   pressure_tol=1e-8)
    for k in range(1, 11):
        sim.step()
        if k in (1, 10):
            assert np.array_equal(sim.u, g[f"u_{k}"])
            assert np.array_equal(sim.v, g[f"v_{k}"])
            assert np.array_equal(sim.p, g[f"p_{k}"])
    assert sim.cg_iters == list(g["cg_iters"][:10])
    assert np.array_equal(sim.hist["kinetic_energy"], g["h_ke"][:10])


@pytest.mark.slow
def test_cavity_c1_full_2000_steps_match_reference():
    g = np.load(GOLDEN / "cavity64_c1.npz")
    sim = O.cavity(re=100, n=64, gamma=0.0, scheme="ab2", tau=0.5, pressure_solver="cg",
                   pressure_tol=1e-8)
    for _ in range(2000):
        sim.step()
    assert np.array_equal(sim.u, g["u_2000"])
    assert np.array_equal(sim.p, g["p_2000"])
    assert sim.t == 3.0517578125


def test_channel_obstacles_scalar_matches_reference():
    g = np.load(GOLDEN / "channel64x16.npz")
    circ = [(0.7, 0.5, 0.12), (1.5, 0.3, 0.08), (2.1, 0.72, 0.1), (2.8, 0.45, 0.06),
            (3.3, 0.2, 0.07)]
    sim = O.obstacle_channel(circ, lx=4.0, ly=1.0, nx=64, ny=16, re=200, L_ref=1.0,
                             inflow_profile="parabolic", pressure_solver="direct")
    O.add_scalar(sim, "c", 1e-3, inflow_band=(0.25, 0.75), limiter="superbee")
    assert np.array_equal(sim.grid.fluid, g["fluid"])
    for k in range(1, 31):
        sim.step()
    assert np.array_equal(sim.u, g["u_30"])
    assert np.array_equal(sim.p, g["p_30"])
    assert np.array_equal(sim.scalars["c"].phi, g["c_30"])
    assert max(sim.hist["divergence"]) < 1e-10


def test_heated_periodic_euler_autogamma_match_reference():
    h = np.load(GOLDEN / "heated24.npz")
    s = O.heated_cavity(ra=1e4, n=24)
    for _ in range(40):
        s.step()
    assert np.array_equal(s.u, h["u"]) and np.array_equal(s.scalars["T"].phi, h["T"])
    assert np.array_equal(s.extra["Nu_left"], h["hx_Nu_left"])

    pz = np.load(GOLDEN / "periodic64x8.npz")
    gr = O.OGrid(64, 8, 1.0, 0.125)
    bcs = dict(left=O.OBC("periodic"), right=O.OBC("periodic"), bottom=O.OBC("slip"),
               top=O.OBC("slip"))
    X, _ = gr.cell_centers()
    sc = O.OPassive("c", 1e-3, limiter="mc", initial=((X > 0.2) & (X < 0.4)).astype(float))
    s = O.OSim(gr, bcs, nu=1e-3, models=[sc], initial_velocity=(1.0, 0.0))
    for _ in range(20):
        s.step()
    assert np.array_equal(s.u, pz["u"]) and np.array_equal(sc.field.phi, pz["c"])

    e = np.load(GOLDEN / "euler16.npz")
    s = O.cavity(re=100, n=16, scheme="euler", pressure_solver="cg", gamma=0.5)
    for _ in range(30):
        s.step()
    assert np.array_equal(s.u, e["u"]) and s.cg_iters == list(e["cg_iters"])

    a = np.load(GOLDEN / "autogamma32.npz")
    s = O.cavity(re=1000, n=32)
    assert s.run(0.05) == str(a["reason"])
    assert s.n == int(a["n"]) and s.t == float(a["t"]) and np.array_equal(s.u, a["u"])


@pytest.mark.parametrize("limiter", ["upwind", "vanleer", "minmod", "superbee", "mc"])
def test_tvd_bounded_and_conservative(limiter):
    """Plug flow in a periodic channel (tests/test_models.py:39-55 of the reference)."""
    g = O.OGrid(64, 8, 1.0, 0.125)
    bcs = dict(left=O.OBC("periodic"), right=O.OBC("periodic"), bottom=O.OBC("slip"),
               top=O.OBC("slip"))
    X, _ = g.cell_centers()
    c0 = ((X > 0.2) & (X < 0.4)).astype(float)
    sc = O.OPassive("c", 0.0, limiter=limiter, initial=c0)
    s = O.OSim(g, bcs, nu=1e-3, models=[sc], initial_velocity=(1.0, 0.0))
    total0 = sc.field.total()
    s.run(0.5)
    c = sc.field.cell()
    assert c.min() >= -1e-12 and c.max() <= 1 + 1e-12
    assert sc.field.total() == pytest.approx(total0, rel=1e-10)
    xc = (c * X).sum() / c.sum()
    assert xc == pytest.approx(0.8, abs=0.02)
