"""Two-phase VOF on the device (csrc/vof.hip) against the reference's outputs
(tests/golden/vof.npz; the oracle reproduces them bit for bit) and the
reference's own VOF tests (tests/test_models.py:135-162)."""

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN / "vof.npz")


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_dam_break_matches_reference(gpu, gold):
    from cfd2d_amd import cases

    case = cases.dam_break(n=32, t_end=1.0)
    sim = case.sim
    m = sim.model("multiphase")
    for k, a in (("fluid", sim.grid.fluid), ("alpha_0", m.alpha), ("rho_0", sim.rho),
                 ("mu_0", sim.nu_eff)):
        assert np.array_equal(a, gold[f"db_{k}"]), k
    worst = {}
    for n in range(1, 26):
        sim.step()
        if n in (1, 5, 25):
            for k, a in (("alpha", m.alpha), ("rho", sim.rho), ("u", sim.u), ("v", sim.v),
                         ("p", sim.p)):
                worst[(n, k)] = _rel(a, gold[f"db_{k}_{n}"])
    print("dam break vs reference", worst, flush=True)
    # the pressure system is solved by PCG to 1e-12 instead of directly
    assert max(worst.values()) < 1e-7
    h = sim.history
    assert _rel(h.dt, gold["db_h_dt"]) < 1e-9
    for k in ("volumen_phase1", "front_x", "hoehe_links"):
        assert _rel(h.extra[k], gold["db_hx_" + k]) < 1e-9, k
    assert abs(h.extra["volumenfehler_rel"][-1]) < 1e-12


def test_reference_hydrostatic_rest(gpu):
    """tests/test_models.py:135-150 of the reference."""
    from cfd2d_amd.boundary import BoundaryCondition as BC
    from cfd2d_amd.grid import Grid
    from cfd2d_amd.solver import Simulation
    from cfd2d_amd.vof import VOF

    g = Grid(20, 20, 0.2, 0.2)
    X, Y = g.cell_centers()
    alpha0 = (Y < 0.1).astype(float)
    bcs = {s: BC("noslip") for s in ("left", "right", "bottom")}
    bcs["top"] = BC("outflow")
    vof = VOF(alpha0, sigma=0.0)
    sim = Simulation(g, bcs, nu=1e-6, models=[vof])
    sim.run(0.2, verbose=False)
    uc, vc = sim.cell_velocity()
    assert np.nanmax(np.hypot(uc, vc)) < 1e-6
    p_bottom = sim.cell_pressure()[10, 0]
    assert p_bottom == pytest.approx(1000 * 9.81 * (0.1 - 0.005), rel=0.02)
    assert abs(vof.volume() / vof.volume0 - 1) < 1e-10


def test_reference_dam_break_conserves_volume(gpu):
    """tests/test_models.py:153-162 of the reference."""
    from cfd2d_amd import cases

    case = cases.dam_break(n=32, obstacle=True, t_end=0.15)
    case.sim.run(case.t_end, verbose=False)
    vof = case.sim.model("multiphase")
    assert abs(vof.volume() / vof.volume0 - 1) < 1e-3
    a = vof.alpha[1:-1, 1:-1]
    assert a.min() >= 0 and a.max() <= 1
    assert case.sim.history.extra["front_x"][-1] > case.info["a"] * 1.3
    assert case.sim.max_divergence() < 1e-8


def _slosh(G, BC, **kw):
    g = G(50, 30, 1.0, 0.6)
    X, Y = g.cell_centers()
    g.add_obstacle((X - 0.5) ** 2 + (Y - 0.18) ** 2 < 0.08 ** 2)
    alpha0 = (Y < 0.3 + 0.05 * np.sin(2 * np.pi * X)).astype(float)
    bcs = dict(left=BC("periodic"), right=BC("periodic"), bottom=BC("noslip"), top=BC("outflow"))
    return g, alpha0, bcs


@pytest.mark.parametrize("multigrid", [True, False])
def test_periodic_slosh_matches_the_oracle(gpu, multigrid):
    """A periodic two-layer slosh past a cylinder, against the oracle (sparse
    direct solve of the variable-density system): the multigrid hierarchy over
    odd level sizes (50x30 -> 25x15 -> 13x8 -> 7x4, the periodic wrap on every
    level) and the Jacobi-PCG."""
    from cfd2d_amd.boundary import BoundaryCondition as BC
    from cfd2d_amd.grid import Grid
    from cfd2d_amd.solver import Simulation
    from cfd2d_amd.vof import VOF
    from oracle import cfd_oracle as O
    from oracle import vof_oracle as V

    g, a0, bcs = _slosh(Grid, BC)
    sim = Simulation(g, bcs, nu=1e-6, gamma=0.5,
                     models=[VOF(a0, sigma=0.07, multigrid=multigrid)])
    og, oa0, obcs = _slosh(O.OGrid, O.OBC)
    osim = O.OSim(og, obcs, nu=1e-6, gamma=0.5, models=[V.OVOF(oa0, sigma=0.07)])
    assert np.array_equal(sim.grid.fluid, osim.grid.fluid)
    worst = {}
    for _ in range(10):
        sim.step()
        osim.step()
    m, om = sim.model("multiphase"), osim.models[-1]
    for k, a, b in (("alpha", m.alpha, om.alpha), ("u", sim.u, osim.u), ("v", sim.v, osim.v),
                    ("p", sim.p, osim.p)):
        worst[k] = _rel(a, b)
    print("periodic slosh, multigrid", multigrid, worst, "iterations", sim.cg_iterations,
          flush=True)
    # PCG to a relative residual of 1e-12 against the direct solve
    assert max(worst.values()) < 1e-8
    assert _rel(sim.history.dt, osim.hist["dt"]) < 1e-10
    assert max(sim.cg_info) == 0


@pytest.mark.parametrize("multigrid", [True, False])
def test_closed_long_tank_matches_the_oracle(gpu, multigrid):
    """A closed 200 x 6 tank (pure Neumann: one pinned cell) with a tilted
    two-layer interface, against the oracle: the multigrid hierarchy coarsens
    to one cell across (200x6 -> 100x3 -> 50x2 -> 25x1 -> 13x1) and carries
    the pinned cell's diagonal."""
    from cfd2d_amd.boundary import BoundaryCondition as BC
    from cfd2d_amd.grid import Grid
    from cfd2d_amd.solver import Simulation
    from cfd2d_amd.vof import VOF
    from oracle import cfd_oracle as O
    from oracle import vof_oracle as V

    def setup(G, B):
        g = G(200, 6, 2.0, 0.06)
        X, Y = g.cell_centers()
        a0 = (Y < 0.03 + 0.02 * (X - 1.0)).astype(float)
        bcs = {s: B("noslip") for s in ("left", "right", "bottom", "top")}
        return g, a0, bcs

    g, a0, bcs = setup(Grid, BC)
    sim = Simulation(g, bcs, nu=1e-6, gamma=0.5,
                     models=[VOF(a0, sigma=0.0, multigrid=multigrid)])
    og, oa0, obcs = setup(O.OGrid, O.OBC)
    osim = O.OSim(og, obcs, nu=1e-6, gamma=0.5, models=[V.OVOF(oa0, sigma=0.0)])
    assert len(sim.pressure._pinned) == 1
    for _ in range(8):
        sim.step()
        osim.step()
    m, om = sim.model("multiphase"), osim.models[-1]
    worst = {k: _rel(a, b) for k, a, b in (("alpha", m.alpha, om.alpha), ("u", sim.u, osim.u),
                                            ("v", sim.v, osim.v), ("p", sim.p, osim.p))}
    print("closed long tank, multigrid", multigrid, worst, "iterations", sim.cg_iterations,
          flush=True)
    assert max(worst.values()) < 1e-8
    assert max(sim.cg_info) == 0
    assert abs(m.volume() / m.volume0 - 1) < 1e-12


def test_inflow_of_the_heavy_phase_matches_the_oracle(gpu):
    """Water entering through an inflow side (alpha_inflow = 1) into a
    channel with a water layer, outflow on the right: the inflow face of the
    alpha transport, the inflow and outflow faces of the variable-density
    system and the mixture properties, against the oracle."""
    from cfd2d_amd.boundary import BoundaryCondition as BC
    from cfd2d_amd.grid import Grid
    from cfd2d_amd.solver import Simulation
    from cfd2d_amd.vof import VOF
    from oracle import cfd_oracle as O
    from oracle import vof_oracle as V

    def setup(G, B):
        g = G(40, 16, 1.0, 0.4)
        X, Y = g.cell_centers()
        a0 = (Y < 0.15).astype(float)
        bcs = dict(left=B("inflow", 0.5), right=B("outflow"), bottom=B("noslip"),
                   top=B("noslip"))
        return g, a0, bcs

    g, a0, bcs = setup(Grid, BC)
    sim = Simulation(g, bcs, nu=1e-6, gamma=0.5,
                     models=[VOF(a0, sigma=0.0, alpha_inflow=1.0)])
    og, oa0, obcs = setup(O.OGrid, O.OBC)
    osim = O.OSim(og, obcs, nu=1e-6, gamma=0.5, models=[V.OVOF(oa0, sigma=0.0, alpha_inflow=1.0)])
    for _ in range(10):
        sim.step()
        osim.step()
    m, om = sim.model("multiphase"), osim.models[-1]
    worst = {k: _rel(a, b) for k, a, b in (("alpha", m.alpha, om.alpha), ("u", sim.u, osim.u),
                                            ("v", sim.v, osim.v), ("p", sim.p, osim.p))}
    print("inflow of water, vs oracle", worst, "iterations", sim.cg_iterations, flush=True)
    assert max(worst.values()) < 1e-8
    assert m.alpha[1, 1:-1].min() > 0.5          # the inflow column fills with water
    assert _rel(sim.history.dt, osim.hist["dt"]) < 1e-10


def test_dam_break_96_matches_the_oracle(gpu):
    """The dam break at its default 96 x 96 (five multigrid levels: a grid-wide
    fine level, then levels held by block 0 down to the 6 x 6 exact solve)
    against the oracle for 10 steps."""
    from cfd2d_amd import cases
    from oracle import vof_oracle as V

    sim = cases.dam_break(n=96, t_end=1.0).sim
    osim = V.dam_break(n=96)
    for _ in range(10):
        sim.step()
        osim.step()
    m, om = sim.model("multiphase"), osim.models[-1]
    worst = {k: _rel(a, b) for k, a, b in (("alpha", m.alpha, om.alpha), ("u", sim.u, osim.u),
                                            ("v", sim.v, osim.v), ("p", sim.p, osim.p))}
    print("dam break 96 vs oracle", worst, "iterations", sim.cg_iterations, flush=True)
    assert max(worst.values()) < 1e-8
    assert max(sim.cg_info) == 0
