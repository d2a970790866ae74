"""The turbulence part of the oracle against the reference's own outputs
(tests/golden/turbulence.npz, written by tests/golden/make_golden.py): the
stress-form momentum kernel (kernels.py:96-103, 131-138), a Smagorinsky
cavity and k-omega SST on the backward-facing step (turbulence.py), all bit
for bit."""

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import cfd_oracle as O
from oracle import turb_oracle as T


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN / "turbulence.npz")


@pytest.mark.parametrize("tag", ["np", "p"])
def test_stress_form_momentum(gold, tag):
    u, v, nu = gold[f"var_u_{tag}"], gold[f"var_v_{tag}"], gold[f"var_nu_{tag}"]
    nx, ny = u.shape[0] - 1, u.shape[1] - 2
    hu, hv = np.full_like(u, 7.0), np.full_like(v, 7.0)
    O.momentum_tendencies_var(u, v, gold[f"var_fluid_{tag}"], nx, ny, 1.0 / nx, 0.75 / ny, nu,
                              0.05, tag == "p", hu, hv)
    assert np.array_equal(hu, gold[f"var_hu_{tag}"])
    assert np.array_equal(hv, gold[f"var_hv_{tag}"])
    # a constant field reduces the stress form to nu (Lap u + grad div u):
    # away from walls it differs from the Laplacian form only by grad div u
    hc, vc = np.zeros_like(u), np.zeros_like(v)
    O.momentum_tendencies(u, v, gold[f"var_fluid_{tag}"], nx, ny, 1.0 / nx, 0.75 / ny, 0.02, 0.05,
                          tag == "p", hc, vc)
    hs, vs = np.zeros_like(u), np.zeros_like(v)
    O.momentum_tendencies_var(u, v, gold[f"var_fluid_{tag}"], nx, ny, 1.0 / nx, 0.75 / ny,
                              np.full_like(nu, 0.02), 0.05, tag == "p", hs, vs)
    dx, dy = 1.0 / nx, 0.75 / ny
    div = (u[1:, 1:-1] - u[:-1, 1:-1]) / dx + (v[1:-1, 1:] - v[1:-1, :-1]) / dy
    fl = gold[f"var_fluid_{tag}"]
    i, j = 5, 8
    assert fl[i:i + 3, j - 1:j + 2].all() and fl[i - 1:i + 2, j:j + 3].all()
    assert hs[i, j] - hc[i, j] == pytest.approx(0.02 * (div[i, j - 1] - div[i - 1, j - 1]) / dx,
                                                rel=1e-9)


def test_smagorinsky_cavity(gold):
    sim = O.cavity(re=1000, n=32, pressure_solver="cg")
    T.add_turbulence(sim, "Smagorinsky")
    m = sim.models[-1]
    assert np.array_equal(m.l2, gold["smag_l2"])
    assert np.array_equal(sim.nu_t, gold["smag_nut_0"])
    for n in range(1, 31):
        sim.step()
        if n in (1, 30):
            for k, a in (("u", sim.u), ("v", sim.v), ("p", sim.p), ("nut", sim.nu_t)):
                assert np.array_equal(a, gold[f"smag_{k}_{n}"]), (n, k)
    assert np.array_equal(sim.hist["dt"], gold["smag_h_dt"])
    assert np.array_equal(sim.extra["nu_t_max"], gold["smag_hx_nu_t_max"])
    assert sim.nu_t[1:-1, 1:-1].max() > 0 and sim.nu_t[2, 0] == 0.0


def test_sst_backward_facing_step(gold):
    sim = T.backward_facing_step(re=2000, length=6, ny=16, pressure_solver="cg")
    T.add_turbulence(sim, "kOmegaSST")
    m = sim.models[-1]
    assert np.array_equal(sim.grid.fluid, gold["sst_fluid"])
    for k, a in (("d", m.d), ("wall_adj", m.wall_adj), ("w_wall", m.w_wall), ("k_0", m.k.phi),
                 ("w_0", m.w.phi), ("nut_0", sim.nu_t)):
        assert np.array_equal(a, gold[f"sst_{k}"]), k
    for n in range(1, 21):
        sim.step()
        if n in (1, 20):
            for k, a in (("u", sim.u), ("v", sim.v), ("p", sim.p), ("nut", sim.nu_t),
                         ("k", m.k.phi), ("w", m.w.phi), ("kdiff", m.k.diff), ("wdiff", m.w.diff)):
                assert np.array_equal(a, gold[f"sst_{k}_{n}"]), (n, k)
    assert np.array_equal(sim.hist["dt"], gold["sst_h_dt"])
    assert np.array_equal(sim.extra["k_mittel"], gold["sst_hx_k_mittel"])
    assert m.k.phi[1:-1, 1:-1][sim.grid.interior].min() >= 0


@pytest.fixture(scope="module")
def flow():
    return np.load(GOLDEN / "flowrate.npz")


def test_sst_turbulent_channel_with_flow_control(flow):
    sim, _ = T.turbulent_channel(re_tau=120, y_plus_1=2.0, pressure_solver="cg")
    m = sim.models[-1]
    assert np.array_equal(sim.nu_t, flow["tc_nut_0"])
    forces = []
    for n in range(1, 41):
        sim.step()
        forces.append(sim.flow_force)
        if n in (1, 40):
            for k, a in (("u", sim.u), ("v", sim.v), ("p", sim.p), ("nut", sim.nu_t),
                         ("k", m.k.phi), ("w", m.w.phi)):
                assert np.array_equal(a, flow[f"tc_{k}_{n}"]), (n, k)
    assert np.array_equal(forces, flow["tc_flow_force"])
    assert np.array_equal(sim.hist["residual"], flow["tc_h_residual"])
    assert np.array_equal(sim.extra["antriebskraft"], flow["tc_hx_antriebskraft"])


def test_periodic_obstacle_flow_control_reprojects(flow):
    g = O.OGrid(32, 32, 1.0, 1.0)
    g.add_obstacle(g.circle(0.5, 0.5, 0.15))
    bcs = dict(left=O.OBC("periodic"), right=O.OBC("periodic"), bottom=O.OBC("slip"),
               top=O.OBC("slip"))
    sim = O.OSim(g, bcs, nu=0.02, flow_rate=1.0, pressure_solver="cg")
    forces = []
    for _ in range(30):
        sim.step()
        forces.append(sim.flow_force)
    for k, a in (("u", sim.u), ("v", sim.v), ("p", sim.p)):
        assert np.array_equal(a, flow[f"po_{k}"]), k
    assert np.array_equal(forces, flow["po_flow_force"])
    assert sim.max_divergence() < 1e-10 and sim.flow_force > 0


def test_flow_control_reaches_poiseuille():
    """Closed form (the reference's tests/test_models.py:17-25): the
    controlled laminar channel settles on u = 6 U y (1 - y) with the driving
    force dp/dx = 12 nu U / H^2."""
    g = O.OGrid(8, 20, 1.0, 1.0)
    bcs = dict(left=O.OBC("periodic"), right=O.OBC("periodic"), bottom=O.OBC("noslip"),
               top=O.OBC("noslip"))
    sim = O.OSim(g, bcs, nu=0.05, flow_rate=1.0)
    sim.run(20, steady_tol=1e-7)
    uc, _ = sim.cell_velocity()
    assert np.abs(uc[3] - 6 * g.yc * (1 - g.yc)).max() < 0.01
    assert sim.flow_force == pytest.approx(12 * 0.05, rel=0.01)
    assert sim.bulk_velocity() == pytest.approx(1.0, abs=1e-10)
