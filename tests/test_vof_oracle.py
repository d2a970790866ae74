"""The two-phase VOF part of the oracle against the reference's own outputs
(tests/golden/vof.npz, written by tests/golden/make_golden.py): the alpha
transport kernel (vof.py:34-133), the CSF surface-tension source and the dam
break with the variable-density pressure solve, all bit for bit; plus the
invariants the transport keeps (boundedness, conservation without inflow)."""

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import cfd_oracle as O
from oracle import vof_oracle as V


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN / "vof.npz")


@pytest.mark.parametrize("tag", ["np", "p"])
@pytest.mark.parametrize("ci,c_alpha", [(0, 0.0), (1, 1.0)])
def test_vof_advect(gold, tag, ci, c_alpha):
    a = gold[f"adv_a_{tag}"]
    nx, ny = a.shape[0] - 2, a.shape[1] - 2
    out = np.full_like(a, 5.0)
    O.vof_advect(a, gold[f"adv_u_{tag}"], gold[f"adv_v_{tag}"], gold[f"adv_fluid_{tag}"], nx, ny,
                 1.0 / nx, 0.75 / ny, 0.013, c_alpha, tag == "p", 0.3, out)
    assert np.array_equal(out, gold[f"adv_out_{tag}_c{ci}"])
    inner = out[1:-1, 1:-1][gold[f"adv_fluid_{tag}"][1:-1, 1:-1]]
    assert inner.min() >= 0.0 and inner.max() <= 1.0


def test_vof_advect_conserves_in_a_closed_periodic_box():
    """Divergence-free periodic velocity, slip walls: the fluxes cancel in
    pairs, so the total is kept to rounding (no clipping for a smooth field)."""
    nx, ny = 24, 16
    g = O.OGrid(nx, ny, 1.0, 1.0)
    g.set_periodic_x()
    X, Y = g.cell_centers()
    a = np.zeros((nx + 2, ny + 2))
    a[1:-1, 1:-1] = 0.5 + 0.3 * np.sin(2 * np.pi * X) * np.sin(np.pi * Y)
    a[0], a[-1] = a[nx], a[1]
    ys = (np.arange(ny + 2) - 0.5) / ny
    u = np.sin(np.pi * np.clip(ys, 0, 1))[None, :] * np.ones((nx + 1, 1)) * 0.7
    u[:, 0] = u[:, -1] = 0.0
    v = np.zeros((nx + 2, ny + 1))
    out = np.zeros_like(a)
    O.vof_advect(a, u, v, g.fluid, nx, ny, 1.0 / nx, 1.0 / ny, 0.01, 1.0, True, 0.0, out)
    assert abs(out[1:-1, 1:-1].sum() - a[1:-1, 1:-1].sum()) < 1e-12


def test_dam_break(gold):
    sim = V.dam_break(n=32)
    m = sim.models[-1]
    for k, a in (("fluid", sim.grid.fluid), ("alpha_0", m.alpha), ("rho_0", sim.rho),
                 ("mu_0", sim.nu_eff), ("u_0", sim.u)):
        assert np.array_equal(a, gold[f"db_{k}"]), k
    hu, hv = np.zeros_like(sim.hu), np.zeros_like(sim.hv)
    m.momentum_source(sim, hu, hv)
    assert np.array_equal(hu, gold["db_csf_hu_0"]) and np.array_equal(hv, gold["db_csf_hv_0"])
    for n in range(1, 26):
        sim.step()
        if n in (1, 5, 25):
            for k, a in (("alpha", m.alpha), ("rho", sim.rho), ("u", sim.u), ("v", sim.v),
                         ("p", sim.p)):
                assert np.array_equal(a, gold[f"db_{k}_{n}"]), (n, k)
    assert np.array_equal(sim.hist["dt"], gold["db_h_dt"])
    for k in ("volumen_phase1", "volumenfehler_rel", "front_x", "hoehe_links"):
        assert np.array_equal(sim.extra[k], gold["db_hx_" + k]), k
    assert abs(sim.extra["volumenfehler_rel"][-1]) < 1e-12
