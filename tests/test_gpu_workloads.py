"""Checkpoint restart, the full config-1 trajectory and the obstacle-channel
workload (config 3) on the GPU."""

import json

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _cmp(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max())


def _channel_spec(solver="direct", numerik=None, **extra):
    spec = {"fall": "custom",
            "parameter": {"obstacles": [{"typ": "kreis", "x": 0.7, "y": 0.5, "r": 0.12},
                                        {"typ": "kreis", "x": 2.1, "y": 0.4, "r": 0.1}],
                          "lx": 4.0, "ly": 1.0, "nx": 64, "ny": 16, "re": 200, "L_ref": 1.0,
                          "inflow_profile": "parabolic"},
            "numerik": dict({"pressure_solver": solver}, **(numerik or {}))}
    spec.update(extra)
    return spec


@pytest.mark.parametrize("solver,numerik,extra", [
    ("direct", None, {}),
    ("cg", None, {"skalar": {"name": "c", "diffusivity": 1e-3, "inflow_band": [0.25, 0.75]}}),
    ("cg_mixed", None, {"skalar": {"name": "c", "diffusivity": 1e-3, "limiter": "superbee"}}),
    ("cg", {"pressure_x0": "extrapolate", "pressure_tol": 1e-6}, {}),
    ("cg_learned", {"pressure_tol": 1e-6, "learned": {"mode": "net", "x0": "extrapolate"}},
     {"skalar": {"name": "c", "diffusivity": 1e-3}}),
])
def test_checkpoint_restart_is_exact(gpu, tmp_path, solver, numerik, extra):
    """tests/test_functions.py:48-70 of the reference: save at step 40 of 80,
    reload into a fresh case, continue; the CG warm start travels in the file,
    and with the extrapolated initial iterate the solution before the last
    one as well (druck_xprev)."""
    from cfd2d_amd import cases, functions

    print(f"restart case: {solver} {numerik}", flush=True)
    spec = _channel_spec(solver, numerik, **extra)
    ref = cases.build_case(spec)
    for _ in range(80):
        ref.sim.step()
    a = cases.build_case(spec)
    for _ in range(40):
        a.sim.step()
    functions.save_checkpoint(a.sim, tmp_path / "cp.npz")
    extrapolated = "extrapolate" in json.dumps(numerik or {})
    assert ("druck_xprev" in np.load(tmp_path / "cp.npz").files) == extrapolated
    b = cases.build_case(spec)
    functions.load_checkpoint(b.sim, tmp_path / "cp.npz")
    for _ in range(40):
        b.sim.step()
    assert b.sim.n == ref.sim.n == 80
    assert np.array_equal(b.sim.u, ref.sim.u)
    assert np.array_equal(b.sim.p, ref.sim.p)
    for name in ref.sim.scalars:
        assert np.array_equal(b.sim.scalars[name].phi, ref.sim.scalars[name].phi)
    assert b.sim.history.kinetic_energy == ref.sim.history.kinetic_energy


@pytest.mark.parametrize("solver,numerik", [
    ("cg", {"pressure_x0": "extrapolate", "pressure_tol": 1e-6}),
    ("cg_learned", {"pressure_tol": 1e-6, "learned": {"mode": "vcycle", "x0": "extrapolate"}}),
])
def test_steps_after_the_end_change_nothing(gpu, tmp_path, solver, numerik):
    """A run to t_end launched in chunks of many steps idles through the steps
    after t_end; its checkpoint, the initial iterates of the extrapolated start
    included, equals that of the same run launched one step at a time."""
    from cfd2d_amd import cases, functions

    print(f"idle steps: {solver}", flush=True)
    files = []
    for one_by_one in (False, True):
        c = cases.build_case(_channel_spec(solver, numerik))
        kw = {"callback": lambda s: None, "callback_every": 1} if one_by_one else {}
        assert c.sim.run(1.0, verbose=False, **kw) == "t_end"
        functions.save_checkpoint(c.sim, tmp_path / f"cp{int(one_by_one)}.npz")
        files.append(np.load(tmp_path / f"cp{int(one_by_one)}.npz"))
    a, b = files
    assert 10 < int(a["n"]) < 200 and int(a["n"]) == int(b["n"])
    assert "druck_xprev" in a.files and sorted(a.files) == sorted(b.files)
    for k in a.files:
        assert np.array_equal(a[k], b[k]), k


def test_checkpoint_layout_matches_reference(gpu, tmp_path):
    from cfd2d_amd import cases, functions

    g = np.load(GOLDEN / "channel64x16.npz")
    layout = json.loads(str(g["checkpoint_layout_json"]))
    obstacles = [{"typ": "kreis", "x": 0.7, "y": 0.5, "r": 0.12},
                 {"typ": "kreis", "x": 1.5, "y": 0.3, "r": 0.08},
                 {"typ": "kreis", "x": 2.1, "y": 0.72, "r": 0.1},
                 {"typ": "kreis", "x": 2.8, "y": 0.45, "r": 0.06},
                 {"typ": "kreis", "x": 3.3, "y": 0.2, "r": 0.07}]
    case = cases.custom_obstacles(obstacles, lx=4.0, ly=1.0, nx=64, ny=16, re=200, L_ref=1.0,
                                  inflow_profile="parabolic", pressure_solver="direct")
    cases.add_scalar(case, "c", 1e-3, inflow_band=(0.25, 0.75), limiter="superbee")
    for _ in range(30):
        case.sim.step()
    d = np.load(functions.save_checkpoint(case.sim, tmp_path / "cp.npz"), allow_pickle=False)
    mine = {k: [str(d[k].dtype), list(d[k].shape)] for k in d.files
            if k not in ("druck_x0", "druck_xprev")}
    # history_json is a unicode scalar whose length depends on the float reprs
    assert mine.pop("history_json")[0].startswith("<U")
    assert layout.pop("history_json")[0].startswith("<U")
    assert mine == layout
    # the reference's own history json of the same run
    h_ref = json.loads(str(g["checkpoint_history_json"]))
    h = json.loads(str(d["history_json"]))
    # "direct" is CG to rtol 1e-12 here (LU in the reference): t agrees to round-off
    assert h["step"] == h_ref["step"]
    assert np.allclose(h["t"], h_ref["t"], rtol=1e-12, atol=0)
    assert np.allclose(h["kinetic_energy"], h_ref["kinetic_energy"], rtol=1e-9)
    assert np.allclose(h["extra"]["c_integral"], h_ref["extra"]["c_integral"], rtol=1e-9)
    assert np.allclose(h["fx"], h_ref["fx"], rtol=1e-6, atol=1e-9)


def test_checkpoint_wrong_geometry_rejected(gpu, tmp_path):
    from cfd2d_amd import cases, functions

    a = cases.lid_driven_cavity(re=100, n=16)
    functions.save_checkpoint(a.sim, tmp_path / "cp.npz")
    b = cases.lid_driven_cavity(re=100, n=20)
    with pytest.raises(ValueError):
        functions.load_checkpoint(b.sim, tmp_path / "cp.npz")


def test_config1_full_trajectory(gpu):
    """Config 1 (64^2, Re 100, gamma 0, AB2, CG rtol 1e-8), 2000 steps, against
    the reference's fields at step 2000."""
    from cfd2d_amd import cases

    g = np.load(GOLDEN / "cavity64_c1.npz")
    sim = cases.lid_driven_cavity(re=100, n=64, gamma=0.0, scheme="ab2", tau=0.5,
                                  pressure_solver="cg", pressure_tol=1e-8, sync_every=250).sim
    for _ in range(2000):
        sim.step()
    sim._flush()
    assert sim.n == 2000 and sim.t == 3.0517578125
    eu, ev, ep = _cmp(sim.u, g["u_2000"]), _cmp(sim.v, g["v_2000"]), _cmp(sim.p, g["p_2000"])
    print(f"config 1 after 2000 steps: max|du|={eu:.2e} max|dv|={ev:.2e} max|dp|={ep:.2e}")
    assert eu < 1e-6 and ev < 1e-6 and ep < 1e-4
    assert np.allclose(sim.history.kinetic_energy, g["h_ke"], rtol=1e-6)


def _config3_case(nx, ny, solver, seed=0, learned=None):
    from cfd2d_amd import cases

    rng = np.random.default_rng(seed)
    xs = rng.uniform(0.3, 3.0, 12)
    ys = rng.uniform(0.1, 0.9, 12)
    rs = rng.uniform(0.03, 0.08, 12)
    obstacles = [{"typ": "kreis", "x": float(x), "y": float(y), "r": float(r)}
                 for x, y, r in zip(xs, ys, rs)]
    case = cases.custom_obstacles(obstacles, lx=4.0, ly=1.0, nx=nx, ny=ny, re=200, L_ref=1.0,
                                  inflow_profile="parabolic", pressure_solver=solver,
                                  pressure_tol=1e-10 if solver == "direct" else 1e-6,
                                  learned=learned)
    cases.add_scalar(case, "c", 1e-3, inflow_band=(0.25, 0.75), limiter="superbee")
    return case, (xs, ys, rs)


def test_config3_reduced_matches_oracle(gpu):
    """Config 3 geometry (12 seeded circles, parabolic inflow U_max 1.5,
    outflow, Superbee scalar on half the inlet) at 256x64 against the oracle."""
    from oracle import cfd_oracle as O

    case, (xs, ys, rs) = _config3_case(256, 64, "cg_mixed")
    sim = case.sim
    ref = O.obstacle_channel(list(zip(xs, ys, rs)), lx=4.0, ly=1.0, nx=256, ny=64, re=200,
                             L_ref=1.0, inflow_profile="parabolic", pressure_solver="direct")
    O.add_scalar(ref, "c", 1e-3, inflow_band=(0.25, 0.75), limiter="superbee")
    assert np.array_equal(sim.grid.fluid, ref.grid.fluid)
    for _ in range(40):
        sim.step()
        ref.step()
    assert _cmp(sim.u, ref.u) < 1e-5
    assert _cmp(sim.v, ref.v) < 1e-5
    assert _cmp(sim.scalars["c"].phi, ref.scalars["c"].phi) < 1e-5
    assert max(sim.cg_relres) <= 1e-6


@pytest.mark.parametrize("solver", ["cg_mixed", "cg_learned_vcycle"])
def test_config3_full_size_invariants(gpu, solver):
    """2048x512 obstacle channel, the configuration's full 1000 steps of the
    mixed-precision solvers (the multigrid FCG, and the learned solver's FCG
    with the complete V-cycle and the extrapolated start): every solve
    reaches the true residual, the scalar stays within [0, 1], inflow equals
    outflow."""
    learned = {"mode": "vcycle", "x0": "extrapolate"} if solver == "cg_learned_vcycle" else None
    case, _ = _config3_case(2048, 512, "cg_learned" if learned else solver, learned=learned)
    sim = case.sim
    sim.sync_every = 100
    print(f"config C3 full run: 2048x512 obstacle channel, 1000 steps, {solver}", flush=True)
    for _ in range(1000):
        sim.step()
    sim._flush()
    assert sim.n == 1000 and len(sim.cg_relres) == 1000
    assert max(sim.cg_relres) <= 1e-6 and all(i == 0 for i in sim.cg_info)
    c = sim.scalars["c"].cell()
    assert np.nanmin(c) >= -1e-9 and np.nanmax(c) <= 1 + 1e-9
    u = sim.u
    g = sim.grid
    q_in = u[0, 1:-1].sum() * g.dy
    q_out = u[g.nx, 1:-1].sum() * g.dy
    assert q_out == pytest.approx(q_in, rel=1e-4)
    print(f"config 3: {sim.n} steps, CG iterations mean {np.mean(sim.cg_iterations):.2f}, "
          f"max |div| {max(sim.history.divergence):.2e}")
