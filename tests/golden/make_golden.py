"""Writes the golden fixtures of this directory by running the reference package.

Run in the development container only (the reference tree is not present on a
GPU host):

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py [/path/to/reference]

The reference (package ``cfd2d`` of kaeptnzursee1910/CFD-) is imported from
the given tree; Numba is absent there, so its kernels run through their pure
Python fallback (kernels.py:13-29), which performs the same IEEE double
arithmetic.  Only inputs and outputs are stored: each fixture is a compressed
``.npz`` of float64/int64/bool arrays plus a small JSON manifest.  Nothing
here is imported by the product or by the tests; the tests only read the
``.npz`` files.

Fixtures (all seeds use ``numpy.random.default_rng``):

* ``kernels_16x12.npz``  single calls of ``momentum_tendencies`` (gamma 0, 0.05,
  1; periodic and not), ``obstacle_bc`` with thin obstacles, and
  ``scalar_transport`` for all six limiters with mixed boundary codes
  (kernels.py:44-391).
* ``cavity64_c1.npz``   lid-driven cavity Re=100, 64x64, gamma 0, AB2, tau 0.5,
  CG with rtol 1e-8 (set on ``sim.pressure.tol``), 2000 steps: u, v, p after
  steps 1, 10 and 2000 plus the per-step history and CG iteration counts.
* ``channel64x16.npz``  obstacle channel (parabolic inflow, outflow, slip
  walls, five circles, direct solver) with a Superbee scalar on half the
  inlet, 30 steps, plus the checkpoint key layout.
* ``heated24.npz``      heated cavity Ra=1e4 (Boussinesq source, fixed-T
  walls), 40 steps, with the thermal monitors.
* ``periodic64x8.npz``  periodic plug flow with an MC-limited scalar, 20 steps
  (periodic index wrap, initial projection with non-zero velocity).
* ``euler16.npz``       cavity 16x16, Euler, gamma 0.5, CG (default rtol
  1e-10), 30 steps.
* ``autogamma32.npz``   cavity Re=1000 32x32 with gamma "auto" and adaptive
  dt, ``run(t_end=0.05)`` (last-step shortening), direct solver.
* ``turbulence.npz``    the stress-form ``momentum_tendencies`` (``variable =
  True``) on random fields with a random viscosity field (periodic and
  not); a Smagorinsky cavity (Re=1000, 32x32, CG, 30 steps, ``l^2`` and
  ``nu_t`` after steps 1 and 30); k-omega SST on the backward-facing step
  (Re=2000, length 6, ny 16, CG, 20 steps: wall distance, wall omega, k,
  omega and ``nu_t`` after the attach and after steps 1 and 20).
* ``flowrate.npz``      flow-rate control (solver.py:405-425): the SST
  turbulent channel (Re_tau 120, first cell at y+ 2, 40 steps) and a periodic
  channel with a circle (32x32, nu 0.02, 30 steps, re-projection), fields,
  flow_force and histories.
* ``vof.npz``           two-phase VOF (vof.py): ``vof_advect`` on random fields
  (periodic and not, with obstacles, compression on and off), the CSF
  surface-tension source on a smoothed column, and the dam break (32x32,
  with the obstacle, 25 steps): alpha, rho, u, v, p and the monitors.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _import_reference(root: str):
    sys.path.insert(0, root)
    import warnings

    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        import cfd2d  # noqa: F401
        from cfd2d import cases, functions, kernels, models
        from cfd2d.boundary import BoundaryCondition
        from cfd2d.grid import Grid
        from cfd2d.solver import Simulation
    return dict(cases=cases, functions=functions, kernels=kernels, models=models,
                BC=BoundaryCondition, Grid=Grid, Simulation=Simulation)


def _hist(sim) -> dict:
    h = sim.history
    out = {
        "h_step": np.asarray(h.step, dtype=np.int64),
        "h_t": np.asarray(h.t, dtype=float),
        "h_dt": np.asarray(h.dt, dtype=float),
        "h_residual": np.asarray(h.residual, dtype=float),
        "h_divergence": np.asarray(h.divergence, dtype=float),
        "h_ke": np.asarray(h.kinetic_energy, dtype=float),
    }
    for k, v in h.extra.items():
        out["hx_" + k] = np.asarray(v, dtype=float)
    return out


# --------------------------------------------------------------------------- #
def kernels_fixture(R) -> dict:
    kernels, Grid = R["kernels"], R["Grid"]
    rng = np.random.default_rng(20240611)
    nx, ny = 16, 12
    dx, dy = 1.0 / nx, 0.75 / ny
    out = {"nx": np.int64(nx), "ny": np.int64(ny), "dx": np.float64(dx), "dy": np.float64(dy)}

    def make_fluid(periodic: bool) -> np.ndarray:
        g = Grid(nx, ny, 1.0, 0.75)
        m = np.zeros((nx, ny), dtype=bool)
        m[4:7, 3:6] = True          # block
        m[10, 2:9] = True           # one-cell-thick wall (thin obstacle)
        m[13:15, 9] = True          # thin horizontal plate
        m[2, 10] = True             # single cell
        g.add_obstacle(m)
        if periodic:
            g.set_periodic_x()
        return g.fluid.copy()

    for periodic in (False, True):
        tag = "p" if periodic else "np"
        fluid = make_fluid(periodic)
        u = rng.standard_normal((nx + 1, ny + 2))
        v = rng.standard_normal((nx + 2, ny + 1))
        out[f"fluid_{tag}"] = fluid
        out[f"u_{tag}"] = u
        out[f"v_{tag}"] = v
        nu = np.full((nx + 2, ny + 2), 0.0137)
        rho = np.ones((nx + 2, ny + 2))
        for gi, gamma in enumerate((0.0, 0.05, 1.0)):
            hu = np.full_like(u, 7.0)
            hv = np.full_like(v, 7.0)
            kernels.momentum_tendencies(u, v, fluid, nx, ny, dx, dy, nu, False, gamma, periodic,
                                        hu, hv, rho, False)
            out[f"mom_hu_{tag}_g{gi}"] = hu
            out[f"mom_hv_{tag}_g{gi}"] = hv
        ub, vb = u.copy(), v.copy()
        kernels.obstacle_bc(ub, vb, fluid, nx, ny, periodic)
        out[f"obc_u_{tag}"] = ub
        out[f"obc_v_{tag}"] = vb

        # scalar transport: random phi, random diffusivity with a few zeros,
        # per-face random boundary codes and values
        phi = rng.uniform(-1.0, 2.0, (nx + 2, ny + 2))
        diff = rng.uniform(0.0, 0.05, (nx + 2, ny + 2))
        diff[rng.uniform(size=diff.shape) < 0.1] = 0.0
        bt = {s: rng.integers(0, 3, n).astype(np.int64)
              for s, n in (("l", ny), ("r", ny), ("b", nx), ("t", nx))}
        bv = {s: rng.uniform(-1.0, 1.0, len(bt[s])) for s in bt}
        out[f"phi_{tag}"] = phi
        out[f"diff_{tag}"] = diff
        for s in bt:
            out[f"bt_{s}_{tag}"] = bt[s]
            out[f"bv_{s}_{tag}"] = bv[s]
        for obs_t, obs_v in ((0, 0.0), (1, 0.37)):
            for lim in range(6):
                o = np.full_like(phi, 3.0)
                kernels.scalar_transport(phi, u, v, fluid, nx, ny, dx, dy, diff, lim, periodic,
                                         bt["l"], bv["l"], bt["r"], bv["r"], bt["b"], bv["b"],
                                         bt["t"], bv["t"], obs_t, obs_v, o)
                out[f"tvd_{tag}_o{obs_t}_l{lim}"] = o
    rs = np.array([-2.0, -0.5, 0.0, 1e-3, 0.3, 0.5, 1.0, 1.5, 2.0, 3.0, 10.0])
    out["lim_r"] = rs
    out["lim_val"] = np.array([[kernels._limiter(r, k) for r in rs] for k in range(6)])
    return out


def cavity_c1(R) -> dict:
    cases = R["cases"]
    case = cases.lid_driven_cavity(re=100, n=64, gamma=0.0, scheme="ab2", tau=0.5,
                                   pressure_solver="cg")
    sim = case.sim
    sim.pressure.tol = 1e-8
    out = {}
    iters = []
    for k in range(1, 2001):
        sim.step()
        iters.append(sim.pressure.last_iterations)
        if k in (1, 10, 2000):
            out[f"u_{k}"] = sim.u.copy()
            out[f"v_{k}"] = sim.v.copy()
            out[f"p_{k}"] = sim.p.copy()
        if k % 200 == 0:
            print(f"  cavity64 step {k}", flush=True)
    out["cg_iters"] = np.asarray(iters, dtype=np.int64)
    out.update(_hist(sim))
    return out


def channel(R, tmpdir: Path) -> dict:
    cases, functions = R["cases"], R["functions"]
    obstacles = [{"typ": "kreis", "x": 0.7, "y": 0.5, "r": 0.12},
                 {"typ": "kreis", "x": 1.5, "y": 0.3, "r": 0.08},
                 {"typ": "kreis", "x": 2.1, "y": 0.72, "r": 0.1},
                 {"typ": "kreis", "x": 2.8, "y": 0.45, "r": 0.06},
                 {"typ": "kreis", "x": 3.3, "y": 0.2, "r": 0.07}]
    case = cases.custom_obstacles(obstacles, lx=4.0, ly=1.0, nx=64, ny=16, re=200, L_ref=1.0,
                                  inflow_profile="parabolic", pressure_solver="direct")
    cases.add_scalar(case, "c", 1e-3, inflow_band=(0.25, 0.75), limiter="superbee")
    sim = case.sim
    out = {"fluid": sim.grid.fluid.copy()}
    for k in range(1, 31):
        sim.step()
        if k in (1, 30):
            out[f"u_{k}"] = sim.u.copy()
            out[f"v_{k}"] = sim.v.copy()
            out[f"p_{k}"] = sim.p.copy()
            out[f"c_{k}"] = sim.scalars["c"].phi.copy()
    out.update(_hist(sim))
    cp = functions.save_checkpoint(sim, tmpdir / "cp.npz")
    d = np.load(cp, allow_pickle=False)
    layout = {k: [str(d[k].dtype), list(d[k].shape)] for k in d.files}
    out["checkpoint_layout_json"] = np.asarray(json.dumps(layout, sort_keys=True))
    out["checkpoint_history_json"] = np.asarray(str(d["history_json"]))
    return out


def heated(R) -> dict:
    cases = R["cases"]
    case = cases.heated_cavity(ra=1e4, n=24)
    sim = case.sim
    out = {}
    for k in range(1, 41):
        sim.step()
    out["u"] = sim.u.copy()
    out["v"] = sim.v.copy()
    out["p"] = sim.p.copy()
    out["T"] = sim.scalars["T"].phi.copy()
    out.update(_hist(sim))
    return out


def periodic(R) -> dict:
    Grid, BC, Simulation, models = R["Grid"], R["BC"], R["Simulation"], R["models"]
    g = Grid(64, 8, 1.0, 0.125)
    bcs = dict(left=BC("periodic"), right=BC("periodic"), bottom=BC("slip"), top=BC("slip"))
    X, _ = g.cell_centers()
    c0 = ((X > 0.2) & (X < 0.4)).astype(float)
    sc = models.PassiveScalar("c", diffusivity=1e-3, limiter="mc", initial=c0)
    s = Simulation(g, bcs, nu=1e-3, models=[sc], initial_velocity=(1.0, 0.0))
    out = {"u_0": s.u.copy(), "v_0": s.v.copy()}
    for _ in range(20):
        s.step()
    out["u"] = s.u.copy()
    out["v"] = s.v.copy()
    out["p"] = s.p.copy()
    out["c"] = sc.field.phi.copy()
    out.update(_hist(s))
    return out


def euler16(R) -> dict:
    cases = R["cases"]
    case = cases.lid_driven_cavity(re=100, n=16, scheme="euler", pressure_solver="cg", gamma=0.5)
    sim = case.sim
    iters = []
    for _ in range(30):
        sim.step()
        iters.append(sim.pressure.last_iterations)
    out = {"u": sim.u.copy(), "v": sim.v.copy(), "p": sim.p.copy(),
           "cg_iters": np.asarray(iters, dtype=np.int64)}
    out.update(_hist(sim))
    return out


def autogamma(R) -> dict:
    cases = R["cases"]
    case = cases.lid_driven_cavity(re=1000, n=32)
    sim = case.sim
    reason = sim.run(0.05, verbose=False)
    out = {"u": sim.u.copy(), "v": sim.v.copy(), "p": sim.p.copy(),
           "gamma": np.float64(sim.gamma), "n": np.int64(sim.n), "t": np.float64(sim.t),
           "reason": np.asarray(reason)}
    out.update(_hist(sim))
    return out


def turbulence(R) -> dict:
    kernels, Grid, cases = R["kernels"], R["Grid"], R["cases"]
    rng = np.random.default_rng(20240917)
    nx, ny = 16, 12
    dx, dy = 1.0 / nx, 0.75 / ny
    out = {}
    for periodic in (False, True):
        tag = "p" if periodic else "np"
        g = Grid(nx, ny, 1.0, 0.75)
        m = np.zeros((nx, ny), dtype=bool)
        m[4:7, 3:6] = True
        m[10, 2:9] = True
        g.add_obstacle(m)
        if periodic:
            g.set_periodic_x()
        fluid = g.fluid.copy()
        u = rng.standard_normal((nx + 1, ny + 2))
        v = rng.standard_normal((nx + 2, ny + 1))
        nu = rng.uniform(0.001, 0.05, (nx + 2, ny + 2))
        hu, hv = np.full_like(u, 7.0), np.full_like(v, 7.0)
        kernels.momentum_tendencies(u, v, fluid, nx, ny, dx, dy, nu, True, 0.05, periodic, hu, hv,
                                    np.ones((nx + 2, ny + 2)), False)
        out.update({f"var_fluid_{tag}": fluid, f"var_u_{tag}": u, f"var_v_{tag}": v,
                    f"var_nu_{tag}": nu, f"var_hu_{tag}": hu, f"var_hv_{tag}": hv})

    case = cases.add_turbulence(cases.lid_driven_cavity(re=1000, n=32, pressure_solver="cg"),
                                "Smagorinsky")
    sim = case.sim
    smag = sim.model("turbulence")
    out["smag_l2"] = smag.l2.copy()
    out["smag_nut_0"] = sim.nu_t.copy()
    for n in range(1, 31):
        sim.step()
        if n in (1, 30):
            out.update({f"smag_u_{n}": sim.u.copy(), f"smag_v_{n}": sim.v.copy(),
                        f"smag_p_{n}": sim.p.copy(), f"smag_nut_{n}": sim.nu_t.copy()})
    out.update({"smag_" + k: v for k, v in _hist(sim).items()})

    case = cases.backward_facing_step(re=2000, length=6, ny=16, t_end=1.0, pressure_solver="cg")
    cases.add_turbulence(case, "kOmegaSST")
    sim = case.sim
    sst = sim.model("turbulence")
    out.update({"sst_d": sst.d.copy(), "sst_wall_adj": sst.wall_adj.copy(),
                "sst_w_wall": sst.w_wall.copy(), "sst_k_0": sst.k.phi.copy(),
                "sst_w_0": sst.w.phi.copy(), "sst_nut_0": sim.nu_t.copy(),
                "sst_fluid": sim.grid.fluid.copy()})
    for n in range(1, 21):
        sim.step()
        if n in (1, 20):
            out.update({f"sst_u_{n}": sim.u.copy(), f"sst_v_{n}": sim.v.copy(),
                        f"sst_p_{n}": sim.p.copy(), f"sst_nut_{n}": sim.nu_t.copy(),
                        f"sst_k_{n}": sst.k.phi.copy(), f"sst_w_{n}": sst.w.phi.copy(),
                        f"sst_kdiff_{n}": sst.k.diff.copy(), f"sst_wdiff_{n}": sst.w.diff.copy()})
    out.update({"sst_" + k: v for k, v in _hist(sim).items()})
    return out


def flowrate(R) -> dict:
    cases, Grid, BC, Simulation = R["cases"], R["Grid"], R["BC"], R["Simulation"]
    out = {}
    case = cases.turbulent_channel(re_tau=120, y_plus_1=2.0, t_end=250, pressure_solver="cg")
    sim = case.sim
    sst = sim.model("turbulence")
    out["tc_nut_0"] = sim.nu_t.copy()
    forces = []
    for n in range(1, 41):
        sim.step()
        forces.append(sim.flow_force)
        if n in (1, 40):
            out.update({f"tc_u_{n}": sim.u.copy(), f"tc_v_{n}": sim.v.copy(),
                        f"tc_p_{n}": sim.p.copy(), f"tc_nut_{n}": sim.nu_t.copy(),
                        f"tc_k_{n}": sst.k.phi.copy(), f"tc_w_{n}": sst.w.phi.copy()})
    out["tc_flow_force"] = np.asarray(forces)
    out.update({"tc_" + k: v for k, v in _hist(sim).items()})

    import cfd2d.geometry as geometry

    g = Grid(32, 32, 1.0, 1.0)
    g.add_obstacle(geometry.circle(g, 0.5, 0.5, 0.15))
    bcs = dict(left=BC("periodic"), right=BC("periodic"), bottom=BC("slip"), top=BC("slip"))
    sim = Simulation(g, bcs, nu=0.02, flow_rate=1.0, pressure_solver="cg")
    forces = []
    for n in range(1, 31):
        sim.step()
        forces.append(sim.flow_force)
    out.update({"po_u": sim.u.copy(), "po_v": sim.v.copy(), "po_p": sim.p.copy(),
                "po_flow_force": np.asarray(forces), "po_fluid": sim.grid.fluid.copy()})
    out.update({"po_" + k: v for k, v in _hist(sim).items()})
    return out


def vof(R) -> dict:
    cases, Grid, BC, Simulation = R["cases"], R["Grid"], R["BC"], R["Simulation"]
    import cfd2d.vof as vofmod

    rng = np.random.default_rng(20241002)
    out = {}
    nx, ny = 16, 12
    dx, dy = 1.0 / nx, 0.75 / ny
    for periodic in (False, True):
        tag = "p" if periodic else "np"
        g = Grid(nx, ny, 1.0, 0.75)
        m = np.zeros((nx, ny), dtype=bool)
        m[4:7, 3:6] = True
        m[10, 2:9] = True
        g.add_obstacle(m)
        if periodic:
            g.set_periodic_x()
        fluid = g.fluid.copy()
        a = np.clip(rng.uniform(-0.2, 1.2, (nx + 2, ny + 2)), 0.0, 1.0)
        u = rng.standard_normal((nx + 1, ny + 2))
        v = rng.standard_normal((nx + 2, ny + 1))
        out.update({f"adv_fluid_{tag}": fluid, f"adv_a_{tag}": a, f"adv_u_{tag}": u,
                    f"adv_v_{tag}": v})
        for ci, c_alpha in enumerate((0.0, 1.0)):
            o = np.full_like(a, 5.0)
            vofmod.vof_advect(a, u, v, fluid, nx, ny, dx, dy, 0.013, c_alpha, periodic, 0.3, o)
            out[f"adv_out_{tag}_c{ci}"] = o

    case = cases.dam_break(n=32, t_end=1.0)
    sim = case.sim
    vm = sim.model("multiphase")
    out.update({"db_alpha_0": vm.alpha.copy(), "db_rho_0": sim.rho.copy(),
                "db_mu_0": sim.nu_eff.copy(), "db_fluid": sim.grid.fluid.copy(),
                "db_u_0": sim.u.copy(), "db_v_0": sim.v.copy()})
    hu, hv = np.zeros_like(sim.hu), np.zeros_like(sim.hv)
    vm.momentum_source(sim, hu, hv)
    out["db_csf_hu_0"], out["db_csf_hv_0"] = hu, hv
    for n in range(1, 26):
        sim.step()
        if n in (1, 5, 25):
            out.update({f"db_alpha_{n}": vm.alpha.copy(), f"db_rho_{n}": sim.rho.copy(),
                        f"db_u_{n}": sim.u.copy(), f"db_v_{n}": sim.v.copy(),
                        f"db_p_{n}": sim.p.copy()})
    out.update({"db_" + k: v for k, v in _hist(sim).items()})
    return out


def main() -> None:
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    R = _import_reference(root)
    import tempfile

    tmp = Path(tempfile.mkdtemp())
    jobs = [("kernels_16x12", lambda: kernels_fixture(R)),
            ("channel64x16", lambda: channel(R, tmp)),
            ("heated24", lambda: heated(R)),
            ("periodic64x8", lambda: periodic(R)),
            ("euler16", lambda: euler16(R)),
            ("autogamma32", lambda: autogamma(R)),
            ("cavity64_c1", lambda: cavity_c1(R)),
            ("turbulence", lambda: turbulence(R)),
            ("flowrate", lambda: flowrate(R)),
            ("vof", lambda: vof(R))]
    only = set(sys.argv[2:])
    for name, fn in jobs:
        if only and name not in only:
            continue
        print(f"writing {name}", flush=True)
        data = fn()
        np.savez_compressed(HERE / f"{name}.npz", **data)
    print("done", flush=True)


if __name__ == "__main__":
    main()
