"""The learned flexible CG (pressure_solver="cg_learned") on the GPU.

The preconditioner is a V-cycle whose coarse correction comes from the U-Net
(shipped weights, trained on the GPU by the library's own kernels).  Checked:
the launch-based V-cycle against the persistent multigrid solver, every mode
against the FP64 solvers and the FP64 oracle, the gain of the trained network
over the same solver without it, the harvested training pairs, the Jacobi-PCG
fallback, and a short harvest -> train -> solve run from scratch.
"""

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cmp(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max())


def _cavity(n, solver, learned=None, re=1000.0, tol=1e-6, maxiter=None):
    from cfd2d_amd import cases

    return cases.lid_driven_cavity(re=re, n=n, pressure_solver=solver, pressure_tol=tol,
                                   pressure_maxiter=maxiter, learned=learned).sim


def _run(sim, steps):
    for _ in range(steps):
        sim.step()
    sim._flush()
    return sim


def test_launch_vcycle_equals_persistent_multigrid(gpu):
    """Teacher mode on level 0 with one inner iteration is the multigrid
    V-cycle of cg_mixed (scaled by a line search, which flexible CG undoes):
    same iteration counts, same fields to round-off."""
    from cfd2d_amd.learned import LearnedConfig

    mg = _run(_cavity(64, "cg_mixed"), 20)
    lv = _run(_cavity(64, "cg_learned", LearnedConfig(mode="teacher", level=0,
                                                      teacher_cycles=1)), 20)
    assert lv.cg_iterations == mg.cg_iterations
    assert _cmp(lv.u, mg.u) < 1e-9 and _cmp(lv.p, mg.p) < 1e-8


@pytest.mark.parametrize("mode", ["net", "none", "teacher"])
def test_learned_modes_match_fp64_cg(gpu, mode):
    from cfd2d_amd.learned import LearnedConfig

    ref = _run(_cavity(64, "cg", tol=1e-10), 30)
    sim = _run(_cavity(64, "cg_learned", LearnedConfig(mode=mode)), 30)
    assert max(sim.cg_relres) <= 1e-6
    assert all(i == 0 for i in sim.cg_info)
    assert sim.pressure.learned.stats()["fallbacks"] == 0
    assert _cmp(sim.u, ref.u) < 1e-6 and _cmp(sim.v, ref.v) < 1e-6


def test_graph_replay_equals_direct_launches(gpu):
    from cfd2d_amd.learned import LearnedConfig

    a = _run(_cavity(128, "cg_learned", LearnedConfig(use_graph=True)), 15)
    b = _run(_cavity(128, "cg_learned", LearnedConfig(use_graph=False)), 15)
    assert a.cg_iterations == b.cg_iterations
    assert np.array_equal(a.u, b.u) and np.array_equal(a.p, b.p)


def test_learned_matches_fp64_oracle(gpu):
    """Velocity within 1e-5 of the FP64 oracle (LU) after 60 steps at 128^2,
    Re 1000, every solve to a true relative residual of 1e-6."""
    from oracle import cfd_oracle as O

    sim = _cavity(128, "cg_learned")
    ref = O.cavity(re=1000, n=128, pressure_solver="direct")
    for _ in range(60):
        sim.step()
        ref.step()
    sim._flush()
    eu, ev = _cmp(sim.u, ref.u), _cmp(sim.v, ref.v)
    print(f"128^2 after 60 steps: max|du| {eu:.2e} max|dv| {ev:.2e}")
    assert max(sim.cg_relres) <= 1e-6
    assert eu < 1e-5 and ev < 1e-5


def _modes_timed(n, modes, steps=10, warm=3, re=1000.0):
    """Iterations per step and ms per step of each learned-solver mode, steps
    warm+1 .. warm+steps from rest, every solve to a true relres <= 1e-6."""
    from cfd2d_amd.learned import LearnedConfig

    out = {}
    for mode in modes:
        sim = _run(_cavity(n, "cg_learned", LearnedConfig(mode=mode, x0="extrapolate"), re=re),
                   warm)
        t0 = time.perf_counter()
        _run(sim, steps)
        out[mode] = (1e3 * (time.perf_counter() - t0) / steps,
                     float(np.mean(sim.cg_iterations[-steps:])))
        assert max(sim.cg_relres[-steps:]) <= 1e-6 and all(i == 0 for i in sim.cg_info)
        sim.close()
    print(f"{n}^2 Re {re:g}: " + ", ".join(f"{m} {it:.2f} it {ms:.2f} ms/step"
                                          for m, (ms, it) in out.items()), flush=True)
    return out


@pytest.mark.parametrize("n", [256, 1024])
def test_network_gain(gpu, n):
    """The trained network against the same solver without a coarse solve
    below its 32x32 level (constant-vector correction only): far fewer
    iterations and faster steps.  Against the complete V-cycle to the dense
    coarsest solve (mode "vcycle") the numbers are printed; the network
    approximates that coarse solve, so it cannot need fewer iterations."""
    out = _modes_timed(n, ("net", "vcycle", "none"))
    assert out["net"][1] < 0.4 * out["none"][1]
    assert out["net"][0] < out["none"][0]
    assert out["vcycle"][1] <= out["net"][1] + 0.5


def test_config2_velocity_error_against_fp64_cg(gpu):
    """The benchmark workload (1024^2, Re 1000, van Leer scalar, 40 steps):
    learned-PCG at rtol 1e-6 against the FP64 Jacobi-PCG path at rtol 1e-10
    (the GPU FP64 path reproduces the oracle's trajectories): velocity within
    1e-5, scalar within 1e-6."""
    from bench import build_case

    a = build_case("cg_learned").sim
    ref_case = build_case("cg")
    ref = ref_case.sim
    ref.pressure.tol = 1e-10
    _run(a, 40)
    _run(ref, 40)
    eu, ev = _cmp(a.u, ref.u), _cmp(a.v, ref.v)
    ec = _cmp(a.scalars["c"].phi, ref.scalars["c"].phi)
    print(f"1024^2 after 40 steps: max|du| {eu:.2e} max|dv| {ev:.2e} max|dc| {ec:.2e}")
    assert max(a.cg_relres) <= 1e-6
    assert eu < 1e-5 and ev < 1e-5 and ec < 1e-6


def test_config2_full_run(gpu):
    """The benchmark workload for all 500 steps of its configuration, on the
    headline solver (flexible CG with the complete V-cycle, extrapolated
    start): every solve ends on a true relative residual <= 1e-6 with no
    fallback; the TVD scalar keeps the bounds of its initial field (van Leer
    with diffusion and closed walls: no new extrema, up to 1e-5 of the range
    for the divergence the solves leave); the velocities, the
    pressure and the scalar after 500 steps agree with the multigrid FCG
    solved to 1e-10 (FP64 vectors and dot products) to 1e-5, 1e-4 of the
    pressure range and 1e-6."""
    from bench import build_case

    steps = 500
    print(f"config C2 full run: 1024^2, {steps} steps, cg_learned_vcycle vs cg_mixed 1e-10",
          flush=True)
    a = build_case("cg_learned_vcycle").sim
    c0 = a.scalars["c"].cell()
    lo, hi = float(np.nanmin(c0)), float(np.nanmax(c0))
    ref = build_case("cg_mixed").sim
    ref.pressure.tol = 1e-10
    a.sync_every = ref.sync_every = 100
    _run(a, steps)
    _run(ref, steps)
    c = a.scalars["c"].cell()
    eu, ev, ep = _cmp(a.u, ref.u), _cmp(a.v, ref.v), _cmp(a.p, ref.p)
    ec = _cmp(a.scalars["c"].phi, ref.scalars["c"].phi)
    prange = float(ref.p.max() - ref.p.min())
    print(f"1024^2 after {steps} steps: iterations mean {np.mean(a.cg_iterations):.2f}, "
          f"max|du| {eu:.2e} max|dv| {ev:.2e} max|dp| {ep:.2e} (p range {prange:.3g}) "
          f"max|dc| {ec:.2e}; scalar in [{np.nanmin(c):.3e}, {np.nanmax(c):.6f}] "
          f"(initial [{lo:.3e}, {hi:.6f}])")
    assert a.n == steps and len(a.cg_relres) == steps
    assert max(a.cg_relres) <= 1e-6 and all(i == 0 for i in a.cg_info)
    tol = 1e-5 * (hi - lo)   # the solves leave a divergence of the order of rtol
    assert np.nanmin(c) >= lo - tol and np.nanmax(c) <= hi + tol
    assert eu < 1e-5 and ev < 1e-5 and ec < 1e-6
    assert ep < 1e-4 * prange
    a.close()
    ref.close()


def test_harvested_pairs_solve_the_coarse_system(gpu):
    """Teacher pairs (input x, target t): A_K t / D = x to the teacher's
    accuracy on the coarse level K (its Galerkin matrix from the hierarchy)."""
    from cfd2d_amd.learned import LearnedConfig

    sim = _run(_cavity(256, "cg_learned", LearnedConfig(mode="teacher", harvest_cap=64)), 10)
    att = sim.pressure.learned
    x, t = att.harvest()
    assert x.shape[0] == min(64, att.stats()["harvested"]) and x.shape[0] > 10
    L = sim.pressure.hierarchy.levels[att.level]
    fl = L["fluid"]
    D = 2.0 / sim.grid.dx ** 2 + 2.0 / sim.grid.dy ** 2
    worst = 0.0
    for k in range(x.shape[0]):
        tv = t[k, :L["nx"], :L["ny"]][fl].astype(np.float64)
        xv = x[k, :L["nx"], :L["ny"]][fl].astype(np.float64)
        worst = max(worst, np.linalg.norm(L["A"] @ tv / D - xv) / np.linalg.norm(xv))
    assert worst < 1e-3
    # the inputs are deflated (orthogonal to the constant vector) and normalised
    assert np.allclose(np.sqrt(np.mean(x[:, :L["nx"], :L["ny"]] ** 2, axis=(1, 2))), 1.0,
                       rtol=1e-3)


def test_fallback_to_jacobi_pcg(gpu):
    """A learned solve that cannot finish in max_iter iterations is continued
    by the FP64 Jacobi-PCG; the result still meets the tolerance."""
    from cfd2d_amd.learned import LearnedConfig

    ref = _run(_cavity(64, "cg", tol=1e-10), 10)
    sim = _run(_cavity(64, "cg_learned", LearnedConfig(mode="none", max_iter=3)), 10)
    assert sim.pressure.learned.stats()["fallbacks"] >= 10
    assert max(sim.cg_relres) <= 1e-6 and all(i == 0 for i in sim.cg_info)
    assert _cmp(sim.u, ref.u) < 1e-6


def test_pressure_solve_learned_equals_dct(gpu):
    import scipy.fft

    n = 128
    sim = _cavity(n, "cg_learned", tol=1e-10)
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


def test_harvest_train_solve_from_scratch(gpu):
    """The whole learned pipeline in miniature: harvest pairs from the
    solver's own solves, train a fresh network on the GPU for a few seconds,
    and use it: fewer iterations than without a network."""
    from cfd2d_amd.learned import LearnedConfig, train

    sim = _run(_cavity(256, "cg_learned", LearnedConfig(mode="teacher", harvest_cap=400)), 60)
    x, t = sim.pressure.learned.harvest()
    sim.close()
    params, sigma, hist = train(x, t, epochs=200, batch=8, lr=1e-3, per_epoch=400, seed=3,
                                time_limit=15.0)
    assert hist[-1] < 0.5 * hist[0]
    its = {}
    for mode in ("net", "none"):
        s = _run(_cavity(256, "cg_learned", LearnedConfig(mode=mode, params=params,
                                                          sigma=sigma)), 8)
        its[mode] = float(np.mean(s.cg_iterations[-5:]))
        assert max(s.cg_relres) <= 1e-6
        s.close()
    print(f"fresh network after {len(hist)} epochs (loss {hist[0]:.3f} -> {hist[-1]:.4f}): "
          f"{its['net']:.1f} iterations vs {its['none']:.1f} without")
    assert its["net"] < 0.7 * its["none"]


def test_network_switched_off_outside_its_training(gpu):
    """The shipped network was trained on cavities; on the obstacle channel
    (outflow, masks, 32 x 8 coarse image) it does not help, so the solver
    switches it off after its check and runs as without it: no Jacobi-PCG
    fallbacks, the same order of iterations as the network-free mode."""
    from cfd2d_amd import cases
    from cfd2d_amd.learned import LearnedConfig

    def channel(mode):
        rng = np.random.default_rng(0)
        xs, ys, rs = rng.uniform(0.3, 3.0, 12), rng.uniform(0.1, 0.9, 12), rng.uniform(0.03, 0.08, 12)
        obst = [{"typ": "kreis", "x": float(x), "y": float(y), "r": float(r)}
                for x, y, r in zip(xs, ys, rs)]
        case = cases.custom_obstacles(obst, lx=4.0, ly=1.0, nx=1024, ny=256, re=200, L_ref=1.0,
                                      inflow_profile="parabolic", pressure_solver="cg_learned",
                                      pressure_tol=1e-6, learned=LearnedConfig(mode=mode))
        return _run(case.sim, 15)

    net, none = channel("net"), channel("none")
    st = net.pressure.learned.stats()
    print(f"channel: network off {st['net_off']}, iterations {np.mean(net.cg_iterations):.1f} "
          f"vs {np.mean(none.cg_iterations):.1f} without, fallbacks {st['fallbacks']}")
    assert max(net.cg_relres) <= 1e-6 and all(i == 0 for i in net.cg_info)
    assert st["fallbacks"] == 0
    assert np.mean(net.cg_iterations) < 1.5 * np.mean(none.cg_iterations) + 8


@pytest.mark.parametrize("mode", ["net", "none"])
def test_fused_coarse_end_matches_launch_sequence(gpu, mode):
    """The coarse end of the V-cycle in one work-group (k_lv_coarse: the
    levels of at most 64^2 cells, the network, the level-K correction)
    against the same preconditioner as separate launches: the same iteration
    counts to within one per solve on average, velocities within the solve
    tolerance.  (The sweeps are the same FP32 arithmetic; the statistics c, s
    and the network's accumulation order differ in rounding.)"""
    from cfd2d_amd.learned import LearnedConfig

    a = _run(_cavity(256, "cg_learned", LearnedConfig(mode=mode, coarse_fused=True,
                                                      fused_legs=False)), 20)
    b = _run(_cavity(256, "cg_learned", LearnedConfig(mode=mode, coarse_fused=False)), 20)
    ia, ib = np.mean(a.cg_iterations), np.mean(b.cg_iterations)
    print(f"{mode}: fused {ia:.2f} iterations, separate launches {ib:.2f}; "
          f"max|du| {_cmp(a.u, b.u):.1e}")
    assert max(a.cg_relres) <= 1e-6 and all(i == 0 for i in a.cg_info)
    assert abs(ia - ib) <= 1.0
    assert _cmp(a.u, b.u) < 1e-6 and _cmp(a.v, b.v) < 1e-6


def test_vcycle_mode_is_the_multigrid_preconditioner(gpu):
    """mode "vcycle" carries the V-cycle on to the dense coarsest solve, which
    is cg_mixed's preconditioner: the same iteration counts as the persistent
    multigrid solver, the same fields to the solve tolerance."""
    from cfd2d_amd.learned import LearnedConfig

    mg = _run(_cavity(128, "cg_mixed"), 20)
    lv = _run(_cavity(128, "cg_learned", LearnedConfig(mode="vcycle", sweeps=2)), 20)
    print(f"128^2: cg_mixed {mg.cg_iterations}, vcycle {lv.cg_iterations}")
    assert lv.cg_iterations == mg.cg_iterations
    assert _cmp(lv.u, mg.u) < 1e-9 and _cmp(lv.p, mg.p) < 1e-7


@pytest.mark.parametrize("sweeps", [2, 3])
def test_pair_legs_equal_single_level_legs(gpu, sweeps):
    """The two levels above the LDS coarse end in one launch per leg
    (k_lv_down2 / k_lv_up2, each work-group recomputing what it needs from the
    other level) against one launch per level and leg: the same FP32
    arithmetic, bit-identical fields and iteration counts (512^2: levels
    256^2 and 128^2 paired)."""
    from cfd2d_amd.learned import LearnedConfig

    out = []
    for pair in (True, False):
        out.append(_run(_cavity(512, "cg_learned", LearnedConfig(mode="vcycle", sweeps=sweeps,
                                                                 fused_pair=pair)), 10))
    a, b = out
    assert a.cg_iterations == b.cg_iterations
    assert np.array_equal(a.u, b.u) and np.array_equal(a.p, b.p)


def test_vcycle_fused_equals_launches(gpu):
    """The complete V-cycle's coarse end in one work-group (levels <= 64^2 to
    the dense solve) is the same FP32 arithmetic as the separate launches:
    bit-identical fields."""
    from cfd2d_amd.learned import LearnedConfig

    a = _run(_cavity(256, "cg_learned", LearnedConfig(mode="vcycle", coarse_fused=True,
                                                      fused_legs=False)), 10)
    b = _run(_cavity(256, "cg_learned", LearnedConfig(mode="vcycle", coarse_fused=False)), 10)
    assert a.cg_iterations == b.cg_iterations
    assert np.array_equal(a.u, b.u) and np.array_equal(a.p, b.p)


def _legs_case(kind, learned):
    from cfd2d_amd import cases

    kw = dict(pressure_solver="cg_learned", pressure_tol=1e-6, learned=learned)
    if kind == "cavity":        # 256^2: 8 x 32 tiles
        return cases.lid_driven_cavity(re=1000, n=256, **kw).sim
    if kind == "cavity2048":    # 16 x 64 tiles on levels 0, 1; several tiles per work-group
        return cases.lid_driven_cavity(re=1000, n=2048, **kw).sim
    if kind == "periodic":      # 258 x 64, periodic in x: wrapped rows, odd level 1 (129 x 32)
        return cases.rayleigh_benard(ra=1e5, ny=64, **kw).sim
    return cases.cylinder_flow(re=100, nx=220, ny=41, **kw).sim   # partial tiles, obstacle


@pytest.mark.parametrize("mode", ["net", "none", "vcycle"])
def test_fused_legs_are_the_launch_sequence(gpu, mode):
    """The legs above the coarse end as one launch per level and leg
    (k_lv_down / k_lv_up: tiles with a two-cell ring in LDS) compute the
    separate sweep, restriction and prolongation launches' FP32 values: with
    the fine level's up-leg (and so the dot partials) left as launches, the
    solve is bit-identical."""
    from cfd2d_amd.learned import LearnedConfig

    a = _run(_cavity(256, "cg_learned", LearnedConfig(mode=mode, fused_fine_up=False)), 12)
    b = _run(_cavity(256, "cg_learned", LearnedConfig(mode=mode, fused_legs=False)), 12)
    assert a.cg_iterations == b.cg_iterations
    assert np.array_equal(a.u, b.u) and np.array_equal(a.p, b.p)


@pytest.mark.parametrize("kind", ["cavity", "cavity2048", "periodic", "cylinder"])
def test_vcycle_fused_legs_match_launches(gpu, kind):
    """The same on other grids (periodic rows wrapped in the tiles' rings, odd
    level sizes, partial tiles, an obstacle): bit-identical with the fine
    up-leg as launches; with it fused as well, the partial sums of r.z and
    r_old.z come in tile order instead of row bands, so the iteration counts
    agree and the fields within the solve tolerance (each solve stops at a
    relative residual of 1e-6, wherever the rounding of the dot products
    leaves it below that)."""
    from cfd2d_amd.learned import LearnedConfig

    steps = 4 if kind == "cavity2048" else 12
    cfg = dict(mode="vcycle", x0="extrapolate", sweeps=2)
    a = _run(_legs_case(kind, LearnedConfig(**cfg)), steps)
    e = _run(_legs_case(kind, LearnedConfig(**cfg, fused_fine_up=False)), steps)
    b = _run(_legs_case(kind, LearnedConfig(**cfg, fused_legs=False)), steps)
    print(f"{kind}: fused legs {a.cg_iterations}, separate launches {b.cg_iterations}; "
          f"max|du| {_cmp(a.u, b.u):.1e} max|dp| {_cmp(a.p, b.p):.1e}")
    assert e.cg_iterations == b.cg_iterations
    assert np.array_equal(e.u, b.u) and np.array_equal(e.p, b.p)
    assert max(a.cg_relres) <= 1e-6 and all(i == 0 for i in a.cg_info)
    assert a.pressure.learned.stats()["fallbacks"] == 0
    assert abs(sum(a.cg_iterations) - sum(b.cg_iterations)) <= 1
    scale = float(np.abs(b.p).max()) + 1e-30
    assert _cmp(a.u, b.u) < 1e-7 and _cmp(a.v, b.v) < 1e-7
    assert _cmp(a.p, b.p) < 1e-6 * scale


@pytest.mark.parametrize("sweeps", [3, 4, "auto"])
def test_more_sweeps_fewer_iterations(gpu, sweeps):
    """3 or 4 damped-Jacobi sweeps per leg above the coarse end (a stronger
    symmetric V-cycle), or the adaptive choice: every solve to the tolerance,
    fewer iterations than with 2 sweeps over the start-up steps (where solves
    take 5 to 8 iterations), the same fields to the solve tolerance; the
    stats report the sweeps of the last solve."""
    from cfd2d_amd.learned import LearnedConfig

    a = _run(_cavity(512, "cg_learned", LearnedConfig(mode="vcycle", x0="extrapolate",
                                                      sweeps=sweeps)), 15)
    b = _run(_cavity(512, "cg_learned", LearnedConfig(mode="vcycle", x0="extrapolate",
                                                      sweeps=2)), 15)
    st = a.pressure.learned.stats()
    print(f"sweeps {sweeps}: {a.cg_iterations} (last solve {st['last_sweeps']} sweeps); "
          f"2 sweeps: {b.cg_iterations}")
    assert max(a.cg_relres) <= 1e-6 and all(i == 0 for i in a.cg_info)
    assert st["fallbacks"] == 0
    assert sum(a.cg_iterations) < sum(b.cg_iterations)
    assert st["last_sweeps"] in ((2, 3) if sweeps == "auto" else (sweeps,))
    assert _cmp(a.u, b.u) < 1e-7 and _cmp(a.v, b.v) < 1e-7


def test_sweeps_need_fused_legs(gpu):
    """More than 2 sweeps exist only in the fused legs: asking for them with
    the fine up-leg as launches is an error, "auto" falls back to 2."""
    from cfd2d_amd.learned import LearnedConfig

    with pytest.raises(ValueError, match="sweeps"):
        _cavity(128, "cg_learned", LearnedConfig(mode="vcycle", sweeps=3, fused_fine_up=False))
    sim = _run(_cavity(128, "cg_learned", LearnedConfig(mode="vcycle", sweeps="auto",
                                                        fused_fine_up=False)), 8)
    assert sim.pressure.learned.stats()["last_sweeps"] == 2


@pytest.mark.parametrize("n,re", [(2048, 1000.0), (1024, 5000.0)])
def test_config4_cavity(gpu, n, re):
    """Config C4's evaluation cases: the learned FCG (shipped network,
    extrapolated start) on the 2048^2 cavity at Re 1000 and the 1024^2 cavity
    at Re 5000 (cases.py:60-70).  Every solve ends on a true relative
    residual <= 1e-6; after 10 steps the velocities agree with the FP64
    Jacobi-PCG path (rtol 1e-10, the path that reproduces the oracle's
    trajectories) to 1e-5 and the pressure to 1e-4 of its range."""
    from cfd2d_amd.learned import LearnedConfig

    steps = 10
    print(f"config C4 case: {n}^2 Re {re:g}", flush=True)
    sim = _run(_cavity(n, "cg_learned", LearnedConfig(mode="net", x0="extrapolate"), re=re),
               steps)
    ref = _run(_cavity(n, "cg", re=re, tol=1e-10, maxiter=200000), steps)
    eu, ev, ep = _cmp(sim.u, ref.u), _cmp(sim.v, ref.v), _cmp(sim.p, ref.p)
    prange = float(ref.p.max() - ref.p.min())
    print(f"{n}^2 Re {re:g} after {steps} steps: iterations {sim.cg_iterations}, "
          f"max|du| {eu:.2e} max|dv| {ev:.2e} max|dp| {ep:.2e} (p range {prange:.3g})")
    assert max(sim.cg_relres) <= 1e-6 and all(i == 0 for i in sim.cg_info)
    assert sim.pressure.learned.stats()["fallbacks"] == 0
    assert eu < 1e-5 and ev < 1e-5
    assert ep < 1e-4 * prange
    sim.close()
    ref.close()
    _modes_timed(n, ("net", "vcycle", "none"), steps=8, warm=4, re=re)


def test_extrapolated_start(gpu):
    """x0 = "extrapolate": each solve starts from x + (dt/dt_old)(x - x_prev);
    the solves still end on a true relative residual <= 1e-6, the velocities
    agree with the reference's warm start (the last solution) to the solve
    tolerance, and they take fewer iterations."""
    from cfd2d_amd.learned import LearnedConfig

    prev = _run(_cavity(256, "cg_learned", LearnedConfig(x0="previous")), 25)
    ext = _run(_cavity(256, "cg_learned", LearnedConfig(x0="extrapolate")), 25)
    ip, ie = np.mean(prev.cg_iterations[5:]), np.mean(ext.cg_iterations[5:])
    print(f"256^2 steps 6-25: {ie:.2f} iterations extrapolated, {ip:.2f} from the last solution; "
          f"max|du| {_cmp(ext.u, prev.u):.1e}")
    assert max(ext.cg_relres) <= 1e-6 and all(i == 0 for i in ext.cg_info)
    assert ie < 0.8 * ip
    assert _cmp(ext.u, prev.u) < 1e-6 and _cmp(ext.v, prev.v) < 1e-6



@pytest.mark.parametrize("solver", ["cg", "cg_mixed"])
def test_extrapolated_start_other_solvers(gpu, solver):
    """The extrapolated initial iterate applies to every constant-density
    solver (pressure_x0="extrapolate"): the FP64 Jacobi-PCG and the multigrid
    FCG reach the same tolerance from it in fewer iterations, with the
    velocities of the last-solution start to the solve tolerance."""
    from cfd2d_amd import cases

    def run(x0):
        sim = cases.lid_driven_cavity(re=1000.0, n=128, pressure_solver=solver,
                                      pressure_tol=1e-8, pressure_maxiter=100000,
                                      pressure_x0=x0).sim
        return _run(sim, 25)

    prev, ext = run("previous"), run("extrapolate")
    ip, ie = np.mean(prev.cg_iterations[5:]), np.mean(ext.cg_iterations[5:])
    print(f"{solver} 128^2 steps 6-25: {ie:.1f} iterations extrapolated, {ip:.1f} previous")
    assert ext._ctx.initial_iterate() == (1, 1)
    assert all(i == 0 for i in ext.cg_info)
    assert ie < ip
    assert _cmp(ext.u, prev.u) < 1e-7 and _cmp(ext.v, prev.v) < 1e-7
