"""Benchmark: MAC projection steps per second on a 1024x1024 lid-driven cavity.

Workload (one GPU): lid-driven cavity Re=1000 on 1024x1024, AB2, gamma "auto",
tau 0.5, with a passive van Leer scalar (D = 1e-3) initialised as 8 Gaussian
blobs (seed 0).  Every step is the whole projection step (BCs, dt/gamma,
predictor, pressure solve to relative residual 1e-6, correction, scalar
transport, monitors), all inputs resident in HBM.  The timed steps are
warmup+1 .. warmup+steps from rest (``timed_steps``).  Solver legs:

* ``cg``                 plain FP64 Jacobi-PCG (the reference's method, the
                         baseline), warm-started from the last solution as the
                         reference's CG is; ``cg_x0extrap`` the same from the
                         extrapolated start of the mixed-precision legs
                         (``speedup_same_x0``: the like-for-like ratio);
* ``cg_learned_vcycle``  flexible CG with FP64 vectors and dots and an FP32
                         aggregation V-cycle down to the dense coarsest solve
                         (coarse levels in LDS): the headline ``value``;
* ``cg_learned``         the same with the BF16 U-Net coarse correction on the
                         32x32 level (shipped trained weights) instead of the
                         levels below it; ``cg_learned_none`` with neither;
* ``cg_mixed``           the V-cycle FCG as one persistent cooperative launch.

``ensemble`` adds the batched-ensemble workload (256 lid / heated cavities of
sides 64-512, FP32 bulk, shared learned PCG) as aggregate cell updates per
second, with and without the network.

Prints one JSON line.  ``roofline`` describes the pressure projection of the
headline leg (it takes nearly all of the step): algorithmic bytes per
projection (per cell: 80 B for the right-hand side, pressure write-back and
correction, plus per iteration the FP64 outer vectors and the FP32 V-cycle,
counted per kernel below) over its mean duration, timed with HIP events on
the launch stream.  ``errors``: max-abs u, v, p, phi of the headline, network
and FP64 legs after the same steps from rest against the FP64 oracle (exact
DCT solves).  ``cpu_baseline`` times the CPU FP64 oracle: one pinned core with
SciPy's CG (``value``) and every core with the OpenMP C CG, on a sample of
the 1024^2 workload and on config C1 (64^2, 2000 steps) in full.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for p in (str(ROOT), str(ROOT / "cfd-_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

N = 1024
RE = 1000.0
RTOL = 1e-6
HBM_PEAK_GBS = 8000.0
# measured HBM traffic (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a 256 MiB
# copy) over the algorithmic bytes of the same projection kernels
TRAFFIC_RATIO_MGCG = 1.10        # profiles/k_project_traffic_cavity1024.md
TRAFFIC_RATIO_VCYCLE = 1.014     # steps 6-25, profiles/learned_vcycle_projection_cavity1024.md
WORKLOAD = ("lid-driven cavity Re=1000, 1024x1024 MAC grid, AB2, gamma auto, tau 0.5, "
            "passive van Leer scalar (D=1e-3, 8 Gaussian blobs, seed 0), "
            "pressure solve to relative residual 1e-6")


def gaussian_blobs(n: int, seed: int = 0, count: int = 8) -> np.ndarray:
    """Initial scalar at cell centres: sum of Gaussians with centres U(0.1,0.9)^2,
    sigma U(0.02,0.08), amplitude U(0,1), drawn in that order."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.1, 0.9, (count, 2))
    s = rng.uniform(0.02, 0.08, count)
    a = rng.uniform(0.0, 1.0, count)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    phi = np.zeros_like(X)
    for k in range(count):
        phi += a[k] * np.exp(-((X - c[k, 0]) ** 2 + (Y - c[k, 1]) ** 2) / (2 * s[k] ** 2))
    return phi


# leg -> (pressure solver, learned-solver settings, initial iterate).  The
# learned legs start each solve from the linear extrapolation in time of the
# last two solutions; cg_learned_x0prev is the headline solver with the
# reference's warm start (the last solution); cg and cg_mixed keep that start,
# and cg_x0extrap is the FP64 baseline with the learned legs' start (the
# like-for-like ratio).  cg_learned_vcycle replaces the network by the rest of
# the V-cycle down to the dense coarsest solve (the method without the
# network); cg_learned_none drops the coarse solve below 32x32 altogether.
LEGS = {"cg": ("cg", None, "previous"),
        "cg_x0extrap": ("cg", None, "extrapolate"),
        "cg_mixed": ("cg_mixed", None, "previous"),
        "cg_learned": ("cg_learned", {"mode": "net"}, "extrapolate"),
        "cg_learned_vcycle": ("cg_learned", {"mode": "vcycle"}, "extrapolate"),
        "cg_learned_none": ("cg_learned", {"mode": "none"}, "extrapolate"),
        "cg_learned_x0prev": ("cg_learned", {"mode": "net"}, "previous")}
# the same workload with a turbulence model on the device (variable viscosity,
# nu_t per step; SST also transports k and omega): the cost of the model
TURB_LEGS = {"cg_learned_smagorinsky": ("cg_learned_vcycle", "Smagorinsky"),
             "cg_learned_sst": ("cg_learned_vcycle", "kOmegaSST")}


USE_GRAPH = True   # learned legs: replay each iteration as a captured HIP graph
LEG_DETAIL = {
    "cg_learned_vcycle": ("flexible CG, FP64 vectors and dots, FP32 aggregation V-cycle whose "
                          "coarse end (levels <= 64x64 to the dense coarsest solve) runs in LDS; "
                          "U-Net not used: it loses to this complete coarse solve"),
    "cg_learned": ("flexible CG, FP64 vectors and dots, FP32 V-cycle with the BF16 U-Net "
                   "coarse correction on the 32x32 level"),
}


def build_case(leg: str, n: int = N):
    from cfd2d_amd import cases

    turb = None
    if leg in TURB_LEGS:
        base, turb = TURB_LEGS[leg]
        solver, learned, x0 = LEGS[base]
    else:
        solver, learned, x0 = LEGS[leg]
    if learned is not None:
        learned = dict(learned, use_graph=USE_GRAPH)
    case = cases.lid_driven_cavity(re=RE, n=n, pressure_solver=solver, pressure_tol=RTOL,
                                   pressure_maxiter=50000, learned=learned, pressure_x0=x0)
    if turb is not None:
        cases.add_turbulence(case, turb)
    cases.add_scalar(case, "c", diffusivity=1e-3, limiter="vanleer", initial=gaussian_blobs(n))
    return case


VOF_N = 256         # the two-phase leg: dam break on 256 x 256 (variable-density solve)


def run_vof_leg(steps: int, warmup: int) -> dict:
    """Dam break (water column, obstacle, CSF, open top) on VOF_N^2: a few
    steps, the variable-density pressure solve dominating."""
    from cfd2d_amd import cases

    sim = cases.dam_break(n=VOF_N, t_end=1.0).sim
    ctx = sim._ctx
    prm = sim._params()
    steps, warmup = min(steps, 40), min(warmup, 10)
    ctx.step(prm, warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(prm, steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    st = ctx.get_state()
    recs = ctx.records(st.n - steps + 1, steps)
    iters = np.array([r.cg_iters for r in recs], dtype=float)
    out = {"solver": "vof_dam_break", "grid": [VOF_N, VOF_N], "steps": steps,
           "steps_per_s": steps / wall, "ms_per_step": 1e3 * wall / steps,
           "cg_iters_mean": float(iters.mean()),
           "not_converged_steps": int(sum(1 for r in recs if r.cg_info != 0)),
           "max_div": float(max(r.divergence for r in recs))}
    sim.close()
    return out


def run_leg(solver: str, steps: int, warmup: int) -> dict:
    if solver == "vof_dam_break":
        return run_vof_leg(steps, warmup)
    case = build_case(solver)
    sim = case.sim
    ctx = sim._ctx
    prm = sim._params()
    t0 = time.perf_counter()
    ctx.step(prm, warmup)
    ctx.sync()
    t_warm = time.perf_counter() - t0
    ctx.set_profiling(True)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(prm, steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    proj_ms, launches = ctx.projection_times()
    ctx.set_profiling(False)
    st = ctx.get_state()
    recs = ctx.records(st.n - steps + 1, steps)
    iters = np.array([r.cg_iters for r in recs], dtype=float)
    relres = np.array([r.cg_relres for r in recs])
    info = [r.cg_info for r in recs]
    divmax = max(r.divergence for r in recs)
    out = {
        "solver": solver,
        "x0": LEGS[TURB_LEGS[solver][0] if solver in TURB_LEGS else solver][2],
        "turbulence": TURB_LEGS[solver][1] if solver in TURB_LEGS else None,
        "steps_per_s": steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "cg_iters_mean": float(iters.mean()),
        "cg_iters_max": int(iters.max()),
        "cg_relres_max": float(relres.max()),
        "not_converged_steps": int(sum(1 for k in info if k != 0)),
        "max_div": float(divmax),
        "proj_ms_mean": proj_ms / max(launches, 1),
        "proj_launches": launches,
        "warmup_s": t_warm,
        "t_end": float(st.t),
        "n_steps_total": int(st.n),
    }
    sim.close()
    return out


def roofline_for(leg: dict) -> dict:
    n_cells = N * N
    s = 8.0
    n_it = leg["cg_iters_mean"]
    if leg["solver"].startswith("cg_learned"):
        # per fine cell, what each kernel must read and write once (FP64 8 B,
        # FP32 4 B, cell codes and component ids 1 B):
        # - start: the extrapolated iterate (x and x_prev read and written,
        #   32 B, with x0 "extrapolate"), the right-hand side (F, G, codes,
        #   components read, b written: 26 B), the initial residual (b, x,
        #   codes, components read; b, r and the V-cycle input written: 42 B);
        # - end: the x sum (10 B), the pressure write-back with the velocity
        #   correction (x, codes, components, F, G, u, v read; p, u, v written:
        #   66 B) and the last iteration's true residual (b, x, codes read; r
        #   and the V-cycle input written: 32 B);
        # - per iteration: the level-0 down-leg (b, codes read; t and the
        #   level-1 right-hand side written: 12 B), the level-0 up-leg (t, b,
        #   codes, the coarse correction, r and r_old read; z written: 30 B),
        #   the direction (z, d_old, codes read; d and q written: 29 B), the
        #   update (x, d, r, q, codes read; x, r_old, r and the V-cycle input
        #   written: 61 B), levels 1-3 with their coefficients (47 B per
        #   coarse cell: 15 B per fine cell) and the coarse end (< 1 B).
        x0_bytes = 32.0 if leg.get("x0") == "extrapolate" else 0.0
        bytes_per_launch = n_cells * (176.0 + x0_bytes + n_it * 148.0)
    elif leg["solver"] == "cg_mixed":
        # per flexible-CG iteration: FP64 outer vectors (r, r_old, x, p, q: 104 B/cell)
        # plus the FP32 V-cycle (fine level 50 B/cell, coarse levels 1/3 more
        # with their 12 B/cell of coefficients: ~66 B/cell)
        bytes_per_launch = n_cells * (8.0 * (3 + 7) + n_it * (104.0 + 66.0))
    else:
        bytes_per_launch = n_cells * s * (3 + 7 + 10 * n_it)
    t = leg["proj_ms_mean"] * 1e-3
    achieved = bytes_per_launch / t / 1e9 if t > 0 else 0.0
    learned = leg["solver"].startswith("cg_learned")
    # measured HBM traffic (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated against a
    # 256 MiB copy) over the algorithmic bytes of the same launches
    if leg["solver"] == "cg_learned_vcycle":
        traffic = round(TRAFFIC_RATIO_VCYCLE * bytes_per_launch)
        source = ("rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE summed over the projection kernels "
                  "of steps 6-25 (the --warmup 5 --steps 20 window): 819.1 MB per step at "
                  "3.8 iterations, 1.014 times the algorithmic bytes; "
                  "profiles/learned_vcycle_projection_cavity1024.md")
    elif leg["solver"] == "cg_mixed":
        traffic = round(TRAFFIC_RATIO_MGCG * bytes_per_launch)
        source = ("rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE of 6-iteration launches "
                  "(ratio 1.10 to algorithmic), profiles/k_project_traffic_cavity1024.md")
    else:
        traffic, source = None, None
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": source,
            "kernel": ("learned projection: the start and end launches (k_lp_*, the "
                       "extrapolated initial iterate in k_lp_rhs) and the flexible-CG "
                       "iteration kernels"
                       if learned else "k_project"),
            "bytes_per_launch": bytes_per_launch,
            "launch_ms": round(leg["proj_ms_mean"], 4), "cg_iters_per_launch": round(n_it, 2)}


def run_ensemble(steps: int = 10, warmup: int = 3, net: bool = True,
                 coarse: str = "network", use_graph: bool | None = None) -> dict:
    """The batched-ensemble workload: 256 lid / heated cavities of sides
    {64, 128, 256, 512} (seed 0), FP32 bulk, the learned solver's flexible CG
    with the coarse end `coarse` ("vcycle": the V-cycle to the dense solve;
    "network": the shared U-Net, or the constant-vector correction alone
    without `net`); aggregate cell updates per second over `steps` steps after
    `warmup`, and the roofline of the step."""
    from cfd2d_amd.ensemble import Ensemble, sample_instances

    ins = sample_instances(256, 0)
    t0 = time.perf_counter()
    graph = USE_GRAPH if use_graph is None else use_graph
    ens = Ensemble(ins, net=net, use_graph=graph, coarse=coarse)
    setup = time.perf_counter() - t0
    ens.step(warmup)
    s0 = ens.stats(0)["sweeps_total"]
    t0 = time.perf_counter()
    ens.step(steps)
    wall = time.perf_counter() - t0
    sweeps = ens.stats(0)["sweeps_total"] - s0
    rel = max(ens.stats(k)["relres"] for k in range(len(ins)))
    # algorithmic bytes per step: per cell 4-byte bulk fields -- predictor 8,
    # right-hand side 3, correction 7, temperature 4 (both families are
    # counted with it) -- plus per lockstep iteration the FP64 outer vectors
    # and the FP32 V-cycle as in the single case (190 B); the lockstep
    # iterations bound the per-instance ones from above
    it = sweeps / steps
    bytes_step = ens.cells * (4.0 * (8 + 3 + 7 + 4) + it * 190.0)
    gbs = bytes_step / (wall / steps) / 1e9
    out = {"coarse": coarse if (net or coarse == "vcycle") else "none",
           "launch": ("one graph launch per step, iterations controlled on the device"
                      if graph else "direct launches, the host polls the solves"),
           "instances": len(ins), "cells": ens.cells, "steps": steps, "warmup": warmup,
           "ms_per_step": round(1e3 * wall / steps, 3),
           "cell_updates_per_s": round(ens.cells * steps / wall, 1),
           "solver_iterations_per_step": round(sweeps / steps, 2),
           "relres_max": rel, "setup_s": round(setup, 2),
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                        "bytes_per_step": bytes_step},
           "predictor_shapes": {str(ins[c[0]].n): [ens.stats(c[0])["pred_tx"],
                                                    ens.stats(c[0])["pred_ty"]]
                                for c in ens.classes},
           "tuning": [ens.tuning(c) for c in range(len(ens.classes))]}
    ens.close()
    return out


# --------------------------------------------------------------------------- #
# CPU baseline and accuracy: the FP64 oracle (NumPy/SciPy + C kernels)
# --------------------------------------------------------------------------- #
# config C1: 64^2 cavity Re 100, gamma 0, AB2, tau 0.5, CG rtol 1e-8, 2000 steps
C1 = {"re": 100.0, "n": 64, "gamma": 0.0, "scheme": "ab2", "tau": 0.5, "tol": 1e-8,
      "steps": 2000}
ERR_STEPS = 25      # steps from rest after which the per-field errors are taken


def host_cores() -> int:
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()


def _oracle_c2(solver: str):
    from oracle import cfd_oracle as O

    sim = O.cavity(re=RE, n=N, pressure_solver=solver, pressure_tol=RTOL)
    sim.pressure.maxiter = 50000
    O.add_scalar(sim, "c", 1e-3, limiter="vanleer", initial=gaussian_blobs(N))
    return sim


def _oracle_c1(solver: str):
    from oracle import cfd_oracle as O

    sim = O.cavity(re=C1["re"], n=C1["n"], gamma=C1["gamma"], scheme=C1["scheme"],
                   tau=C1["tau"], pressure_solver=solver, pressure_tol=C1["tol"])
    sim.pressure.maxiter = 50000
    return sim


def cpu_child(mode: str, arg: str, budget_s: float) -> dict:
    """One CPU measurement in a child process (its thread count and pinning
    set by the parent).  c2: steps of the 1024^2 workload from rest until the
    budget (at least `arg` steps); c1: config C1 in full, then the max-abs
    error of the GPU's C1 fields (npz at `arg`, if given) against it;
    errors: the oracle with exact (DCT) solves for the steps of the npz at
    `arg`, and each GPU leg's max-abs field errors against it."""
    t0 = time.perf_counter()
    if mode in ("c2_1", "c2_all"):
        sim = _oracle_c2("cg" if mode == "c2_1" else "cg_omp")
        setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        k = 0
        while k < int(arg) or time.perf_counter() - t0 < budget_s:
            sim.step()
            k += 1
        return {"steps": k, "wall_s": time.perf_counter() - t0, "setup_s": setup,
                "cg_iters": sim.cg_iters}
    if mode in ("c1_1", "c1_all"):
        sim = _oracle_c1("cg" if mode == "c1_1" else "cg_omp")
        t0 = time.perf_counter()
        for _ in range(C1["steps"]):
            sim.step()
        out = {"steps": C1["steps"], "wall_s": time.perf_counter() - t0,
               "cg_iters_mean": float(np.mean(sim.cg_iters))}
        if arg != "-":
            g = np.load(arg)
            out["errors"] = {f: float(np.abs(g[f] - getattr(sim, f)).max())
                             for f in ("u", "v", "p")}
        return out
    if mode == "errors":
        g = np.load(arg)
        steps = int(g["steps"])
        sim = _oracle_c2("dct")
        for _ in range(steps):
            sim.step()
        ref = {"u": sim.u, "v": sim.v, "p": sim.p, "phi": sim.scalars["c"].phi}
        out = {"steps": steps, "wall_s": time.perf_counter() - t0}
        for leg in str(g["legs"]).split(","):
            out[leg] = {f: float(np.abs(g[f"{leg}_{f}"] - ref[f]).max()) for f in ref}
        return out
    raise ValueError(mode)


def run_cpu_child(mode: str, arg: str, budget_s: float, threads: int, pin: bool,
                  timeout: float = 900.0) -> dict | None:
    env = dict(os.environ)
    t = str(threads)
    env.update({"OMP_NUM_THREADS": t, "OPENBLAS_NUM_THREADS": t, "MKL_NUM_THREADS": t})
    cmd = [sys.executable, str(Path(__file__).resolve()), "--cpu-child", mode, arg, str(budget_s),
           "1" if pin else "0"]
    try:
        res = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=timeout)
        line = [l for l in res.stdout.splitlines() if l.startswith("{")][-1]
        return json.loads(line)
    except Exception as e:  # pragma: no cover - reported, not fatal
        print(f"cpu child {mode} failed: {e}", file=sys.stderr, flush=True)
        return None


def gpu_fields(leg: str, steps: int) -> dict:
    """u, v, p and the scalar of a GPU leg after `steps` steps from rest."""
    sim = build_case(leg).sim
    ctx = sim._ctx
    ctx.step(sim._params(), steps)
    ctx.sync()
    sim._flush()
    out = {"u": sim.u.copy(), "v": sim.v.copy(), "p": sim.p.copy(),
           "phi": sim.scalars["c"].phi.copy()}
    sim.close()
    return out


def gpu_c1() -> tuple[dict, float]:
    """Config C1 on the GPU (FP64 Jacobi-PCG, the reference's method): the
    final fields and the steps per second."""
    from cfd2d_amd import cases

    sim = cases.lid_driven_cavity(re=C1["re"], n=C1["n"], gamma=C1["gamma"],
                                  scheme=C1["scheme"], tau=C1["tau"], pressure_solver="cg",
                                  pressure_tol=C1["tol"], pressure_maxiter=50000).sim
    ctx = sim._ctx
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(sim._params(), C1["steps"])
    ctx.sync()
    wall = time.perf_counter() - t0
    sim._flush()
    out = {"u": sim.u.copy(), "v": sim.v.copy(), "p": sim.p.copy()}
    sim.close()
    return out, C1["steps"] / wall


def cpu_baseline(gpu_c1_path: str | None, c1_gpu_rate: float | None) -> dict | None:
    """The FP64 oracle on the host: (i) one pinned core with SciPy's CG (the
    reference's single-threaded configuration), (ii) every core of this
    process with the C Jacobi-CG and the gather-form kernels on OpenMP.  C2
    (the 1024^2 workload) is sampled from rest; C1 runs in full."""
    cores = host_cores()
    c2_1 = run_cpu_child("c2_1", "1", 20.0, 1, True)
    c2_all = run_cpu_child("c2_all", "3", 20.0, cores, False)
    c1_1 = run_cpu_child("c1_1", gpu_c1_path or "-", 0.0, 1, True)
    c1_all = run_cpu_child("c1_all", "-", 0.0, cores, False)
    if c2_1 is None:
        return None

    def rate(r):
        return round(r["steps"] / r["wall_s"], 4) if r else None

    out = {"value": rate(c2_1), "unit": "steps/s", "cores": 1, "kind": "port",
           "sample": (f"FP64 oracle (C kernels + SciPy Jacobi-CG, rtol 1e-6), first "
                      f"{c2_1['steps']} step(s) of the same 1024x1024 workload from rest, "
                      f"{c2_1['wall_s']:.1f} s on one pinned core; CG iterations "
                      f"{c2_1['cg_iters']}.  steps/s = steps / wall time of the sample; "
                      f"the first steps from rest need the most CG iterations, so the "
                      f"sample understates the rate of later steps"),
           "host_cores": cores}
    if c2_all:
        out["all_cores"] = {"value": rate(c2_all), "unit": "steps/s", "cores": cores,
                            "sample": (f"first {c2_all['steps']} steps from rest, "
                                       f"{c2_all['wall_s']:.1f} s, C Jacobi-CG and gather-form "
                                       f"kernels on {cores} OpenMP threads; CG iterations "
                                       f"{c2_all['cg_iters']}")}
    c1 = {"workload": "config C1: 64x64 cavity Re 100, gamma 0, AB2, tau 0.5, CG rtol 1e-8, "
                      "2000 steps from rest"}
    if c1_1:
        c1["one_core_steps_per_s"] = rate(c1_1)
        c1["cg_iters_mean"] = round(c1_1["cg_iters_mean"], 2)
        if "errors" in c1_1:
            c1["gpu_fp64_errors"] = c1_1["errors"]
    if c1_all:
        c1["all_cores_steps_per_s"] = rate(c1_all)
    if c1_gpu_rate:
        c1["gpu_fp64_steps_per_s"] = round(c1_gpu_rate, 2)
    out["config1"] = c1
    return out


def field_errors(legs: list[str], steps: int, tmpdir: Path) -> dict | None:
    """Max-abs errors of u, v, p, phi after `steps` steps from rest: each GPU
    leg against the FP64 oracle with exact (DCT) pressure solves, and the
    first leg against the GPU FP64 Jacobi-PCG leg ("cg")."""
    data = {"steps": np.int64(steps), "legs": np.array(",".join(legs))}
    fields = {}
    for leg in legs:
        fields[leg] = gpu_fields(leg, steps)
        for f, a in fields[leg].items():
            data[f"{leg}_{f}"] = a
    path = tmpdir / "gpu_fields.npz"
    np.savez(path, **data)
    r = run_cpu_child("errors", str(path), 0.0, host_cores(), False)
    if r is None:
        return None
    out = {"steps_from_rest": steps, "reference": "FP64 oracle, exact (DCT) pressure solves",
           "vs_oracle": {leg: r[leg] for leg in legs}}
    if "cg" in fields and legs[0] != "cg":
        out["vs_gpu_fp64_cg"] = {f: float(np.abs(fields[legs[0]][f] - fields["cg"][f]).max())
                                 for f in fields["cg"]}
    return out


def copy_calibration(nbytes: int = 256 << 20) -> None:
    """One device-to-device copy of `nbytes` (rocclr's copy kernel reads and
    writes exactly that many bytes): the reference dispatch for scaling the
    HBM byte counters of the projection kernel."""
    import ctypes

    from cfd2d_amd import _lib

    L = _lib.lib()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(L.cfd_dev_alloc(nbytes, ctypes.byref(a)), "alloc")
    _lib.check(L.cfd_dev_alloc(nbytes, ctypes.byref(b)), "alloc")
    _lib.check(L.cfd_memcpy(b, a, nbytes, 3), "copy")
    L.cfd_dev_free(a)
    L.cfd_dev_free(b)
    print(f"copy calibration: {nbytes} bytes device-to-device", file=sys.stderr, flush=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--legs", default="cg,cg_x0extrap,cg_learned,cg_learned_vcycle,"
                                      "cg_learned_none,cg_learned_x0prev,cg_mixed,"
                                      "cg_learned_smagorinsky,cg_learned_sst,vof_dam_break")
    ap.add_argument("--cpu-child", nargs=4, default=None,
                    metavar=("MODE", "ARG", "BUDGET_S", "PIN"))
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device-to-device (a known-size kernel for "
                         "scaling rocprof FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--ensemble-only", choices=("graph", "direct"), default=None,
                    help="run only the 256-cavity ensemble with the V-cycle coarse end (one graph "
                         "launch per step, or direct launches for profilers) and print its JSON")
    ap.add_argument("--no-graph", action="store_true",
                    help="learned legs launch their iteration kernels directly instead of "
                         "replaying a captured graph (for profilers that do not follow graphs)")
    args = ap.parse_args()
    global USE_GRAPH
    USE_GRAPH = not args.no_graph
    if args.cpu_child:
        mode, arg, budget, pin = args.cpu_child
        if pin == "1" and hasattr(os, "sched_setaffinity"):
            os.sched_setaffinity(0, {min(os.sched_getaffinity(0))})
        print(json.dumps(cpu_child(mode, arg, float(budget))))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the projection "
              "step and its benchmark use a single GPU", file=sys.stderr)
        sys.exit(2)

    if args.ensemble_only:
        print(json.dumps(run_ensemble(coarse="vcycle", use_graph=args.ensemble_only == "graph")))
        return
    if args.copy_calibration:
        copy_calibration()
    legs = {}
    for solver in args.legs.split(","):
        try:
            legs[solver] = run_leg(solver, args.steps, args.warmup)
        except ValueError as e:
            legs[solver] = {"solver": solver, "error": str(e)}
        print(f"leg {solver}: {json.dumps(legs[solver])}", file=sys.stderr, flush=True)
    def good(name):
        leg = legs.get(name)
        return leg if leg is not None and "error" not in leg else None

    # the headline: the fastest mixed-precision flexible CG of the learned
    # solver's family -- with the complete V-cycle coarse end (which beats the
    # network's approximate coarse solve), else with the network
    head = (good("cg_learned_vcycle") or good("cg_learned") or good("cg_mixed") or good("cg")
            or next((good(k) for k in legs if good(k)), None))
    if head is None:
        print("no solver leg ran", file=sys.stderr)
        sys.exit(1)
    base = good("cg")
    base_x0 = good("cg_x0extrap")
    none = good("cg_learned_none")
    vcyc = good("cg_learned_vcycle")
    net = good("cg_learned")
    import tempfile

    tmp = Path(tempfile.mkdtemp(prefix="cfd_bench_"))
    errors = c1_path = c1_rate = None
    if not args.no_cpu:
        err_legs = [k for k in (head["solver"], "cg_learned", "cg") if good(k)]
        err_legs = list(dict.fromkeys(err_legs))
        errors = field_errors(err_legs, min(ERR_STEPS, args.warmup + args.steps), tmp)
        print(f"errors: {json.dumps(errors)}", file=sys.stderr, flush=True)
        c1_fields, c1_rate = gpu_c1()
        c1_path = str(tmp / "gpu_c1.npz")
        np.savez(c1_path, **c1_fields)
    ensemble = None
    if not args.no_ensemble:
        try:
            ensemble = run_ensemble(coarse="vcycle")
            if USE_GRAPH:
                ensemble["direct_launches"] = run_ensemble(coarse="vcycle", use_graph=False)
            ensemble["with_network"] = run_ensemble(net=True)
            ensemble["without_network"] = run_ensemble(net=False)
        except (ValueError, RuntimeError) as e:
            ensemble = {"error": str(e)}
        print(f"ensemble: {json.dumps(ensemble)}", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline(c1_path, c1_rate)
    print(f"cpu baseline: {json.dumps(cpu)}", file=sys.stderr, flush=True)
    line = {
        "metric": "MAC steps/s @1024² cavity: learned-PCG fp32 vs plain fp64 CG "
                  "(resid 1e-6)",
        "value": round(head["steps_per_s"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "timed_steps": [args.warmup + 1, args.warmup + args.steps],
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"cg_learned": "f64+f32+bf16", "cg_learned_vcycle": "f64+f32",
                  "cg_mixed": "f64+f32"}.get(head["solver"], "f64"),
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "grid": [N, N], "solver": head["solver"],
                   "solver_detail": LEG_DETAIL.get(head["solver"], head["solver"]),
                   "rtol": RTOL},
        "roofline": roofline_for(head) if "proj_ms_mean" in head else None,
        "cpu_baseline": cpu,
        "fp64_cg_steps_per_s": round(base["steps_per_s"], 3) if base and "error" not in base
        else None,
        "speedup_vs_fp64_cg": (round(head["steps_per_s"] / base["steps_per_s"], 3)
                               if base and "error" not in base else None),
        "speedup_same_x0": (round(head["steps_per_s"] / base_x0["steps_per_s"], 3)
                            if base_x0 and head.get("x0") == base_x0["x0"] else None),
        "errors": errors,
        # the U-Net coarse correction against the same solver with the V-cycle
        # carried on to the dense coarsest solve (the complete method without
        # the network) and against no coarse solve below 32x32 at all
        "network_gain": ({"iters_net": round(net["cg_iters_mean"], 2),
                          "steps_per_s_net": round(net["steps_per_s"], 3),
                          "speedup_net_vs_fp64_cg": (round(net["steps_per_s"]
                                                           / base["steps_per_s"], 3)
                                                     if base else None),
                          "iters_vcycle": round(vcyc["cg_iters_mean"], 2) if vcyc else None,
                          "steps_per_s_vcycle": round(vcyc["steps_per_s"], 3) if vcyc else None,
                          "iters_none": round(none["cg_iters_mean"], 2) if none else None,
                          "steps_per_s_none": round(none["steps_per_s"], 3) if none else None}
                         if net and (none or vcyc) else None),
        "ensemble": ensemble,
        "legs": legs,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
