"""Benchmark: MAC projection steps per second on a 1024x1024 lid-driven cavity.

Workload (one GPU): lid-driven cavity Re=1000 on 1024x1024, AB2, gamma "auto",
tau 0.5, with a passive van Leer scalar (D = 1e-3) initialised as 8 Gaussian
blobs (seed 0).  Every step is the whole projection step (BCs, dt/gamma,
predictor, pressure solve to relative residual 1e-6, correction, scalar
transport, monitors), all inputs resident in HBM.  Solver legs:

* ``cg``               plain FP64 Jacobi-PCG (the reference's method, the baseline);
* ``cg_learned``       the learned flexible CG (FP64 outer, FP32 V-cycle, BF16
                       U-Net coarse correction; the shipped trained weights): the
                       headline ``value``;
* ``cg_learned_none``  the same solver without the network (the gain it brings);
* ``cg_mixed``         FP64 flexible CG with the FP32 multigrid V-cycle.

``ensemble`` adds the batched-ensemble workload (256 lid / heated cavities of
sides 64-512, FP32 bulk, shared learned PCG) as aggregate cell updates per
second, with and without the network.

Prints one JSON line.  ``roofline`` describes the pressure projection of the
headline leg (it takes nearly all of the step): algorithmic bytes per
projection (per cell: 80 B for the right-hand side, pressure write-back and
correction, plus per iteration the FP64 outer vectors and the FP32 V-cycle,
counted per kernel below) over its mean duration, timed with HIP events on
the launch stream.  ``cpu_baseline`` times the CPU FP64 oracle (SciPy CG) on
one host core for a bounded number of steps of the same workload.
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
TRAFFIC_RATIO_MGCG = 1.10
TRAFFIC_RATIO_LEARNED = 1.20
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
TURB_LEGS = {"cg_learned_smagorinsky": ("cg_learned", "Smagorinsky"),
             "cg_learned_sst": ("cg_learned", "kOmegaSST")}


USE_GRAPH = True   # learned legs: replay each iteration as a captured HIP graph


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
        # per learned-FCG iteration and fine cell: V-cycle sweeps on level 0
        # (down sweep 13 B, restriction 11 B, prolongation sweep 14 B, last sweep
        # with the r.z and r_old.z partials 29 B), FP64 direction + A d (29 B),
        # FP64 update of x, r, r_old plus the next V-cycle input (73 B); coarse
        # levels add a third of the level-0 float traffic with their
        # coefficients (21 B); the U-Net on the 32x32 level is negligible.
        # The extrapolated initial iterate reads and writes x and x_prev (32 B)
        x0_bytes = 32.0 if leg.get("x0") == "extrapolate" else 0.0
        bytes_per_launch = n_cells * (8.0 * (3 + 7) + x0_bytes + n_it * (169.0 + 21.0))
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
    if learned:
        traffic = round(TRAFFIC_RATIO_LEARNED * bytes_per_launch)
        source = ("rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE summed over the projection kernels "
                  "of 10 steps (ratio 1.20 to algorithmic), "
                  "profiles/learned_projection_cavity1024.md")
    elif leg["solver"] == "cg_mixed":
        traffic = round(TRAFFIC_RATIO_MGCG * bytes_per_launch)
        source = ("rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE of 6-iteration launches "
                  "(ratio 1.10 to algorithmic), profiles/k_project_traffic_cavity1024.md")
    else:
        traffic, source = None, None
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": source,
            "kernel": ("learned projection: the start and end launches (k_lp_*), the "
                       "extrapolated initial iterate and the flexible-CG iteration kernels"
                       if learned else "k_project"),
            "bytes_per_launch": bytes_per_launch,
            "launch_ms": round(leg["proj_ms_mean"], 4), "cg_iters_per_launch": round(n_it, 2)}


def run_ensemble(steps: int = 10, warmup: int = 3, net: bool = True) -> dict:
    """The batched-ensemble workload: 256 lid / heated cavities of sides
    {64, 128, 256, 512} (seed 0), FP32 bulk, shared learned PCG; aggregate
    cell updates per second over `steps` steps after `warmup`."""
    from cfd2d_amd.ensemble import Ensemble, sample_instances

    ins = sample_instances(256, 0)
    t0 = time.perf_counter()
    ens = Ensemble(ins, net=net, use_graph=USE_GRAPH)
    setup = time.perf_counter() - t0
    ens.step(warmup)
    s0 = ens.stats(0)["sweeps_total"]
    t0 = time.perf_counter()
    ens.step(steps)
    wall = time.perf_counter() - t0
    sweeps = ens.stats(0)["sweeps_total"] - s0
    rel = max(ens.stats(k)["relres"] for k in range(len(ins)))
    out = {"instances": len(ins), "cells": ens.cells, "steps": steps, "warmup": warmup,
           "ms_per_step": round(1e3 * wall / steps, 3),
           "cell_updates_per_s": round(ens.cells * steps / wall, 1),
           "solver_iterations_per_step": round(sweeps / steps, 2),
           "relres_max": rel, "setup_s": round(setup, 2),
           "predictor_shapes": {str(ins[c[0]].n): [ens.stats(c[0])["pred_tx"],
                                                    ens.stats(c[0])["pred_ty"]]
                                for c in ens.classes}}
    ens.close()
    return out


# --------------------------------------------------------------------------- #
# CPU baseline: the FP64 oracle (NumPy/SciPy + serial C kernels) on one core
# --------------------------------------------------------------------------- #
def cpu_baseline_child(budget_s: float, max_steps: int) -> dict:
    from oracle import cfd_oracle as O

    t0 = time.perf_counter()
    sim = O.cavity(re=RE, n=N, pressure_solver="cg", pressure_tol=RTOL)
    sim.pressure.maxiter = 50000
    O.add_scalar(sim, "c", 1e-3, limiter="vanleer", initial=gaussian_blobs(N))
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    k = 0
    while k < max_steps and time.perf_counter() - t0 < budget_s:
        sim.step()
        k += 1
    wall = time.perf_counter() - t0
    return {"steps": k, "wall_s": wall, "setup_s": setup, "cg_iters": sim.cg_iters}


def cpu_baseline(budget_s: float = 20.0, max_steps: int = 30) -> dict | None:
    env = dict(os.environ)
    env.update({"OMP_NUM_THREADS": "1", "OPENBLAS_NUM_THREADS": "1", "MKL_NUM_THREADS": "1"})
    cmd = [sys.executable, str(Path(__file__).resolve()), "--cpu-child", str(budget_s),
           str(max_steps)]
    try:
        res = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
        line = [l for l in res.stdout.splitlines() if l.startswith("{")][-1]
        r = json.loads(line)
    except Exception as e:  # pragma: no cover - reported, not fatal
        print(f"cpu baseline failed: {e}", file=sys.stderr)
        return None
    return {"value": round(r["steps"] / r["wall_s"], 4), "unit": "steps/s", "cores": 1,
            "kind": "port",
            "sample": (f"FP64 oracle (serial C kernels + SciPy Jacobi-CG, rtol 1e-6), "
                       f"first {r['steps']} steps of the same 1024x1024 workload from rest, "
                       f"{r['wall_s']:.1f} s, one thread; CG iterations "
                       f"{r['cg_iters']}")}


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
    ap.add_argument("--cpu-child", nargs=2, default=None)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device-to-device (a known-size kernel for "
                         "scaling rocprof FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--no-graph", action="store_true",
                    help="learned legs launch their iteration kernels directly instead of "
                         "replaying a captured graph (for profilers that do not follow graphs)")
    args = ap.parse_args()
    global USE_GRAPH
    USE_GRAPH = not args.no_graph
    if args.cpu_child:
        print(json.dumps(cpu_baseline_child(float(args.cpu_child[0]), int(args.cpu_child[1]))))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path is "
              "single-GPU in this build", file=sys.stderr)
        sys.exit(2)

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

    head = (good("cg_learned") or good("cg_mixed") or good("cg")
            or next((good(k) for k in legs if good(k)), None))
    if head is None:
        print("no solver leg ran", file=sys.stderr)
        sys.exit(1)
    base = good("cg")
    base_x0 = good("cg_x0extrap")
    none = good("cg_learned_none")
    vcyc = good("cg_learned_vcycle")
    ensemble = None
    if not args.no_ensemble:
        try:
            ensemble = run_ensemble(net=True)
            ensemble["without_network"] = run_ensemble(net=False)
        except (ValueError, RuntimeError) as e:
            ensemble = {"error": str(e)}
        print(f"ensemble: {json.dumps(ensemble)}", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline()
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
        "dtype": {"cg_learned": "f64+f32+bf16", "cg_mixed": "f64+f32"}.get(head["solver"], "f64"),
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "grid": [N, N], "solver": head["solver"],
                   "rtol": RTOL},
        "roofline": roofline_for(head) if "proj_ms_mean" in head else None,
        "cpu_baseline": cpu,
        "fp64_cg_steps_per_s": round(base["steps_per_s"], 3) if base and "error" not in base
        else None,
        "speedup_vs_fp64_cg": (round(head["steps_per_s"] / base["steps_per_s"], 3)
                               if base and "error" not in base else None),
        "speedup_same_x0": (round(head["steps_per_s"] / base_x0["steps_per_s"], 3)
                            if base_x0 and head.get("x0") == base_x0["x0"] else None),
        # the network against the same solver with the V-cycle carried on to
        # the dense coarsest solve (the complete method without the network),
        # and against no coarse solve below 32x32 at all
        "network_gain": ({"iters_with": round(head["cg_iters_mean"], 2),
                          "steps_per_s_with": round(head["steps_per_s"], 3),
                          "iters_vcycle": round(vcyc["cg_iters_mean"], 2) if vcyc else None,
                          "steps_per_s_vcycle": round(vcyc["steps_per_s"], 3) if vcyc else None,
                          "iters_none": round(none["cg_iters_mean"], 2) if none else None,
                          "steps_per_s_none": round(none["steps_per_s"], 3) if none else None}
                         if head["solver"] == "cg_learned" and (none or vcyc) else None),
        "ensemble": ensemble,
        "legs": legs,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
