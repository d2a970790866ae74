"""The batched cavity ensemble on the GPU: every instance against its
single-case FP64 run, instances independent of their batch, graph replay,
the shared network's gain, and the full 256-instance workload."""

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mixed():
    from cfd2d_amd.ensemble import Instance

    return [Instance("lid", 64, re=100.0), Instance("heated", 64, ra=1e5),
            Instance("lid", 128, re=1000.0), Instance("heated", 128, ra=1e4),
            Instance("lid", 256, re=3200.0)]


def test_instances_match_single_case_fp64(gpu):
    """FP32 bulk against the FP64 single-case path (parity-checked against the
    oracle) after 30 steps: same t and dt, velocity and temperature within
    1e-5 (FP32 rounding and the 1e-6 solver tolerance)."""
    from cfd2d_amd.ensemble import Ensemble

    ins = _mixed()
    ens = Ensemble(ins)
    ens.step(30)
    for k, inst in enumerate(ins):
        st = ens.stats(k)
        sim = inst.single_case(pressure_solver="cg", pressure_tol=1e-10).sim
        for _ in range(30):
            sim.step()
        sim._flush()
        assert st["steps"] == 30 and st["t"] == pytest.approx(sim.t, rel=1e-12)
        assert st["relres"] <= 1e-6 and st["info"] == 0
        du = np.abs(ens.field(k, "u") - sim.u).max()
        dv = np.abs(ens.field(k, "v") - sim.v).max()
        msg = f"{inst.kind} n={inst.n}: max|du| {du:.2e} max|dv| {dv:.2e}"
        assert du < 1e-5 and dv < 1e-5, msg
        if inst.kind == "heated":
            dT = np.abs(ens.field(k, "T")[1:-1, 1:-1] - sim.scalars["T"].phi[1:-1, 1:-1]).max()
            msg += f" max|dT| {dT:.2e}"
            assert dT < 1e-5, msg
        assert st["kinetic_energy"] == pytest.approx(sim.history.kinetic_energy[-1], rel=1e-4)
        print(msg)
        sim.close()
    ens.close()


def test_instance_independent_of_batch_and_graph(gpu):
    from cfd2d_amd.ensemble import Ensemble

    ins = _mixed()
    full = Ensemble(ins)
    full.step(12)
    alone = Ensemble([ins[3]], use_graph=False)
    alone.step(12)
    assert np.array_equal(alone.field(0, "u"), full.field(3, "u"))
    assert np.array_equal(alone.field(0, "T"), full.field(3, "T"))
    full.close()
    alone.close()


@pytest.mark.parametrize("coarse", ["vcycle", "network"])
def test_autotuned_shapes_change_no_value(gpu, coarse):
    """The autotuner's choices (the predictor's and the V-cycle kernels'
    work-group shapes, the FP64 vector kernels' cells per thread) leave every
    field bit for bit as with the default shapes, and the tuning record holds
    a valid choice for every class and level."""
    from cfd2d_amd.ensemble import Ensemble

    ins = _mixed() + [_heated_512()]
    tuned = Ensemble(ins, coarse=coarse)
    plain = Ensemble(ins, coarse=coarse, autotune=False)
    tuned.step(6)
    plain.step(6)
    for c in range(len(tuned.classes)):
        t = tuned.tuning(c)
        print(t)
        assert t["predictor"][0] * t["predictor"][1] in (256, 512)
        assert all(tx * ty == 256 for tx, ty in t["sweeps"])
        assert t["vector_batch"] in (2, 4, 8, 16)
        assert t["vector_us"] <= t["vector_default_us"]
        assert all(a <= b for a, b in zip(t["sweeps_us"], t["sweeps_default_us"]))
        d = plain.tuning(c)
        assert d["sweeps"] == [[256, 1]] * len(d["sweeps"]) and d["vector_batch"] == 8
    for k in range(len(ins)):
        for f in ("u", "v", "p"):
            assert np.array_equal(tuned.field(k, f), plain.field(k, f)), (k, f)
        assert tuned.stats(k)["iters"] == plain.stats(k)["iters"]
    tuned.close()
    plain.close()


def _heated_512():
    from cfd2d_amd.ensemble import Instance

    return Instance("heated", 512, ra=1e6)


def test_network_gain_in_ensemble(gpu):
    from cfd2d_amd.ensemble import Ensemble

    its = {}
    for net in (True, False):
        ens = Ensemble(_mixed(), net=net)
        ens.step(15)
        its[net] = ens.stats(0)["sweeps_total"]
        ens.close()
    print(f"solver iterations over 15 steps: {its[True]} with the network, {its[False]} without")
    assert its[True] < 0.6 * its[False]


def test_full_workload_256_instances(gpu):
    """The 256-instance workload (24.7 M cells): every solve converges, the
    autotuner picked a shape per class, throughput is reported."""
    from cfd2d_amd.ensemble import Ensemble, sample_instances

    ins = sample_instances(256, 0)
    ens = Ensemble(ins)
    ens.step(2)
    t0 = time.perf_counter()
    ens.step(5)
    wall = time.perf_counter() - t0
    rel = [ens.stats(k)["relres"] for k in range(0, 256, 17)]
    assert max(rel) <= 1e-6
    assert all(ens.stats(k)["info"] == 0 for k in range(0, 256, 17))
    tx = {ens.stats(c[0])["pred_tx"] * ens.stats(c[0])["pred_ty"] for c in ens.classes}
    assert tx and all(t in (256, 512) for t in tx)
    for c in range(len(ens.classes)):
        print(ens.tuning(c))
    print(f"256 instances, {ens.cells / 1e6:.2f} M cells: {5 * ens.cells / wall / 1e9:.3f} G "
          f"cell-updates/s")
    ens.close()


def test_extrapolated_start_saves_iterations(gpu):
    """Each instance's solve from the extrapolation of its last two solutions
    against the last solution alone: fewer lockstep iterations, fields within
    the solve tolerance."""
    from cfd2d_amd.ensemble import Ensemble

    ins = _mixed()
    out = {}
    for x0 in ("extrapolate", "previous"):
        ens = Ensemble(ins, x0=x0)
        ens.step(6)
        s0 = ens.stats(0)["sweeps_total"]
        ens.step(10)
        out[x0] = (ens.stats(0)["sweeps_total"] - s0, [ens.field(k, "u") for k in range(len(ins))],
                   max(ens.stats(k)["relres"] for k in range(len(ins))))
        ens.close()
    print(f"iterations over steps 7-16: {out['extrapolate'][0]} extrapolated, "
          f"{out['previous'][0]} from the last solution")
    assert out["extrapolate"][2] <= 1e-6
    assert out["extrapolate"][0] < 0.8 * out["previous"][0]
    for ue, up in zip(out["extrapolate"][1], out["previous"][1]):
        assert np.abs(ue - up).max() < 1e-5


def _wide():
    """All four sides, both families, the ends of both parameter ranges."""
    from cfd2d_amd.ensemble import Instance

    return [Instance("lid", 64, re=100.0), Instance("heated", 64, ra=1e6),
            Instance("heated", 128, ra=1e4), Instance("lid", 256, re=5000.0),
            Instance("lid", 512, re=1000.0), Instance("heated", 512, ra=1e6)]


@pytest.mark.parametrize("coarse", ["network", "vcycle"])
def test_instances_match_fp64_oracle(gpu, coarse):
    """Every instance of a batch covering sides 64-512, Re 100-5000 and Ra
    1e4-1e6 against the FP64 oracle (cases.py:60-70, 207-223; pressure by the
    exact DCT solve of the closed box) after 20 steps: the same t and dt,
    velocity and temperature within 1e-5 (the FP32 bulk and the 1e-6 solver
    tolerance), every solve converged."""
    from cfd2d_amd.ensemble import Ensemble
    from oracle import cfd_oracle as O

    ins = _wide()
    ens = Ensemble(ins, coarse=coarse)
    steps = 20
    ens.step(steps)
    for k, inst in enumerate(ins):
        print(f"ensemble vs oracle ({coarse}): {inst.kind} n={inst.n}", flush=True)
        if inst.kind == "lid":
            ref = O.cavity(re=inst.re, n=inst.n, pressure_solver="dct")
        else:
            ref = O.heated_cavity(ra=inst.ra, pr=inst.pr, n=inst.n, pressure_solver="dct")
        for _ in range(steps):
            ref.step()
        st = ens.stats(k)
        assert st["steps"] == steps and st["t"] == pytest.approx(ref.t, rel=1e-12)
        assert st["relres"] <= 1e-6 and st["info"] == 0
        du = np.abs(ens.field(k, "u") - ref.u).max()
        dv = np.abs(ens.field(k, "v") - ref.v).max()
        msg = f"{inst.kind} n={inst.n}: max|du| {du:.2e} max|dv| {dv:.2e}"
        assert du < 1e-5 and dv < 1e-5, msg
        if inst.kind == "heated":
            T = ref.scalars["T"].phi
            dT = np.abs(ens.field(k, "T")[1:-1, 1:-1] - T[1:-1, 1:-1]).max()
            msg += f" max|dT| {dT:.2e}"
            assert dT < 1e-5, msg
        print(msg)
    ens.close()


def test_vcycle_coarse_end_is_faster_and_converges(gpu):
    """The complete V-cycle coarse end against the shared network on the
    full 256-instance workload: fewer lockstep iterations per step."""
    from cfd2d_amd.ensemble import Ensemble, sample_instances

    ins = sample_instances(256, 0)
    out = {}
    for coarse in ("network", "vcycle"):
        ens = Ensemble(ins, coarse=coarse)
        ens.step(3)
        s0 = ens.stats(0)["sweeps_total"]
        t0 = time.perf_counter()
        ens.step(5)
        wall = time.perf_counter() - t0
        out[coarse] = ((ens.stats(0)["sweeps_total"] - s0) / 5, 1e3 * wall / 5)
        assert max(ens.stats(k)["relres"] for k in range(len(ins))) <= 1e-6
        ens.close()
    print("256 instances: " + ", ".join(f"{c} {it:.1f} iterations {ms:.1f} ms/step"
                                        for c, (it, ms) in out.items()))
    assert out["vcycle"][0] < out["network"][0]
