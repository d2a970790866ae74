"""Function objects evaluated on the device (probes, field average): their
records against the host evaluation of the reference (functions.py:24-39,
59-114, 180-210) on the fields downloaded after the same step, bit for bit,
and the reference's own function-object tests (tests/test_functions.py)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POINTS = [(0.5, 0.5), (0.7, 0.52), (0.05, 0.95), (2.1, 0.4), (1.0, 0.0), (3.99, 0.999),
          (0.7 + 0.12, 0.5)]


def _channel(solver="cg", **extra):
    from cfd2d_amd import cases

    spec = {"fall": "custom",
            "parameter": {"obstacles": [{"typ": "kreis", "x": 0.7, "y": 0.5, "r": 0.12},
                                        {"typ": "kreis", "x": 2.1, "y": 0.4, "r": 0.1}],
                          "lx": 4.0, "ly": 1.0, "nx": 64, "ny": 16, "re": 200, "L_ref": 1.0,
                          "inflow_profile": "parabolic"},
            "numerik": {"pressure_solver": solver}}
    spec.update(extra)
    return cases.build_case(spec)


def _host_sample(sim, field, pts):
    """The reference's evaluation on host copies of the fields (scipy)."""
    from scipy.interpolate import RegularGridInterpolator

    from cfd2d_amd import functions as F

    x = np.array([p[0] for p in pts])
    y = np.array([p[1] for p in pts])
    if field == "u":
        return sim.sample_u(x, y)
    if field == "v":
        return sim.sample_v(x, y)
    g = sim.grid
    f = RegularGridInterpolator((g.xc, g.yc), F.field_data(sim, field), bounds_error=False,
                                fill_value=None)
    return f(np.column_stack([np.clip(x, g.xc[0], g.xc[-1]), np.clip(y, g.yc[0], g.yc[-1])]))


@pytest.mark.parametrize("solver", ["direct", "cg_mixed"])
def test_device_probes_match_host_sampling(gpu, solver):
    from cfd2d_amd import functions as F

    case = _channel(solver, skalar={"name": "c", "diffusivity": 1e-3, "inflow_band": [0.25, 0.75]})
    sim = case.sim
    fields = ["u", "v", "pressure", "speed", "vorticity", "c"]
    pr = F.Probes(POINTS, fields=fields)
    sim.function_objects.append(pr)
    for k in range(12):
        sim.step()
        assert pr.on_device and len(pr.t) == k + 1 == sim.n
        assert pr.t[-1] == sim.t
        for f in fields:
            want = _host_sample(sim, f, POINTS)
            got = pr.data[f][-1]
            print(solver, k, f, got, want, flush=True)
            assert np.array_equal(got, want, equal_nan=True), (f, got, want)
    # a point in the cylinder: NaN for the cell fields, finite for the faces
    assert np.isnan(pr.data["pressure"][-1][1]) and np.isfinite(pr.data["u"][-1][1])


def test_device_probes_need_no_per_step_sync(gpu, tmp_path):
    """Chunked run: every step recorded (reference tests/test_functions.py:20-30),
    only every 'alle'-th with every > 1, and a host-only field next to them."""
    from cfd2d_amd import cases, functions as F

    case = cases.lid_driven_cavity(re=100, n=16)
    sim = case.sim
    pr = F.Probes([(0.5, 0.5), (0.5, 0.9)], fields=("u", "pressure", "vorticity"))
    ls = F.LineSample((0.5, 0.0), (0.5, 1.0), n=21, fields=("u", "v"), label="mitte")
    sim.function_objects += [pr, ls]
    assert sim._host_function_objects() == []
    sim.run(0.5, verbose=False)
    files = pr.write(tmp_path) + ls.write(tmp_path)
    assert all(f.exists() for f in files)
    assert len(pr.t) == sim.n and np.all(np.diff(pr.t) > 0)
    assert pr.t[-1] == sim.t
    for f in pr.fields:
        assert np.array_equal(pr.data[f][-1], _host_sample(sim, f, pr.points), equal_nan=True)
    assert ls.result["u"][0] == pytest.approx(0.0, abs=1e-12)
    assert ls.result["u"][-1] == pytest.approx(1.0, abs=1e-12)

    case = cases.lid_driven_cavity(re=100, n=16)
    sim = case.sim
    p3 = F.Probes([(0.3, 0.3)], fields=("u",), every=3)
    ps = F.Probes([(0.3, 0.3)], fields=("streamfunction",), every=3)
    sim.function_objects += [p3, ps]
    assert sim._host_function_objects() == [ps]
    sim.run(1e9, max_steps=20, verbose=False)
    assert p3.on_device and not ps.on_device
    assert len(p3.t) == len(ps.t) == 6 and p3.t == ps.t


def test_probe_buffer_overflow_is_counted(gpu):
    from cfd2d_amd import cases, functions as F

    case = cases.lid_driven_cavity(re=100, n=16)
    sim = case.sim
    pr = F.Probes([(0.5, 0.5)], fields=("u",))
    sim.function_objects.append(pr)
    sim._host_function_objects()
    sim._probe_set = F.DeviceProbeSet(sim, [pr], capacity=4)
    with pytest.warns(UserWarning, match="6 Sondenwerte verloren"):
        sim.run(1e9, max_steps=10, verbose=False)
    assert len(pr.t) == 4 and sim._probe_set.lost == 6
    with pytest.raises(ValueError):
        bad = F.Probes([(0.5, 0.5)], fields=("u",))
        bad.samples = lambda s: [(np.array([9], np.int32), np.zeros((1, 2), np.int32),
                                  np.zeros((1, 2)))]
        F.DeviceProbeSet(sim, [bad])


def test_device_field_average_matches_host_accumulation(gpu, tmp_path):
    from cfd2d_amd import functions as F

    case = _channel("cg")
    sim = case.sim
    fa = F.FieldAverage(t_start=0.02)
    sim.function_objects.append(fa)
    T, s1, s2 = 0.0, None, None
    for _ in range(15):
        sim.step()
        if sim.t < fa.t_start:
            continue
        uc, vc = sim.cell_velocity()
        vals = np.stack([uc, vc, sim.cell_pressure()])
        if s1 is None:
            s1, s2 = np.zeros_like(vals), np.zeros_like(vals)
        s1 += vals * sim.dt
        s2 += vals * vals * sim.dt
        T += sim.dt
    assert fa.on_device and s1 is not None
    dT, d1, d2 = fa.sums()
    assert dT == T
    assert np.array_equal(d1, s1, equal_nan=True) and np.array_equal(d2, s2, equal_nan=True)
    (path,) = fa.write(tmp_path)
    z = np.load(path)
    assert float(z["T"]) == T
    assert np.array_equal(z["u_mittel"], s1[0] / T, equal_nan=True)
    assert np.isnan(z["p_rms"][sim.grid.solid]).all()


def test_case_function_objects_run(gpu, tmp_path):
    """The reference's case.json round trip with every function object type
    (tests/test_functions.py:99-116), without the report."""
    from cfd2d_amd import cases, functions as F

    spec = {"fall": "cylinder", "parameter": {"re": 100, "nx": 88, "ny": 17, "t_end": 0.2},
            "function_objects": [{"typ": "sonden", "punkte": [[0.5, 0.2]]},
                                 {"typ": "mittelwerte", "t_start": 0.0},
                                 {"typ": "vtk", "intervall": 0.1},
                                 {"typ": "checkpoint", "intervall": 0.1}]}
    cases.save_case(spec, tmp_path / "case.json")
    loaded = cases.load_case(tmp_path / "case.json")
    assert loaded == spec
    case = cases.build_case(loaded)
    out = tmp_path / "out"
    case.sim.function_objects = F.build_function_objects(loaded["function_objects"], out)
    case.sim.run(case.t_end, verbose=False)
    for fo in case.sim.function_objects:
        fo.write(out)
    for f in ("sonden.csv", "mittelwerte.npz", "checkpoint.npz", "vtk/feld.vtk.series",
              "vtk/feld_00000.vtk"):
        assert (out / f).exists(), f
    pr = case.sim.function_objects[0]
    assert pr.on_device and len(pr.t) == case.sim.n


@pytest.mark.parametrize("shape", ["channel", "wide"])
def test_vorticity_and_streamfunction_on_device(gpu, shape):
    """Node vorticity and streamfunction computed by the device kernels equal
    the reference's numpy evaluation on the downloaded velocity; "wide" runs
    the bottom-row sum over more than one 1024-lane stage."""
    from hostfields import node_vorticity, streamfunction

    from cfd2d_amd.boundary import BoundaryCondition as BC
    from cfd2d_amd.grid import Grid
    from cfd2d_amd.solver import Simulation

    if shape == "channel":
        sim = _channel("cg").sim
        steps = 10
    else:
        grid = Grid(2100, 6, 35.0, 0.1)
        m = np.zeros((2100, 6), bool)
        m[1500:1510, 2:4] = True
        grid.add_obstacle(m)
        grid.finalize()
        bcs = dict(left=BC("inflow", 1.0), right=BC("outflow"), bottom=BC("noslip"),
                   top=BC("noslip"))
        sim = Simulation(grid, bcs, nu=1e-2, pressure_solver="cg")
        steps = 3
    for _ in range(steps):
        sim.step()
    u, v = sim.u, sim.v
    w = sim.vorticity()
    print(shape, "vorticity", float(np.abs(w).max()), flush=True)
    assert np.array_equal(w, node_vorticity(sim.grid, u, v))
    psi = sim.streamfunction()
    print(shape, "streamfunction", float(np.abs(psi).max()), flush=True)
    assert np.array_equal(psi, streamfunction(sim.grid, u, v))
    assert np.abs(psi[:, -1]).max() > 0
    wc = sim.cell_vorticity()
    assert np.isnan(wc[sim.grid.solid]).all() and np.isfinite(wc[~sim.grid.solid]).all()
