"""Function objects on the host: the bilinear sampling rule the device probes
use (pinned against scipy's RegularGridInterpolator, the reference's
interpolator, functions.py:24-39), the field table, the VTK layout
(tests/test_functions.py:33-45 of the reference) and the case.json builder
(functions.py:413-430).  No GPU: the fields live in host arrays."""

import json

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from cfd2d_amd import functions as F
from cfd2d_amd.grid import Grid
from cfd2d_amd.solver import Simulation
from hostfields import node_vorticity, streamfunction


class HostFields:
    """Simulation's field accessors (solver.py:517-596) over host arrays."""

    cell_velocity = Simulation.cell_velocity
    cell_pressure = Simulation.cell_pressure
    cell_vorticity = Simulation.cell_vorticity
    sample_u = Simulation.sample_u
    sample_v = Simulation.sample_v

    def __init__(self, grid, seed=0):
        rng = np.random.default_rng(seed)
        g = grid
        self.grid = g
        self.u = rng.standard_normal((g.nx + 1, g.ny + 2))
        self.v = rng.standard_normal((g.nx + 2, g.ny + 1))
        self.p = rng.standard_normal((g.nx + 2, g.ny + 2))
        self.t, self.n, self.dt = 0.0, 0, 0.01
        self.models = []

    def vorticity(self):
        return node_vorticity(self.grid, self.u, self.v)

    def streamfunction(self):
        return streamfunction(self.grid, self.u, self.v)


def _grid(nx=12, ny=9, obstacle=True):
    g = Grid(nx, ny, 1.3, 1.0)
    if obstacle:
        m = np.zeros((nx, ny), bool)
        m[4:7, 3:5] = True
        g.add_obstacle(m)
    g.finalize()
    return g


def _points(g, rng, n=300):
    """Random points (some outside the domain) plus points on grid lines."""
    x = np.concatenate([rng.uniform(-0.3, g.lx + 0.3, n), g.xc, g.xn[: len(g.xc)],
                        rng.choice(g.xc, 40)])
    y = np.concatenate([rng.uniform(-0.3, g.ly + 0.3, n), rng.choice(g.yc, len(g.xc)),
                        rng.choice(g.yn, len(g.xc)), rng.choice(g.yc, 40)])
    return x, y


def _scipy(data, axes, x, y, clip):
    f = RegularGridInterpolator(axes, data, bounds_error=False, fill_value=None)
    if clip:
        x = np.clip(x, axes[0][0], axes[0][-1])
        y = np.clip(y, axes[1][0], axes[1][-1])
    return f(np.column_stack([x, y]))


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("kind", ["u", "v", "cell"])
def test_probe_rule_is_scipy_bit_for_bit(kind):
    """The corner/distance rule and the evaluation order of the device probes
    give exactly scipy's values, NaN corners (solid cells) and extrapolation
    included."""
    g = _grid()
    rng = np.random.default_rng(3)
    axes = F.field_axes(g, kind)
    data = rng.standard_normal((len(axes[0]), len(axes[1])))
    if kind == "cell":
        data = np.where(g.solid, np.nan, data)
    x, y = _points(g, rng)
    corner, frac = F.probe_cells(g, kind, x, y)
    ref = _scipy(data, axes, x, y, clip=kind == "cell")
    assert np.isnan(ref).any() == (kind == "cell")
    assert _same(F.interpolate(data, corner, frac), ref)


def test_probe_rule_catches_left_intervals():
    """A point on a grid line next to a solid cell: scipy takes the interval
    to the right; the left one would pull the NaN in (the rule is not
    interchangeable)."""
    g = _grid()
    data = np.where(g.solid, np.nan, 1.0)
    x, y = np.array([g.xc[7]]), np.array([g.yc[3]])     # right of the solid block
    corner, frac = F.probe_cells(g, "cell", x, y)
    assert tuple(corner[0]) == (7, 3) and tuple(frac[0]) == (0.0, 0.0)
    assert F.interpolate(data, corner, frac)[0] == 1.0
    assert _scipy(data, F.field_axes(g, "cell"), x, y, True)[0] == 1.0


def test_sample_field_matches_reference_interpolation():
    g = _grid()
    sim = HostFields(g)
    rng = np.random.default_rng(5)
    x, y = _points(g, rng, 100)
    assert _same(F.sample_field(sim, "u", x, y), sim.sample_u(x, y))
    assert _same(F.sample_field(sim, "v", x, y), sim.sample_v(x, y))
    for name in ("pressure", "speed", "vorticity", "streamfunction"):
        ref = _scipy(F.field_data(sim, name), (g.xc, g.yc), x, y, True)
        assert _same(F.sample_field(sim, name, x, y), ref), name


def test_field_table():
    g = _grid()
    sim = HostFields(g)
    uc, vc = sim.cell_velocity()
    assert _same(F.field_data(sim, "speed"), np.sqrt(uc ** 2 + vc ** 2))
    assert np.isnan(F.field_data(sim, "pressure")[g.solid]).all()
    psi = F.field_data(sim, "streamfunction")
    assert np.isnan(psi[g.solid]).all() and np.isfinite(psi[~g.solid]).all()
    assert F.available_fields(sim) == ["speed", "vorticity", "pressure", "u", "v", "streamfunction"]
    with pytest.raises(ValueError):
        F.field_data(sim, "nope")
    assert F.device_field(sim, "streamfunction") is None
    assert [F.device_field(sim, k) for k in ("u", "v", "pressure", "speed", "vorticity")] == \
        [0, 1, 2, 3, 4]


def test_host_probes_and_csv(tmp_path):
    g = _grid()
    sim = HostFields(g)
    pr = F.Probes([(0.5, 0.5), (0.2, 0.9)], fields=("u", "streamfunction"), every=2)
    for n in range(1, 6):
        sim.n, sim.t = n, 0.1 * n
        pr.execute(sim)
    assert pr.t == pytest.approx([0.2, 0.4])
    (path,) = pr.write(tmp_path)
    rows = path.read_text(encoding="utf-8").splitlines()
    assert rows[0] == "t;u@(0.5,0.5);u@(0.2,0.9);streamfunction@(0.5,0.5);streamfunction@(0.2,0.9)"
    first = [float(c) for c in rows[1].split(";")]
    want = np.concatenate([[0.2], F.sample_field(sim, "u", [0.5, 0.2], [0.5, 0.9]),
                           F.sample_field(sim, "streamfunction", [0.5, 0.2], [0.5, 0.9])])
    assert first == pytest.approx(want, rel=1e-5, nan_ok=True)
    assert len(rows) == 3


def test_line_sample_hits_wall_values(tmp_path):
    """The reference's check (tests/test_functions.py:20-30): u on the wall is
    the mean of the interior value and its mirrored ghost."""
    g = _grid(obstacle=False)
    sim = HostFields(g)
    sim.u[:, 0] = -sim.u[:, 1]
    sim.u[:, -1] = 2.0 - sim.u[:, -2]
    ls = F.LineSample((0.5, 0.0), (0.5, 1.0), n=21, fields=("u", "v"), label="mitte")
    ls.execute(sim)
    (path,) = ls.write(tmp_path)
    assert path.name == "linie_mitte.csv"
    assert ls.result["u"][0] == pytest.approx(0.0, abs=1e-12)
    assert ls.result["u"][-1] == pytest.approx(1.0, abs=1e-12)
    assert ls.result["s"][-1] == pytest.approx(1.0)


def test_host_field_average_is_time_weighted():
    g = _grid()
    sim = HostFields(g)
    fa = F.FieldAverage(t_start=0.15)
    snaps = []
    for n, dt in enumerate([0.1, 0.05, 0.1, 0.2], start=1):
        sim.u = sim.u + dt
        sim.p = sim.p * 0.9
        sim.dt, sim.t, sim.n = dt, sim.t + dt, n
        fa.execute(sim)
        if sim.t >= 0.15:
            snaps.append((dt, sim.cell_velocity()[0].copy(), sim.cell_pressure().copy()))
    T = sum(s[0] for s in snaps)
    assert fa.T == pytest.approx(T)
    mu = sum(dt * a for dt, a, _ in snaps) / T
    assert np.allclose(fa.mean("u"), mu, equal_nan=True)
    varp = sum(dt * p * p for dt, _, p in snaps) / T - (sum(dt * p for dt, _, p in snaps) / T) ** 2
    assert np.allclose(fa.rms("p"), np.sqrt(np.maximum(varp, 0)), equal_nan=True)


def test_vtk_layout(tmp_path):
    """Header, sizes and the vector block (the reference's
    tests/test_functions.py:33-45 on this grid)."""
    g = _grid()
    sim = HostFields(g)
    path = F.write_vtk(sim, tmp_path / "f.vtk")
    raw = path.read_bytes()
    assert raw.startswith(b"# vtk DataFile Version 3.0\ncfd2d t=0\nBINARY\n")
    assert b"DIMENSIONS 13 10 1" in raw and b"CELL_DATA 108" in raw
    i = raw.index(b"X_COORDINATES 13 float\n") + len(b"X_COORDINATES 13 float\n")
    assert np.array_equal(np.frombuffer(raw[i:i + 52], ">f4"), g.xn.astype(np.float32))
    i = raw.index(b"VECTORS U float\n") + len(b"VECTORS U float\n")
    vec = np.frombuffer(raw[i:i + 108 * 12], dtype=">f4").reshape(-1, 3)
    uc, vc = sim.cell_velocity()
    assert np.array_equal(vec[:, 0], np.nan_to_num(uc).T.ravel().astype(np.float32))
    assert np.array_equal(vec[:, 1], np.nan_to_num(vc).T.ravel().astype(np.float32))
    names = [k for k in F.available_fields(sim) if k not in ("u", "v")] + ["fluid"]
    pos = [raw.index(f"SCALARS {k} float 1\nLOOKUP_TABLE default\n".encode()) for k in names]
    assert pos == sorted(pos)
    hdr = b"SCALARS fluid float 1\nLOOKUP_TABLE default\n"
    i = raw.index(hdr) + len(hdr)
    assert np.array_equal(np.frombuffer(raw[i:i + 108 * 4], ">f4"),
                          g.fluid_interior.T.ravel().astype(np.float32))
    assert raw.endswith(b"\n") and len(raw) == i + 108 * 4 + 1


def test_vtk_writer_series(tmp_path):
    g = _grid()
    sim = HostFields(g)
    w = F.VTKWriter(0.1, tmp_path)
    for n in range(1, 8):
        sim.n, sim.t = n, 0.04 * n
        w.execute(sim)
    assert [t for t, _ in w.files] == pytest.approx([0.04, 0.16, 0.28])
    (series,) = w.write(tmp_path)
    idx = json.loads(series.read_text())
    assert idx["files"][0] == {"name": "feld_00000.vtk", "time": 0.04}
    assert (tmp_path / "vtk" / "feld_00002.vtk").exists()


def test_build_function_objects_from_case_spec(tmp_path):
    specs = [{"typ": "sonden", "punkte": [[0.5, 0.2]]},
             {"typ": "linie", "p0": [0, 0], "p1": [1, 1], "name": "diag"},
             {"typ": "mittelwerte", "t_start": 2.0},
             {"typ": "vtk", "intervall": 0.1},
             {"typ": "checkpoint", "intervall": 0.5}]
    fos = F.build_function_objects(specs, tmp_path)
    assert [type(f).__name__ for f in fos] == ["Probes", "LineSample", "FieldAverage", "VTKWriter",
                                               "Checkpoint"]
    assert fos[0].fields == ["u", "v", "pressure"] and fos[0].every == 1
    assert fos[1].fields == ["u", "v"] and fos[1].n == 100 and fos[1].label == "diag"
    assert fos[2].t_start == 2.0
    assert [f.spec()["typ"] for f in fos] == [s["typ"] for s in specs]
    with pytest.raises(ValueError):
        F.build_function_objects([{"typ": "unbekannt"}], tmp_path)
    assert F.build_function_objects(None, tmp_path) == []
