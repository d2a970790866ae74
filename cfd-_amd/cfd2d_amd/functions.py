"""Function objects, VTK export and checkpoints (cfd2d/functions.py).

* Probes, FieldAverage: evaluated on the device in the step's launch
  sequence (csrc/monitors.hip), so a run with them synchronises with the host
  no more often than without; the host collects their records when it reads
  them.  Probes of fields the device does not sample (speed, vorticity,
  streamfunction, ...) fall back to the reference's host evaluation.
* LineSample, VTKWriter: on the host from the fields (functions.py:117-178,
  244-322).
* Checkpoints in the reference's ``checkpoint.npz`` layout
  (functions.py:326-410).  Keys: u, v, p, t, n, dt, flow_force, dt_old, fluid,
  [hu_old, hv_old], skalar_<name>, modell<k>_<key>, history_json -- identical
  to the reference's, so files move both ways.  One extra key, ``druck_x0``,
  stores the pressure solver's warm start (the reference keeps it in memory
  only, pressure.py:140), which makes a CG restart continue bit for bit; with
  the extrapolated initial iterate a second one, ``druck_xprev``, stores the
  solution before the last, so that restart is exact too.  The reference's
  loader ignores both.
"""

from __future__ import annotations

import csv
import ctypes
import json
from pathlib import Path

import numpy as np

from . import _lib


# --------------------------------------------------------------------------- #
# fields at the cell centres (visualization.py:11-18, 50-73)
# --------------------------------------------------------------------------- #
# the reference's field table, in its order (the VTK file lists them so)
BASE_FIELDS = ("speed", "vorticity", "pressure", "u", "v", "streamfunction")


def _node_to_cell(w: np.ndarray) -> np.ndarray:
    return 0.25 * (w[:-1, :-1] + w[1:, :-1] + w[:-1, 1:] + w[1:, 1:])


def field_data(sim, field: str) -> np.ndarray:
    """A field at the cell centres, shape (nx, ny), NaN in solid cells."""
    solid = sim.grid.solid
    if field in ("speed", "u", "v"):
        uc, vc = sim.cell_velocity()
        return {"speed": lambda: np.sqrt(uc ** 2 + vc ** 2), "u": lambda: uc, "v": lambda: vc}[field]()
    if field == "vorticity":
        return sim.cell_vorticity()
    if field == "pressure":
        return sim.cell_pressure()
    if field == "streamfunction":
        return np.where(solid, np.nan, _node_to_cell(sim.streamfunction()))
    for m in getattr(sim, "models", []):
        if field in m.field_keys():
            return m.fields(sim)[field][1]
    raise ValueError(f"Unbekanntes Feld {field!r}, erlaubt: {available_fields(sim)}")


def available_fields(sim) -> list[str]:
    names = list(BASE_FIELDS)
    for m in getattr(sim, "models", []):
        names += [k for k in m.field_keys() if k not in names]
    return names


# --------------------------------------------------------------------------- #
# bilinear sampling (functions.py:24-39 through scipy's RegularGridInterpolator)
# --------------------------------------------------------------------------- #
def field_axes(grid, kind: str):
    """Coordinates of the sample points of a field along x and y: the u faces,
    the v faces (each with the ghost row/column outside the wall) or the cell
    centres ("cell")."""
    g = grid
    if kind == "u":
        return np.arange(g.nx + 1) * g.dx, (np.arange(g.ny + 2) - 0.5) * g.dy
    if kind == "v":
        return (np.arange(g.nx + 2) - 0.5) * g.dx, np.arange(g.ny + 1) * g.dy
    return g.xc, g.yc


def _interval(pts: np.ndarray, x: np.ndarray):
    """Interval [pts[i], pts[i+1]) holding x, the first or last one for points
    outside, and the normalised distance from its left end (outside [0, 1]
    when extrapolating).  A point on a grid line belongs to the interval to
    its right, which decides which neighbours a NaN can come from."""
    i = np.clip(np.searchsorted(pts, x, side="right") - 1, 0, len(pts) - 2)
    return i, (x - pts[i]) / (pts[i + 1] - pts[i])


def probe_cells(grid, kind: str, x, y):
    """Lower corner (i, j) in the field's host layout and the distances
    (a, b) of every point, for `kind` "u", "v" or "cell"; cell fields clip
    the points to the outermost cell centres (functions.py:36-38)."""
    gx, gy = field_axes(grid, kind)
    x = np.atleast_1d(np.asarray(x, dtype=float))
    y = np.atleast_1d(np.asarray(y, dtype=float))
    if kind == "cell":
        x, y = np.clip(x, gx[0], gx[-1]), np.clip(y, gy[0], gy[-1])
    i, a = _interval(gx, x)
    j, b = _interval(gy, y)
    return np.stack([i, j], 1).astype(np.int32), np.stack([a, b], 1)


def interpolate(data: np.ndarray, corner: np.ndarray, frac: np.ndarray) -> np.ndarray:
    """f00 (1-a)(1-b) + f01 (1-a) b + f10 a (1-b) + f11 a b, left to right,
    which reproduces scipy's linear RegularGridInterpolator bit for bit (a NaN
    corner gives NaN, whatever its weight)."""
    i, j = corner[:, 0], corner[:, 1]
    a, b = frac[:, 0], frac[:, 1]
    ca, cb = 1 - a, 1 - b
    return (data[i, j] * ca * cb + data[i, j + 1] * ca * b + data[i + 1, j] * a * cb
            + data[i + 1, j + 1] * a * b)


def sample_field(sim, field: str, x, y) -> np.ndarray:
    """A field at arbitrary points (functions.py:24-39): u and v on their face
    grids, extrapolated linearly outside, the others from the cell centres."""
    if field in ("u", "v"):
        corner, frac = probe_cells(sim.grid, field, x, y)
        return interpolate(sim.u if field == "u" else sim.v, corner, frac)
    corner, frac = probe_cells(sim.grid, "cell", x, y)
    return interpolate(field_data(sim, field), corner, frac)


# field codes of the device probes (CFD_PROBE_* of the C interface)
DEVICE_PROBE_FIELDS = {"u": 0, "v": 1, "pressure": 2, "speed": 3, "vorticity": 4}
PROBE_SCALAR = 5
DEVICE_SCALARS = 2    # transported scalars the device step carries


def device_field(sim, field: str):
    """The device code of a field the probes can sample in the step's launch
    sequence, or None (streamfunction needs a prefix sum over the domain)."""
    if field in DEVICE_PROBE_FIELDS:
        return DEVICE_PROBE_FIELDS[field]
    if field in BASE_FIELDS:
        return None
    for m in getattr(sim, "models", []):
        if field in m.field_keys():
            sid = getattr(getattr(m, "field", None), "sid", None)
            return PROBE_SCALAR + sid if sid is not None and sid < DEVICE_SCALARS else None
    return None


class FunctionObject:
    """Evaluated after each step (solver.py:402-403) and written at the end.
    Objects with ``on_device`` set run inside the step's launch sequence and
    cost the host nothing per step."""

    name = "function"
    on_device = False

    def execute(self, sim) -> None:
        pass

    def write(self, outdir: Path) -> list[Path]:
        return []

    def describe(self) -> str:
        return self.name

    def spec(self) -> dict:
        return {}


def _csv(path: Path, header: list, rows) -> Path:
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w", newline="", encoding="utf-8") as fh:
        out = csv.writer(fh, delimiter=";")
        out.writerow(header)
        out.writerows(rows)
    return path


# --------------------------------------------------------------------------- #
class Probes(FunctionObject):
    """Time series of fields at fixed points after every ``every``-th step
    (functions.py:59-114).  Fields the device samples (u, v, pressure, speed,
    vorticity, transported scalars) are recorded on the device; a probe with
    another field runs on the host after every step."""

    def __init__(self, points, fields=("u", "v", "pressure"), every: int = 1):
        self.points = [tuple(float(c) for c in p) for p in points]
        self.fields = list(fields)
        self.every = max(1, int(every))
        self.t: list = []
        self.data: dict = {f: [] for f in self.fields}
        self.name = "probes"
        self.on_device = False
        self.sim = None

    def device_ok(self, sim) -> bool:
        return all(device_field(sim, f) is not None for f in self.fields)

    def samples(self, sim):
        """(field code, corner, distances) of every (field, point), field major."""
        xs = np.array([p[0] for p in self.points])
        ys = np.array([p[1] for p in self.points])
        out = []
        for f in self.fields:
            kind = f if f in ("u", "v") else "cell"
            corner, frac = probe_cells(sim.grid, kind, xs, ys)
            out.append((np.full(len(xs), device_field(sim, f), np.int32), corner, frac))
        return out

    def take(self, times: np.ndarray, block: np.ndarray) -> None:
        """Device records: times [k] and values [k][fields x points]."""
        npt = len(self.points)
        self.t += [float(t) for t in times]
        for q, f in enumerate(self.fields):
            self.data[f] += list(block[:, q * npt:(q + 1) * npt].copy())

    def execute(self, sim) -> None:
        self.sim = sim
        if self.on_device or sim.n % self.every:
            return
        xs = [p[0] for p in self.points]
        ys = [p[1] for p in self.points]
        self.t.append(sim.t)
        for f in self.fields:
            self.data[f].append(sample_field(sim, f, xs, ys))

    def write(self, outdir: Path) -> list[Path]:
        if self.on_device and self.sim is not None:
            self.sim.collect_function_objects()
        if not self.t:
            return []
        header = ["t"] + [f"{f}@({x:g},{y:g})" for f in self.fields for (x, y) in self.points]
        table = np.column_stack([np.asarray(self.t)] + [np.asarray(self.data[f]) for f in self.fields])
        rows = ([f"{v:.6g}" for v in r] for r in table)
        return [_csv(Path(outdir) / "sonden.csv", header, rows)]

    def describe(self) -> str:
        return f"Sonden an {len(self.points)} Punkten: {', '.join(self.fields)}"

    def spec(self) -> dict:
        return {"typ": "sonden", "punkte": self.points, "felder": self.fields, "alle": self.every}


class DeviceProbeSet:
    """The context's device probe set: the samples of one or more Probes of
    the same period, recorded into a device buffer of ``capacity`` records
    that the host drains when it synchronises."""

    def __init__(self, sim, members: list, capacity: int = 8192):
        self.members = members
        self.capacity = int(capacity)
        parts = [p.samples(sim) for p in members]
        self.slices = []
        off = 0
        for p, part in zip(members, parts):
            cnt = sum(len(s[0]) for s in part)
            self.slices.append((p, off, cnt))
            off += cnt
        flat = [s for part in parts for s in part]
        src = np.ascontiguousarray(np.concatenate([s[0] for s in flat]), dtype=np.int32)
        corner = np.ascontiguousarray(np.concatenate([s[1] for s in flat]), dtype=np.int32)
        frac = np.ascontiguousarray(np.concatenate([s[2] for s in flat]), dtype=np.float64)
        self.nsample = len(src)
        d = _lib.ProbeDesc()
        d.nsample, d.every, d.capacity = self.nsample, members[0].every, self.capacity
        d.src, d.corner, d.frac = _lib.ptr(src), _lib.ptr(corner), _lib.ptr(frac)
        sim._ctx.call("cfd_probes_set", ctypes.byref(d))
        self.lost = 0

    def collect(self, sim) -> None:
        times = np.empty(self.capacity)
        vals = np.empty((self.capacity, self.nsample))
        n, lost = ctypes.c_int32(), ctypes.c_int64()
        sim._ctx.call("cfd_probes_read", self.capacity, _lib.ptr(times), _lib.ptr(vals),
                      ctypes.byref(n), ctypes.byref(lost))
        if lost.value:
            import warnings

            self.lost += int(lost.value)
            warnings.warn(f"{lost.value} Sondenwerte verloren (Puffer {self.capacity} Einträge)")
        k = n.value
        for p, off, cnt in self.slices:
            p.take(times[:k], vals[:k, off:off + cnt])


# --------------------------------------------------------------------------- #
class LineSample(FunctionObject):
    """A profile along p0 -> p1, sampled when written (functions.py:117-178);
    nothing per step."""

    on_device = True

    def __init__(self, p0, p1, n: int = 100, fields=("u", "v", "pressure"), label: str = "linie"):
        self.p0 = tuple(float(c) for c in p0)
        self.p1 = tuple(float(c) for c in p1)
        self.n = int(n)
        self.fields = list(fields)
        self.label = label
        self.name = "linie"
        self.result = None
        self.sim = None

    def execute(self, sim) -> None:
        self.sim = sim

    def sample(self, sim) -> dict:
        s = np.linspace(0, 1, self.n)
        (x0, y0), (x1, y1) = self.p0, self.p1
        x, y = x0 + s * (x1 - x0), y0 + s * (y1 - y0)
        out = {"s": s * np.hypot(x1 - x0, y1 - y0), "x": x, "y": y}
        out.update({f: sample_field(sim, f, x, y) for f in self.fields})
        return out

    def write(self, outdir: Path) -> list[Path]:
        if self.sim is None:
            return []
        self.result = d = self.sample(self.sim)
        cols = ["s", "x", "y"] + self.fields
        rows = ([f"{d[c][k]:.6g}" for c in cols] for k in range(self.n))
        return [_csv(Path(outdir) / f"linie_{self.label}.csv", cols, rows)]

    def describe(self) -> str:
        return f"Linienprobe {self.p0} → {self.p1} ({self.n} Punkte): {', '.join(self.fields)}"

    def spec(self) -> dict:
        return {"typ": "linie", "p0": self.p0, "p1": self.p1, "n": self.n, "felder": self.fields,
                "name": self.label}


# --------------------------------------------------------------------------- #
class FieldAverage(FunctionObject):
    """Time averages and RMS fluctuations of the cell-centred u, v, p from the
    first step ending at t >= t_start (functions.py:180-240): sums of a dt and
    a^2 dt.  On the device when it is the context's only one; otherwise on
    the host after every step."""

    KEYS = ("u", "v", "p")

    def __init__(self, t_start: float = 0.0):
        self.t_start = float(t_start)
        self.name = "mittelwerte"
        self.on_device = False
        self.sim = None
        self._T = 0.0
        self._sum = self._sum2 = None

    def attach(self, sim) -> None:
        sim._ctx.call("cfd_favg_set", 1, self.t_start)
        self.on_device = True
        self.sim = sim

    def execute(self, sim) -> None:
        self.sim = sim
        if self.on_device or sim.t < self.t_start:
            return
        uc, vc = sim.cell_velocity()
        vals = np.stack([uc, vc, sim.cell_pressure()])
        if self._sum is None:
            self._sum, self._sum2 = np.zeros_like(vals), np.zeros_like(vals)
        self._sum += vals * sim.dt
        self._sum2 += vals * vals * sim.dt
        self._T += sim.dt

    def sums(self):
        """(T, sums [3][nx][ny], sums of squares) for u, v, p."""
        if not self.on_device:
            return self._T, self._sum, self._sum2
        sim = self.sim
        sim._flush()
        g = sim.grid
        T = ctypes.c_double()
        s1, s2 = np.empty((3, g.nx, g.ny)), np.empty((3, g.nx, g.ny))
        sim._ctx.call("cfd_favg_read", ctypes.byref(T), _lib.ptr(s1), _lib.ptr(s2))
        return float(T.value), s1, s2

    @property
    def T(self) -> float:
        return self.sums()[0] if self.sim is not None else 0.0

    def _stats(self):
        T, s1, s2 = self.sums()
        if s1 is None:
            return T, None, None
        Tn = max(T, 1e-300)
        mean = s1 / Tn
        return T, mean, np.sqrt(np.maximum(s2 / Tn - mean * mean, 0.0))

    def mean(self, k: str) -> np.ndarray:
        return self._stats()[1][self.KEYS.index(k)]

    def rms(self, k: str) -> np.ndarray:
        return self._stats()[2][self.KEYS.index(k)]

    def write(self, outdir: Path) -> list[Path]:
        if self.sim is None:
            return []
        T, mean, rms = self._stats()
        if mean is None or T <= 0:
            return []
        path = Path(outdir) / "mittelwerte.npz"
        path.parent.mkdir(parents=True, exist_ok=True)
        arrays = {f"{k}_mittel": mean[q] for q, k in enumerate(self.KEYS)}
        arrays.update({f"{k}_rms": rms[q] for q, k in enumerate(self.KEYS)})
        np.savez_compressed(path, T=T, **arrays)
        return [path]

    def describe(self) -> str:
        return f"Zeitmittelung ab t = {self.t_start:g}"

    def spec(self) -> dict:
        return {"typ": "mittelwerte", "t_start": self.t_start}


def attach_function_objects(sim, fos: list) -> list:
    """Put on the device what it can evaluate in the step's launch sequence:
    the Probes of the first device-capable period and the first FieldAverage.
    Returns the function objects the host must execute after every step."""
    probes = [f for f in fos if isinstance(f, Probes) and f.device_ok(sim)]
    group = [p for p in probes if p.every == probes[0].every] if probes else []
    host = []
    for f in fos:
        if isinstance(f, Probes):
            f.on_device = any(f is p for p in group)
            f.sim = sim
        if isinstance(f, FieldAverage) and f.sim is None and not any(
                isinstance(o, FieldAverage) and o.on_device for o in fos):
            f.attach(sim)
        if isinstance(f, LineSample):
            f.sim = sim
        if not f.on_device:
            host.append(f)
    if group:
        sim._probe_set = DeviceProbeSet(sim, group)
    elif sim._probe_set is not None:
        d = _lib.ProbeDesc()
        sim._ctx.call("cfd_probes_set", ctypes.byref(d))
        sim._probe_set = None
    return host


# --------------------------------------------------------------------------- #
def write_vtk(sim, path) -> Path:
    """All fields in one legacy binary VTK RECTILINEAR_GRID file: the cell
    velocity as vector U, every field of the field table and the fluid mask
    as float scalars (functions.py:244-287)."""
    g = sim.grid
    nx, ny = g.nx, g.ny

    def f4(a) -> bytes:
        return np.asarray(a, dtype=">f4").tobytes()

    def cells(a) -> bytes:
        return f4(np.nan_to_num(np.asarray(a, dtype=float), nan=0.0).T.ravel())

    uc, vc = sim.cell_velocity()
    vectors = np.column_stack([np.nan_to_num(uc).T.ravel(), np.nan_to_num(vc).T.ravel(),
                               np.zeros(nx * ny)])
    parts = [f"# vtk DataFile Version 3.0\ncfd2d t={sim.t:.8g}\nBINARY\n"
             f"DATASET RECTILINEAR_GRID\nDIMENSIONS {nx + 1} {ny + 1} 1\n",
             f"X_COORDINATES {nx + 1} float\n", f4(g.xn), "\n",
             f"Y_COORDINATES {ny + 1} float\n", f4(g.yn), "\n",
             "Z_COORDINATES 1 float\n", f4([0.0]), "\n",
             f"CELL_DATA {nx * ny}\nVECTORS U float\n", f4(vectors), "\n"]
    scalars = [(k, field_data(sim, k)) for k in available_fields(sim) if k not in ("u", "v")]
    scalars.append(("fluid", g.fluid_interior.astype(float)))
    for k, data in scalars:
        parts += [f"SCALARS {k.replace(' ', '_')} float 1\nLOOKUP_TABLE default\n", cells(data), "\n"]
    path = Path(path)
    path.write_bytes(b"".join(p.encode("ascii") if isinstance(p, str) else p for p in parts))
    return path


class VTKWriter(FunctionObject):
    """VTK files every `interval` of simulated time plus a ParaView
    ``.vtk.series`` index (functions.py:290-322); it reads the fields, so it
    runs on the host after every step."""

    def __init__(self, interval: float, outdir):
        self.interval = float(interval)
        self.outdir = Path(outdir) / "vtk"
        self.next_t = 0.0
        self.files: list = []
        self.name = "vtk"

    def execute(self, sim) -> None:
        if sim.t + 1e-12 < self.next_t:
            return
        self.outdir.mkdir(parents=True, exist_ok=True)
        name = f"feld_{len(self.files):05d}.vtk"
        write_vtk(sim, self.outdir / name)
        self.files.append((sim.t, name))
        self.next_t = sim.t + self.interval

    def write(self, outdir: Path) -> list[Path]:
        if not self.files:
            return []
        index = {"file-series-version": "1.0",
                 "files": [{"name": name, "time": t} for t, name in self.files]}
        path = self.outdir / "feld.vtk.series"
        path.write_text(json.dumps(index, indent=1), encoding="utf-8")
        return [path]

    def describe(self) -> str:
        return f"VTK-Export alle Δt = {self.interval:g} nach {self.outdir}"

    def spec(self) -> dict:
        return {"typ": "vtk", "intervall": self.interval}


def save_checkpoint(sim, path) -> Path:
    path = Path(path)
    hu_old, hv_old = sim.hu_old, sim.hv_old
    data = {"u": sim.u, "v": sim.v, "p": sim.p, "t": sim.t, "n": sim.n, "dt": sim.dt,
            "flow_force": sim.flow_force, "dt_old": sim.dt_old if sim.dt_old else 0.0,
            "fluid": sim.grid.fluid}
    if hu_old is not None:
        data["hu_old"] = hu_old
        data["hv_old"] = hv_old
    for name, f in sim.scalars.items():
        data[f"skalar_{name}"] = f.phi
    for k, m in enumerate(sim.models):
        for key, arr in (m.state() if hasattr(m, "state") else {}).items():
            data[f"modell{k}_{key}"] = arr
    hist = sim.history
    data["history_json"] = json.dumps(
        {k: list(map(float, v)) for k, v in hist.__dict__.items() if k != "extra"}
        | {"extra": {k: list(map(float, v)) for k, v in hist.extra.items()}})
    data["druck_x0"] = sim._ctx.download("x")
    if sim._ctx.initial_iterate()[1]:
        # the solution before the last one: an extrapolated initial iterate
        # continues from it after a restart
        data["druck_xprev"] = sim._ctx.download("xprev")
    np.savez_compressed(path, **data)
    return path


def load_checkpoint(sim, path) -> None:
    """Load a checkpoint into a simulation of the same geometry; ValueError
    otherwise (functions.py:348-379)."""
    d = np.load(Path(path), allow_pickle=False)
    g = sim.grid
    if d["u"].shape != (g.nx + 1, g.ny + 2) or not np.array_equal(d["fluid"], g.fluid):
        raise ValueError("Der Speicherstand passt nicht zum Gitter bzw. zur Geometrie dieses Falls.")
    sim._flush()
    sim.t = float(d["t"])
    sim.n = int(d["n"])
    sim.dt = float(d["dt"])
    sim.flow_force = float(d["flow_force"])
    sim.dt_old = float(d["dt_old"]) or None
    sim._recorded = sim.n
    sim._push_state()
    sim.u = d["u"]
    sim.v = d["v"]
    sim.p = d["p"]
    if "hu_old" in d and sim.scheme == "ab2":
        sim.hu_old = d["hu_old"]
        sim.hv_old = d["hv_old"]
    else:
        sim.hu_old = None
    for name, f in sim.scalars.items():
        key = f"skalar_{name}"
        if key in d:
            f.phi = d[key]
    for k, m in enumerate(sim.models):
        if hasattr(m, "set_state"):
            st = {key.split("_", 1)[1]: d[key] for key in d.files if key.startswith(f"modell{k}_")}
            if st:
                m.set_state(sim, st)
    if "druck_x0" in d and d["druck_x0"].shape == (g.nx, g.ny):
        sim._ctx.upload("x", d["druck_x0"])
        if ("druck_xprev" in d and d["druck_xprev"].shape == (g.nx, g.ny)
                and sim._ctx.initial_iterate()[0] == 1):
            sim._ctx.upload("xprev", d["druck_xprev"])
    h = json.loads(str(d["history_json"]))
    extra = h.pop("extra", {})
    for k, v in h.items():
        setattr(sim.history, k, v)
    sim.history.extra = extra
    sim._apply_bc()


class Checkpoint(FunctionObject):
    """Writes ``checkpoint.npz`` every ``interval`` of simulated time
    (functions.py:382-410)."""

    def __init__(self, interval: float, outdir):
        self.interval = float(interval)
        self.path = Path(outdir) / "checkpoint.npz"
        self.next_t = self.interval
        self.name = "checkpoint"
        self.sim = None

    def execute(self, sim) -> None:
        self.sim = sim
        if sim.t + 1e-12 >= self.next_t:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            save_checkpoint(sim, self.path)
            self.next_t = sim.t + self.interval

    def write(self, outdir: Path) -> list[Path]:
        if self.sim is not None:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            save_checkpoint(self.sim, self.path)
            return [self.path]
        return []

    def describe(self) -> str:
        return f"Speicherstand alle Δt = {self.interval:g}"

    def spec(self) -> dict:
        return {"typ": "checkpoint", "intervall": self.interval}


# case.json "function_objects" entries -> objects (functions.py:413-430)
_FO_TYPES = {
    "sonden": lambda s, out: Probes(s["punkte"], s.get("felder", ("u", "v", "pressure")),
                                    s.get("alle", 1)),
    "linie": lambda s, out: LineSample(s["p0"], s["p1"], s.get("n", 100), s.get("felder", ("u", "v")),
                                       s.get("name", "linie")),
    "mittelwerte": lambda s, out: FieldAverage(s.get("t_start", 0.0)),
    "vtk": lambda s, out: VTKWriter(s["intervall"], out),
    "checkpoint": lambda s, out: Checkpoint(s["intervall"], out),
}


def build_function_objects(specs: list[dict], outdir) -> list:
    made = []
    for s in specs or []:
        make = _FO_TYPES.get(s["typ"])
        if make is None:
            raise ValueError(f"Unbekanntes Function Object {s['typ']!r}")
        made.append(make(s, outdir))
    return made
