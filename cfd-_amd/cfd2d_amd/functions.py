"""Checkpoints in the reference's ``checkpoint.npz`` layout
(cfd2d/functions.py:326-410).

Keys: u, v, p, t, n, dt, flow_force, dt_old, fluid, [hu_old, hv_old],
skalar_<name>, modell<k>_<key>, history_json -- identical to the reference's,
so files move both ways.  One extra key, ``druck_x0``, stores the pressure
solver's warm start (the reference keeps it in memory only, pressure.py:140),
which makes a CG restart continue bit for bit; the reference's loader ignores
it.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np


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
    h = json.loads(str(d["history_json"]))
    extra = h.pop("extra", {})
    for k, v in h.items():
        setattr(sim.history, k, v)
    sim.history.extra = extra
    sim._apply_bc()


class Checkpoint:
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
