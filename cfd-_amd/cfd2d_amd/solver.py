"""Projection time stepper on the GPU (interface of cfd2d/solver.py:33-631).

Chorin projection on the staggered MAC grid, as in the reference:
boundary conditions, dt and gamma control, momentum predictor ``F = u + dt H``
(Euler or variable-step AB2), pressure Poisson solve, ``u = F - dt grad p``,
then scalar transport.  All of it runs on the device; one step is a short
sequence of kernel launches with no host synchronisation (dt, gamma and the
CG convergence test are decided on the device).

The host sees the simulation through per-step records (t, dt, residual,
max |div u|, kinetic energy, model monitors, CG statistics) that the device
writes into a ring buffer.  ``step()`` synchronises every ``sync_every``
steps (default 1, which reproduces the reference's behaviour exactly,
including raising ``FloatingPointError`` from the step that became
unstable); ``run()`` advances in batches up to the next callback or print
step and stops on the device at ``t_end``, ``max_steps`` or ``steady_tol``.

Field attributes (``u``, ``v``, ``p``, ``hu_old``, ``hv_old``) are host copies
made on access; assign to them (``sim.u = arr``) to upload.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .boundary import BoundaryCondition, DomainBoundary
from .device import DeviceContext
from .grid import Grid
from .pressure import PressureSolver

RUN_CHUNK = 256


class StopSimulation(Exception):
    """Raised to stop a running simulation cleanly."""


@dataclass
class History:
    step: list = field(default_factory=list)
    t: list = field(default_factory=list)
    dt: list = field(default_factory=list)
    residual: list = field(default_factory=list)
    divergence: list = field(default_factory=list)
    kinetic_energy: list = field(default_factory=list)
    fx: list = field(default_factory=list)
    fy: list = field(default_factory=list)
    extra: dict = field(default_factory=dict)

    def as_arrays(self) -> dict:
        out = {k: np.asarray(v) for k, v in self.__dict__.items() if k != "extra"}
        for k, v in self.extra.items():
            out[k] = np.asarray(v)
        return out


class Simulation:
    """2D incompressible flow on the GPU.

    Parameters as cfd2d.solver.Simulation (solver.py:58-100), plus
    ``pressure_tol`` / ``pressure_maxiter`` (the reference fixes 1e-10 / 5000),
    ``pressure_x0`` (``"previous"``: each solve starts from the last solution,
    as the reference's CG does; ``"extrapolate"``: from the linear
    extrapolation in time of the last two solutions), ``sync_every`` and
    ``learned`` (settings of the learned solver).  ``pressure_solver`` is
    ``"direct"`` (default), ``"cg"``, ``"cg_mixed"`` or ``"cg_learned"``.
    """

    def __init__(self, grid: Grid, bcs: dict[str, BoundaryCondition], nu: float,
                 scheme: str = "ab2", gamma: float | str = "auto", tau: float = 0.5,
                 dt: float | None = None, pressure_solver: str = "direct",
                 compute_forces: bool = False, initial_velocity: tuple = (0.0, 0.0),
                 models: list | None = None, body_force: tuple = (0.0, 0.0),
                 flow_rate: float | None = None, pressure_tol: float | None = None,
                 pressure_maxiter: int | None = None, sync_every: int = 1, learned=None,
                 pressure_x0: str = "previous"):
        if nu <= 0:
            raise ValueError("Die Viskosität muss positiv sein.")
        if scheme not in ("ab2", "euler"):
            raise ValueError("scheme muss 'ab2' oder 'euler' sein.")
        if not (0 < tau <= 1):
            raise ValueError("tau muss in (0, 1] liegen.")
        if gamma != "auto" and not (0.0 <= float(gamma) <= 1.0):
            raise ValueError("gamma muss in [0, 1] liegen oder 'auto' sein.")

        self.removed_cells = grid.finalize()
        self.grid = grid
        self.nu = float(nu)
        self.scheme = scheme
        self.gamma_setting = gamma
        self.tau = float(tau)
        self.dt_fixed = dt
        self.compute_forces = compute_forces
        self.sync_every = max(1, int(sync_every))
        self.boundary = DomainBoundary(grid, bcs)
        self.bcs = bcs
        self.pressure = PressureSolver(grid, self.boundary, method=pressure_solver,
                                       tol=1e-10 if pressure_tol is None else pressure_tol,
                                       maxiter=pressure_maxiter, learned=learned,
                                       x0=pressure_x0)
        self._ctx = DeviceContext(grid, self.boundary, self.pressure)
        self.pressure.attach(self._ctx)

        self.u_fluid = grid.u_fluid_faces()
        self.v_fluid = grid.v_fluid_faces()
        self.periodic = self.boundary.periodic
        if self.periodic:
            self.u_fluid[0, :] = False
        self._n_fluid_cells = int(grid.fluid_interior.sum())
        self.variable_viscosity = False
        self.variable_density = False
        self.body_force = (float(body_force[0]), float(body_force[1]))
        self.flow_rate = None if flow_rate is None else float(flow_rate)
        self.flow_force = 0.0
        self.scalars: dict = {}
        self.function_objects: list = []
        self._fo_key = None
        self._fo_host: list = []
        self._probe_set = None
        self._recorded = 0
        self._since_sync = 0
        self._pending = False
        self.models = []
        for m in models or []:
            self.add_model(m)

        u0, v0 = initial_velocity
        if u0 or v0:
            u = np.zeros((grid.nx + 1, grid.ny + 2))
            v = np.zeros((grid.nx + 2, grid.ny + 1))
            if u0:
                u[self.u_fluid] = u0
            if v0:
                v[self.v_fluid] = v0
            self._ctx.upload("u", u)
            self._ctx.upload("v", v)

        self.t = 0.0
        self.n = 0
        self.dt = 0.0
        self.dt_old = None
        self.gamma = 0.0 if gamma == "auto" else float(gamma)
        self.residual = np.inf
        self.history = History()
        self.wall_time = 0.0
        self.cg_iterations: list[int] = []
        self.cg_info: list[int] = []      # per step: 0 converged, >0 iterations without, <0 breakdown
        self.cg_relres: list[float] = []

        self._apply_bc()
        self._project_initial()

    # ------------------------------------------------------------------ #
    def add_model(self, model) -> None:
        model.attach(self)
        self.models.append(model)

    @property
    def nu_t(self) -> np.ndarray:
        """Turbulent viscosity (nx+2, ny+2) as the model set it for the last
        step (zero when laminar); a host copy."""
        g = self.grid
        if not self.variable_viscosity:
            return np.zeros((g.nx + 2, g.ny + 2))
        self._flush()
        return self._ctx.download("nut")

    @property
    def nu_eff(self) -> np.ndarray:
        """nu + nu_t (solver.py:332-333); with two phases the mixture's
        dynamic viscosity (vof.py:205-210)."""
        if self.variable_density:
            self._flush()
            return self._ctx.download("mu")
        if not self.variable_viscosity:
            return np.full((self.grid.nx + 2, self.grid.ny + 2), self.nu)
        return np.add(self.nu, self.nu_t)

    @property
    def rho(self) -> np.ndarray:
        """Density (nx+2, ny+2): 1, or the two-phase mixture's."""
        if not self.variable_density:
            return np.ones((self.grid.nx + 2, self.grid.ny + 2))
        self._flush()
        return self._ctx.download("rho")

    def model(self, kind: str):
        for m in self.models:
            if getattr(m, "kind", None) == kind:
                return m
        return None

    # device fields as host arrays -------------------------------------------
    @property
    def u(self) -> np.ndarray:
        self._flush()
        return self._ctx.download("u")

    @u.setter
    def u(self, arr) -> None:
        self._flush()
        self._ctx.upload("u", arr)

    @property
    def v(self) -> np.ndarray:
        self._flush()
        return self._ctx.download("v")

    @v.setter
    def v(self, arr) -> None:
        self._flush()
        self._ctx.upload("v", arr)

    @property
    def p(self) -> np.ndarray:
        self._flush()
        return self._ctx.download("p")

    @p.setter
    def p(self, arr) -> None:
        self._flush()
        self._ctx.upload("p", arr)

    @property
    def hu_old(self):
        self._flush()
        if self.scheme != "ab2" or not self._ctx.get_state().has_old:
            return None
        return self._ctx.download("hu_old")

    @hu_old.setter
    def hu_old(self, arr) -> None:
        self._set_history_field("hu_old", arr)

    @property
    def hv_old(self):
        self._flush()
        if self.scheme != "ab2" or not self._ctx.get_state().has_old:
            return None
        return self._ctx.download("hv_old")

    @hv_old.setter
    def hv_old(self, arr) -> None:
        self._set_history_field("hv_old", arr)

    def _set_history_field(self, name, arr) -> None:
        self._flush()
        st = self._ctx.get_state()
        if arr is None:
            st.has_old = 0
            self._ctx.set_state(st)
            return
        self._ctx.upload(name, arr)
        st.has_old = 1 if self.scheme == "ab2" else 0
        self._ctx.set_state(st)

    def _push_state(self) -> None:
        """Write the host scalars (t, n, dt, dt_old, flow_force) to the device."""
        st = self._ctx.get_state()
        st.t, st.n, st.dt = float(self.t), int(self.n), float(self.dt)
        st.dt_old = float(self.dt_old) if self.dt_old else 0.0
        st.flow_force = float(self.flow_force)
        st.gamma = float(self.gamma)
        self._ctx.set_state(st)

    # ------------------------------------------------------------------ #
    def _params(self, run: bool = False, t_end: float = 0.0, max_steps=None,
                steady_tol=None) -> _lib.StepParams:
        p = _lib.StepParams()
        p.nu, p.tau = self.nu, self.tau
        p.scheme_ab2 = 1 if self.scheme == "ab2" else 0
        p.gamma_auto = 1 if self.gamma_setting == "auto" else 0
        p.gamma_fixed = 0.0 if self.gamma_setting == "auto" else float(self.gamma_setting)
        p.dt_fixed = float(self.dt_fixed) if self.dt_fixed else 0.0
        lims = [m.dt_limit(self) for m in self.models if not getattr(m, "device_dt_limit", False)]
        lims = [x for x in lims if x is not None]
        p.dt_model = min(lims) if lims else 1e308
        p.fx = self.body_force[0]
        p.fy = self.body_force[1]
        p.solver = self.pressure.solver_id
        p.rtol = self.pressure.rtol
        p.maxiter = self.pressure.maxiter
        p.reliable_delta = 0.1
        if self.flow_rate is not None:
            p.flow_control = 1
            p.flow_rate = self.flow_rate
        if run:
            p.check_t_end = 1
            p.t_end = float(t_end)
            p.clamp_t_end = 1 if not self.dt_fixed else 0
            p.max_steps = int(max_steps) if max_steps is not None else 0
            p.steady_tol = float(steady_tol) if steady_tol is not None else 0.0
        return p

    def _apply_bc(self) -> None:
        self._ctx.apply_bc()

    def _project_initial(self) -> None:
        u, v = self._ctx.download("u"), self._ctx.download("v")
        if not (np.any(u) or np.any(v)):
            return
        self._ctx.project_initial(self._params())

    def compute_dt(self) -> float:
        """Host evaluation of the step size rule (solver.py:209-231)."""
        if self.dt_fixed:
            return float(self.dt_fixed)
        g = self.grid
        umax = max(np.abs(self.u).max(), 1e-12)
        vmax = max(np.abs(self.v).max(), 1e-12)
        if self.variable_density:
            nu_max = float((self.nu_eff / self.rho)[g.fluid].max())
        else:
            nu_max = float(self.nu_eff.max()) if self.variable_viscosity else self.nu
        dt_visc = 0.5 / (nu_max * (1.0 / g.dx ** 2 + 1.0 / g.dy ** 2))
        dt_conv = min(g.dx / umax, g.dy / vmax)
        if self.scheme == "ab2":
            dt_visc *= 0.5
            dt_conv *= 0.8
        dt = self.tau * min(dt_visc, dt_conv)
        for m in self.models:
            lim = m.dt_limit(self)
            if lim is not None:
                dt = min(dt, self.tau * lim)
        return dt

    # ------------------------------------------------------------------ #
    def _flush(self) -> None:
        """Collect the records of finished steps into the history."""
        if not self._pending:
            return
        self._pending = False
        self._since_sync = 0
        st = self._ctx.get_state()
        new = st.n - self._recorded
        recs = self._ctx.records(self._recorded + 1, new) if new > 0 else []
        self._recorded = st.n
        for r in recs:
            if r.nmon < 0:   # the unstable step: counted, not recorded
                continue
            self._append(r)
        if self._probe_set is not None:
            self._probe_set.collect(self)
        self.t, self.n, self.dt = float(st.t), int(st.n), float(st.dt)
        self.flow_force = float(st.flow_force)
        self.dt_old = float(st.dt_old) if st.n > 0 else None
        self.gamma = float(st.gamma)
        if recs:
            last = recs[-1]
            self.residual = float(last.residual)
            self.pressure.last_iterations = int(last.cg_iters)
            self.pressure.last_relres = float(last.cg_relres)
            for r in recs:
                if r.cg_info > 0:
                    import warnings

                    warnings.warn(f"CG-Drucklöser nicht konvergiert ({r.cg_info} Iterationen)")
        if st.error:
            st.error = 0
            st.done = 0
            self._ctx.set_state(st)
            raise FloatingPointError(
                f"Simulation instabil in Schritt {self.n} (t = {self.t:.4g}). "
                "Zeitschritt verkleinern (tau), Gitter verfeinern oder gamma erhöhen.")

    def _append(self, r) -> None:
        h = self.history
        h.step.append(int(r.n))
        h.t.append(float(r.t))
        h.dt.append(float(r.dt))
        h.residual.append(float(r.residual))
        h.divergence.append(float(r.divergence))
        h.kinetic_energy.append(float(r.kinetic_energy))
        self.cg_iterations.append(int(r.cg_iters))
        self.cg_info.append(int(r.cg_info))
        self.cg_relres.append(float(r.cg_relres))
        if self.compute_forces:
            fx, fy = self._forces_from(r)
            h.fx.append(fx)
            h.fy.append(fy)
        if self.flow_rate is not None:
            h.extra.setdefault("antriebskraft", []).append(float(r.mon[_lib.REC_FLOW_FORCE]))
        for m in self.models:
            for k, val in (m.record_monitors(self, r) or {}).items():
                h.extra.setdefault(k, []).append(val)

    def _forces_from(self, r) -> tuple[float, float]:
        g, m = self.grid, r.mon
        fx = (m[2] - m[3]) * g.dy
        fy = (m[4] - m[5]) * g.dx
        fx += self.nu * (m[6] + m[7]) / (0.5 * g.dy) * g.dx
        fy += self.nu * (m[8] + m[9]) / (0.5 * g.dx) * g.dy
        return float(fx), float(fy)

    def _host_function_objects(self) -> list:
        """Set the function objects up (once per list): probes and the field
        average run on the device in the step's launch sequence; the returned
        rest runs on the host after every step."""
        key = tuple(id(f) for f in self.function_objects)
        if key != self._fo_key:
            from .functions import attach_function_objects

            self.collect_function_objects()
            self._fo_host = attach_function_objects(self, self.function_objects)
            self._fo_key = key
        return self._fo_host

    def collect_function_objects(self) -> None:
        """Bring the device records of the function objects to the host."""
        self._pending = True
        self._flush()

    def step(self) -> None:
        """One projection step (solver.py:326-403), enqueued on the device."""
        t0 = time.perf_counter()
        host_fo = self._host_function_objects()
        self._ctx.step(self._params(), 1)
        self._pending = True
        self._since_sync += 1
        if host_fo or self._since_sync >= self.sync_every:
            self._flush()
            for fo in host_fo:
                fo.execute(self)
        self.wall_time += time.perf_counter() - t0

    def run(self, t_end: float, max_steps: int | None = None, steady_tol: float | None = None,
            callback=None, callback_every: int = 50, verbose: bool = True,
            print_every: int = 200) -> str:
        """Run to ``t_end`` or convergence (solver.py:448-505); returns the reason
        ``"t_end"``, ``"konvergiert"``, ``"max_steps"`` or ``"abgebrochen"``."""
        reason = "t_end"
        t_start = time.perf_counter()
        self._flush()
        try:
            while self.t < t_end - 1e-12:
                if max_steps is not None and self.n >= max_steps:
                    reason = "max_steps"
                    break
                chunk = RUN_CHUNK
                if callback is not None:
                    chunk = min(chunk, callback_every - self.n % callback_every)
                if verbose:
                    chunk = min(chunk, print_every - self.n % print_every)
                host_fo = self._host_function_objects()
                if host_fo:
                    chunk = 1
                if max_steps is not None:
                    chunk = min(chunk, max_steps - self.n)
                t0 = time.perf_counter()
                self._ctx.step(self._params(True, t_end, max_steps, steady_tol), chunk)
                self._pending = True
                try:
                    self._flush()
                finally:
                    self.wall_time += time.perf_counter() - t0
                    st = self._ctx.get_state()
                    if st.done:
                        st.done = 0
                        self._ctx.set_state(st)
                for fo in host_fo:
                    fo.execute(self)
                if verbose and self.n % print_every == 0:
                    print(self.status_line())
                if callback is not None and self.n % callback_every == 0:
                    if callback(self) is False:
                        reason = "abgebrochen"
                        break
                if steady_tol is not None and self.n > 10 and self.residual < steady_tol:
                    reason = "konvergiert"
                    break
        except (KeyboardInterrupt, StopSimulation):
            reason = "abgebrochen"
        if callback is not None:
            try:
                callback(self)
            except StopSimulation:
                pass
        if verbose:
            print(self.status_line())
            el = time.perf_counter() - t_start
            print(f"Ende: {reason}. {self.n} Schritte, Rechenzeit {el:.1f} s "
                  f"({1e3 * self.wall_time / max(self.n, 1):.2f} ms/Schritt)")
        return reason

    def status_line(self) -> str:
        s = (f"Schritt {self.n:7d} | t = {self.t:9.4f} | dt = {self.dt:.3e} | "
             f"Residuum = {self.residual:.3e} | max|div u| = {self.max_divergence():.2e}")
        if self.compute_forces and self.history.fx:
            s += f" | Fx = {self.history.fx[-1]:+.4f} | Fy = {self.history.fy[-1]:+.4f}"
        return s

    # ------------------------------------------------------------------ #
    # evaluation on host copies (solver.py:517-615)
    # ------------------------------------------------------------------ #
    def bulk_velocity(self) -> float:
        """Mean u over the fluid faces 1..nx (solver.py:405-408)."""
        nx = self.grid.nx
        return float(self.u[1:nx + 1, 1:-1][self.u_fluid[1:nx + 1, 1:-1]].mean())

    def divergence(self) -> np.ndarray:
        g = self.grid
        u, v = self.u, self.v
        div = (u[1:, 1:-1] - u[:-1, 1:-1]) / g.dx + (v[1:-1, 1:] - v[1:-1, :-1]) / g.dy
        return np.where(g.fluid_interior, div, 0.0)

    def max_divergence(self) -> float:
        return float(np.abs(self.divergence()).max())

    def cell_velocity(self) -> tuple[np.ndarray, np.ndarray]:
        u, v = self.u, self.v
        uc = 0.5 * (u[:-1, 1:-1] + u[1:, 1:-1])
        vc = 0.5 * (v[1:-1, :-1] + v[1:-1, 1:])
        solid = self.grid.solid
        return np.where(solid, np.nan, uc), np.where(solid, np.nan, vc)

    def cell_pressure(self) -> np.ndarray:
        return np.where(self.grid.solid, np.nan, self.p[1:-1, 1:-1])

    def kinetic_energy(self) -> float:
        uc, vc = self.cell_velocity()
        return float(0.5 * np.nansum(uc ** 2 + vc ** 2) * self.grid.dx * self.grid.dy)

    def obstacle_forces(self) -> tuple[float, float]:
        g = self.grid
        fl = g.fluid
        solid = ~fl
        solid[0, :] = solid[-1, :] = solid[:, 0] = solid[:, -1] = False
        f = fl[1:-1, 1:-1]
        east, west = f & solid[2:, 1:-1], f & solid[:-2, 1:-1]
        north, south = f & solid[1:-1, 2:], f & solid[1:-1, :-2]
        p = self.p[1:-1, 1:-1]
        u, v = self.u, self.v
        uc = 0.5 * (u[:-1, 1:-1] + u[1:, 1:-1])
        vc = 0.5 * (v[1:-1, :-1] + v[1:-1, 1:])
        fx = (p[east].sum() - p[west].sum()) * g.dy
        fy = (p[north].sum() - p[south].sum()) * g.dx
        fx += self.nu * (uc[north].sum() + uc[south].sum()) / (0.5 * g.dy) * g.dx
        fy += self.nu * (vc[east].sum() + vc[west].sum()) / (0.5 * g.dx) * g.dy
        return float(fx), float(fy)

    def _derived(self, kind: int) -> np.ndarray:
        self._flush()
        g = self.grid
        out = np.empty((g.nx + 1, g.ny + 1))
        self._ctx.call("cfd_download_derived", kind, _lib.ptr(out))
        return out

    def vorticity(self) -> np.ndarray:
        """dv/dx - du/dy at the grid nodes, (nx+1, ny+1) (solver.py:569-573),
        computed on the device."""
        return self._derived(0)

    def cell_vorticity(self) -> np.ndarray:
        w = self.vorticity()
        wc = 0.25 * (w[:-1, :-1] + w[1:, :-1] + w[:-1, 1:] + w[1:, 1:])
        return np.where(self.grid.solid, np.nan, wc)

    def streamfunction(self) -> np.ndarray:
        """psi at the grid nodes with u = dpsi/dy, v = -dpsi/dx, psi(0, 0) = 0
        (solver.py:580-596), computed on the device."""
        return self._derived(1)

    def sample_u(self, x, y) -> np.ndarray:
        from scipy.interpolate import RegularGridInterpolator

        g = self.grid
        xs = np.arange(g.nx + 1) * g.dx
        ys = (np.arange(g.ny + 2) - 0.5) * g.dy
        f = RegularGridInterpolator((xs, ys), self.u, bounds_error=False, fill_value=None)
        return f(np.column_stack([np.ravel(x), np.ravel(y)]))

    def sample_v(self, x, y) -> np.ndarray:
        from scipy.interpolate import RegularGridInterpolator

        g = self.grid
        xs = (np.arange(g.nx + 2) - 0.5) * g.dx
        ys = np.arange(g.ny + 1) * g.dy
        f = RegularGridInterpolator((xs, ys), self.v, bounds_error=False, fill_value=None)
        return f(np.column_stack([np.ravel(x), np.ravel(y)]))

    def summary(self) -> str:
        g = self.grid
        solver = {"direct": "CG (Jacobi, FP64, rtol 1e-12)", "cg": "CG (Jacobi, FP64)",
                  "cg_mixed": "FCG (multigrid, FP32/FP64)",
                  "cg_learned": "FCG (learned U-Net V-cycle, FP32/FP64)"}[self.pressure.method]
        lines = [
            f"Gitter: {g.nx} x {g.ny} Zellen, Gebiet {g.lx:g} x {g.ly:g}, "
            f"dx = {g.dx:.4g}, dy = {g.dy:.4g}",
            f"Fluidzellen: {g.fluid_interior.sum()} / {g.nx * g.ny}"
            + (f" ({self.removed_cells} eingeschlossene Zellen entfernt)"
               if self.removed_cells else ""),
            f"Viskosität nu = {self.nu:.4g}",
            f"Zeitschema: {'Adams-Bashforth 2' if self.scheme == 'ab2' else 'Euler explizit'}, "
            f"tau = {self.tau}, gamma = {self.gamma_setting}",
            f"Drucklöser: {solver} auf der GPU",
            "Randbedingungen:",
            self.boundary.describe(),
        ]
        return "\n".join(lines)

    def close(self) -> None:
        self._ctx.close()
