"""Pressure Poisson system ``A p = -rhs`` with ``A = -Laplacian`` (interface of
cfd2d/pressure.py:29-185).

The device solves the system matrix-free.  This host part describes it once:

* per fluid cell an index into a table of distinct diagonal values, computed
  with the reference's order of additions (x couplings, y couplings, periodic
  couplings, +2/h^2 for outflow Dirichlet faces, +1/dx^2 for the pinned cell of
  each pure-Neumann component; pressure.py:53-125), plus two bits marking a
  Dirichlet x-face or y-face;
* the pure-Neumann components (``connected_components`` over the couplings)
  and their cell counts, for the compatibility shift and the pressure mean.

Methods: ``"cg"`` is Jacobi-PCG in FP64 with the reference's semantics (warm
start, recursive-residual stop, ``maxiter`` 5000); ``"direct"`` (the
reference's LU) runs the same solver to rtol 1e-12 with a large iteration cap;
``"cg_mixed"`` is flexible CG with FP64 vectors and dot products,
preconditioned by an FP32 aggregation-multigrid V-cycle
(:mod:`cfd2d_amd.multigrid`), ending on the FP64 true residual;
``"cg_learned"`` is the same flexible CG with the learned preconditioner of
:mod:`cfd2d_amd.learned` (a V-cycle whose coarse correction is a U-Net).
"""

from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

from . import _lib
from .boundary import DomainBoundary
from .grid import Grid

MAX_COMPONENTS = 8
MAX_DIAG_VALUES = 63


class PressureSolver:
    METHODS = ("direct", "cg", "cg_mixed", "cg_learned")
    X0 = ("previous", "extrapolate")

    def __init__(self, grid: Grid, boundary: DomainBoundary, method: str = "direct",
                 tol: float = 1e-10, maxiter: int | None = None, learned=None,
                 x0: str = "previous"):
        if method not in self.METHODS:
            raise ValueError(f"Unbekannter Drucklöser {method!r}, erlaubt: {self.METHODS}")
        if x0 not in self.X0:
            raise ValueError(f"x0 must be one of {self.X0}, got {x0!r}")
        self.x0 = x0
        self.grid, self.method = grid, method
        self.tol = float(tol)
        nx, ny = grid.nx, grid.ny
        ax, ay = 1.0 / grid.dx ** 2, 1.0 / grid.dy ** 2
        fl = grid.fluid_interior
        self.fluid = fl
        n = int(fl.sum())
        if n == 0:
            raise ValueError("Das Gebiet enthält keine Fluidzellen.")
        self.n = n
        num = -np.ones((nx, ny), dtype=np.int64)
        num[fl] = np.arange(n)
        diag = np.zeros((nx, ny))
        src, dst = [], []
        cx = fl[:-1, :] & fl[1:, :]
        src += [num[:-1, :][cx], num[1:, :][cx]]
        diag[:-1, :][cx] += ax
        diag[1:, :][cx] += ax
        cy = fl[:, :-1] & fl[:, 1:]
        src += [num[:, :-1][cy], num[:, 1:][cy]]
        diag[:, :-1][cy] += ay
        diag[:, 1:][cy] += ay
        self.periodic = boundary.periodic
        if self.periodic:
            cp = fl[-1, :] & fl[0, :]
            src += [num[-1, :][cp], num[0, :][cp]]
            diag[-1, :][cp] += ax
            diag[0, :][cp] += ax
        self.dirichlet = {s: boundary.outflow_mask(s) for s in ("left", "right", "bottom", "top")}
        diag[0, :][self.dirichlet["left"]] += 2 * ax
        diag[-1, :][self.dirichlet["right"]] += 2 * ax
        diag[:, 0][self.dirichlet["bottom"]] += 2 * ay
        diag[:, -1][self.dirichlet["top"]] += 2 * ay
        dir_x = np.zeros((nx, ny), dtype=bool)
        dir_y = np.zeros((nx, ny), dtype=bool)
        dir_x[0, :] |= self.dirichlet["left"]
        dir_x[-1, :] |= self.dirichlet["right"]
        dir_y[:, 0] |= self.dirichlet["bottom"]
        dir_y[:, -1] |= self.dirichlet["top"]

        # coupling graph: pairs (a, b) for a < b in the order added above
        pairs_a = np.concatenate(src[0::2]) if src else np.zeros(0, np.int64)
        pairs_b = np.concatenate(src[1::2]) if src else np.zeros(0, np.int64)
        graph = sp.coo_matrix((np.ones(pairs_a.size), (pairs_a, pairs_b)), shape=(n, n)).tocsr()
        ncomp, comp_of = connected_components(graph, directed=False)
        labels = np.zeros((nx, ny), dtype=np.int64)
        labels[fl] = comp_of + 1
        self.labels = labels
        with_dir = set(np.unique(labels[(dir_x | dir_y) & fl]).tolist())
        self.neumann_components, self._pinned = [], []
        for c in range(1, ncomp + 1):
            if c in with_dir:
                continue
            first = int(np.argmax(comp_of == c - 1))
            pi, pj = np.argwhere(num == first)[0]
            diag[pi, pj] += ax
            self.neumann_components.append(c)
            self._pinned.append((int(pi), int(pj)))
        if len(self.neumann_components) > MAX_COMPONENTS:
            raise ValueError(f"{len(self.neumann_components)} reine Neumann-Gebiete "
                             f"(höchstens {MAX_COMPONENTS} werden unterstützt)")
        self.diag = diag

        # device description: diagonal table and per-cell codes
        values, index = np.unique(diag[fl], return_inverse=True)
        if values.size > MAX_DIAG_VALUES:
            raise ValueError("zu viele verschiedene Diagonalwerte im Drucksystem")
        self.diag_table = np.zeros(64)
        self.inv_table = np.zeros(64)
        self.diag_table[1:values.size + 1] = values
        self.inv_table[1:values.size + 1] = 1.0 / values
        code = np.zeros((nx, ny), dtype=np.uint8)
        code[fl] = (index + 1).astype(np.uint8)
        code |= (dir_x & fl).astype(np.uint8) << 6
        code |= (dir_y & fl).astype(np.uint8) << 7
        self.code = code
        comp = np.zeros((nx, ny), dtype=np.uint8)
        for k, c in enumerate(self.neumann_components):
            comp[labels == c] = k + 1
        self.comp = comp
        self.comp_count = np.array([float((labels == c).sum()) for c in self.neumann_components]
                                   + [1.0] * (MAX_COMPONENTS - len(self.neumann_components)))
        if maxiter is None:
            maxiter = 5000 if method != "direct" else max(20000, 4 * (nx + ny) * 8)
        self.maxiter = int(maxiter)
        self.last_iterations = 0
        self.last_relres = 0.0
        self._ctx = None
        self.learned_spec = learned
        self.learned = None

    # ------------------------------------------------------------------ #
    @property
    def solver_id(self) -> int:
        return {"cg_mixed": _lib.SOLVERS["mgcg"],
                "cg_learned": _lib.SOLVERS["learned"]}.get(self.method, _lib.SOLVERS["pcg64"])

    @property
    def rtol(self) -> float:
        return min(self.tol, 1e-12) if self.method == "direct" else self.tol

    def attach(self, ctx) -> None:
        self._ctx = ctx
        if self.x0 == "extrapolate":
            ctx.call("cfd_set_initial_iterate", 1)
        if self.method in ("cg_mixed", "cg_learned"):
            from .multigrid import Hierarchy

            g = self.grid
            self.hierarchy = Hierarchy(self.fluid, self.diag, 1.0 / g.dx ** 2, 1.0 / g.dy ** 2,
                                       self.periodic)
            self.hierarchy.upload(ctx)
        if self.method == "cg_learned":
            from .learned import Attached, LearnedConfig

            self.learned = Attached(ctx, self.hierarchy, LearnedConfig.from_spec(self.learned_spec),
                                    x0=self.x0)

    def solve(self, rhs: np.ndarray) -> np.ndarray:
        """Solve ``Lap(p) = rhs`` for rhs of shape (nx, ny) on the device
        (pressure.py:143-185); warm-started from the previous solution."""
        if self._ctx is None:
            raise RuntimeError("PressureSolver is not attached to a device context")
        g = self.grid
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        if rhs.shape != (g.nx, g.ny):
            raise ValueError(f"rhs has shape {rhs.shape}, expected {(g.nx, g.ny)}")
        out = np.zeros((g.nx, g.ny))
        it = ctypes.c_int32(0)
        rel = ctypes.c_double(0.0)
        rc = _lib.lib().cfd_pressure_solve(self._ctx.handle, _lib.ptr(rhs), _lib.ptr(out),
                                           self.solver_id, self.rtol, self.maxiter,
                                           ctypes.byref(it), ctypes.byref(rel))
        _lib.check(rc, "pressure solve")
        self.last_iterations = int(it.value)
        self.last_relres = float(rel.value)
        return out
