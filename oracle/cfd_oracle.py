"""CPU FP64 oracle of the cfd2d projection step.

TEST INFRASTRUCTURE ONLY.  This module restates, on the CPU in IEEE double,
the reference algorithm of kaeptnzursee1910/CFD- (package ``cfd2d``) for the
path ``Simulation.step``.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it, and only as the checker (or
as the timed CPU baseline); the product package never imports it.

The O(N) stencil loops run in the serial C of ``oracle/kernels.c`` (built by
``oracle/Makefile`` into ``oracle/liboracle.so``); setup, boundary conditions
and the pressure solve use NumPy and SciPy with the reference's own call
arguments (``splu(permc_spec="COLAMD")`` and Jacobi-preconditioned
``scipy.sparse.linalg.cg(rtol=tol, maxiter=5000)``).

Which reference lines each piece follows:

=========================  =====================================================
``OGrid``                  cfd2d/grid.py:26-152
``inflow_profile``         cfd2d/boundary.py:63-89
``OBoundary``              cfd2d/boundary.py:92-242
``OPressure``              cfd2d/pressure.py:32-185
``OScalar``                cfd2d/models.py:73-250 (sources, clipping)
``OBoussinesq``            cfd2d/models.py:339-426 (momentum_source 386-396)
``OSim``                   cfd2d/solver.py:95-540 (constant density; the
                           variable viscosity of turb_oracle.py's models)
``cavity`` etc.            cfd2d/cases.py:51-70, 161-223, 340-379
``save/load_checkpoint``   cfd2d/functions.py:326-379
=========================  =====================================================

Parity pins (see tests/test_oracle.py): golden vectors written by the
reference itself (tests/golden/), the DCT-II exact Poisson solve, the
closed-form time step, the limiter formulas, divergence-freedom and
conservation invariants.
"""

from __future__ import annotations

import ctypes
import json
import subprocess
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy import ndimage
from scipy.sparse.csgraph import connected_components

_HERE = Path(__file__).resolve().parent
_LIB = None

LIMITERS = {"upwind": 0, "vanleer": 1, "minmod": 2, "superbee": 3, "linear": 4, "mc": 5}
SIDES = ("left", "right", "bottom", "top")


# --------------------------------------------------------------------------- #
# C kernels
# --------------------------------------------------------------------------- #
def build() -> Path:
    """Compile oracle/kernels.c (gcc, -ffp-contract=off) into oracle/liboracle.so."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _HERE / "liboracle.so"


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build()
        L = ctypes.CDLL(str(path))
        d = ctypes.c_double
        i = ctypes.c_int
        vp = ctypes.c_void_p
        L.oracle_momentum_tendencies.argtypes = [vp, vp, vp, i, i, d, d, d, d, i, vp, vp]
        L.oracle_momentum_tendencies_var.argtypes = [vp, vp, vp, i, i, d, d, vp, d, i, vp, vp, vp]
        L.oracle_vof_advect.argtypes = [vp, vp, vp, vp, i, i, d, d, d, d, i, d, vp]
        L.oracle_obstacle_bc.argtypes = [vp, vp, vp, i, i, i]
        L.oracle_scalar_transport.argtypes = ([vp, vp, vp, vp, i, i, d, d, vp, i, i]
                                              + [vp] * 8 + [i, d, vp])
        L.oracle_limiter.argtypes = [d, i]
        L.oracle_limiter.restype = d
        L.oracle_pcg_csr.argtypes = [ctypes.c_long, vp, vp, vp, vp, vp, vp, d, i, vp,
                                     ctypes.POINTER(i)]
        L.oracle_pcg_csr.restype = i
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _chk(a: np.ndarray, shape, dtype) -> np.ndarray:
    if a.shape != shape or a.dtype != dtype or not a.flags.c_contiguous:
        raise ValueError(f"array {a.shape} {a.dtype} does not match {shape} {dtype}")
    return a


def momentum_tendencies(u, v, fluid, nx, ny, dx, dy, nu0, gamma, periodic, hu, hv) -> None:
    """kernels.py:44-155, constant viscosity ``nu0`` (the value of ``nu[1, 1]``)."""
    fl = np.ascontiguousarray(fluid, dtype=np.uint8)
    _chk(u, (nx + 1, ny + 2), np.float64)
    _chk(v, (nx + 2, ny + 1), np.float64)
    _chk(hu, (nx + 1, ny + 2), np.float64)
    _chk(hv, (nx + 2, ny + 1), np.float64)
    lib().oracle_momentum_tendencies(_p(u), _p(v), _p(fl), nx, ny, dx, dy, nu0, gamma,
                                     int(bool(periodic)), _p(hu), _p(hv))


def momentum_tendencies_var(u, v, fluid, nx, ny, dx, dy, nu, gamma, periodic, hu, hv,
                            rho=None) -> None:
    """kernels.py:44-155 with ``variable = True``: stress-form diffusion of the
    cell viscosity field ``nu`` (nu_eff, (nx+2, ny+2)); with ``rho`` (variable
    density) ``nu`` is the dynamic viscosity and the diffusion is divided by
    the face density."""
    fl = np.ascontiguousarray(fluid, dtype=np.uint8)
    _chk(u, (nx + 1, ny + 2), np.float64)
    _chk(v, (nx + 2, ny + 1), np.float64)
    _chk(nu, (nx + 2, ny + 2), np.float64)
    _chk(hu, (nx + 1, ny + 2), np.float64)
    _chk(hv, (nx + 2, ny + 1), np.float64)
    if rho is not None:
        _chk(rho, (nx + 2, ny + 2), np.float64)
    lib().oracle_momentum_tendencies_var(_p(u), _p(v), _p(fl), nx, ny, dx, dy, _p(nu), gamma,
                                         int(bool(periodic)), _p(hu), _p(hv),
                                         None if rho is None else _p(rho))


def vof_advect(a, u, v, fluid, nx, ny, dx, dy, dt, c_alpha, periodic, a_in, out) -> None:
    """vof.py:34-133, one explicit alpha step into ``out``."""
    fl = np.ascontiguousarray(fluid, dtype=np.uint8)
    for arr, shape in ((a, (nx + 2, ny + 2)), (out, (nx + 2, ny + 2)), (u, (nx + 1, ny + 2)),
                       (v, (nx + 2, ny + 1))):
        _chk(arr, shape, np.float64)
    lib().oracle_vof_advect(_p(a), _p(u), _p(v), _p(fl), nx, ny, dx, dy, dt, c_alpha,
                            int(bool(periodic)), a_in, _p(out))


def obstacle_bc(u, v, fluid, nx, ny, periodic) -> None:
    """kernels.py:158-209."""
    fl = np.ascontiguousarray(fluid, dtype=np.uint8)
    _chk(u, (nx + 1, ny + 2), np.float64)
    _chk(v, (nx + 2, ny + 1), np.float64)
    lib().oracle_obstacle_bc(_p(u), _p(v), _p(fl), nx, ny, int(bool(periodic)))


def scalar_transport(phi, u, v, fluid, nx, ny, dx, dy, diff, limiter, periodic,
                     bt_l, bv_l, bt_r, bv_r, bt_b, bv_b, bt_t, bv_t, obs_t, obs_v, out) -> None:
    """kernels.py:254-391."""
    fl = np.ascontiguousarray(fluid, dtype=np.uint8)
    cell = (nx + 2, ny + 2)
    _chk(phi, cell, np.float64)
    _chk(diff, cell, np.float64)
    _chk(out, cell, np.float64)
    bts = [np.ascontiguousarray(b, dtype=np.int64) for b in (bt_l, bt_r, bt_b, bt_t)]
    bvs = [np.ascontiguousarray(b, dtype=np.float64) for b in (bv_l, bv_r, bv_b, bv_t)]
    lib().oracle_scalar_transport(_p(phi), _p(u), _p(v), _p(fl), nx, ny, dx, dy, _p(diff),
                                  int(limiter), int(bool(periodic)),
                                  _p(bts[0]), _p(bvs[0]), _p(bts[1]), _p(bvs[1]),
                                  _p(bts[2]), _p(bvs[2]), _p(bts[3]), _p(bvs[3]),
                                  int(obs_t), float(obs_v), _p(out))


def limiter(r: float, kind: int) -> float:
    return lib().oracle_limiter(float(r), int(kind))


# --------------------------------------------------------------------------- #
# Grid and boundary conditions
# --------------------------------------------------------------------------- #
class OGrid:
    """MAC grid with a ghost layer; ``fluid`` is (nx+2, ny+2) (grid.py:26-152)."""

    def __init__(self, nx, ny, lx, ly):
        self.nx, self.ny = int(nx), int(ny)
        self.lx, self.ly = float(lx), float(ly)
        self.dx, self.dy = self.lx / self.nx, self.ly / self.ny
        self.fluid = np.zeros((self.nx + 2, self.ny + 2), dtype=bool)
        self.fluid[1:-1, 1:-1] = True
        self.periodic_x = False

    @property
    def xc(self):
        return (np.arange(self.nx) + 0.5) * self.dx

    @property
    def yc(self):
        return (np.arange(self.ny) + 0.5) * self.dy

    def cell_centers(self):
        return np.meshgrid(self.xc, self.yc, indexing="ij")

    @property
    def interior(self):
        return self.fluid[1:-1, 1:-1]

    def add_obstacle(self, mask):
        self.fluid[1:-1, 1:-1] &= ~np.asarray(mask, dtype=bool)

    def circle(self, xc, yc, r):
        X, Y = self.cell_centers()
        return (X - xc) ** 2 + (Y - yc) ** 2 <= r ** 2

    def finalize(self) -> int:
        inner = self.fluid[1:-1, 1:-1]
        lab, n = ndimage.label(inner)
        if n <= 1:
            return 0
        edge = np.zeros_like(inner)
        edge[0, :] = edge[-1, :] = edge[:, 0] = edge[:, -1] = True
        keep = np.unique(lab[edge & inner])
        keep = keep[keep > 0]
        enclosed = inner & ~np.isin(lab, keep)
        if enclosed.any():
            self.fluid[1:-1, 1:-1] &= ~enclosed
        return int(enclosed.sum())

    def set_periodic_x(self):
        self.periodic_x = True
        self.fluid[0, :] = self.fluid[self.nx, :]
        self.fluid[self.nx + 1, :] = self.fluid[1, :]
        self.fluid[:, 0] = False
        self.fluid[:, -1] = False

    def u_faces(self):
        return self.fluid[:-1, :] & self.fluid[1:, :]

    def v_faces(self):
        return self.fluid[:, :-1] & self.fluid[:, 1:]


def inflow_profile(coords, open_mask, h, mean_velocity, profile):
    """Per open run along a side: block profile or 6 U s (L-s)/L^2 (boundary.py:63-89)."""
    vals = np.zeros_like(coords)
    n = len(open_mask)
    k = 0
    while k < n:
        if not open_mask[k]:
            k += 1
            continue
        e = k
        while e + 1 < n and open_mask[e + 1]:
            e += 1
        a = coords[k] - 0.5 * h
        b = coords[e] + 0.5 * h
        s = coords[k:e + 1]
        if profile == "parabolic":
            vals[k:e + 1] = 6.0 * mean_velocity * (s - a) * (b - s) / (b - a) ** 2
        else:
            vals[k:e + 1] = mean_velocity
        k = e + 1
    return vals


class OBC:
    def __init__(self, kind="noslip", velocity=0.0, profile="uniform"):
        self.kind, self.velocity, self.profile = kind, float(velocity), profile


class OBoundary:
    """Domain boundary conditions on u, v (boundary.py:92-242)."""

    def __init__(self, grid: OGrid, bcs: dict):
        self.grid, self.bcs = grid, bcs
        self.periodic = bcs["left"].kind == "periodic"
        if self.periodic:
            grid.set_periodic_x()
        f, nx, ny = grid.fluid, grid.nx, grid.ny
        self.open = {"left": f[1, 1:ny + 1].copy(), "right": f[nx, 1:ny + 1].copy(),
                     "bottom": f[1:nx + 1, 1].copy(), "top": f[1:nx + 1, ny].copy()}
        self.inflow = {}
        for side, bc in bcs.items():
            if bc.kind != "inflow":
                continue
            lr = side in ("left", "right")
            vals = inflow_profile(grid.yc if lr else grid.xc, self.open[side],
                                  grid.dy if lr else grid.dx, bc.velocity, bc.profile)
            self.inflow[side] = (1.0 if side in ("left", "bottom") else -1.0) * vals

    def outflow_mask(self, side):
        if self.bcs[side].kind != "outflow":
            return np.zeros_like(self.open[side])
        return self.open[side]

    def apply(self, u, v):
        nx, ny = self.grid.nx, self.grid.ny
        if self.periodic:
            u[0, :] = u[nx, :]
            v[0, :] = v[nx, :]
            v[nx + 1, :] = v[1, :]
        # x sides: normal u on the boundary face, tangential v in the ghost column
        for side, fi, gi, ii in (("left", 0, 0, 1), ("right", nx, nx + 1, nx)):
            bc, op = self.bcs[side], self.open[side]
            k = bc.kind
            if k == "noslip":
                u[fi, 1:ny + 1] = 0.0
                v[gi, :] = 2.0 * bc.velocity - v[ii, :]
            elif k == "slip":
                u[fi, 1:ny + 1] = 0.0
                v[gi, :] = v[ii, :]
            elif k == "inflow":
                u[fi, 1:ny + 1] = self.inflow[side]
                v[gi, :] = -v[ii, :]
            elif k == "outflow":
                u[fi, 1:ny + 1] = np.where(op, u[fi, 1:ny + 1], 0.0)
                v[gi, :] = v[ii, :]
        # y sides: normal v on the boundary face, tangential u in the ghost row
        for side, fj, gj, jj in (("bottom", 0, 0, 1), ("top", ny, ny + 1, ny)):
            bc, op = self.bcs[side], self.open[side]
            k = bc.kind
            if k == "noslip":
                v[1:nx + 1, fj] = 0.0
                u[:, gj] = 2.0 * bc.velocity - u[:, jj]
            elif k == "slip":
                v[1:nx + 1, fj] = 0.0
                u[:, gj] = u[:, jj]
            elif k == "inflow":
                v[1:nx + 1, fj] = self.inflow[side]
                u[:, gj] = -u[:, jj]
            elif k == "outflow":
                v[1:nx + 1, fj] = np.where(op, v[1:nx + 1, fj], 0.0)
                u[:, gj] = u[:, jj]
        if not self.periodic:
            u[0, 1:ny + 1][~self.open["left"]] = 0.0
            u[nx, 1:ny + 1][~self.open["right"]] = 0.0
        v[1:nx + 1, 0][~self.open["bottom"]] = 0.0
        v[1:nx + 1, ny][~self.open["top"]] = 0.0

    def extrapolate_outflow_predictor(self, F, G):
        nx, ny = self.grid.nx, self.grid.ny
        if self.periodic:
            F[0, :] = F[nx, :]
        if self.bcs["left"].kind == "outflow":
            m = self.open["left"]
            F[0, 1:ny + 1][m] = F[1, 1:ny + 1][m]
        if self.bcs["right"].kind == "outflow":
            m = self.open["right"]
            F[nx, 1:ny + 1][m] = F[nx - 1, 1:ny + 1][m]
        if self.bcs["bottom"].kind == "outflow":
            m = self.open["bottom"]
            G[1:nx + 1, 0][m] = G[1:nx + 1, 1][m]
        if self.bcs["top"].kind == "outflow":
            m = self.open["top"]
            G[1:nx + 1, ny][m] = G[1:nx + 1, ny - 1][m]


# --------------------------------------------------------------------------- #
# Pressure Poisson solve
# --------------------------------------------------------------------------- #
class OPressure:
    """Assembled 5-point -Laplacian on fluid cells, LU or Jacobi-CG (pressure.py:32-185).

    Two methods beyond the reference's: ``"cg_omp"`` runs the same Jacobi-CG
    loop in the C library on OMP_NUM_THREADS threads (the all-cores CPU
    baseline); ``"dct"`` solves a closed, uniform, obstacle-free box exactly by
    the DCT-II diagonalisation of the Neumann Laplacian (an FP64 reference
    for large grids, where splu is slow: it equals the LU solve after the
    reference's mean removal, tests/test_oracle.py)."""

    def __init__(self, grid: OGrid, boundary: OBoundary, method="direct", tol=1e-10,
                 maxiter=5000):
        nx, ny = grid.nx, grid.ny
        ax, ay = 1.0 / grid.dx ** 2, 1.0 / grid.dy ** 2
        self._ax, self._ay = ax, ay
        fl = grid.interior
        self.fluid, self.method, self.tol, self.maxiter = fl, method, tol, maxiter
        n = int(fl.sum())
        if n == 0:
            raise ValueError("no fluid cells")
        num = -np.ones((nx, ny), dtype=np.int64)
        num[fl] = np.arange(n)
        self.n = n
        diag = np.zeros((nx, ny))
        R, C, V = [], [], []

        def couple(ma, mb, pair, coef):
            a, b = ma[pair], mb[pair]
            R.extend([a, b])
            C.extend([b, a])
            V.extend([np.full(a.size, -coef), np.full(a.size, -coef)])

        px = fl[:-1, :] & fl[1:, :]
        couple(num[:-1, :], num[1:, :], px, ax)
        diag[:-1, :][px] += ax
        diag[1:, :][px] += ax
        py = fl[:, :-1] & fl[:, 1:]
        couple(num[:, :-1], num[:, 1:], py, ay)
        diag[:, :-1][py] += ay
        diag[:, 1:][py] += ay
        if boundary.periodic:
            pp = fl[-1, :] & fl[0, :]
            couple(num[-1, :], num[0, :], pp, ax)
            diag[-1, :][pp] += ax
            diag[0, :][pp] += ax
        self.dirichlet = {s: boundary.outflow_mask(s) for s in SIDES}
        diag[0, :][self.dirichlet["left"]] += 2 * ax
        diag[-1, :][self.dirichlet["right"]] += 2 * ax
        diag[:, 0][self.dirichlet["bottom"]] += 2 * ay
        diag[:, -1][self.dirichlet["top"]] += 2 * ay
        on_dir = np.zeros((nx, ny), dtype=bool)
        on_dir[0, :] |= self.dirichlet["left"]
        on_dir[-1, :] |= self.dirichlet["right"]
        on_dir[:, 0] |= self.dirichlet["bottom"]
        on_dir[:, -1] |= self.dirichlet["top"]

        rows = np.concatenate(R) if R else np.zeros(0, np.int64)
        cols = np.concatenate(C) if C else np.zeros(0, np.int64)
        graph = sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(n, n))
        ncomp, comp = connected_components(graph, directed=False)
        labels = np.zeros((nx, ny), dtype=np.int64)
        labels[fl] = comp + 1
        self.labels = labels
        with_dir = set(np.unique(labels[on_dir & fl]).tolist())
        self.neumann, self.pinned = [], []
        for c in range(1, ncomp + 1):
            if c in with_dir:
                continue
            first = int(np.argmax(comp == c - 1))
            pi, pj = np.argwhere(num == first)[0]
            diag[pi, pj] += ax
            self.neumann.append(c)
            self.pinned.append((int(pi), int(pj)))
        d = diag[fl]
        self.diag = d
        self._pairs = (px, py, (fl[-1, :] & fl[0, :]) if boundary.periodic else None)
        self._rc = (np.concatenate(R + [np.arange(n)]), np.concatenate(C + [np.arange(n)]))
        A = sp.csc_matrix((np.concatenate(V + [d]),
                           (np.concatenate(R + [np.arange(n)]), np.concatenate(C + [np.arange(n)]))),
                          shape=(n, n))
        self.A = A
        if method == "direct":
            self._lu = spla.splu(A, permc_spec="COLAMD")
        elif method == "dct":
            if (n != nx * ny or boundary.periodic or len(self.neumann) != 1
                    or any(m.any() for m in self.dirichlet.values())):
                raise ValueError("the DCT solve needs a closed box without obstacles")
            lx = (2.0 - 2.0 * np.cos(np.pi * np.arange(nx) / nx)) * ax
            ly = (2.0 - 2.0 * np.cos(np.pi * np.arange(ny) / ny)) * ay
            lam = lx[:, None] + ly[None, :]
            lam[0, 0] = 1.0
            self._lam = lam
        else:
            self._M = sp.diags(1.0 / d)
            if method == "cg_omp":
                csr = A.tocsr()
                csr.sort_indices()
                self._csr = (csr.indptr.astype(np.int64), csr.indices.astype(np.int64),
                             csr.data.astype(np.float64))
                self._inv_diag = np.ascontiguousarray(1.0 / d)
                self._work = np.zeros(4 * n)
        self.x0 = np.zeros(n)
        self.last_iterations = 0

    def solve(self, rhs):
        b = rhs[self.fluid].astype(float)
        if self.neumann:
            lab = self.labels[self.fluid]
            for c in self.neumann:
                m = lab == c
                b[m] -= b[m].mean()
        b = -b
        if self.method == "direct":
            x = self._lu.solve(b)
            self.last_iterations = 1
        elif self.method == "dct":
            import scipy.fft

            r = scipy.fft.dctn(b.reshape(self.fluid.shape), type=2, norm="ortho") / self._lam
            r[0, 0] = 0.0
            x = scipy.fft.idctn(r, type=2, norm="ortho").ravel()
            self.last_iterations = 1
        elif self.method == "cg_omp":
            x = np.ascontiguousarray(self.x0, dtype=np.float64).copy()
            bb = np.ascontiguousarray(b)
            it = ctypes.c_int(0)
            ptr, ind, val = self._csr
            info = lib().oracle_pcg_csr(self.n, _p(ptr), _p(ind), _p(val), _p(self._inv_diag),
                                        _p(bb), _p(x), self.tol, self.maxiter, _p(self._work),
                                        ctypes.byref(it))
            self.last_iterations = int(it.value)
            self.last_info = int(info)
            self.x0 = x
        else:
            count = [0]

            def cb(_):
                count[0] += 1

            x, info = spla.cg(self.A, b, x0=self.x0, rtol=self.tol, maxiter=self.maxiter,
                              M=self._M, callback=cb)
            self.last_iterations = count[0]
            self.last_info = info
            self.x0 = x
        p = np.zeros(self.fluid.shape)
        p[self.fluid] = x
        for c in self.neumann:
            m = self.labels == c
            p[m] -= p[m].mean()
        return p


    def solve_variable(self, rhs, bx, by):
        """div(b grad p) = rhs with face coefficients b = 1/rho: bx (nx+1, ny)
        on the x faces, by (nx, ny+1); assembled anew and solved directly
        (pressure.py:187-244)."""
        fl = self.fluid
        nx, ny = fl.shape
        dx2, dy2 = self._ax, self._ay
        px, py, pp = self._pairs
        diag = np.zeros((nx, ny))
        vals = []
        c = (bx[1:nx, :] * dx2)[px]
        vals += [-c, -c]
        diag[:-1, :][px] += c
        diag[1:, :][px] += c
        c = (by[:, 1:ny] * dy2)[py]
        vals += [-c, -c]
        diag[:, :-1][py] += c
        diag[:, 1:][py] += c
        if pp is not None:
            c = (bx[nx, :] * dx2)[pp]
            vals += [-c, -c]
            diag[-1, :][pp] += c
            diag[0, :][pp] += c
        d = self.dirichlet
        diag[0, :][d["left"]] += 2 * dx2 * bx[0, :][d["left"]]
        diag[-1, :][d["right"]] += 2 * dx2 * bx[nx, :][d["right"]]
        diag[:, 0][d["bottom"]] += 2 * dy2 * by[:, 0][d["bottom"]]
        diag[:, -1][d["top"]] += 2 * dy2 * by[:, ny][d["top"]]
        for fi, fj in self.pinned:
            diag[fi, fj] += dx2 * max(float(bx.max()), 1e-30)
        A = sp.csc_matrix((np.concatenate(vals + [diag[fl]]), self._rc), shape=(self.n, self.n))
        b = rhs[fl].astype(float)
        if self.neumann:
            lab = self.labels[fl]
            for comp in self.neumann:
                m = lab == comp
                b[m] -= b[m].mean()
        x = spla.spsolve(A, -b)
        p = np.zeros(fl.shape)
        p[fl] = x
        for comp in self.neumann:
            m = self.labels == comp
            p[m] -= p[m].mean()
        return p


# --------------------------------------------------------------------------- #
# Scalar transport
# --------------------------------------------------------------------------- #
_SCODE = {"zerogradient": 0, "fixed": 1, "flux": 2}


class OScalar:
    """Explicit-Euler TVD transport of one cell field (models.py:73-250)."""

    def __init__(self, name, diffusivity, bcs=None, obstacle=("zerogradient", 0.0),
                 limiter="vanleer", initial=0.0, source=None, lower=None, upper=None):
        self.name, self.D0, self.bcs_in = name, float(diffusivity), bcs or {}
        self.obstacle, self.limiter, self.initial = obstacle, limiter, initial
        self.source, self.lower, self.upper = source, lower, upper

    def attach(self, sim):
        self.sim = sim
        g = sim.grid
        nx, ny = g.nx, g.ny
        self.phi = np.zeros((nx + 2, ny + 2))
        if np.isscalar(self.initial):
            self.phi[:] = float(self.initial)
        else:
            init = np.asarray(self.initial, dtype=float)
            if init.shape == (nx, ny):
                self.phi[1:-1, 1:-1] = init
            else:
                self.phi[:] = init
        self.diff = np.full((nx + 2, ny + 2), self.D0)
        self.rhs = np.zeros_like(self.phi)
        kinds = {s: (("fixed", 0.0) if sim.bcs[s].kind == "inflow" else ("zerogradient", 0.0))
                 for s in SIDES}
        kinds.update(self.bcs_in)
        coords = {"left": g.yc, "right": g.yc, "bottom": g.xc, "top": g.xc}
        self.bt, self.bv = {}, {}
        for s in SIDES:
            kind, val = kinds[s]
            c = coords[s]
            self.bt[s] = np.full(len(c), _SCODE[kind], dtype=np.int64)
            if callable(val):
                self.bv[s] = np.asarray(val(c), dtype=float)
            else:
                self.bv[s] = np.broadcast_to(np.asarray(val, dtype=float), c.shape).copy()
        self.fill_ghosts()

    def fill_ghosts(self):
        phi, g = self.phi, self.sim.grid
        nx, ny = g.nx, g.ny
        D = np.maximum(self.diff, 1e-300)

        def ghost(inner, code, val, h, Dl):
            out = inner.copy()
            f = code == 1
            out[f] = 2 * val[f] - inner[f]
            q = code == 2
            out[q] = inner[q] + val[q] * h / Dl[q]
            return out

        if self.sim.periodic:
            phi[0, :] = phi[nx, :]
            phi[nx + 1, :] = phi[1, :]
        else:
            phi[0, 1:-1] = ghost(phi[1, 1:-1], self.bt["left"], self.bv["left"], g.dx, D[1, 1:-1])
            phi[nx + 1, 1:-1] = ghost(phi[nx, 1:-1], self.bt["right"], self.bv["right"], g.dx,
                                      D[nx, 1:-1])
        phi[1:-1, 0] = ghost(phi[1:-1, 1], self.bt["bottom"], self.bv["bottom"], g.dy, D[1:-1, 1])
        phi[1:-1, ny + 1] = ghost(phi[1:-1, ny], self.bt["top"], self.bv["top"], g.dy,
                                  D[1:-1, ny])
        solid = ~g.fluid.copy()
        solid[0, :] = solid[-1, :] = solid[:, 0] = solid[:, -1] = False
        if solid.any():
            w = g.fluid.astype(float)
            pf = np.where(g.fluid, phi, 0.0)
            s = np.zeros_like(phi)
            c = np.zeros_like(phi)
            s[1:-1, 1:-1] = pf[2:, 1:-1] + pf[:-2, 1:-1] + pf[1:-1, 2:] + pf[1:-1, :-2]
            c[1:-1, 1:-1] = w[2:, 1:-1] + w[:-2, 1:-1] + w[1:-1, 2:] + w[1:-1, :-2]
            near = solid & (c > 0)
            avg = np.where(c > 0, s / np.maximum(c, 1), 0.0)
            if self.obstacle[0] == "fixed":
                phi[near] = 2 * float(self.obstacle[1]) - avg[near]
                phi[solid & ~near] = float(self.obstacle[1])
            else:
                phi[near] = avg[near]

    def step(self, dt):
        sim, g = self.sim, self.sim.grid
        fixed = self.obstacle[0] == "fixed"
        scalar_transport(self.phi, sim.u, sim.v, g.fluid, g.nx, g.ny, g.dx, g.dy, self.diff,
                         LIMITERS[self.limiter], sim.periodic,
                         self.bt["left"], self.bv["left"], self.bt["right"], self.bv["right"],
                         self.bt["bottom"], self.bv["bottom"], self.bt["top"], self.bv["top"],
                         1 if fixed else 0, float(self.obstacle[1]) if fixed else 0.0, self.rhs)
        inner = self.phi[1:-1, 1:-1]
        new = inner + dt * self.rhs[1:-1, 1:-1]
        if self.source is not None:       # S = S_exp - S_imp phi, implicit part pointwise
            s_exp, s_imp = self.source(sim, self)
            new = (new + dt * s_exp) / (1.0 + dt * np.maximum(s_imp, 0.0))
        if self.lower is not None:
            new = np.maximum(new, self.lower)
        if self.upper is not None:
            new = np.minimum(new, self.upper)
        self.phi[1:-1, 1:-1] = np.where(g.interior, new, inner)
        self.fill_ghosts()

    def dt_limit(self):
        g = self.sim.grid
        dmax = float(self.diff[g.fluid].max()) if g.fluid.any() else self.D0
        if dmax <= 0:
            return np.inf
        return 0.5 / (dmax * (1.0 / g.dx ** 2 + 1.0 / g.dy ** 2))

    def cell(self):
        return np.where(~self.sim.grid.interior, np.nan, self.phi[1:-1, 1:-1])

    def total(self):
        g = self.sim.grid
        return float(np.nansum(self.cell()) * g.dx * g.dy)

    def wall_flux(self, side):
        g, phi = self.sim.grid, self.phi
        code, val = self.bt[side], self.bv[side]
        if side == "left":
            inner, D, h = phi[1, 1:-1], self.diff[1, 1:-1], g.dx
        elif side == "right":
            inner, D, h = phi[g.nx, 1:-1], self.diff[g.nx, 1:-1], g.dx
        elif side == "bottom":
            inner, D, h = phi[1:-1, 1], self.diff[1:-1, 1], g.dy
        else:
            inner, D, h = phi[1:-1, g.ny], self.diff[1:-1, g.ny], g.dy
        q = np.zeros_like(inner)
        f = code == 1
        q[f] = -D[f] * (val[f] - inner[f]) / (0.5 * h)
        q[code == 2] = -val[code == 2]
        return q


class OPassive:
    """Passive scalar model (models.py:293-333)."""

    kind = "scalar"

    def __init__(self, name="c", diffusivity=1e-3, bcs=None, obstacle=("zerogradient", 0.0),
                 limiter="vanleer", initial=0.0):
        self.field = OScalar(name, diffusivity, bcs, obstacle, limiter, initial)
        self.name = name

    def attach(self, sim):
        self.field.attach(sim)
        sim.scalars[self.name] = self.field

    def momentum_source(self, sim, hu, hv):
        pass

    def post_step(self, sim, dt):
        self.field.step(dt)

    def dt_limit(self, sim):
        lim = self.field.dt_limit()
        return lim if np.isfinite(lim) else None

    def monitors(self, sim):
        return {f"{self.name}_integral": self.field.total()}


class OBoussinesq:
    """Temperature with Boussinesq buoyancy (models.py:339-426)."""

    kind = "thermal"

    def __init__(self, alpha, bcs=None, gravity=(0.0, -1.0), T_ref=0.0, initial=0.0,
                 obstacle=("zerogradient", 0.0), limiter="vanleer", buoyancy=True):
        self.field = OScalar("T", alpha, bcs, obstacle, limiter, initial)
        self.alpha, self.gravity, self.T_ref = float(alpha), gravity, float(T_ref)
        self.buoyancy = buoyancy

    def attach(self, sim):
        self.field.attach(sim)
        sim.scalars["T"] = self.field

    def momentum_source(self, sim, hu, hv):
        if not self.buoyancy:
            return
        T = self.field.phi
        gx, gy = self.gravity
        if gx:
            Tf = 0.5 * (T[:-1, :] + T[1:, :]) - self.T_ref
            hu[sim.u_fluid] += -gx * Tf[sim.u_fluid]
        if gy:
            Tf = 0.5 * (T[:, :-1] + T[:, 1:]) - self.T_ref
            hv[sim.v_fluid] += -gy * Tf[sim.v_fluid]

    def post_step(self, sim, dt):
        self.field.step(dt)

    def dt_limit(self, sim):
        lim = self.field.dt_limit()
        return lim if np.isfinite(lim) else None

    def nusselt(self, side, L=1.0, dT=1.0):
        if not np.any(self.field.bt[side] == 1):
            return None
        q = self.field.wall_flux(side)
        return abs(float(np.mean(q))) * L / (self.alpha * dT) if q.size else None

    def monitors(self, sim):
        out = {"T_mittel": float(np.nanmean(self.field.cell()))}
        for s in SIDES:
            nu = self.nusselt(s)
            if nu is not None:
                out[f"Nu_{s}"] = nu
        return out


# --------------------------------------------------------------------------- #
# Time stepper
# --------------------------------------------------------------------------- #
class OSim:
    """Projection time stepper (solver.py:95-540), laminar and constant density."""

    def __init__(self, grid, bcs, nu, scheme="ab2", gamma="auto", tau=0.5, dt=None,
                 pressure_solver="direct", initial_velocity=(0.0, 0.0), models=None,
                 body_force=(0.0, 0.0), pressure_tol=1e-10, flow_rate=None):
        self.removed_cells = grid.finalize()
        self.grid, self.nu, self.scheme = grid, float(nu), scheme
        self.gamma_setting, self.tau, self.dt_fixed = gamma, float(tau), dt
        nx, ny = grid.nx, grid.ny
        self.boundary = OBoundary(grid, bcs)
        self.bcs = bcs
        self.pressure = OPressure(grid, self.boundary, pressure_solver, pressure_tol)
        self.u = np.zeros((nx + 1, ny + 2))
        self.v = np.zeros((nx + 2, ny + 1))
        self.p = np.zeros((nx + 2, ny + 2))
        self.F = np.zeros_like(self.u)
        self.G = np.zeros_like(self.v)
        self.hu = np.zeros_like(self.u)
        self.hv = np.zeros_like(self.v)
        self.hu_old = self.hv_old = self.dt_old = None
        self.u_fluid = grid.u_faces()
        self.v_fluid = grid.v_faces()
        self.periodic = self.boundary.periodic
        if self.periodic:
            self.u_fluid[0, :] = False
        self.body_force = (float(body_force[0]), float(body_force[1]))
        self.flow_rate = flow_rate
        self.flow_force = 0.0
        self.nu_t = np.zeros((nx + 2, ny + 2))
        self.nu_eff = np.full((nx + 2, ny + 2), self.nu)
        self.variable_viscosity = False
        self.variable_density = False
        self.rho = np.ones((nx + 2, ny + 2))
        self.scalars, self.models = {}, []
        for m in models or []:
            m.attach(self)
            self.models.append(m)
        u0, v0 = initial_velocity
        if u0:
            self.u[self.u_fluid] = u0
        if v0:
            self.v[self.v_fluid] = v0
        self.t, self.n, self.dt = 0.0, 0, 0.0
        self.gamma = 0.0 if gamma == "auto" else float(gamma)
        self.residual = np.inf
        self.hist = {"step": [], "t": [], "dt": [], "residual": [], "divergence": [],
                     "kinetic_energy": []}
        self.extra = {}
        self.cg_iters = []
        self.apply_bc()
        self.project_initial()

    def add_model(self, m):
        m.attach(self)
        self.models.append(m)

    def apply_bc(self):
        self.boundary.apply(self.u, self.v)
        obstacle_bc(self.u, self.v, self.grid.fluid, self.grid.nx, self.grid.ny, self.periodic)

    def project_initial(self):
        if not (np.any(self.u) or np.any(self.v)):
            return
        self.F[:] = self.u
        self.G[:] = self.v
        self.boundary.extrapolate_outflow_predictor(self.F, self.G)
        self.solve_pressure_and_correct(1.0)
        self.p[:] = 0.0
        self.apply_bc()

    def compute_dt(self):
        if self.dt_fixed:
            return float(self.dt_fixed)
        g = self.grid
        umax = max(np.abs(self.u).max(), 1e-12)
        vmax = max(np.abs(self.v).max(), 1e-12)
        if self.variable_density:
            nu_max = float((self.nu_eff / self.rho)[self.grid.fluid].max())
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

    def auto_gamma(self, dt):
        g = self.grid
        umax = np.abs(self.u).max()
        vmax = np.abs(self.v).max()
        if self.scheme == "euler":
            cfl = max(umax * dt / g.dx, vmax * dt / g.dy)
            return float(min(1.0, max(0.05, 1.2 * cfl)))
        if self.variable_density:
            nu_ref = float((self.nu_eff / self.rho)[g.fluid].min())
        else:
            nu_ref = float(self.nu_eff[g.fluid].min()) if self.variable_viscosity else self.nu
        re_cell = max(umax * g.dx, vmax * g.dy) / nu_ref
        return float(min(1.0, max(0.05, 0.015 * re_cell)))

    def solve_pressure_and_correct(self, dt):
        g = self.grid
        nx, ny, dx, dy = g.nx, g.ny, g.dx, g.dy
        F, G = self.F, self.G
        if self.periodic:
            F[0, :] = F[nx, :]
        div = (F[1:, 1:-1] - F[:-1, 1:-1]) / dx + (G[1:-1, 1:] - G[1:-1, :-1]) / dy
        if self.variable_density:     # face coefficients 1/rho (solver.py:267-273)
            bx = 2.0 / (self.rho[:-1, :] + self.rho[1:, :])
            by = 2.0 / (self.rho[:, :-1] + self.rho[:, 1:])
            if self.periodic:
                bx[nx, :] = 2.0 / (self.rho[nx, :] + self.rho[1, :])
            pin = self.pressure.solve_variable(div / dt, bx[:, 1:-1], by[1:-1, :])
        else:
            bx = by = None
            pin = self.pressure.solve(div / dt)

        def cx(sl):
            return 1.0 if bx is None else bx[sl]

        def cy(sl):
            return 1.0 if by is None else by[sl]

        p = self.p
        p[:] = 0.0
        p[1:-1, 1:-1] = pin
        if self.periodic:
            p[0, :] = p[nx, :]
            p[nx + 1, :] = p[1, :]
        uf = self.u_fluid[1:nx, 1:ny + 1]
        corr = F[1:nx, 1:ny + 1] - dt * cx((slice(1, nx), slice(1, ny + 1))) * (
            p[2:nx + 1, 1:ny + 1] - p[1:nx, 1:ny + 1]) / dx
        self.u[1:nx, 1:ny + 1] = np.where(uf, corr, F[1:nx, 1:ny + 1])
        vf = self.v_fluid[1:nx + 1, 1:ny]
        corr = G[1:nx + 1, 1:ny] - dt * cy((slice(1, nx + 1), slice(1, ny))) * (
            p[1:nx + 1, 2:ny + 1] - p[1:nx + 1, 1:ny]) / dy
        self.v[1:nx + 1, 1:ny] = np.where(vf, corr, G[1:nx + 1, 1:ny])
        if self.periodic:
            uf = self.u_fluid[nx, 1:ny + 1]
            corr = F[nx, 1:ny + 1] - dt * cx((nx, slice(1, ny + 1))) * (
                p[1, 1:ny + 1] - p[nx, 1:ny + 1]) / dx
            self.u[nx, 1:ny + 1] = np.where(uf, corr, F[nx, 1:ny + 1])
            self.u[0, :] = self.u[nx, :]
            return
        d = self.pressure.dirichlet
        self.u[0, 1:ny + 1] = np.where(
            d["left"], F[0, 1:ny + 1] - 2 * dt * cx((0, slice(1, ny + 1))) * p[1, 1:ny + 1] / dx,
            F[0, 1:ny + 1])
        self.u[nx, 1:ny + 1] = np.where(
            d["right"], F[nx, 1:ny + 1] + 2 * dt * cx((nx, slice(1, ny + 1))) * p[nx, 1:ny + 1] / dx,
            F[nx, 1:ny + 1])
        self.v[1:nx + 1, 0] = np.where(
            d["bottom"], G[1:nx + 1, 0] - 2 * dt * cy((slice(1, nx + 1), 0)) * p[1:nx + 1, 1] / dy,
            G[1:nx + 1, 0])
        self.v[1:nx + 1, ny] = np.where(
            d["top"], G[1:nx + 1, ny] + 2 * dt * cy((slice(1, nx + 1), ny)) * p[1:nx + 1, ny] / dy,
            G[1:nx + 1, ny])

    def step(self):
        g = self.grid
        self.apply_bc()
        for m in self.models:
            if hasattr(m, "pre_momentum"):
                m.pre_momentum(self)
        if self.variable_viscosity and not self.variable_density:
            np.add(self.nu, self.nu_t, out=self.nu_eff)
        dt = self.compute_dt()
        self.gamma = (self.auto_gamma(dt) if self.gamma_setting == "auto"
                      else float(self.gamma_setting))
        if self.variable_viscosity:
            momentum_tendencies_var(self.u, self.v, g.fluid, g.nx, g.ny, g.dx, g.dy, self.nu_eff,
                                    self.gamma, self.periodic, self.hu, self.hv,
                                    self.rho if self.variable_density else None)
        else:
            momentum_tendencies(self.u, self.v, g.fluid, g.nx, g.ny, g.dx, g.dy, self.nu,
                                self.gamma, self.periodic, self.hu, self.hv)
        for m in self.models:
            m.momentum_source(self, self.hu, self.hv)
        if self.scheme == "ab2" and self.hu_old is not None:
            r = dt / self.dt_old
            a, b = 1.0 + 0.5 * r, -0.5 * r
            Hu = a * self.hu + b * self.hu_old
            Hv = a * self.hv + b * self.hv_old
        else:
            Hu, Hv = self.hu, self.hv
        u_prev, v_prev = self.u.copy(), self.v.copy()
        self.F[:] = self.u
        self.G[:] = self.v
        self.F[self.u_fluid] += dt * Hu[self.u_fluid]
        self.G[self.v_fluid] += dt * Hv[self.v_fluid]
        fx = self.body_force[0] + self.flow_force
        fy = self.body_force[1]
        if fx:
            self.F[self.u_fluid] += dt * fx
        if fy:
            self.G[self.v_fluid] += dt * fy
        self.boundary.extrapolate_outflow_predictor(self.F, self.G)
        self.solve_pressure_and_correct(dt)
        self.apply_bc()
        if self.flow_rate is not None:
            self.control_flow_rate(dt)
        for m in self.models:
            m.post_step(self, dt)
        if self.scheme == "ab2":
            self.hu_old = self.hu.copy()
            self.hv_old = self.hv.copy()
        self.dt_old = dt
        self.t += dt
        self.n += 1
        self.dt = dt
        self.cg_iters.append(self.pressure.last_iterations)
        du = np.abs(self.u[self.u_fluid] - u_prev[self.u_fluid]).max(initial=0.0)
        dv = np.abs(self.v[self.v_fluid] - v_prev[self.v_fluid]).max(initial=0.0)
        self.residual = max(du, dv) / dt
        if not np.isfinite(self.residual) or self.residual > 1e8:
            raise FloatingPointError(f"unstable at step {self.n}")
        h = self.hist
        h["step"].append(self.n)
        h["t"].append(self.t)
        h["dt"].append(self.dt)
        h["residual"].append(self.residual)
        h["divergence"].append(self.max_divergence())
        h["kinetic_energy"].append(self.kinetic_energy())
        if self.flow_rate is not None:
            self.extra.setdefault("antriebskraft", []).append(self.flow_force)
        for m in self.models:
            for k, val in m.monitors(self).items():
                self.extra.setdefault(k, []).append(val)

    def bulk_velocity(self):
        """Mean u over the fluid faces 1..nx (solver.py:405-408)."""
        nx = self.grid.nx
        return float(self.u[1:nx + 1, 1:-1][self.u_fluid[1:nx + 1, 1:-1]].mean())

    def control_flow_rate(self, dt):
        """Shift u on the fluid faces to the target bulk velocity and add the
        shift / dt to the driving force; with obstacles the shifted field is
        projected again with dt = 1, keeping p (solver.py:410-425)."""
        du = self.flow_rate - self.bulk_velocity()
        self.u[self.u_fluid] += du
        if self.periodic:
            self.u[0, :] = self.u[self.grid.nx, :]
        self.flow_force += du / dt
        if (~self.grid.interior).any():
            self.F[:] = self.u
            self.G[:] = self.v
            self.boundary.extrapolate_outflow_predictor(self.F, self.G)
            keep = self.p.copy()
            self.solve_pressure_and_correct(1.0)
            self.p[:] = keep
        self.apply_bc()

    def run(self, t_end, max_steps=None, steady_tol=None):
        reason = "t_end"
        while self.t < t_end - 1e-12:
            if max_steps is not None and self.n >= max_steps:
                reason = "max_steps"
                break
            if self.dt_fixed is None:
                dt_next = self.compute_dt()
                if self.t + dt_next > t_end:
                    self.dt_fixed = t_end - self.t
                    try:
                        self.step()
                    finally:
                        self.dt_fixed = None
                else:
                    self.step()
            else:
                self.step()
            if steady_tol is not None and self.n > 10 and self.residual < steady_tol:
                reason = "konvergiert"
                break
        return reason

    def divergence(self):
        g = self.grid
        div = ((self.u[1:, 1:-1] - self.u[:-1, 1:-1]) / g.dx
               + (self.v[1:-1, 1:] - self.v[1:-1, :-1]) / g.dy)
        return np.where(g.interior, div, 0.0)

    def max_divergence(self):
        return float(np.abs(self.divergence()).max())

    def cell_velocity(self):
        uc = 0.5 * (self.u[:-1, 1:-1] + self.u[1:, 1:-1])
        vc = 0.5 * (self.v[1:-1, :-1] + self.v[1:-1, 1:])
        solid = ~self.grid.interior
        return np.where(solid, np.nan, uc), np.where(solid, np.nan, vc)

    def kinetic_energy(self):
        uc, vc = self.cell_velocity()
        return float(0.5 * np.nansum(uc ** 2 + vc ** 2) * self.grid.dx * self.grid.dy)


# --------------------------------------------------------------------------- #
# Cases (cases.py) and checkpoints (functions.py)
# --------------------------------------------------------------------------- #
def cavity(re=100.0, n=64, **num) -> OSim:
    """cases.py:60-70."""
    g = OGrid(n, n, 1.0, 1.0)
    bcs = {"left": OBC("noslip"), "right": OBC("noslip"), "bottom": OBC("noslip"),
           "top": OBC("noslip", 1.0)}
    return OSim(g, bcs, nu=1.0 / re, **num)


def obstacle_channel(circles, lx=4.0, ly=1.0, nx=320, ny=80, re=100.0, U=1.0, L_ref=1.0,
                     inflow_profile="uniform", walls="slip", **num) -> OSim:
    """cases.py:161-201 for circle obstacles ("kreis"), L_ref given explicitly."""
    g = OGrid(nx, ny, lx, ly)
    for (x, y, r) in circles:
        g.add_obstacle(g.circle(x, y, r))
    bcs = {"left": OBC("inflow", U, inflow_profile), "right": OBC("outflow"),
           "bottom": OBC(walls), "top": OBC(walls)}
    return OSim(g, bcs, nu=U * L_ref / re, **num)


def heated_cavity(ra=1e4, pr=0.71, n=48, limiter="vanleer", **num) -> OSim:
    """cases.py:207-223."""
    g = OGrid(n, n, 1.0, 1.0)
    bcs = {s: OBC("noslip") for s in SIDES}
    th = OBoussinesq(1.0 / np.sqrt(ra * pr), limiter=limiter,
                     bcs={"left": ("fixed", 0.5), "right": ("fixed", -0.5),
                          "bottom": ("zerogradient", 0.0), "top": ("zerogradient", 0.0)})
    return OSim(g, bcs, nu=np.sqrt(pr / ra), models=[th], **num)


def add_scalar(sim: OSim, name="c", diffusivity=1e-3, inflow_value=1.0, inflow_band=None,
               limiter="vanleer", initial=0.0) -> OPassive:
    """cases.py:359-379 (no obstacle value)."""
    bcs = {}
    for s in SIDES:
        if sim.bcs[s].kind == "inflow":
            if inflow_band is None:
                bcs[s] = ("fixed", inflow_value)
            else:
                a, b = inflow_band
                bcs[s] = ("fixed", lambda x, a=a, b=b: np.where((x >= a) & (x <= b),
                                                                 inflow_value, 0.0))
    m = OPassive(name, diffusivity, bcs, limiter=limiter, initial=initial)
    sim.add_model(m)
    return m


def gaussian_blobs(grid: OGrid, seed=0, count=8) -> np.ndarray:
    """Scalar initial field: sum of Gaussian blobs at cell centres, drawn in the order
    centres U(0.1,0.9)^2, sigma U(0.02,0.08), amplitude U(0,1)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.1, 0.9, (count, 2))
    s = rng.uniform(0.02, 0.08, count)
    a = rng.uniform(0.0, 1.0, count)
    X, Y = grid.cell_centers()
    phi = np.zeros_like(X)
    for k in range(count):
        phi += a[k] * np.exp(-((X - c[k, 0]) ** 2 + (Y - c[k, 1]) ** 2) / (2 * s[k] ** 2))
    return phi


def save_checkpoint(sim: OSim, path) -> Path:
    """functions.py:326-345 (keys u, v, p, t, n, dt, flow_force, dt_old, fluid, hu_old,
    hv_old, skalar_<name>, history_json)."""
    path = Path(path)
    data = {"u": sim.u, "v": sim.v, "p": sim.p, "t": sim.t, "n": sim.n, "dt": sim.dt,
            "flow_force": sim.flow_force, "dt_old": sim.dt_old if sim.dt_old else 0.0,
            "fluid": sim.grid.fluid}
    if sim.hu_old is not None:
        data["hu_old"] = sim.hu_old
        data["hv_old"] = sim.hv_old
    for name, f in sim.scalars.items():
        data[f"skalar_{name}"] = f.phi
    h = dict(sim.hist)
    data["history_json"] = json.dumps({k: list(map(float, v)) for k, v in h.items()}
                                      | {"extra": {k: list(map(float, v))
                                                   for k, v in sim.extra.items()}})
    np.savez_compressed(path, **data)
    return path
