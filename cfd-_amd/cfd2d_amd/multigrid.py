"""Aggregation multigrid hierarchy for the pressure system (host setup).

The preconditioner of the ``cg_mixed`` pressure solver is an FP32 V-cycle on a
hierarchy built once per geometry:

* level 0 is the fine system ``A`` (5-point -Laplacian on fluid cells with the
  Dirichlet and pinned-cell diagonal terms, applied matrix-free on the device);
* level l+1 aggregates 2x2 cells of level l: ``A_{l+1} = s P^T A_l P`` with
  ``P`` the piecewise-constant prolongation and ``s = 0.5`` (the factor that
  maps the aggregation Galerkin operator onto the rediscretised coarse
  Laplacian); aggregates of face-adjacent cells stay face-adjacent, so every
  level is again a 5-point stencil (diagonal, east and north couplings) and
  masks, outflow Dirichlet faces, periodic wrap and pinned cells carry over;
* the last level (at most 64 cells) is solved with a dense inverse.

Smoothing is damped Jacobi (omega = 0.8), two sweeps before and after the
coarse correction, which keeps the V-cycle symmetric.
"""

from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib

COARSE_MAX = 64
SCALE = 0.5
OMEGA = 0.8
WC = 1.0


def assemble(fluid: np.ndarray, diag: np.ndarray, ax: float, ay: float,
             periodic: bool) -> sp.csr_matrix:
    """The fine pressure matrix over fluid cells (row-major numbering)."""
    n = int(fluid.sum())
    num = -np.ones(fluid.shape, dtype=np.int64)
    num[fluid] = np.arange(n)
    rows, cols, vals = [num[fluid]], [num[fluid]], [diag[fluid]]
    pairs = [(fluid[:-1, :] & fluid[1:, :], num[:-1, :], num[1:, :], ax),
             (fluid[:, :-1] & fluid[:, 1:], num[:, :-1], num[:, 1:], ay)]
    if periodic:
        pairs.append((fluid[-1, :] & fluid[0, :], num[-1, :], num[0, :], ax))
    for m, a, b, c in pairs:
        rows += [a[m], b[m]]
        cols += [b[m], a[m]]
        vals += [np.full(int(m.sum()), -c)] * 2
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


def _aggregate(fluid: np.ndarray):
    nx, ny = fluid.shape
    cnx, cny = (nx + 1) // 2, (ny + 1) // 2
    agg = (np.arange(nx) // 2)[:, None] * cny + (np.arange(ny) // 2)[None, :]
    cfl = np.zeros(cnx * cny, dtype=bool)
    cfl[np.unique(agg[fluid])] = True
    cnum = -np.ones(cnx * cny, dtype=np.int64)
    cnum[cfl] = np.arange(int(cfl.sum()))
    cols = cnum[agg[fluid]]
    P = sp.csr_matrix((np.ones(cols.size), (np.arange(cols.size), cols)),
                      shape=(int(fluid.sum()), int(cfl.sum())))
    return P, cfl.reshape(cnx, cny)


def five_point(A: sp.csr_matrix, fluid: np.ndarray, periodic: bool):
    """diag, east and north coupling magnitudes of a 5-point operator."""
    nx, ny = fluid.shape
    num = -np.ones(fluid.shape, dtype=np.int64)
    num[fluid] = np.arange(int(fluid.sum()))
    A = A.tocsr()
    d = np.zeros(fluid.shape)
    d[fluid] = A.diagonal()
    ce = np.zeros(fluid.shape)
    cn = np.zeros(fluid.shape)
    east = np.roll(num, -1, axis=0) if periodic else np.vstack([num[1:], -np.ones((1, ny), np.int64)])
    m = fluid & (east >= 0)
    ce[m] = -np.asarray(A[num[m], east[m]]).ravel()
    north = np.hstack([num[:, 1:], -np.ones((nx, 1), np.int64)])
    m = fluid & (north >= 0)
    cn[m] = -np.asarray(A[num[m], north[m]]).ravel()
    # the stencil must reproduce A exactly
    R = sp.diags(d[fluid])
    for c, nb in ((ce, east), (cn, north)):
        mm = fluid & (nb >= 0) & (c != 0)
        off = sp.csr_matrix((-c[mm], (num[mm], nb[mm])), shape=A.shape)
        R = R + off + off.T
    if abs(R - A).max() > 1e-9 * abs(A).max():
        raise ValueError("coarse operator is not a 5-point stencil")
    return d, ce, cn


class Hierarchy:
    def __init__(self, fluid: np.ndarray, diag: np.ndarray, ax: float, ay: float,
                 periodic: bool):
        self.periodic = periodic
        A = assemble(fluid, diag, ax, ay, periodic)
        self.levels = [{"nx": fluid.shape[0], "ny": fluid.shape[1], "fluid": fluid, "A": A}]
        fl = fluid
        while fl.size > COARSE_MAX:
            cnx = (fl.shape[0] + 1) // 2
            if periodic and cnx < 3:
                break
            if len(self.levels) >= 15:
                break
            P, cfl = _aggregate(fl)
            A = (SCALE * (P.T @ A @ P)).tocsr()
            d, ce, cn = five_point(A, cfl, periodic)
            self.levels.append({"nx": cfl.shape[0], "ny": cfl.shape[1], "fluid": cfl, "A": A,
                                "diag": d, "ce": ce, "cn": cn})
            fl = cfl
        if len(self.levels) < 2:
            raise ValueError("grid too small for the multigrid pressure solver")
        last = self.levels[-1]
        n = last["nx"] * last["ny"]
        if n > 4096:
            raise ValueError("multigrid coarsest level too large (grid aspect ratio)")
        full = np.eye(n)
        idx = np.flatnonzero(last["fluid"].ravel())
        full[np.ix_(idx, idx)] = last["A"].toarray()
        self.cinv = np.linalg.inv(full).astype(np.float32)
        self.cinv[:, np.flatnonzero(~last["fluid"].ravel())] = 0.0

    def upload(self, ctx) -> None:
        d = _lib.MGDesc()
        d.nlev = len(self.levels)
        d.omega, d.wc = OMEGA, WC
        last = self.levels[-1]
        d.ncoarse = last["nx"] * last["ny"]
        keep = [np.ascontiguousarray(self.cinv)]
        d.cinv = _lib.ptr(keep[0])
        for k, L in enumerate(self.levels):
            d.lev[k].nx, d.lev[k].ny = L["nx"], L["ny"]
            if k == 0:
                continue
            arrs = [np.ascontiguousarray(L[n], dtype=np.float32) for n in ("diag", "ce", "cn")]
            keep += arrs
            d.lev[k].diag, d.lev[k].ce, d.lev[k].cn = (_lib.ptr(a) for a in arrs)
        _lib.check(_lib.lib().cfd_set_multigrid(ctx.handle, ctypes.byref(d)), "multigrid setup")
