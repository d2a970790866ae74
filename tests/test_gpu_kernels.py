"""Stateless HIP kernels (C ABI) against the reference's own outputs.

The golden vectors in tests/golden/kernels_16x12.npz were written by the
reference's kernels.py on a 16x12 grid with block, one-cell-thick and single
cell obstacles, random fields, every boundary code and all six limiters.
The kernels keep the reference's operation order and are built without FMA
contraction, so the comparison is bit-exact.
"""

import ctypes

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _dev(a, dtype=None):
    from cfd2d_amd._lib import DeviceArray

    return DeviceArray(np.ascontiguousarray(a if dtype is None else a.astype(dtype)))


def _host(d):
    return d.get()


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN / "kernels_16x12.npz")


@pytest.mark.parametrize("tag,periodic", [("np", False), ("p", True)])
@pytest.mark.parametrize("gi,gamma", [(0, 0.0), (1, 0.05), (2, 1.0)])
def test_momentum_tendencies_bitexact(gpu, gold, tag, periodic, gi, gamma):
    from cfd2d_amd import _lib

    nx, ny = int(gold["nx"]), int(gold["ny"])
    u, v = _dev(gold[f"u_{tag}"]), _dev(gold[f"v_{tag}"])
    fl = _dev(gold[f"fluid_{tag}"], np.uint8)
    hu = _dev(np.full((nx + 1, ny + 2), 7.0))
    hv = _dev(np.full((nx + 2, ny + 1), 7.0))
    rc = _lib.lib().cfd_momentum_tendencies(u.ptr, v.ptr, fl.ptr, nx, ny,
                                            float(gold["dx"]), float(gold["dy"]), 0.0137, gamma,
                                            int(periodic), hu.ptr, hv.ptr, None)
    _lib.check(rc, "momentum")
    assert np.array_equal(_host(hu), gold[f"mom_hu_{tag}_g{gi}"])
    assert np.array_equal(_host(hv), gold[f"mom_hv_{tag}_g{gi}"])


@pytest.mark.parametrize("tag,periodic", [("np", False), ("p", True)])
def test_obstacle_bc_bitexact(gpu, gold, tag, periodic):
    from cfd2d_amd import _lib

    nx, ny = int(gold["nx"]), int(gold["ny"])
    u, v = _dev(gold[f"u_{tag}"]), _dev(gold[f"v_{tag}"])
    fl = _dev(gold[f"fluid_{tag}"], np.uint8)
    _lib.check(_lib.lib().cfd_obstacle_bc(u.ptr, v.ptr, fl.ptr, nx, ny,
                                          int(periodic), None), "obstacle_bc")
    assert np.array_equal(_host(u), gold[f"obc_u_{tag}"])
    assert np.array_equal(_host(v), gold[f"obc_v_{tag}"])


@pytest.mark.parametrize("tag,periodic", [("np", False), ("p", True)])
@pytest.mark.parametrize("obs", [0, 1])
@pytest.mark.parametrize("lim", range(6))
def test_scalar_transport_bitexact(gpu, gold, tag, periodic, obs, lim):
    from cfd2d_amd import _lib

    nx, ny = int(gold["nx"]), int(gold["ny"])
    args = [_dev(gold[f"phi_{tag}"]), _dev(gold[f"u_{tag}"]), _dev(gold[f"v_{tag}"]),
            _dev(gold[f"fluid_{tag}"], np.uint8)]
    diff = _dev(gold[f"diff_{tag}"])
    bts = [_dev(gold[f"bt_{s}_{tag}"], np.int64) for s in "lrbt"]
    bvs = [_dev(gold[f"bv_{s}_{tag}"]) for s in "lrbt"]
    out = _dev(np.full((nx + 2, ny + 2), 3.0))
    obs_v = 0.37 if obs else 0.0
    rc = _lib.lib().cfd_scalar_transport(
        args[0].ptr, args[1].ptr, args[2].ptr, args[3].ptr, nx, ny,
        float(gold["dx"]), float(gold["dy"]), diff.ptr, 0.0, lim, int(periodic),
        bts[0].ptr, bvs[0].ptr, bts[1].ptr, bvs[1].ptr,
        bts[2].ptr, bvs[2].ptr, bts[3].ptr, bvs[3].ptr,
        obs, obs_v, out.ptr, None)
    _lib.check(rc, "scalar_transport")
    assert np.array_equal(_host(out), gold[f"tvd_{tag}_o{obs}_l{lim}"])


def test_bad_arguments_raise(gpu):
    from cfd2d_amd import _lib

    rc = _lib.lib().cfd_momentum_tendencies(None, None, None, 16, 12, 0.1, 0.1, 0.01, 0.0, 0,
                                            None, None, None)
    assert rc == _lib.CFD_ERR_ARG
    with pytest.raises(ValueError):
        _lib.check(rc, "momentum")
