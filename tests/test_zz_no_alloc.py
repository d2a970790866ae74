"""The allocation-free entry point cfd_download_derived_noalloc against
cfd_download_derived, through the C ABI on the GPU: bit-identical node fields
(vorticity and streamfunction) on two calls in a row, the second on a smaller
grid (the shared staging then holds the first call's values beyond the second
field's end), and the size limit's error."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KINDS = {"vorticity": 0, "streamfunction": 1}


def _flow(nx, ny, steps, obstacle):
    from cfd2d_amd import cases

    if obstacle:
        case = cases.custom_obstacles([{"typ": "kreis", "x": 0.9, "y": 0.45, "r": 0.12}],
                                      lx=3.0, ly=1.0, nx=nx, ny=ny, re=100, L_ref=1.0,
                                      pressure_solver="cg", pressure_tol=1e-10)
    else:
        case = cases.lid_driven_cavity(re=400, n=nx, pressure_solver="cg", pressure_tol=1e-10)
    for _ in range(steps):
        case.sim.step()
    case.sim._flush()
    return case.sim


def _derived(sim, fn, kind):
    from cfd2d_amd import _lib

    out = np.full((sim.grid.nx + 1, sim.grid.ny + 1), np.nan)
    rc = getattr(_lib.lib(), fn)(sim._ctx.handle, KINDS[kind], _lib.ptr(out))
    _lib.check(rc, fn)
    return out


@pytest.mark.parametrize("kind", list(KINDS))
def test_noalloc_equals_allocating_entry_point(gpu, kind):
    big = _flow(192, 64, 15, obstacle=True)      # 193 x 65 nodes
    small = _flow(48, 48, 15, obstacle=False)    # 49 x 49 nodes: the second, smaller call
    for sim in (big, small):
        print(f"noalloc {kind}: {sim.grid.nx} x {sim.grid.ny}", flush=True)
        a = _derived(sim, "cfd_download_derived_noalloc", kind)
        b = _derived(sim, "cfd_download_derived", kind)
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b)
    assert np.abs(_derived(small, "cfd_download_derived_noalloc", kind)).max() > 0
    big.close()
    small.close()


def test_noalloc_rejects_grids_beyond_its_staging(gpu):
    from cfd2d_amd import cases

    sim = cases.lid_driven_cavity(re=100, n=2100, pressure_solver="cg").sim
    with pytest.raises(ValueError, match="allocation-free staging"):
        _derived(sim, "cfd_download_derived_noalloc", "vorticity")
    sim.close()
