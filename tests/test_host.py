"""Host-side logic of cfd2d_amd and the C ABI library, without a GPU."""

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT
from oracle import cfd_oracle as O


def test_library_exports_every_declared_symbol():
    from cfd2d_amd import _lib

    header = (ROOT / "include" / "cfd2d_hip.h").read_text()
    declared = sorted(set(re.findall(r"^(?:int|int64_t|const char\*)\s+(cfd_\w+)\(", header, re.M)))
    assert len(declared) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (cfd_\w+)$", out, re.M))
    assert set(declared) <= exported, set(declared) - exported
    assert set(declared) == set(_lib.exported_symbols())
    L = _lib.lib()
    for name in declared:
        assert getattr(L, name) is not None


def test_no_device_means_no_simulation():
    from cfd2d_amd import _lib, cases

    if _lib.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(RuntimeError, match="GPU"):
        cases.lid_driven_cavity(n=16)


def _pressure_pair(kind):
    from cfd2d_amd.boundary import BoundaryCondition as BC
    from cfd2d_amd.boundary import DomainBoundary
    from cfd2d_amd.grid import Grid
    from cfd2d_amd import geometry
    from cfd2d_amd.pressure import PressureSolver

    if kind == "cavity":
        g = Grid(20, 16, 1.0, 0.8)
        og = O.OGrid(20, 16, 1.0, 0.8)
        bcs = {s: BC("noslip") for s in ("left", "right", "bottom", "top")}
        obcs = {s: O.OBC("noslip") for s in ("left", "right", "bottom", "top")}
    elif kind == "channel":
        g = Grid(40, 12, 4.0, 1.0)
        og = O.OGrid(40, 12, 4.0, 1.0)
        for gg in (g,):
            gg.add_obstacle(geometry.circle(gg, 1.0, 0.5, 0.2))
            gg.add_obstacle(geometry.rectangle(gg, 2.5, 0.0, 2.8, 1.0))  # splits the domain
        og.add_obstacle(og.circle(1.0, 0.5, 0.2))
        X, Y = og.cell_centers()
        og.add_obstacle((X >= 2.5) & (X <= 2.8) & (Y >= 0.0) & (Y <= 1.0))
        bcs = {"left": BC("inflow", 1.0, "parabolic"), "right": BC("outflow"),
               "bottom": BC("slip"), "top": BC("slip")}
        obcs = {"left": O.OBC("inflow", 1.0, "parabolic"), "right": O.OBC("outflow"),
                "bottom": O.OBC("slip"), "top": O.OBC("slip")}
    else:
        g = Grid(16, 10, 1.0, 0.5)
        og = O.OGrid(16, 10, 1.0, 0.5)
        bcs = {"left": BC("periodic"), "right": BC("periodic"), "bottom": BC("noslip"),
               "top": BC("outflow")}
        obcs = {"left": O.OBC("periodic"), "right": O.OBC("periodic"),
                "bottom": O.OBC("noslip"), "top": O.OBC("outflow")}
    g.finalize()
    og.finalize()
    ps = PressureSolver(g, DomainBoundary(g, bcs), method="cg")
    ops = O.OPressure(og, O.OBoundary(og, obcs), method="cg")
    return ps, ops


@pytest.mark.parametrize("kind", ["cavity", "channel", "periodic"])
def test_pressure_description_matches_oracle_assembly(kind):
    ps, ops = _pressure_pair(kind)
    fl = ps.fluid
    idx = ps.code & 63
    assert np.array_equal(idx > 0, fl)
    diag = ps.diag_table[idx]
    assert np.array_equal(diag[fl], ops.diag)                 # exact, reference order
    assert np.array_equal(ps.inv_table[idx][fl], 1.0 / ops.diag)
    assert ps._pinned == ops.pinned
    assert len(ps.neumann_components) == len(ops.neumann)
    # matrix-free stencil with these codes reproduces the assembled matrix
    nx, ny = fl.shape
    rng = np.random.default_rng(0)
    x = np.where(fl, rng.standard_normal(fl.shape), 0.0)
    xp = np.zeros((nx + 2, ny + 2))
    xp[1:-1, 1:-1] = x
    if ps.periodic:
        xp[0, 1:-1], xp[-1, 1:-1] = x[-1], x[0]
    ax, ay = 1 / ps.grid.dx ** 2, 1 / ps.grid.dy ** 2
    Ax = diag * x - ax * (xp[2:, 1:-1] + xp[:-2, 1:-1]) - ay * (xp[1:-1, 2:] + xp[1:-1, :-2])
    assert np.allclose(Ax[fl], ops.A @ x[fl], rtol=1e-13, atol=1e-9 * ax)
    dirx = (ps.code & 64) > 0
    assert np.array_equal(dirx[-1, :], ops.dirichlet["right"] & fl[-1, :])


def test_inflow_profile_matches_oracle():
    from cfd2d_amd.boundary import inflow_profile

    y = (np.arange(40) + 0.5) / 40
    open_mask = np.ones(40, bool)
    open_mask[10:14] = False
    open_mask[30] = False
    for prof in ("uniform", "parabolic"):
        a = inflow_profile(y, open_mask, 1 / 40, 1.3, prof)
        b = O.inflow_profile(y, open_mask, 1 / 40, 1.3, prof)
        assert np.array_equal(a, b)
    p = inflow_profile(y, np.ones(40, bool), 1 / 40, 1.0, "parabolic")
    assert p.sum() / 40 == pytest.approx(1.0, rel=1e-3)


def test_grid_finalize_removes_enclosed_fluid():
    from cfd2d_amd import geometry
    from cfd2d_amd.grid import Grid

    g = Grid(40, 40, 1.0, 1.0)
    ring = geometry.circle(g, 0.5, 0.5, 0.3) & ~geometry.circle(g, 0.5, 0.5, 0.2)
    g.add_obstacle(ring)
    assert g.finalize() > 0
    assert g.solid[20, 20]


def _pockets_by_search(inner):
    """Fluid cells with no face-neighbour path through fluid to the outer ring,
    by breadth-first search from every ring fluid cell; with the count of
    fluid regions (a second search per unvisited cell)."""
    nx, ny = inner.shape
    seen = np.zeros_like(inner)
    queue = [(i, j) for i in range(nx) for j in range(ny)
             if inner[i, j] and (i in (0, nx - 1) or j in (0, ny - 1))]
    for c in queue:
        seen[c] = True
    while queue:
        i, j = queue.pop()
        for a, b in ((i + 1, j), (i - 1, j), (i, j + 1), (i, j - 1)):
            if 0 <= a < nx and 0 <= b < ny and inner[a, b] and not seen[a, b]:
                seen[a, b] = True
                queue.append((a, b))
    label = np.zeros(inner.shape, int)
    regions = 0
    for i in range(nx):
        for j in range(ny):
            if inner[i, j] and not label[i, j]:
                regions += 1
                label[i, j] = regions
                stack = [(i, j)]
                while stack:
                    p, q = stack.pop()
                    for a, b in ((p + 1, q), (p - 1, q), (p, q + 1), (p, q - 1)):
                        if 0 <= a < nx and 0 <= b < ny and inner[a, b] and not label[a, b]:
                            label[a, b] = regions
                            stack.append((a, b))
    return inner & ~seen, regions


def test_grid_finalize_matches_search():
    """finalize against a breadth-first search on random obstacle masks, plus
    the closed-region cases: one enclosed region stays, two enclosed regions
    both go, diagonal contact is no connection."""
    from cfd2d_amd.grid import Grid

    rng = np.random.default_rng(5)
    masks = [rng.random((n, m)) < f for n, m, f in
             [(8, 8, 0.3), (9, 12, 0.45), (12, 7, 0.55), (10, 10, 0.6)] for _ in range(25)]
    one = np.ones((8, 8), bool)
    one[2:6, 2:6] = False                      # a single closed region
    two = one.copy()
    two[:, 4] = True                           # two closed regions of 8 and 4 cells
    diag = np.zeros((6, 6), bool)
    diag[1:4, 1:4] = True
    diag[2, 2] = False                         # pocket at (2, 2) ...
    diag[1, 1] = diag[3, 3] = False            # ... touching fluid only diagonally
    masks += [one, two, diag]
    for mask in masks:
        g = Grid(*mask.shape, 1.0, 1.0)
        g.add_obstacle(mask)
        inner = g.fluid_interior.copy()
        pocket, regions = _pockets_by_search(inner)
        if regions <= 1:
            pocket[:] = False
        assert g.finalize() == int(pocket.sum())
        assert np.array_equal(g.fluid_interior, inner & ~pocket)
    g = Grid(8, 8, 1.0, 1.0)
    g.add_obstacle(one)
    assert g.finalize() == 0
    g = Grid(8, 8, 1.0, 1.0)
    g.add_obstacle(two)
    assert g.finalize() == 12 and not g.fluid_interior.any()


def test_validation_errors():
    from cfd2d_amd.boundary import BoundaryCondition as BC
    from cfd2d_amd.grid import Grid
    from cfd2d_amd.solver import Simulation

    with pytest.raises(ValueError):
        BC("wand")
    with pytest.raises(ValueError):
        Grid(3, 8, 1, 1)
    g = Grid(8, 8, 1, 1)
    bcs = {s: BC("noslip") for s in ("left", "right", "bottom", "top")}
    for kw in ({"nu": 0.0}, {"nu": 1.0, "scheme": "rk4"}, {"nu": 1.0, "tau": 1.5},
               {"nu": 1.0, "gamma": 2.0}, {"nu": 1.0, "pressure_solver": "lu"}):
        with pytest.raises(ValueError):
            Simulation(g, bcs, **kw)


def test_case_json_roundtrip(tmp_path):
    from cfd2d_amd import cases

    spec = {"fall": "custom", "parameter": {"obstacles": [{"typ": "kreis", "x": 1, "y": 0.5,
                                                           "r": 0.1}], "nx": 64, "ny": 16},
            "numerik": {"scheme": "ab2", "pressure_solver": "cg_mixed", "pressure_tol": 1e-6},
            "skalar": {"name": "c", "diffusivity": 1e-3, "inflow_band": [0.25, 0.75],
                       "limiter": "superbee"}}
    cases.save_case(spec, tmp_path / "case.json")
    assert cases.load_case(tmp_path / "case.json") == spec
    with pytest.raises(ValueError):
        cases.build_case({"fall": "warp_drive"})


def test_ctypes_structs_match_the_c_header(tmp_path):
    """Every struct of include/cfd2d_hip.h has the size its ctypes mirror has."""
    import ctypes

    from cfd2d_amd import _lib

    pairs = {"cfd_grid_desc": _lib.GridDesc, "cfd_scalar_desc": _lib.ScalarDesc,
             "cfd_step_params": _lib.StepParams, "cfd_state": _lib.State,
             "cfd_record": _lib.Record, "cfd_mg_desc": _lib.MGDesc,
             "cfd_learned_desc": _lib.LearnedDesc, "cfd_learned_stats_t": _lib.LearnedStats,
             "cfd_ens_desc": _lib.EnsDesc, "cfd_ens_stats_t": _lib.EnsStats,
             "cfd_ens_tuning_t": _lib.EnsTuning,
             "cfd_probe_desc": _lib.ProbeDesc, "cfd_turb_desc": _lib.TurbDesc,
             "cfd_vof_desc": _lib.VofDesc}
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "cfd2d_hip.h"\nint main(void){\n' +
                   "".join(f'printf("%zu\\n", sizeof({k}));\n' for k in pairs) + "return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert got == [ctypes.sizeof(v) for v in pairs.values()], dict(zip(pairs, got))


def test_ensemble_workload_sampler():
    from cfd2d_amd.ensemble import sample_instances

    a = sample_instances(256, 0)
    b = sample_instances(256, 0)
    assert a == b and len(a) == 256
    assert {i.n for i in a} <= {64, 128, 256, 512}
    lid = [i for i in a if i.kind == "lid"]
    hot = [i for i in a if i.kind == "heated"]
    assert len(lid) == len(hot) == 128
    assert all(100 <= i.re <= 5000 for i in lid) and all(1e4 <= i.ra <= 1e6 for i in hot)
    assert hot[0].nu == pytest.approx(np.sqrt(0.71 / hot[0].ra))
