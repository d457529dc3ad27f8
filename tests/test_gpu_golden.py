"""The device library reproduces the golden fixtures of the wave-equation step bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import assert_bitwise
from test_cpu_golden import G, load_fixture

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", G.CASES)
def test_library_reproduces_fixture(ctx, name):
    from gschism_amd import _lib as L
    from gschism_amd import core

    print(f"[gpu] golden fixture {name}", flush=True)
    fx = load_fixture(name)
    p = G.params_from(fx["params"])
    mesh = core.Mesh(ctx, fx["xy"], fx["elnode"], fx["depth"], fx["sigma"], open_bnd=fx["open_bnd"], kb=fx["kb"],
                     renumber=False)
    mesh.set(L.F_ETA2, fx["eta"])
    mesh.set(L.F_NODE_UVW, fx["uvw"])
    mesh.set(L.F_SIDE_U2, fx["su2"])
    mesh.update_levels(p)
    solver = core.Solver(mesh)
    for s in range(G.STEPS):
        mesh.backtrack(p)
        mesh.rhs_producers(p)
        solver.assemble(p)
        x, it, st = solver.solve(None, rtol=1e-12, maxiter=2000)
        assert st & L.GSH_W_MAXITER == 0
        assert it == int(fx[f"iters_{s}"]), f"step {s} iterations"
        assert_bitwise(mesh.get(L.F_FOOT), fx[f"foot_{s}"], f"{name} step {s} foot")
        assert_bitwise(mesh.get(L.F_BSTATUS), fx[f"bstatus_{s}"], f"{name} step {s} status")
        assert_bitwise(solver.get_rhs(), fx[f"b_{s}"], f"{name} step {s} b")
        assert_bitwise(x, fx[f"x_{s}"], f"{name} step {s} x")
        if s == 0:
            assert_bitwise(mesh.get(L.F_BTU), fx["btu_0"], f"{name} btu")
            assert_bitwise(mesh.get(L.F_SIDE_HHAT), fx["hhat_0"], f"{name} hhat")
            assert_bitwise(solver.get_values(), fx["val_0"], f"{name} matrix")
        solver.update_elevation()
        mesh.momentum(p)
        mesh.vertical_velocity(p)
        assert_bitwise(mesh.get(L.F_SIDE_U2), fx[f"su2_{s}"], f"{name} step {s} side velocity")
        assert_bitwise(mesh.get(L.F_ELEM_W), fx[f"ew_{s}"], f"{name} step {s} element w")
        mesh.update_levels(p)
        mesh.nodal_velocity(p)
        assert_bitwise(mesh.get(L.F_NODE_UVW), fx[f"uvw_{s}"], f"{name} step {s} nodal velocities")
    assert_bitwise(mesh.get(L.F_ZC), fx["zc_end"], f"{name} zc")
    solver.close()
    mesh.close()
