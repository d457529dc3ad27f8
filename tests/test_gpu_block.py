"""GPU parity of the multi-RHS block solve (gsh_block_*): ensemble members of one mesh with their
own bathymetry and right-hand side, solved together on the shared CSR pattern, against the
oracle's block restatement bit for bit and against single solves of the same systems."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import assert_bitwise
from gschism_amd import _lib as L
from gschism_amd import cases, core
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _ensemble(ctx, members, nx=96, ny=64, seed0=100):
    spec = cases.CaseSpec(f"ensemble-{nx}x{ny}", nx, ny, 1000.0, nvrt=6, depth=None, substeps=5)
    data = cases.make_case(spec)
    mesh, p, _ = cases.load_mesh(ctx, data)
    depth0 = mesh.topology(L.T_DEPTH)
    blk = core.Block(mesh, members)
    solo = core.Solver(mesh)
    vals, bs = [], []
    for m in range(members):
        d = depth0 * (1.0 + 0.05 * cases.normal(seed0 + m, mesh.n))
        mesh.set(L.F_DEPTH, d)
        mesh.update_levels(p)
        mesh.backtrack(p)
        mesh.rhs_producers(p)
        blk.assemble(m, p)
        solo.assemble(p)
        vals.append(solo.get_values())
        bs.append(solo.get_rhs())
    return mesh, p, blk, solo, np.array(vals), np.array(bs)


@pytest.mark.parametrize("members", [5, 16])
def test_block_matches_oracle_and_single_solves(ctx, members):
    mesh, p, blk, solo, vals, bs = _ensemble(ctx, members)
    ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    print(f"[gpu] block of {members} members, n={mesh.n}", flush=True)
    it, st = blk.solve(rtol=1e-12, maxiter=3000, check_every=10)
    assert (st == 0).all()
    xo, ito, _ = O.pcg_block(ip, ix, vals, bs, rtol=1e-12, maxiter=3000, nthreads=8)
    assert np.array_equal(it, ito), (it, ito)
    for m in range(members):
        x = blk.solution(m)
        assert_bitwise(x, xo[m], f"block member {m}")
        solo.set_values(vals[m])
        solo.reset_history()
        xs, its, sts = solo.solve(bs[m], rtol=1e-12, maxiter=3000)
        assert sts == 0 and its == it[m], (m, its, it[m])
        assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-10
    blk.close()
    solo.close()


def test_block_set_member_zero_rhs_and_maxiter(ctx):
    mesh, p, blk, solo, vals, bs = _ensemble(ctx, 4, nx=48, ny=40)
    ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    # replace member 1's system from the host, zero member 3's right-hand side
    v1 = vals[1] * 1.5
    b1 = bs[1][::-1].copy()
    blk.set_member(1, v1, b1)
    blk.set_member(3, None, np.zeros(mesh.n))
    vals[1], bs[1], bs[3] = v1, b1, 0.0
    it, st = blk.solve(rtol=1e-12, maxiter=3000)
    xo, ito, _ = O.pcg_block(ip, ix, vals, bs, rtol=1e-12, maxiter=3000, nthreads=8)
    assert np.array_equal(it, ito) and it[3] == 0
    for m in range(4):
        assert_bitwise(blk.solution(m), xo[m], f"member {m} after set_member")
    assert not blk.solution(3).any()
    # stopping at maxiter: every non-trivial member flagged
    it, st = blk.solve(rtol=1e-30, maxiter=25, check_every=10)
    assert (it[[0, 1, 2]] == 25).all() and (st[[0, 1, 2]] == L.GSH_W_MAXITER).all() and it[3] == 0
    xo, ito, _ = O.pcg_block(ip, ix, vals, bs, rtol=1e-30, maxiter=25, check_every=10)
    for m in range(4):
        assert_bitwise(blk.solution(m), xo[m], f"member {m} at maxiter")
    with pytest.raises(L.GshError) as ei:
        core.Block(mesh, 300)
    assert ei.value.code == L.GSH_E_ARG
    blk.close()
    solo.close()
