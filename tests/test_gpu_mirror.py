"""The reference-named solver interface (PCGSolverMultiGPU, solveWaveEquation, hot_path_step)
drives the device path and reproduces the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import assert_bitwise
from gschism_amd import _lib as L
from gschism_amd import cases, core
from gschism_amd import pcg_solver as ps
from oracle import pyoracle as O
from test_gpu_parity import _setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precond", ["jacobi", "ssor"])
def test_mirror_solver_on_mesh(ctx, precond):
    spec = cases.CaseSpec("mirror-50x30", 50, 30, 500.0, nvrt=5, depth=None, substeps=4)
    mesh, om, p = _setup(ctx, spec)
    print(f"[gpu] mirror solver precond={precond}", flush=True)
    solver = ps.PCGSolverMultiGPU(ps.CudaDeviceConfig([0]), [mesh.n], [mesh.nnz], mesh=mesh, precond=precond)
    ps.backtrackAllNodes(mesh, p)
    ps.computeRHSProducers(mesh, p)
    om.backtrack(p, nthreads=8)
    om.rhs(p)
    ps.setupCoefficientMatrixAndRHS(mesh, solver, p)
    val, b = om.assemble(p)
    assert_bitwise(solver.csr_data(0).get(), val, "csr_data")
    ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    assert np.array_equal(solver.csr_indptr(0).get(), ip) and np.array_equal(solver.csr_indices(0).get(), ix)
    iters = ps.solveWaveEquation(solver, 1500, 1e-12)
    if precond == "ssor":
        xo, ito, _, _, _ = O.pcg_ssor(ip, ix, val, b, rtol=1e-12, maxiter=1500, rows_per_thread=1)
    else:
        xo, ito, _, _ = O.pcg(ip, ix, val, b, rtol=1e-12, maxiter=1500)
    assert iters == ito
    assert_bitwise(solver.solution(), xo, "mirror solution")
    assert solver.checkSolution() < 1e-10
    ps.updateElevation(mesh, solver)
    assert_bitwise(mesh.get(L.F_ETA2), xo, "eta2 after updateElevation")


def test_mirror_csr_solver_and_set_values(ctx):
    rng = np.random.default_rng(2)
    n = 800
    import scipy.sparse as sp

    A = sp.random(n, n, density=6.0 / n, random_state=8, format="csr")
    A = A + A.T
    A = (A + sp.diags(np.abs(A).sum(1).A1 + 1.0)).tocsr()
    A.sort_indices()
    b = rng.normal(size=n)
    solver = ps.PCGSolverMultiGPU(ps.CudaDeviceConfig([0]), [n], [A.nnz], indptr=A.indptr, indices=A.indices,
                                  ctx=ctx)
    solver.csr_data(0).set(A.data)
    solver.setRHS(b)
    it = solver.solve(rtol=1e-12, maxiter=500)
    xo, ito, _, _ = O.pcg(A.indptr, A.indices, A.data, b, rtol=1e-12, maxiter=500)
    assert it == ito
    assert_bitwise(solver.solution(), xo, "mirror CSR solution")
    with pytest.raises(ValueError):
        ps.PCGSolverMultiGPU(ps.CudaDeviceConfig([0]), [n], [A.nnz + 1], indptr=A.indptr, indices=A.indices, ctx=ctx)


def test_hot_path_step_matches_manual_sequence(ctx):
    spec = cases.CaseSpec("step-40x24", 40, 24, 600.0, nvrt=4, depth=None, substeps=4)
    mesh_a, _, p = _setup(ctx, spec)
    mesh_b, _, _ = _setup(ctx, spec)
    sa = ps.PCGSolverMultiGPU(ps.CudaDeviceConfig([0]), [mesh_a.n], [mesh_a.nnz], mesh=mesh_a)
    sb = core.Solver(mesh_b)
    for step in range(2):
        ia = ps.hot_path_step(mesh_a, sa, p, maxIter=1500, tol=1e-12)
        mesh_b.backtrack(p)
        mesh_b.rhs_producers(p)
        sb.assemble(p)
        _, ib, _ = sb.solve(None, rtol=1e-12, maxiter=1500)
        sb.update_elevation()
        mesh_b.momentum(p)
        mesh_b.vertical_velocity(p)
        mesh_b.update_levels(p)
        mesh_b.nodal_velocity(p)
        assert ia == ib
        assert_bitwise(mesh_a.get(L.F_ETA2), mesh_b.get(L.F_ETA2), f"step {step} eta2")
        assert_bitwise(mesh_a.get(L.F_SIDE_U2), mesh_b.get(L.F_SIDE_U2), f"step {step} side velocity")
        assert_bitwise(mesh_a.get(L.F_NODE_UVW), mesh_b.get(L.F_NODE_UVW), f"step {step} nodal velocities")


def test_native_c_caller_runs_a_step(ctx):
    """tests/c_abi/step_harness (plain C, built by build()) runs one whole step through the C ABI
    in a process of its own: every call succeeds, the solve converges, the state stays finite."""
    import subprocess
    from pathlib import Path

    exe = Path(__file__).resolve().parent / "c_abi" / "_build" / "step_harness"
    assert exe.exists(), "run __graft_entry__.build() first"
    print("[gpu] native C caller: step_harness 40 24 6", flush=True)
    r = subprocess.run([str(exe), "40", "24", "6"], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip(), r.stderr.strip(), flush=True)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "status=0" in r.stdout and "finite=1" in r.stdout
