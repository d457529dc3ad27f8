"""Drop-in mirror of the reference solver and step interface, backed by libgsh.

Reference interfaces replaced here (paths relative to the reference checkout):
  PCGSolverMultiGPU(configCuda, N, NNZ)           src/solver/pcg_solver.py:538-671
    .solve(rtol, maxiter, ensureSymmetric, checkEvery) -> iterations    pcg_solver.py:1075-1124
    .checkSolution(label) -> ||b - A x|| / ||b||                         pcg_solver.py:1127-1143
    .csr_data(m) / .csr_indices(m) / .csr_indptr(m)                      pcg_solver.py:1146-1148
  solveWaveEquation(solver, maxIter, tol)         src/python/schism_step.py:29-47
  schism_step hot-path calls (cufunc.*)           src/python/schism_step.py:183-275

Device arrays are exposed as `DeviceArray` views (pointer, length, dtype) with host copies in
and out; they also carry `__cuda_array_interface__` so array libraries can wrap them.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .core import Context, Mesh, Solver


class CudaDeviceConfig:
    """The reference's device configuration (src/iop/cudaKernelWrapper.py:20-47): a list of devices."""

    def __init__(self, devices=(0,)):
        self.devices = list(devices)


class DeviceArray:
    """A device buffer owned by the library (pointer + length + dtype)."""

    def __init__(self, ptr: int, size: int, dtype, owner):
        self.ptr = int(ptr)
        self.size = int(size)
        self.dtype = np.dtype(dtype)
        self._owner = owner  # keeps the solver alive

    @property
    def __cuda_array_interface__(self):
        return {"shape": (self.size,), "typestr": self.dtype.str, "data": (self.ptr, False), "version": 3}

    def get(self) -> np.ndarray:
        """Host copy."""
        s = self._owner
        if self.dtype == np.float64 and self.ptr == s._csr[0]:
            return s._solver.get_values()
        if self.ptr == s._csr[1]:
            return s._indices.copy()
        if self.ptr == s._csr[2]:
            return s._indptr.copy()
        raise ValueError("no host copy for this array")

    def set(self, host) -> None:
        s = self._owner
        if self.ptr != s._csr[0]:
            raise ValueError("only the CSR values are writable")
        s._solver.set_values(np.ascontiguousarray(host, dtype=np.float64))


class PCGSolverMultiGPU:
    """Preconditioned CG on one MI355X (one row block; numParts must be 1).

    Constructed either on a mesh (the wave-equation pattern, as main.py:281 does after the
    grid is read) or on a caller CSR pattern (indptr, indices). `precision` is "fp64" (the
    reference arithmetic) or "mixed" (FP32 inner solves, FP64 residual and refinement);
    `precond` is "jacobi" (the reference's) or "ssor" (multicolour symmetric SOR).
    """

    TPB = 256
    nativeTimeout = 60.0

    def __init__(self, configCuda: CudaDeviceConfig, N, NNZ, *, mesh: Mesh | None = None, indptr=None, indices=None,
                 ctx: Context | None = None, precision: str = "fp64", precond: str = "jacobi"):
        N, NNZ = list(N), list(NNZ)
        if len(NNZ) != len(N):
            raise ValueError(f"NNZ must have length numParts={len(N)}, got {len(NNZ)}")
        if len(configCuda.devices) != len(N):
            raise ValueError(f"Number of devices ({len(configCuda.devices)}) must match numParts ({len(N)})")
        if len(N) != 1:
            raise ValueError("this build runs one GPU per solver (numParts must be 1)")
        prec = {"fp64": L.PREC_FP64, "mixed": L.PREC_MIXED}[precision]
        pc = {"jacobi": L.PRECOND_JACOBI, "ssor": L.PRECOND_SSOR}[precond]
        if mesh is not None:
            if mesh.n != N[0] or mesh.nnz != NNZ[0]:
                raise ValueError("N/NNZ do not match the mesh pattern")
            self._solver = Solver(mesh, prec, precond=pc)
            self._indptr = mesh.topology(L.T_CSR_INDPTR)
            self._indices = mesh.topology(L.T_CSR_INDICES)
        else:
            if indptr is None or indices is None:
                raise ValueError("a solver without a mesh needs indptr and indices")
            self._ctx = ctx or Context(configCuda.devices[0])
            self._indptr = np.ascontiguousarray(indptr, dtype=np.int32)
            self._indices = np.ascontiguousarray(indices, dtype=np.int32)
            if self._indptr.size != N[0] + 1 or self._indices.size != NNZ[0]:
                raise ValueError("N/NNZ do not match indptr/indices")
            self._solver = Solver(None, prec, ctx=self._ctx, indptr=self._indptr, indices=self._indices,
                                  precond=pc)
        self.configCuda = configCuda
        self.numParts = 1
        self.N_m = N
        self.NNZ_m = NNZ
        self.N_total = N[0]
        self._csr = self._solver.csr_ptrs()
        self._b = None

    # ------------------------------------------------------------------ CSR arrays
    def csr_data(self, m: int) -> DeviceArray:
        self._part(m)
        return DeviceArray(self._csr[0], self.NNZ_m[0], np.float64, self)

    def csr_indices(self, m: int) -> DeviceArray:
        self._part(m)
        return DeviceArray(self._csr[1], self.NNZ_m[0], np.int32, self)

    def csr_indptr(self, m: int) -> DeviceArray:
        self._part(m)
        return DeviceArray(self._csr[2], self.N_m[0] + 1, np.int32, self)

    def _part(self, m):
        if m != 0:
            raise IndexError("partition index out of range")

    # ------------------------------------------------------------------ right-hand side
    def setRHS(self, b) -> None:
        """b for the next solve (the reference gathers it from the mesh's qel, gatherBFromMesh)."""
        self._b = np.ascontiguousarray(b, dtype=np.float64)
        self._solver.set_rhs(self._b)

    # ------------------------------------------------------------------ public API
    def solve(self, rtol=1e-5, maxiter=None, ensureSymmetric=False, checkEvery=10):
        """Returns the iteration count (maxiter when not converged, as the reference does)."""
        if maxiter is None:
            maxiter = self.N_total * 10
        b, self._b = self._b, None
        _, it, _ = self._solver.solve(b, rtol=rtol, maxiter=int(maxiter), check_every=int(checkEvery),
                                      timeout_s=self.nativeTimeout, want_x=False)
        return it

    def checkSolution(self, label=""):
        rel = self._solver.check()
        print(f"[PCG check{(' ' + label) if label else ''}] ||r||/||b||={rel:.6e}")
        return rel

    def solution(self) -> np.ndarray:
        """Host copy of the last solution (the reference scatters it to the mesh, scatterXToMesh)."""
        return self._solver.get_solution()


def solveWaveEquation(solver: PCGSolverMultiGPU, maxIter, tol):
    """schism_step.solveWaveEquation (schism_step.py:29-47): solve, then checkSolution."""
    iters = solver.solve(rtol=tol, maxiter=maxIter)
    solver.checkSolution()
    return iters


# ---------------------------------------------------------------------- step pieces
def backtrackAllNodes(mesh: Mesh, params) -> None:
    """cufunc.backtrackAllNodes (meshdata.cu:3547-3660): every side x level."""
    mesh.backtrack(params)


def computeNodalVelocities(mesh: Mesh, params) -> None:
    """zeroNodalVelocities + computeNodalVerticalVelocities1/2 + computeNodalHorizontalVelocities1/2
    (schism_step.py:327-331, meshdata.cu:5765-5929): nodal u2/w2 from side u2 and element w."""
    mesh.nodal_velocity(params)


def computeRHSProducers(mesh: Mesh, params) -> None:
    """computeElevationBoundaryConditions + computeVerticallyIntegratedVelocity +
    computeVegetationAndDragParameters + computeElemAverageGhat (schism_step.py:214-232)."""
    mesh.rhs_producers(params)


def setupCoefficientMatrixAndRHS(mesh: Mesh, solver: PCGSolverMultiGPU, params, time: float = 0.0) -> None:
    """setupCoefficientMatrixAndRHS + rollUpElevation + setEssentialBC (schism_step.py:237-243)."""
    solver._solver.assemble(params, time)


def solveMomentumEquations(mesh: Mesh, params, time: float = 0.0) -> None:
    """computeHorGradientAtSide + solveMomentumEquations + applyMomentumBC (schism_step.py:280-288,
    meshdata.cu:5045-5446): the side velocities of the next step from the new elevation."""
    mesh.momentum(params)


def computeVerticalVelocity(mesh: Mesh, params) -> None:
    """computeVerticalVelocity (schism_step.py:294, meshdata.cu:5978-6102): element w from the
    side velocities of the momentum step."""
    mesh.vertical_velocity(params)


def updateElevation(mesh: Mesh, solver: PCGSolverMultiGPU) -> None:
    """rollUpElevation + updateElevation (meshdata.cu:3538-3544, 4449-4456)."""
    solver._solver.update_elevation()


def hot_path_step(mesh: Mesh, solver: PCGSolverMultiGPU, params, maxIter=1500, tol=1e-12) -> int:
    """One time step of the rebuilt path in schism_step order (schism_step.py:183-331): ELM
    backtracking, RHS producers, assembly, the wave-equation solve, the elevation update, the
    momentum solve, the element vertical velocity, the level update and the nodal velocities the
    next step backtracks through."""
    backtrackAllNodes(mesh, params)
    computeRHSProducers(mesh, params)
    setupCoefficientMatrixAndRHS(mesh, solver, params)
    iters = solver.solve(rtol=tol, maxiter=maxIter)
    updateElevation(mesh, solver)
    solveMomentumEquations(mesh, params)
    computeVerticalVelocity(mesh, params)
    mesh.update_levels(params)
    computeNodalVelocities(mesh, params)
    return iters
