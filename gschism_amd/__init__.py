"""gschism_amd: MI355X-native SCHISM wave-equation hot path (ELM backtracking, assembly,
Jacobi PCG) behind the C ABI of include/gsh.h."""
from . import _lib
from ._lib import GshError, StepParams, default_step_params
from .core import Context, Mesh, Solver, batch_solve

__all__ = ["Context", "Mesh", "Solver", "batch_solve", "GshError", "StepParams", "default_step_params", "_lib"]
