"""jfem: MI355X elasticity path of JuliaFEM (host side of libjfem_hip.so)."""
from . import mesh
from ._lib import Context, JfemError, load
from .problems import (Assembly, Dirichlet, Elasticity, LinearSolver, NonlinearSolver, Problem, assemble, run,
                       run_nonlinear, solve)

__all__ = ["mesh", "Context", "JfemError", "load", "Assembly", "Dirichlet", "Elasticity", "LinearSolver",
           "NonlinearSolver", "Problem", "assemble", "run", "run_nonlinear", "solve"]
