"""Device linear solvers with the interface of ``pymoto.solvers`` (``pymoto/solvers/__init__.py``)."""
from .solvers import LinearSolver, LDAWrapper
from .iterative import (CG, ChebyshevSmoother, DampedJacobi, GeometricMultigrid, Preconditioner,
                        auto_determine_solver, auto_levels, build_hierarchy)

__all__ = ["LinearSolver", "LDAWrapper", "Preconditioner", "DampedJacobi", "ChebyshevSmoother", "GeometricMultigrid", "CG",
           "auto_determine_solver", "auto_levels", "build_hierarchy"]
