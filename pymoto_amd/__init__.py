"""MI355X-native FE design-iteration hot path with pyMOTO's Module API.

Density filter -> SIMP -> stiffness operator -> MG-preconditioned CG -> compliance -> adjoint
sensitivities, each a ``Module`` with pyMOTO's constructor signature, running in hand-written HIP
kernels for gfx950 behind the C-ABI of ``include/pymoto_amd.h`` (``pymoto_amd/lib``).
"""
import warnings

__version__ = "0.1.0"

from .core import Module, Network, Signal, SignalSlice, make_signals
from .domain import VoxelDomain
from .dyadic import DyadicMatrix
from . import solvers
from .assembly import AssembleGeneral, AssemblePoisson, AssembleStiffness, StiffnessMatrix
from .filter import DensityFilter, FilterConv
from .generic import EinSum, MathExpression
from .linalg import LinSolve, SystemOfEquations
from .optim import OC

__all__ = ["Signal", "SignalSlice", "Module", "Network", "make_signals", "VoxelDomain", "DyadicMatrix", "solvers",
           "AssembleGeneral", "AssembleStiffness", "AssemblePoisson", "StiffnessMatrix", "FilterConv", "DensityFilter", "MathExpression",
           "EinSum", "LinSolve", "SystemOfEquations", "OC"]

# deprecated aliases kept by the reference until 2.2 (pymoto/__init__.py:102-117)
DEPRECATED_NAMES = [("DomainDefinition", "VoxelDomain", "2.2"), ("DyadCarrier", "DyadicMatrix", "2.2"),
                    ("MathGeneral", "MathExpression", "2.2")]


def __getattr__(name):
    for old, new, v in DEPRECATED_NAMES:
        if name == old:
            warnings.warn(f"`{old}` is renamed `{new}` and will be removed after version {v}", DeprecationWarning,
                          stacklevel=2)
            return globals()[new]
    raise AttributeError(f"module {__name__} has no attribute {name}")


def __dir__():
    return sorted(__all__ + [n[0] for n in DEPRECATED_NAMES])
