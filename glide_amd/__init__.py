"""MI355X (gfx950) implementation of GLIDE's coupled SSA momentum + mass step
and its enthalpy column solve.

Drop-in for the reference's forward/adjoint API of that path:
``IceDynamics.forward/backward`` (glide/model.py), the multigrid hierarchy and
FAS solvers (glide/multigrid.py), the per-level operators
(glide/operators.py), the enthalpy solver and ``ThermalModel``
(glide/enthalpy.py, glide/model.py) and ``GlideStep`` (glide/torch.py).  Fields are float64
torch tensors on the ROCm device; the compute runs in hand-written HIP kernels
loaded from ``libglide_hip.so`` through a C ABI (include/glide_hip.h,
include/glide_enthalpy.h).
"""
__version__ = "0.1.0"

from .field import BroadcastOption, Constant, Field, GridEntity, LocalOption, SubgridField  # noqa: F401
from .grid import Grid  # noqa: F401
from .enthalpy import ColumnSmootherConfig, EnthalpyOperators, EnthalpyTermFlags  # noqa: F401
from .model import IceDynamics, ThermalModel  # noqa: F401
from .multigrid import (FASAdjointConfig, FASAdjointSolver, FASCDConfig, FASCDSolver, Multigrid,  # noqa: F401
                        VankaOptions)
from .operators import AdjointOperators, ForwardOperators, NewtonConfig, VankaConfig  # noqa: F401
