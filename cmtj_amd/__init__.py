"""cmtj_amd -- MI355X-native macrospin (s)LLGS engine with cmtj's Python API.

``import cmtj_amd as cmtj`` gives the names of the reference module
(``Junction``, ``Layer``, ``CVector``, ``ScalarDriver``, ``AxialDriver``,
``NullDriver``, the driver aliases, the enums and their exported values, and
``constants.PhysicalConstants``).  ``cmtj_amd.batch`` runs many junctions in
one launch; ``cmtj_amd.install_as_cmtj()`` makes ``import cmtj`` resolve here.
"""
from __future__ import annotations

import sys as _sys

from . import _lib  # noqa: F401  (loads libcmtj_amd.so; raises if it is missing)
from . import constants
from .core import (  # noqa: F401
    AdaptiveIntegrationParams,
    AxialDriver,
    Axis,
    CVector,
    Junction,
    Layer,
    NullDriver,
    Reference,
    ScalarDriver,
    SolverMode,
    UpdateType,
    c_dot,
    constantDriver,
    gaussianImpulseDriver,
    gaussianStepDriver,
    posSineDriver,
    pulseDriver,
    sineDriver,
    stepDriver,
    trapezoidDriver,
)

# enum values exported at module level, as pybind11's export_values() does
# (python/cmtj.cpp:129-176); later enums win on name clashes
for _enum in (Axis, Reference, SolverMode, UpdateType):
    for _m in _enum:
        globals()[_m.name] = _m
del _enum, _m

__version__ = "0.1.0"


def install_as_cmtj() -> None:
    """Registers this package under the module name ``cmtj`` so that scripts
    written for the reference import it unchanged."""
    mod = _sys.modules[__name__]
    _sys.modules.setdefault("cmtj", mod)
    _sys.modules.setdefault("cmtj.constants", constants)
    from . import reservoir, stack, utils

    _sys.modules.setdefault("cmtj.utils", utils)
    _sys.modules.setdefault("cmtj.stack", stack)
    _sys.modules.setdefault("cmtj.reservoir", reservoir)


__all__ = [
    "AdaptiveIntegrationParams", "AxialDriver", "Axis", "CVector", "Junction", "Layer", "NullDriver", "Reference", "ScalarDriver",
    "SolverMode", "UpdateType", "c_dot", "constantDriver", "gaussianImpulseDriver", "gaussianStepDriver",
    "posSineDriver", "pulseDriver", "sineDriver", "stepDriver", "trapezoidDriver", "constants",
    "xaxis", "yaxis", "zaxis", "all", "none", "fixed", "top", "bottom", "RK4", "Heun", "EulerHeun",
    "DormandPrince", "constant", "pulse", "sine", "step", "posine", "halfsine", "trapezoid",
    "gaussimpulse", "gaussstep", "custom",
]
