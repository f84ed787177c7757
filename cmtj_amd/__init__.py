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
    written for the reference import it unchanged: ``cmtj`` and its compiled
    core ``cmtj._cmtj`` (also bare ``_cmtj``, which the reference's package
    imports), ``cmtj.constants``, ``cmtj.stack``, ``cmtj.reservoir``,
    ``cmtj.utils`` and its submodules ``procedures``, ``resistance``,
    ``linear``, ``filters`` and ``constants``."""
    mod = _sys.modules[__name__]
    from . import reservoir, stack, utils
    from .utils import constants as uconstants
    from .utils import filters, linear, procedures, resistance

    names = {
        "cmtj": mod,
        "cmtj._cmtj": mod,
        "_cmtj": mod,
        "cmtj.constants": constants,
        "cmtj.stack": stack,
        "cmtj.reservoir": reservoir,
        "cmtj.utils": utils,
        "cmtj.utils.procedures": procedures,
        "cmtj.utils.resistance": resistance,
        "cmtj.utils.linear": linear,
        "cmtj.utils.filters": filters,
        "cmtj.utils.constants": uconstants,
    }
    for name, module in names.items():
        _sys.modules.setdefault(name, module)


__all__ = [
    "AdaptiveIntegrationParams", "AxialDriver", "Axis", "CVector", "Junction", "Layer", "NullDriver", "Reference", "ScalarDriver",
    "SolverMode", "UpdateType", "c_dot", "constantDriver", "gaussianImpulseDriver", "gaussianStepDriver",
    "posSineDriver", "pulseDriver", "sineDriver", "stepDriver", "trapezoidDriver", "constants",
    "xaxis", "yaxis", "zaxis", "all", "none", "fixed", "top", "bottom", "RK4", "Heun", "EulerHeun",
    "DormandPrince", "constant", "pulse", "sine", "step", "posine", "halfsine", "trapezoid",
    "gaussimpulse", "gaussstep", "custom",
]
