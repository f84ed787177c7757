"""Scripts written for the reference import their modules unchanged after
install_as_cmtj(): the package, its compiled core and the utils submodules
that the reference's examples use."""
from __future__ import annotations

import importlib
import sys

import pytest


def test_reference_import_paths_resolve():
    import cmtj_amd

    cmtj_amd.install_as_cmtj()
    from cmtj.utils.procedures import PIMM_procedure, ResistanceParameters, VSD_procedure  # noqa: F401
    from cmtj.utils.resistance import calculate_resistance_parallel, compute_sd  # noqa: F401
    from cmtj.utils.linear import FieldScan  # noqa: F401
    from cmtj.utils.filters import Filters  # noqa: F401
    from cmtj.utils import FieldScan as FS, TtoAm, mu0  # noqa: F401
    from cmtj import CVector, Junction, Layer  # noqa: F401
    from cmtj._cmtj import ScalarDriver  # noqa: F401
    from cmtj.constants import PhysicalConstants  # noqa: F401
    from cmtj.stack import SeriesStack  # noqa: F401
    from cmtj.reservoir import GroupInteraction  # noqa: F401
    for name in ("cmtj", "cmtj._cmtj", "cmtj.utils", "cmtj.utils.procedures", "cmtj.utils.resistance",
                 "cmtj.utils.linear", "cmtj.utils.filters", "cmtj.utils.constants", "cmtj.stack",
                 "cmtj.reservoir", "cmtj.constants"):
        assert importlib.import_module(name) is sys.modules[name]
    assert sys.modules["cmtj.utils.procedures"].PIMM_procedure is PIMM_procedure


def test_procedures_chain_points_by_default():
    """The reference chains sweep points (each starts from the previous
    point's final state); the batched independent mode is opt-in."""
    import inspect

    from cmtj_amd.utils.procedures import PIMM_procedure, VSD_procedure

    for fn in (PIMM_procedure, VSD_procedure):
        assert inspect.signature(fn).parameters["independent"].default is False


def test_utils_constants_share_the_engine_store():
    from cmtj_amd.constants import PhysicalConstants
    from cmtj_amd.utils.constants import Constants

    try:
        Constants.set_mu0(1.3e-6)
        assert PhysicalConstants.magnetic_permeability() == 1.3e-6
        assert Constants.echarge() == -PhysicalConstants.elementary_charge()
        Constants.set_OetoAm(80.0)
        assert Constants.AmtoOe() == pytest.approx(1 / 80.0)
    finally:
        Constants.reset_to_defaults()
    assert Constants.mu0() == 12.566e-7
