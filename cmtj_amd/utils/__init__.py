"""Host-side helpers of the reference's sweep callers (cmtj/utils/): field
scans, resistance post-processing, the low-pass filter, constants, and the
PIMM/VSD procedures.  Only the procedures integrate (through the engine);
the rest shape inputs and read outputs.  Submodules keep the reference's
import paths (``cmtj_amd.utils.procedures`` for ``cmtj.utils.procedures``)."""
from __future__ import annotations

from .constants import (AmtoOe, AmtoT, Constants, OetoAm, TtoAm, bohr_magneton, boltzmann, echarge, gamma,
                        gamma_rad, gyromagnetic_ratio, hbar, hplanck, me, mu0)
from .filters import Filters
from .linear import FieldScan
from .resistance import (calculate_magnetoresistance, calculate_resistance_parallel, calculate_resistance_series,
                         compute_gmr, compute_resistance, compute_sd)

__all__ = ["Filters", "FieldScan", "compute_sd", "compute_resistance", "calculate_magnetoresistance",
           "calculate_resistance_series", "calculate_resistance_parallel", "compute_gmr", "Constants", "mu0",
           "gyromagnetic_ratio", "gamma", "gamma_rad", "TtoAm", "AmtoT", "OetoAm", "AmtoOe", "hplanck", "hbar",
           "echarge", "me", "boltzmann", "bohr_magneton"]
