"""Physical constants and unit conversions of the Python helpers
(cmtj/utils/constants.py).  Constants that the engine also uses (mu0, the
gyromagnetic ratio, T->A/m, hbar, e, k_B) are read from and written to the
engine's runtime store (cmtj_amd.constants.PhysicalConstants), so a change
made here reaches the next run; the others are held here.  Module-level
names are the values at import time, as in the reference."""
from __future__ import annotations

import math

from ..constants import PhysicalConstants as _Engine

# (engine getter, engine setter, sign applied between the two views)
_ENGINE = {
    "mu0": ("magnetic_permeability", "set_magnetic_permeability", 1.0),
    "gyromagnetic_ratio": ("gyromagnetic_ratio", "set_gyromagnetic_ratio", 1.0),
    "TtoAm": ("TtoAm", "set_TtoAm", 1.0),
    "hbar": ("hbar", "set_hbar", 1.0),
    # the helpers carry the electron's (negative) charge, the engine its magnitude
    "echarge": ("elementary_charge", "set_elementary_charge", -1.0),
    "boltzmann": ("boltzmann_constant", "set_boltzmann_constant", 1.0),
}
_DEFAULTS = {"gamma": 28024e6, "gamma_rad": 1.76e11, "OetoAm": 79.57747, "me": 9.109e-31}


class Constants:
    """Getters and setters by name, e.g. Constants.mu0(), Constants.set_mu0(v)."""

    _held = dict(_DEFAULTS)

    @classmethod
    def _get(cls, name: str) -> float:
        if name in _ENGINE:
            getter, _, sign = _ENGINE[name]
            return sign * getattr(_Engine, getter)()
        return cls._held[name]

    @classmethod
    def _set(cls, name: str, value: float) -> None:
        if name in _ENGINE:
            _, setter, sign = _ENGINE[name]
            getattr(_Engine, setter)(sign * value)
        else:
            cls._held[name] = float(value)

    @classmethod
    def hplanck(cls) -> float:
        return 2 * math.pi * cls._get("hbar")

    @classmethod
    def set_hplanck(cls, value: float) -> None:
        cls._set("hbar", value / (2 * math.pi))

    @classmethod
    def AmtoT(cls) -> float:
        return 1.0 / cls._get("TtoAm")

    @classmethod
    def AmtoOe(cls) -> float:
        return 1.0 / cls._get("OetoAm")

    @classmethod
    def bohr_magneton(cls) -> float:
        return cls._get("echarge") * cls._get("hbar") / (2 * cls._get("me"))

    @classmethod
    def reset_to_defaults(cls) -> None:
        cls._held = dict(_DEFAULTS)
        _Engine.resetToDefaults()

    @classmethod
    def get_all_constants(cls) -> dict:
        names = list(_ENGINE) + list(_DEFAULTS) + ["hplanck", "AmtoT", "AmtoOe", "bohr_magneton"]
        return {k: getattr(cls, k)() for k in names}


def _accessors(name: str):
    return classmethod(lambda cls: cls._get(name)), classmethod(lambda cls, v: cls._set(name, v))


for _name in list(_ENGINE) + list(_DEFAULTS):
    _g, _s = _accessors(_name)
    setattr(Constants, _name, _g)
    setattr(Constants, "set_" + _name, _s)
del _name, _g, _s

mu0 = Constants.mu0()
gyromagnetic_ratio = Constants.gyromagnetic_ratio()
gamma = Constants.gamma()
gamma_rad = Constants.gamma_rad()
TtoAm = Constants.TtoAm()
AmtoT = Constants.AmtoT()
OetoAm = Constants.OetoAm()
AmtoOe = Constants.AmtoOe()
hplanck = Constants.hplanck()
hbar = Constants.hbar()
echarge = Constants.echarge()
me = Constants.me()
boltzmann = Constants.boltzmann()
bohr_magneton = Constants.bohr_magneton()

__all__ = ["Constants", "mu0", "gyromagnetic_ratio", "gamma", "gamma_rad", "TtoAm", "AmtoT", "OetoAm", "AmtoOe",
           "hplanck", "hbar", "echarge", "me", "boltzmann", "bohr_magneton"]
