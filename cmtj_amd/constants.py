"""Runtime-mutable physical constants (reference python/cmtj.cpp:496-511,
core/constants.hpp:6-53).  The values live in libcmtj_amd.so; every run takes
a snapshot of them at launch."""
from __future__ import annotations

from . import _lib


class PhysicalConstants:
    @staticmethod
    def _get(name: str) -> float:
        return float(_lib.constants_get()[name])

    @staticmethod
    def _set(name: str, value: float) -> None:
        c = _lib.constants_get()
        c[name] = float(value)
        _lib.constants_set(c)

    @staticmethod
    def set_magnetic_permeability(value: float) -> None:
        PhysicalConstants._set("mu0", value)

    @staticmethod
    def set_gyromagnetic_ratio(value: float) -> None:
        PhysicalConstants._set("gyro", value)

    @staticmethod
    def set_TtoAm(value: float) -> None:
        PhysicalConstants._set("TtoAm", value)

    @staticmethod
    def set_hbar(value: float) -> None:
        PhysicalConstants._set("hbar", value)

    @staticmethod
    def set_elementary_charge(value: float) -> None:
        PhysicalConstants._set("echarge", value)

    @staticmethod
    def set_boltzmann_constant(value: float) -> None:
        PhysicalConstants._set("kB", value)

    @staticmethod
    def magnetic_permeability() -> float:
        return PhysicalConstants._get("mu0")

    @staticmethod
    def gyromagnetic_ratio() -> float:
        return PhysicalConstants._get("gyro")

    @staticmethod
    def TtoAm() -> float:
        return PhysicalConstants._get("TtoAm")

    @staticmethod
    def hbar() -> float:
        return PhysicalConstants._get("hbar")

    @staticmethod
    def elementary_charge() -> float:
        return PhysicalConstants._get("echarge")

    @staticmethod
    def boltzmann_constant() -> float:
        return PhysicalConstants._get("kB")

    @staticmethod
    def resetToDefaults() -> None:
        _lib.constants_reset()


__all__ = ["PhysicalConstants"]
