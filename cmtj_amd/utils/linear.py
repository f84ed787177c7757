"""Field sweeps for the sweep procedures (cmtj/utils/linear.py:6-122).

A field of amplitude A at polar angle theta and azimuth phi (degrees) is
A * u(theta, phi) with the spherical unit vector
u = (sin(theta) cos(phi), sin(theta) sin(phi), cos(theta)).  Every scan
below varies one of (A, theta, phi) along an inclusive linear grid and
returns (grid, fields[points, 3]).
"""
from __future__ import annotations

import numpy as np

from ..core import CVector


def _unit(theta_deg, phi_deg) -> np.ndarray:
    """Spherical unit vectors, components on the last axis."""
    th = np.radians(np.asarray(theta_deg, dtype=np.float64))
    ph = np.radians(np.asarray(phi_deg, dtype=np.float64))
    th, ph = np.broadcast_arrays(th, ph)
    s = np.sin(th)
    return np.stack([s * np.cos(ph), s * np.sin(ph), np.cos(th)], axis=-1)


def _grid(start: float, stop: float, steps: int) -> np.ndarray:
    return np.linspace(start, stop, num=steps, endpoint=True)


class FieldScan:
    """Field vectors for amplitude and angle sweeps; angles in degrees."""

    @staticmethod
    def angle2vector(theta, phi, amplitude=1) -> CVector:
        """The field of the given amplitude along (theta, phi)."""
        ux, uy, uz = _unit(theta, phi)
        return CVector(ux * amplitude, uy * amplitude, uz * amplitude)

    @staticmethod
    def vector2angle(x, y, z) -> tuple:
        """(theta, phi, r) of a Cartesian vector, angles in degrees; works
        element-wise on arrays."""
        rho = np.hypot(x, y)
        return np.degrees(np.arctan2(rho, z)), np.degrees(np.arctan2(y, x)), np.sqrt(rho * rho + np.square(z))

    @staticmethod
    def cvector2angle(vector: CVector) -> tuple:
        return FieldScan.vector2angle(vector.x, vector.y, vector.z)

    @staticmethod
    def amplitude_scan(start: float, stop: float, steps: int, theta: float, phi: float,
                       back: bool = False) -> tuple[np.ndarray, np.ndarray]:
        """Amplitudes start..stop along a fixed direction; with back=True the
        sweep returns to start (the turning point appears once)."""
        amps = _grid(start, stop, steps)
        if back:
            amps = np.concatenate([amps, amps[-2::-1]])
        return amps, amps[:, None] * _unit(theta, phi)[None, :]

    @staticmethod
    def theta_scan(start: float, stop: float, steps: int, amplitude: float,
                   phi: float) -> tuple[np.ndarray, np.ndarray]:
        """Polar angle start..stop at fixed amplitude and azimuth."""
        thetas = _grid(start, stop, steps)
        return thetas, _unit(thetas, phi) * amplitude

    @staticmethod
    def phi_scan(start: float, stop: float, steps: int, amplitude: float,
                 theta: float) -> tuple[np.ndarray, np.ndarray]:
        """Azimuth start..stop at fixed amplitude and polar angle."""
        phis = _grid(start, stop, steps)
        return phis, _unit(theta, phis) * amplitude


__all__ = ["FieldScan"]
