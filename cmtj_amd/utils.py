"""Host-side helpers of the reference's sweep callers (cmtj/utils/): field
scans, resistance post-processing and the low-pass filter.  They shape inputs
for, and read outputs of, the integrator; none of them integrates."""
from __future__ import annotations

import numpy as np

from .core import CVector

# cmtj/utils/constants.py:48-54 (Python-side constants of the reference)
mu0 = 12.566e-7
TtoAm = 795774.715459
AmtoT = 1.0 / TtoAm
OetoAm = 79.57747
AmtoOe = 1.0 / OetoAm


class FieldScan:
    """Field sweeps (cmtj/utils/linear.py:6-122); angles in degrees."""

    @staticmethod
    def _trig_compute(theta, phi):
        st = np.sin(np.deg2rad(theta))
        ct = np.cos(np.deg2rad(theta))
        sp = np.sin(np.deg2rad(phi))
        cp = np.cos(np.deg2rad(phi))
        return st, ct, sp, cp

    @staticmethod
    def angle2vector(theta, phi, amplitude=1) -> CVector:
        st, ct, sp, cp = FieldScan._trig_compute(theta, phi)
        return CVector(st * cp * amplitude, st * sp * amplitude, ct * amplitude)

    @staticmethod
    def vector2angle(x, y, z):
        r = np.sqrt(x**2 + y**2 + z**2)
        theta = np.rad2deg(np.arctan2(np.sqrt(x**2 + y**2), z))
        phi = np.rad2deg(np.arctan2(y, x))
        return theta, phi, r

    @staticmethod
    def amplitude_scan(start, stop, steps, theta, phi, back=False):
        Hspan = np.linspace(start, stop, endpoint=True, num=steps)
        st, ct, sp, cp = FieldScan._trig_compute(theta, phi)
        Hx = st * cp * Hspan
        Hy = st * sp * Hspan
        Hz = ct * Hspan
        if back:
            forward = np.vstack((Hx, Hy, Hz)).T
            return (np.concatenate((Hspan[:-1], Hspan[::-1]), axis=0),
                    np.concatenate((forward[:-1], forward[::-1]), axis=0))
        return Hspan, np.vstack((Hx, Hy, Hz)).T

    @staticmethod
    def theta_scan(start, stop, steps, amplitude, phi):
        theta_span = np.linspace(start, stop, endpoint=True, num=steps)
        st, ct, sp, cp = FieldScan._trig_compute(theta_span, phi)
        return theta_span, np.vstack((st * cp * amplitude, st * sp * amplitude, ct * amplitude)).T

    @staticmethod
    def phi_scan(start, stop, steps, amplitude, theta):
        phi_span = np.linspace(start, stop, endpoint=True, num=steps)
        st, ct, sp, cp = FieldScan._trig_compute(theta, phi_span)
        return phi_span, np.vstack((st * cp * amplitude, st * sp * amplitude, ct * amplitude * np.ones_like(sp))).T


def compute_resistance(Rx0, Ry0, AMR, AHE, SMR, m, l, w):
    """Per-layer Rxx/Rxy (cmtj/utils/resistance.py:25-66); m is [L, 3] or [L, 3, T]."""
    n = len(Rx0)
    m = np.asarray(m)
    shape = (n,) if m.ndim == 2 else (n, m.shape[2])
    Sx = np.zeros(shape)
    Sy = np.zeros(shape)
    for i in range(n):
        w_l = w[i] / l[i]
        Sx[i] = Rx0[i] + (AMR[i] * m[i, 0] ** 2 + SMR[i] * m[i, 1] ** 2)
        Sy[i] = Ry0[i] + 0.5 * AHE[i] * m[i, 2] + (w_l) * (SMR[i] + AMR[i]) * m[i, 0] * m[i, 1]
    return Sx, Sy


def calculate_resistance_parallel(Rx0, Ry0, AMR, AHE, SMR, m, l, w):
    Sx, Sy = compute_resistance(Rx0, Ry0, AMR, AHE, SMR, m, l, w)
    return 1.0 / np.sum(1.0 / Sx, axis=0), 1.0 / np.sum(1.0 / Sy, axis=0)


def calculate_resistance_series(Rx0, Ry0, AMR, AHE, SMR, m, l, w):
    Sx, Sy = compute_resistance(Rx0, Ry0, AMR, AHE, SMR, m, l, w)
    return np.sum(Sx, axis=0), np.sum(Sy, axis=0)


def compute_gmr(Rp, Rap, m1, m2):
    """cmtj/utils/resistance.py:69-75."""
    return Rp + 0.5 * (Rap - Rp) * (1 - np.sum(m1 * m2, axis=0))


class Filters:
    @staticmethod
    def butter_lowpass_filter(data, cutoff, fs, order=5):
        """Butterworth low-pass (cmtj/utils/filters.py:33-43)."""
        from scipy.signal import butter, lfilter

        nyq = 0.5 * fs
        b, a = butter(order, cutoff / nyq, btype="low", analog=False)
        return lfilter(b, a, data, zi=None)


def compute_sd(dynamic_r, dynamic_i, integration_step):
    """Spin-diode mixing voltage (cmtj/utils/resistance.py:12-22)."""
    SD = -dynamic_i * dynamic_r
    fs = 1.0 / integration_step
    return np.mean(Filters.butter_lowpass_filter(SD, cutoff=10e6, fs=fs, order=3))
