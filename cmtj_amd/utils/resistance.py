"""Resistance post-processing of logged magnetisations
(cmtj/utils/resistance.py:12-120).

Per layer i with unit magnetisation m = (mx, my, mz), the STRIP model gives
a longitudinal and a transverse resistance

    Rxx_i = Rx0_i + AMR_i mx^2 + SMR_i my^2
    Rxy_i = Ry0_i + AHE_i mz / 2 + (w_i / l_i)(AMR_i + SMR_i) mx my

and layers combine in series (sum) or in parallel (reciprocal of the summed
conductances).  Two-layer GMR/TMR follows the angle between the layers:
R = Rp + (Rap - Rp)(1 - cos)/2.  m is [layers, 3] for one state or
[layers, 3, T] for a time series; the outputs keep the trailing time axis.
"""
from __future__ import annotations

import numpy as np

from .filters import Filters


def _per_layer(values, n: int, ndim: int) -> np.ndarray:
    """Layer parameters as a column that broadcasts over a trailing time axis."""
    return np.asarray(values, dtype=np.float64)[:n].reshape((n,) + (1,) * (ndim - 2))


def compute_resistance(Rx0, Ry0, AMR, AHE, SMR, m, l, w) -> tuple[np.ndarray, np.ndarray]:
    """Per-layer (Rxx, Rxy) of the STRIP model, shape [layers] or [layers, T]."""
    m = np.asarray(m, dtype=np.float64)
    n = len(Rx0)
    col = lambda v: _per_layer(v, n, m.ndim)  # noqa: E731
    mx, my, mz = m[:n, 0], m[:n, 1], m[:n, 2]
    rxx = col(Rx0) + col(AMR) * np.square(mx) + col(SMR) * np.square(my)
    aspect = col(w) / col(l)
    rxy = col(Ry0) + col(AHE) * mz / 2 + aspect * (col(AMR) + col(SMR)) * (mx * my)
    return rxx, rxy


def calculate_resistance_series(Rx0, Ry0, AMR, AHE, SMR, m, l, w) -> tuple[np.ndarray, np.ndarray]:
    """Layers in series: the per-layer resistances add."""
    rxx, rxy = compute_resistance(Rx0, Ry0, AMR, AHE, SMR, m, l, w)
    return rxx.sum(axis=0), rxy.sum(axis=0)


def calculate_resistance_parallel(Rx0, Ry0, AMR, AHE, SMR, m, l, w) -> tuple[np.ndarray, np.ndarray]:
    """Layers in parallel: the per-layer conductances add."""
    rxx, rxy = compute_resistance(Rx0, Ry0, AMR, AHE, SMR, m, l, w)
    return np.reciprocal(np.reciprocal(rxx).sum(axis=0)), np.reciprocal(np.reciprocal(rxy).sum(axis=0))


def compute_gmr(Rp: float, Rap: float, m1: np.ndarray, m2: np.ndarray) -> np.ndarray:
    """GMR of two layers: Rp when parallel, Rap when antiparallel; the cosine
    is the dot product over the first (component) axis."""
    cos_angle = np.einsum("i...,i...->...", np.asarray(m1, dtype=np.float64), np.asarray(m2, dtype=np.float64))
    return Rp + (Rap - Rp) * (1.0 - cos_angle) / 2


def calculate_magnetoresistance(Rp: float, Rap: float, m: np.ndarray) -> np.ndarray:
    """The reference's two-layer helper (resistance.py:78-89) for m of shape
    [2, 3, T]: it adds (Rap - Rp)/2 times the cosine itself, not (1 - cos),
    and that is what it returns here too."""
    m = np.asarray(m, dtype=np.float64)
    if m.shape[0] != 2:
        raise ValueError(f"The magnetoresistance can only be computed for 2 layers. Current shape {m.shape}")
    return Rp + (Rap - Rp) * np.einsum("i...,i...->...", m[0], m[1]) / 2


def compute_sd(dynamic_r: np.ndarray, dynamic_i: np.ndarray, integration_step: float) -> float:
    """Spin-diode (mixing) voltage: the DC part of -I(t) R(t), taken as the
    mean of its 10 MHz order-3 Butterworth low-pass."""
    product = -np.multiply(dynamic_i, dynamic_r)
    return np.mean(Filters.butter_lowpass_filter(product, cutoff=10e6, fs=1.0 / integration_step, order=3))


__all__ = ["compute_resistance", "calculate_resistance_series", "calculate_resistance_parallel", "compute_gmr",
           "calculate_magnetoresistance", "compute_sd"]
