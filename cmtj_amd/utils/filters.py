"""Digital low-pass used by the spin-diode mixing voltage
(cmtj/utils/filters.py:33-43): a Butterworth design applied forward only,
as a direct-form IIR recursion with zero initial state (scipy's lfilter), so
the output at sample k depends on samples <= k."""
from __future__ import annotations

import numpy as np


class Filters:
    @staticmethod
    def butter_lowpass_filter(data: np.ndarray, cutoff: float, fs: float, order: int = 5) -> np.ndarray:
        """Order-`order` Butterworth low-pass at `cutoff` Hz for samples taken
        at `fs` Hz, run causally over the last axis of `data`."""
        from scipy.signal import butter, lfilter

        num, den = butter(order, cutoff, btype="lowpass", fs=fs)
        return lfilter(num, den, data)


__all__ = ["Filters"]
