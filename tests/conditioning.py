"""Sensitivity of the oracle's own trajectory to a tiny initial change."""
from __future__ import annotations

import numpy as np

from cmtj_amd import CVector


def sensitivity(builder, J, eps: float = 1e-14) -> float:
    """max |m - m'| over the log when layer 0 starts eps away in m_y."""
    j1, kw = builder(J)
    j2, _ = builder(J)
    j1.setNoiseSeed(5)
    j2.setNoiseSeed(5)
    l = j2.layers[0]
    l.mag = CVector(l.mag.x, l.mag.y + eps, l.mag.z)
    j1.runSimulation(**kw)
    j2.runSimulation(**kw)
    a, b = j1.getLog(), j2.getLog()
    return max(float(np.max(np.abs(np.asarray(a[k]) - np.asarray(b[k])))) for k in a if "_m" in k)
