"""Post-processing of batched logs on the device: oscillation spectra and
switching statistics.

The reference post-processes one log at a time on the host with scipy
(cmtj/utils/procedures.py: spectra from the FFT of the m_x log; switching
outcomes from the final state).  Here the logs of a whole batch stay on the
GPU: one batched real FFT over [instance, entry]."""
from __future__ import annotations

import math


def spectrum_peaks(x, dt: float, discard: float = 0.2):
    """Peak frequency and full width at half maximum of the power spectrum of
    each row of x ([n, T] device tensor sampled every dt seconds), after
    dropping the first `discard` fraction (transient) and the mean.  Returns
    (freq [n], fwhm [n], power [n, F]) as tensors; the FWHM interpolates the
    half-maximum crossings linearly between bins."""
    import torch

    x = x.double()
    t0 = int(x.shape[1] * discard)
    y = x[:, t0:]
    y = y - y.mean(dim=1, keepdim=True)
    n = y.shape[1]
    p = torch.fft.rfft(y, dim=1).abs() ** 2
    p[:, 0] = 0.0
    df = 1.0 / (n * dt)
    k = p.argmax(dim=1)
    pk = p.gather(1, k[:, None])[:, 0]
    half = pk / 2
    F = p.shape[1]
    idx = torch.arange(F, device=p.device)[None, :]
    above = p >= half[:, None]
    # walk outwards from the peak to the first bins below half maximum
    left_mask = (~above) & (idx < k[:, None])
    right_mask = (~above) & (idx > k[:, None])
    left = torch.where(left_mask, idx, torch.full_like(idx, -1)).max(dim=1).values
    right = torch.where(right_mask, idx, torch.full_like(idx, F)).min(dim=1).values

    def cross(lo, hi):
        lo = lo.clamp(0, F - 1)
        hi = hi.clamp(0, F - 1)
        plo = p.gather(1, lo[:, None])[:, 0]
        phi = p.gather(1, hi[:, None])[:, 0]
        w = torch.where(phi != plo, (half - plo) / (phi - plo), torch.zeros_like(plo))
        return lo.double() + w * (hi - lo).double()

    fl = cross(left, left + 1)
    fr = cross(right - 1, right)
    fwhm = (fr - fl) * df
    return k.double() * df, fwhm, p


def switching_probability(final_mz, n_groups: int):
    """Fraction of instances whose final m_z is negative, per consecutive
    group of instances ([n] -> [n_groups]), with the binomial standard
    error."""
    import torch

    s = (final_mz < 0).double().view(n_groups, -1)
    p = s.mean(dim=1)
    m = s.shape[1]
    return p, torch.sqrt(p * (1 - p) / max(m, 1))


__all__ = ["spectrum_peaks", "switching_probability"]
_ = math
