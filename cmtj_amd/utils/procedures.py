"""PIMM and VSD procedures with the whole sweep in one device launch.

The reference's procedures (cmtj/utils/procedures.py:42-196, 199-331) run
one ``runSimulation`` per field point (and per frequency for VSD), each
point starting from the final state of the previous one, then post-process
every log on the host.  The default here, ``independent=False``, keeps
those chained semantics (point k starts where point k-1 ended, so hysteretic
sweeps follow their branch), one drop-in run per point.  With
``independent=True`` (opt-in) every point starts from the junction's present
magnetisation plus its own random disturbance, all points run as one batch,
and the spectra are taken on the device (one batched FFT); a sweep that
crosses a hysteretic region then gives different results from the chained
one.  Signatures, defaults, validation messages and outputs follow the
reference.
"""
from __future__ import annotations

from collections import defaultdict
from dataclasses import dataclass
from typing import Any, Callable

import numpy as np

from ..batch import JunctionBatch
from ..core import AxialDriver, Axis, CVector, NullDriver, ScalarDriver, SolverMode
from .resistance import calculate_resistance_series


@dataclass
class ResistanceParameters:
    """Resistance parameters of one layer (procedures.py:15-25); unused ones stay 0."""

    Rxx0: float = 0
    Rxy0: float = 0
    Rahe: float = 0
    Rsmr: float = 0
    Ramr: float = 0
    w: float = 0
    l: float = 0


def compute_spectrum_strip(input_m: np.ndarray, int_step: float, max_frequency: float):
    """|FFT| of a trajectory, positive half, frequencies <= max_frequency
    (procedures.py:28-39)."""
    x = np.asarray(input_m)
    n = x.shape[-1]
    yf = np.abs(np.fft.fft(x, axis=-1))[..., : n // 2]
    freqs = np.fft.fftfreq(n, int_step)[: n // 2]
    keep = np.flatnonzero(freqs <= max_frequency)
    return yf[..., keep], freqs[keep]


def _on_axis(direction, driver) -> AxialDriver:
    parts = [NullDriver(), NullDriver(), NullDriver()]
    parts[{Axis.xaxis: 0, Axis.yaxis: 1, Axis.zaxis: 2}[Axis(direction)]] = driver
    return AxialDriver(*parts)


def _disturbed(junction, n: int, disturbance: float, rng) -> dict:
    """Per-point initial magnetisation of every layer: the present one plus a
    normal disturbance, normalised (as setLayerMagnetisation does); with no
    disturbance the present one as stored (the reference then leaves the
    magnetisation alone, procedures.py:130-139)."""
    out = {}
    for lid in junction.getLayerIds():
        m = junction.getLayerMagnetisation(lid)
        base = np.array([m.x, m.y, m.z], dtype=np.float64)
        v = np.repeat(base[None, :], n, axis=0)
        if disturbance:
            v = v + rng.normal(0.0, disturbance, size=(n, 3))
            v = v / np.linalg.norm(v, axis=1, keepdims=True)
        out[lid] = v
    return out


def _weights(junction) -> np.ndarray:
    w = np.array([l.thickness * l.Ms for l in junction.layers], dtype=np.float64)
    return w / w.sum()


def PIMM_procedure(junction, Hvecs: np.ndarray, int_step: float, resistance_params: list,
                   Hoe_direction: Axis = Axis.zaxis, Hoe_excitation: float = 50, Hoe_duration: int = 3,
                   simulation_duration: float = 5e-9, wait_time: float = 0e-9, max_frequency: float = 80e9,
                   resistance_fn: Callable = calculate_resistance_series, disturbance: float = 1e-3,
                   take_last_n: int = 100, full_output: bool = False, disable_tqdm: bool = False,
                   static_only: bool = False, *, independent: bool = False, seed=None,
                   precision: str = "fp64") -> tuple[np.ndarray, np.ndarray, dict[str, Any]]:
    """Pulse-induced microwave magnetometry over the field points Hvecs
    (procedures.py:42-196): a step Oersted pulse of Hoe_excitation A/m for
    Hoe_duration steps along Hoe_direction, m logged every step; the spectrum
    of the thickness*Ms-weighted sum of the layers' m component along the
    pulse, after wait_time; Rx, Ry from resistance_fn over the last
    take_last_n entries.  Returns (spectrum [points, freqs], freqs, output)
    with output keys H, Rx, Ry, m_avg (and m_traj plus per-layer spectra
    with full_output)."""
    if wait_time > simulation_duration:
        raise ValueError("wait_time must be smaller than simulation_duration!")
    layer_ids = junction.getLayerIds()
    if len(layer_ids) != len(resistance_params):
        raise ValueError("The number of layers in the junction must match the number of resistance parameters!")
    comp = {Axis.xaxis: 0, Axis.yaxis: 1, Axis.zaxis: 2}[Axis(Hoe_direction)]
    oedriver = _on_axis(Hoe_direction, ScalarDriver.getStepDriver(0, Hoe_excitation, 0, int_step * Hoe_duration))
    Hvecs = np.asarray(Hvecs, dtype=np.float64).reshape(-1, 3)
    n = len(Hvecs)
    rng = np.random.default_rng(seed)
    wts = _weights(junction)
    nl = len(layer_ids)
    junction.setLayerOerstedFieldDriver("all", oedriver)
    if independent:
        b = JunctionBatch.replicate(junction, n, simulation_duration, int_step, int_step, SolverMode.RK4,
                                    precision=precision, log="m", log_dtype="f64")
        b.set_external_field("all", Hvecs)
        if disturbance:  # otherwise every point starts from the stored m, as in the reference
            for lid, v in _disturbed(junction, n, disturbance, rng).items():
                b.set_magnetisation(lid, v)
        r = b.run()
        ne = int(b.js[0]["n_entries"])
        traj = r.log[: n * 3 * nl * ne].view(n, nl, 3, ne)  # device, [point, layer, xyz, entry]
        final = r.final_m.cpu().numpy()[:, :nl]
        for q, l in enumerate(junction.layers):
            l.mag = CVector(*final[-1, q])
    else:
        trajs = []
        for H in Hvecs:
            junction.clearLog()
            junction.setLayerExternalFieldDriver("all", AxialDriver(*[ScalarDriver.getConstantDriver(float(h))
                                                                      for h in H]))
            if disturbance:
                for lid, v in _disturbed(junction, 1, disturbance, rng).items():
                    junction.setLayerMagnetisation(lid, CVector(*v[0]))
            junction.runSimulation(simulation_duration, int_step, int_step)
            lg = junction.getLog()
            trajs.append([[lg[f"{lid}_m{a}"] for a in "xyz"] for lid in layer_ids])
        import torch

        traj = torch.as_tensor(np.asarray(trajs, dtype=np.float64))
        ne = traj.shape[-1]
    import torch

    wt = torch.as_tensor(wts, dtype=torch.float64, device=traj.device)[None, :, None, None]
    m_traj = traj * wt  # weighted by thickness * Ms, as the reference does
    t = np.arange(ne, dtype=np.float64) * int_step
    first = int(np.argmax(t >= wait_time)) if np.any(t >= wait_time) else ne
    output = defaultdict(list)
    tail = m_traj[..., -take_last_n:].cpu().numpy()
    last = m_traj[..., -1].sum(dim=1).cpu().numpy()
    spectrum, freqs = [], None
    if not static_only:
        mixed = m_traj[:, :, comp, first:]                      # [point, layer, t]
        ns = mixed.shape[-1]
        spec_l = torch.fft.fft(mixed, dim=-1).abs()[..., : ns // 2]
        spec = torch.fft.fft(mixed.sum(dim=1), dim=-1).abs()[..., : ns // 2]
        f_all = np.fft.fftfreq(ns, int_step)[: ns // 2]
        keep = np.flatnonzero(f_all <= max_frequency)
        freqs = f_all[keep].reshape(-1, 1)
        spectrum = spec[:, keep].cpu().numpy()
        if full_output:
            spec_l = spec_l[..., keep].cpu().numpy()
    for p in range(n):
        Rx, Ry = resistance_fn([r.Rxx0 for r in resistance_params], [r.Rxy0 for r in resistance_params],
                               [r.Ramr for r in resistance_params], [r.Rahe for r in resistance_params],
                               [r.Rsmr for r in resistance_params], tail[p],
                               l=[r.l for r in resistance_params], w=[r.w for r in resistance_params])
        output["H"].append(Hvecs[p])
        output["Rx"].append(Rx)
        output["Ry"].append(Ry)
        output["m_avg"].append(last[p])
        if full_output and not static_only:
            output["m_traj"].append(m_traj[p].cpu().numpy())
            for li, lid in enumerate(layer_ids):
                output[lid].append(spec_l[p, li][:, None])
    if full_output:
        for lid in layer_ids:
            output[lid] = np.asarray(output[lid]).squeeze()
    return np.squeeze(np.asarray(spectrum)), freqs, output


def VSD_procedure(junction, Hvecs: np.ndarray, frequencies: np.ndarray, int_step: float,
                  resistance_params: list = None, Hoe_direction: Axis = Axis.yaxis, Hoe_excitation: float = 50,
                  simulation_duration: float = 30e-9, disturbance: float = 1e-3, Rtype: str = "Rz",
                  resistance_fn: Callable = calculate_resistance_series, disable_tqdm: bool = False, *,
                  independent: bool = False, seed=None, precision: str = "fp64") -> np.ndarray:
    """Voltage spin-diode spectrum (procedures.py:199-331): a sine Oersted
    drive at each frequency for each field point; V_mix = mean of the 10 MHz
    low-passed -sin(2 pi f t) * R(t) (resistance.py:12-22).  Rtype "Rz" uses
    the junction's own resistance log (R for one layer, R_a_b for two),
    "Rx"/"Ry" resistance_fn over the m logs.  With independent=True all
    (field, frequency) pairs run as one batch; by default they chain, each
    run starting from the previous run's final state.  Returns [fields, freqs]."""
    if resistance_params is None:
        resistance_params = []
    layer_ids = junction.getLayerIds()
    if Rtype == "Rz" and len(layer_ids) > 2:
        raise ValueError("Rz can only be used for 2 layer junctions. Use Rx or Ry instead.")
    elif len(resistance_params) != len(layer_ids):
        raise ValueError("The number of layers in the junction must match the number of resistance parameters!")
    if Rtype not in ("Rx", "Ry", "Rz"):
        raise ValueError("Rtype must be either Rx or Ry or Rz")
    Hvecs = np.asarray(Hvecs, dtype=np.float64).reshape(-1, 3)
    frequencies = np.asarray(frequencies, dtype=np.float64).ravel()
    nH, nF = len(Hvecs), len(frequencies)
    rng = np.random.default_rng(seed)
    nl = len(layer_ids)
    junction.setLayerOerstedFieldDriver(
        "all", _on_axis(Hoe_direction, ScalarDriver.getSineDriver(0, Hoe_excitation, float(frequencies[0]), 0)))

    def resistance(m):  # m: [n, layer, 3, T] numpy
        if Rtype == "Rz":
            return None
        out = []
        for q in range(m.shape[0]):
            Rx, Ry = resistance_fn([r.Rxx0 for r in resistance_params], [r.Rxy0 for r in resistance_params],
                                   [r.Ramr for r in resistance_params], [r.Rahe for r in resistance_params],
                                   [r.Rsmr for r in resistance_params], m[q],
                                   l=[r.l for r in resistance_params], w=[r.w for r in resistance_params])
            out.append(Rx if Rtype == "Rx" else Ry)
        return np.asarray(out)

    rkey = "R" if nl == 1 else f"R_{layer_ids[0]}_{layer_ids[1]}"
    if independent:
        n = nH * nF
        b = JunctionBatch.replicate(junction, n, simulation_duration, int_step, int_step, SolverMode.RK4,
                                    precision=precision, log="r" if Rtype == "Rz" else "m", log_dtype="f64")
        b.set_external_field("all", np.repeat(Hvecs, nF, axis=0))
        fq = np.zeros((n, 3))
        fq[:] = np.tile(frequencies, nH)[:, None]
        b.set_driver("all", "oersted", "frequency", fq)
        if disturbance:  # otherwise every point starts from the stored m, as in the reference
            for lid, v in _disturbed(junction, n, disturbance, rng).items():
                b.set_magnetisation(lid, v)
        r = b.run()
        ne = int(b.js[0]["n_entries"])
        if Rtype == "Rz":
            R = r.channel(rkey)  # stays on the device
        else:
            R = resistance(r.log[: n * 3 * nl * ne].view(n, nl, 3, ne).cpu().numpy())
        fcol = np.tile(frequencies, nH)
    else:
        Rs, fcol = [], []
        for H in Hvecs:
            for f in frequencies:
                junction.clearLog()
                junction.setLayerExternalFieldDriver(
                    "all", AxialDriver(*[ScalarDriver.getConstantDriver(float(h)) for h in H]))
                junction.setLayerOerstedFieldDriver(
                    "all", _on_axis(Hoe_direction, ScalarDriver.getSineDriver(0, Hoe_excitation, float(f), 0)))
                if disturbance:
                    for lid, v in _disturbed(junction, 1, disturbance, rng).items():
                        junction.setLayerMagnetisation(lid, CVector(*v[0]))
                junction.runSimulation(simulation_duration, int_step, int_step)
                lg = junction.getLog()
                if Rtype == "Rz":
                    Rs.append(np.asarray(lg[rkey]))
                else:
                    Rs.append(resistance(np.asarray([[[lg[f"{lid}_m{a}"] for a in "xyz"] for lid in layer_ids]]))[0])
                fcol.append(f)
        R = np.asarray(Rs)
        fcol = np.asarray(fcol)
        ne = R.shape[1]
    return _mixing_voltage(R, fcol, int_step).reshape(nH, nF)


def _mixing_voltage(R, freqs, int_step: float) -> np.ndarray:
    """compute_sd (resistance.py:12-22) of every series at once on the
    device: -sin(2 pi f t) R through the order-3 10 MHz Butterworth low-pass
    (filters.py:33-43; coefficients from scipy.signal.butter as there), then
    the mean."""
    import torch
    from scipy.signal import butter

    from .. import _lib

    fs = 1.0 / int_step
    nyq = 0.5 * fs
    b, a = butter(3, 10e6 / nyq, btype="low", analog=False)
    if not isinstance(R, torch.Tensor):
        R = torch.as_tensor(np.ascontiguousarray(R, dtype=np.float64), device="cuda")
    R = R.to(torch.float64)
    if R.dim() == 2 and R.shape[1] > 1 and R.stride(1) != 1:
        R = R.contiguous()
    return _lib.vsd_mix(R, freqs, int_step, b, a).cpu().numpy()


__all__ = ["ResistanceParameters", "PIMM_procedure", "VSD_procedure", "compute_spectrum_strip"]
