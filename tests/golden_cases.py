"""The reference's curated-example regressions, restated as functions of the
Junction class that runs them (tests/test_curated_examples.py of the
reference).  Parameters and call sequences are the reference test's -- they
are the inputs that produced the golden JSON files under tests/golden/ --
written here in this project's own code.  Each case returns the dict the
reference compares (3 decimals)."""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np

from cmtj_amd import AxialDriver, CVector, Layer, NullDriver, ScalarDriver, SolverMode
from cmtj_amd import constantDriver, sineDriver, stepDriver
from cmtj_amd.utils import (FieldScan, Filters, OetoAm, TtoAm, calculate_resistance_parallel,
                            calculate_resistance_series, compute_gmr)

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name: str) -> dict:
    return json.loads((GOLDEN / f"{name}.json").read_text())


def assert_matches(name: str, data: dict) -> None:
    """compare_with_reference of the reference test (3 decimals)."""
    ref = load(name)
    for key, val in data.items():
        assert key in ref, key
        if isinstance(val, list) and val and isinstance(val[0], np.ndarray):
            assert len(val) == len(ref[key])
            for i, (c, r) in enumerate(zip(val, ref[key])):
                np.testing.assert_array_almost_equal(np.round(c, 3), np.asarray(r), decimal=3,
                                                     err_msg=f"{name}:{key}[{i}]")
        elif isinstance(val, np.ndarray):
            r = np.asarray(ref[key])
            assert val.shape == r.shape, (key, val.shape, r.shape)
            np.testing.assert_array_almost_equal(np.round(val, 3), r, decimal=3, err_msg=f"{name}:{key}")
        else:
            assert list(val) == list(ref[key])


def _m(log, ids):
    return np.asarray([[log[f"{i}_mx"], log[f"{i}_my"], log[f"{i}_mz"]] for i in ids])


# reference tests/test_curated_examples.py:884-948
def case_sto(J) -> dict:
    demag = [CVector(0.0, 0.0, 0.0), CVector(0.0, 0.0, 0.0), CVector(0.0, 0.0, 1.0)]
    free = Layer.createSTTLayer(
        id="free", mag=CVector(0.0, 0.0, 1.0), anis=CVector(0, 0.0, 1.0), Ms=1.0, thickness=1.4e-9,
        cellSurface=7e-10 * 7e-10, demagTensor=demag, damping=0.3, SlonczewskiSpacerLayerParameter=1.0,
        spinPolarisation=1.0, beta=1,
    )
    free.setReferenceLayer(CVector(0, 1.0, 1.0))
    j = J([free], 100, 200)
    j.setLayerAnisotropyDriver("free", constantDriver(350e3))
    j.setLayerCurrentDriver("free", constantDriver(-6e9))
    j.runSimulation(50e-9, 1e-13, 1e-13)
    log = j.getLog()
    return {k: np.array(log[src][-100:]) for k, src in
            (("mx", "free_mx"), ("my", "free_my"), ("mz", "free_mz"), ("R", "R"))}


# reference tests/test_curated_examples.py:670-776
def case_fmr_basic(J, n_fields: int = 5) -> dict:
    from scipy.fft import fft, fftfreq

    Kdir = CVector(0, 0, 1)
    demag = [CVector(0, 0, 0), CVector(0, 0, 0), CVector(0, 0, 1.0)]
    l1 = Layer("free", mag=CVector(0, 0, 0.99), anis=Kdir, Ms=1.03, thickness=1.0e-9, damping=0.01,
               demagTensor=demag, cellSurface=1e-9)
    l2 = Layer("perpendicular", mag=CVector(0, 0, 0.99), anis=Kdir, Ms=1.03, thickness=1.0e-9, damping=0.01,
               cellSurface=1e-9, demagTensor=demag)
    l3 = Layer("buffer", mag=CVector(1, 0, 0), anis=CVector(1, 0, 0), Ms=0.12, thickness=1.0e-9, damping=0.015,
               cellSurface=1e-9, demagTensor=demag)
    l1.setAnisotropyDriver(constantDriver(489e3))
    l2.setAnisotropyDriver(constantDriver(514e3))
    l3.setAnisotropyDriver(constantDriver(17e3))
    j = J([l1, l2, l3])
    Hmax = -700e3
    _, Hvecs = FieldScan.amplitude_scan(start=-Hmax, stop=Hmax, steps=20, theta=0, phi=0)
    dt, sim_time, wait_time = 5e-13, 60e-9, 4e-9
    j.setLayerExternalFieldDriver("all", AxialDriver(*Hvecs[0]))
    j.runSimulation(sim_time, dt, dt)
    spectrum, ffreqs = [], None
    for Hv in Hvecs[:n_fields]:
        j.clearLog()
        j.setLayerExternalFieldDriver("all", AxialDriver(*Hv))
        j.setLayerOerstedFieldDriver("all", AxialDriver(NullDriver(), stepDriver(0, 500, 0, 2e-9), NullDriver()))
        j.runSimulation(sim_time, dt, dt)
        log = j.getLog()
        tm = np.asarray(log["time"])
        idx = np.argwhere(tm > wait_time).ravel()
        mixed = (1 / 3) * (np.asarray(log["free_mx"])[idx] + np.asarray(log["perpendicular_mx"])[idx]
                           + np.asarray(log["buffer_mx"])[idx])
        y = np.abs(fft(mixed))
        f = fftfreq(len(mixed), dt)
        y, f = y[1: len(y) // 2], f[1: len(f) // 2]
        keep = np.where(f < 30e9)
        spectrum.append(y[keep])
        ffreqs = f[keep]
    return {"spectrum": spectrum, "ffreqs": ffreqs}


# reference tests/test_curated_examples.py:779-881
def case_harmonic_hall(J) -> dict:
    Rx0, SMR, AMR, AHE, w, l = [304.306], [-0.464], [-0.053], [-5.71], [3e-5], [2e-5]
    demag = [
        CVector(0.00024164288391924, 2.71396011566517e-10, 5.95503928124313e-14),
        CVector(2.71396011566517e-10, 0.000160046006320031, 1.32504057070646e-14),
        CVector(5.95503928124313e-14, 1.32504057070646e-14, 0.999598310229469),
    ]
    free = Layer(id="free", mag=CVector(0, 0, 1), anis=CVector(0.0, 0, 1), Ms=0.525, thickness=1.45e-9,
                 cellSurface=w[0] * l[0], damping=0.03, demagTensor=demag)
    j = J([free])
    j.setLayerOerstedFieldDriver("all", AxialDriver(NullDriver(), NullDriver(),
                                                    ScalarDriver.getStepDriver(0, 50, 0.0, 1e-11)))
    j.setLayerAnisotropyDriver("all", ScalarDriver.getConstantDriver(1.54e5))
    Hscan, Hvecs = FieldScan.amplitude_scan(theta=92, phi=0, start=-300e3, stop=300e3, steps=20, back=False)
    mags = [CVector(0, 0, 1)]
    Rxx = []
    for H in Hvecs:
        j.clearLog()
        j.setLayerExternalFieldDriver("all", AxialDriver(*(ScalarDriver.getConstantDriver(h) for h in H)))
        j.setLayerMagnetisation("free", mags[0])
        j.runSimulation(10e-9, 1e-12, 1e-12)
        mags[0] = j.getLayerMagnetisation("free")
        log = j.getLog()
        rx, _ = calculate_resistance_parallel(Rx0, [0], AMR=AMR, AHE=AHE, SMR=SMR, m=_m(log, ["free"]), l=l, w=w)
        Rxx.append(rx[-1])
    return {"Hscan": Hscan, "simulated": np.asarray(Rxx)}


# reference tests/test_curated_examples.py:217-334
def case_ahe_loops(J) -> dict:
    t_fm, Ms, alpha = 1e-9, 1.1, 5e-2
    Hdl, Hfl = 90 * OetoAm * 0.45, 90 * OetoAm
    w, l = 10e-6, 80e-6
    jden = 4.24e10
    Kdir1 = FieldScan.angle2vector(1, 0)
    Kdir2 = FieldScan.angle2vector(0.01, 0)
    demag = [CVector(0.0, 0.0, 0.0), CVector(0.0, 0.0, 0.0), CVector(0.0, 0.0, 1)]
    free = Layer.createSOTLayer(id="free", mag=Kdir1, anis=Kdir1, Ms=Ms, thickness=t_fm, cellSurface=w * l,
                                demagTensor=demag, damping=alpha, dampingLikeTorque=Hdl / jden,
                                fieldLikeTorque=Hfl / jden)
    bottom = Layer.createSOTLayer(id="bottom", mag=Kdir2, anis=Kdir2, Ms=Ms, thickness=t_fm, cellSurface=w * l,
                                  demagTensor=demag, damping=alpha, dampingLikeTorque=Hdl / jden,
                                  fieldLikeTorque=Hfl / jden)
    j = J([free, bottom])
    j.setLayerAnisotropyDriver("free", constantDriver(0.55e6))
    j.setLayerAnisotropyDriver("bottom", constantDriver(0.3e6))
    j.setLayerReferenceLayer("free", CVector(0, 1.0, 0.0))
    j.setLayerReferenceLayer("bottom", CVector(0, 1.0, 0.0))
    j.setIECDriver("free", "bottom", constantDriver(3e-3))
    j.runSimulation(4e-9, 1e-13, 1e-13)
    Hscan = np.linspace(-2000 * OetoAm, 2000 * OetoAm, 10)
    R = []
    for H in Hscan:
        j.clearLog()
        j.setLayerExternalFieldDriver("all", AxialDriver(0, 0, H))
        j.runSimulation(15e-9, 1e-13, 1e-13, calculateEnergies=False)
        log = j.getLog()
        _, Rxy = calculate_resistance_series([100] * 2, [0] * 2, AMR=[0.41] * 2, AHE=[5, 5], SMR=[5.11] * 2,
                                             m=_m(log, ("free", "bottom")), l=[l] * 2, w=[w] * 2)
        R.append(Rxy[-100:].mean())
    return {"Hscan": np.asarray(Hscan), "R": np.asarray(R)}


# reference tests/test_curated_examples.py:337-439
def case_cims_hysteresis(J) -> dict:
    demag = [CVector(0.0, 0.0, 0.0), CVector(0.0, 0.0, 0.0), CVector(0.0, 0.0, 1)]
    Hdl, Hfl = 0.000371 * TtoAm, -0.000433 * TtoAm
    jden1 = 6e10
    last_N, dt_pulse, tstart, tstep, sim_time, t_hm = 100, 6e-9, 1e-9, 5e-13, 15e-9, 6e-9
    free = Layer.createSOTLayer(id="free", mag=CVector(1, 0, 0), anis=CVector(0, 0, 1.0), Ms=1.45,
                                thickness=1.7e-9, cellSurface=0.0, demagTensor=demag, damping=0.0064,
                                dampingLikeTorque=Hdl / jden1, fieldLikeTorque=Hfl / jden1)
    j = J([free])
    j.setLayerAnisotropyDriver("free", ScalarDriver.getConstantDriver(0.3e6))
    j.setLayerReferenceLayer("free", CVector(0.0, 1.0, 0.0))
    j.clearLog()
    j.runSimulation(14e-9, tstep, tstep)
    scales = []
    for Hoescale in (1, 0, -1):
        v = CVector(0.001, 0.99, 0)
        v.normalize()
        j.setLayerMagnetisation("free", v)
        jscan = np.linspace(-0.1e12, 0.1e12, 20)
        jscan = np.concatenate([jscan, jscan[::-1][1:]])
        hyst = []
        for current in jscan:
            j.clearLog()
            j.setLayerCurrentDriver("all", stepDriver(0, current, tstart, tstart + dt_pulse))
            j.setLayerOerstedFieldDriver("all", AxialDriver(constantDriver(0),
                                                            stepDriver(0, Hoescale * current * t_hm / 2, tstart,
                                                                       tstart + dt_pulse),
                                                            constantDriver(0)))
            j.runSimulation(sim_time, tstep, tstep)
            m = _m(j.getLog(), ("free",))
            hyst.append(compute_gmr(100, 200, m.squeeze()[:, -last_N:],
                                    np.asarray([[0, 1, 0]] * last_N).T).mean())
        scales.append(np.asarray(hyst))
    return {"hysteresis_scales": scales, "Hoescales": [1, 0, -1]}


def _vsd_junction(J, Ms, Ku, alpha):
    demag = [CVector(0.0, 0.0, 0.0), CVector(0.0, 0.0, 0.0), CVector(0.0, 0.0, 1.0)]
    free = Layer(id="free", mag=CVector(0.0, 0.0, 1.0), anis=CVector(0, 0.0, 1.0), Ms=Ms, thickness=1.45e-9,
                 cellSurface=0, demagTensor=demag, damping=alpha)
    j = J([free])
    j.setLayerAnisotropyDriver("free", constantDriver(Ku))
    return j


def _vmix(j, H, theta, phi, f, Irf, use_rx, dt=1e-13):
    Rx0, Ry0, SMR, AMR, AHE, w, l = 304.306, 1.008, -0.464, -0.053, -5.71, 3e-5, 2e-5
    j.clearLog()
    j.setLayerExternalFieldDriver("all", AxialDriver(constantDriver(H * math.sin(theta) * math.cos(phi)),
                                                     constantDriver(H * math.sin(theta) * math.sin(phi)),
                                                     constantDriver(H * math.cos(theta))))
    j.setLayerOerstedFieldDriver("all", AxialDriver(NullDriver(), NullDriver(), sineDriver(0, -5000, f, 0)))
    j.runSimulation(40e-9, dt, dt, solverMode=SolverMode.RK4)
    log = j.getLog()
    rx, ry = calculate_resistance_parallel(Rx0=[Rx0], Ry0=[Ry0], AMR=[AMR], AHE=[AHE], SMR=[SMR],
                                           m=_m(log, ["free"]), l=[l], w=[w])
    R = rx if use_rx else ry
    I = Irf * np.sin(2 * math.pi * f * np.asarray(log["time"]))
    SD = -I * R
    return np.mean(Filters.butter_lowpass_filter(SD, cutoff=10e6, fs=1.0 / dt, order=3))


# reference tests/test_curated_examples.py:442-555 (orientation "4p")
def case_vsd_field_scan(J) -> dict:
    j = _vsd_junction(J, 0.525, 1.54e5, 30e-3)
    Hspace = np.linspace(350e3, 630e3, num=20)
    v = [_vmix(j, H, np.deg2rad(91), 0.0, 14e9, 0.75e-3, use_rx=False) for H in Hspace]
    return {"hscan": Hspace, "vscan": np.asarray(v)}


# reference tests/test_curated_examples.py:558-667 (orientation "4p")
def case_vsd_freq_scan(J) -> dict:
    j = _vsd_junction(J, 0.525, 1.54e5, 30e-3)
    fspace = np.arange(10, 16, 1) * 1e9
    v = [_vmix(j, 500e3, np.deg2rad(91), 0.0, f, 0.75e-3, use_rx=True) for f in fspace]
    return {"fspace": fspace, "fsweep": np.asarray(v)}


CASES = {
    "test_sto": case_sto,
    "test_fmr_basic": case_fmr_basic,
    "test_harmonic_hall": case_harmonic_hall,
    "test_ahe_loops": case_ahe_loops,
    "test_cims_hysteresis": case_cims_hysteresis,
    "test_vsd_field_scan": case_vsd_field_scan,
    "test_vsd_freq_scan": case_vsd_freq_scan,
}
