"""Synthetic sweep workloads of the benchmark, as batches.

  r_h_loop      1-layer R(H) loop, 101 fields, STRIP Rx, RK4 dt=1e-13, 10 ns
  pimm          2-layer MTJ pulse-induced spectra, fields 0..600 kA/m at 30 deg
                in-plane, per-point duration U(2, 20) ns, 1 kA/m Hoe_y step
  cims_pbit     2-layer static-SOT switching ensemble at 300 K, Euler-Heun,
                trapezoid current pulses, perpendicular free layer
  sto           2-layer STT oscillator, DC current sweep, RK4

Parameters the workload descriptions leave open are fixed here and listed in
each builder's docstring.
"""
from __future__ import annotations

import math

import numpy as np

from .batch import JunctionBatch
from .core import AxialDriver, CVector, Junction, Layer, NullDriver, ScalarDriver, SolverMode, constantDriver

THIN = [CVector(0, 0, 0), CVector(0, 0, 0), CVector(0, 0, 1)]


def r_h_loop(n_fields: int = 101, total_time: float = 10e-9, dt: float = 1e-13, wf: float = 1e-12,
             precision: str = "fp64", log: str = "full") -> JunctionBatch:
    """Free layer Ms 1.2 T, t 2 nm, S 1e-16 m^2, alpha 0.01, Ku 5e3 J/m^3
    along x, N = diag(0.00024, 0.0022, 0.99756), m0 = (1,0,0); STRIP with
    Rx0 = 100, AMR_X = 1; Hx swept -800..800 kA/m.  Each field point starts
    from m0 (independent points)."""
    l = Layer("free", mag=CVector(1, 0, 0), anis=CVector(1, 0, 0), Ms=1.2, thickness=2e-9, cellSurface=1e-16,
              demagTensor=[CVector(0.00024, 0, 0), CVector(0, 0.0022, 0), CVector(0, 0, 0.99756)], damping=0.01)
    j = Junction([l], [100.0], [0.0], [1.0], [0.0], [0.0], [0.0], [0.0])
    j.setLayerAnisotropyDriver("free", constantDriver(5e3))
    b = JunctionBatch.replicate(j, n_fields, total_time, dt, wf, SolverMode.RK4, precision=precision, log=log)
    H = np.zeros((n_fields, 3))
    H[:, 0] = np.linspace(-800e3, 800e3, n_fields)
    b.set_external_field("all", H)
    return b


def pimm_junction() -> Junction:
    """Free: Ms 1.0 T, t 1 nm, Ku 1.5e5 J/m^3 along z, alpha 0.01; reference:
    Ms 1.6 T, t 3 nm, Ku 1e4 J/m^3 along x, alpha 0.01; thin-film demag
    diag(0,0,1); S = 1e-16 m^2; J = -1e-4 J/m^2; Rp 100, Rap 200 Ohm; Hoe_y
    step of 1 kA/m over the first 3 steps (the PIMM_procedure default,
    cmtj/utils/procedures.py:96-108)."""
    free = Layer("free", mag=CVector(math.cos(math.pi / 6), math.sin(math.pi / 6), 0.05), anis=CVector(0, 0, 1),
                 Ms=1.0, thickness=1e-9, cellSurface=1e-16, demagTensor=THIN, damping=0.01)
    ref = Layer("ref", mag=CVector(1, 0, 0), anis=CVector(1, 0, 0), Ms=1.6, thickness=3e-9, cellSurface=1e-16,
                demagTensor=THIN, damping=0.01)
    j = Junction([free, ref], 100, 200)
    j.setLayerAnisotropyDriver("free", constantDriver(1.5e5))
    j.setLayerAnisotropyDriver("ref", constantDriver(1e4))
    j.setIECDriver("free", "ref", constantDriver(-1e-4))
    dt = 1e-13
    j.setLayerOerstedFieldDriver("all", AxialDriver(NullDriver(), ScalarDriver.getStepDriver(0, 1e3, 0, 3 * dt),
                                                    NullDriver()))
    return j


def pimm(n_points: int = 4096, seed: int = 0, t_min: float = 2e-9, t_max: float = 20e-9, dt: float = 1e-13,
         precision: str = "fp32", log: str = "mx", log_dtype=None) -> JunctionBatch:
    """PIMM sweep: bias 0..600 kA/m at 30 deg in-plane, per-point duration
    drawn U(t_min, t_max) from `seed`, logging every step (the FFT input)."""
    rng = np.random.default_rng(seed)
    durations = rng.uniform(t_min, t_max, n_points)
    j = pimm_junction()
    b = JunctionBatch.replicate(j, n_points, durations, dt, dt, SolverMode.RK4, precision=precision, log=log,
                                log_dtype=log_dtype)
    Hmag = np.linspace(0, 600e3, n_points)
    phi = math.radians(30)
    H = np.stack([Hmag * math.cos(phi), Hmag * math.sin(phi), np.zeros(n_points)], axis=1)
    b.set_external_field("all", H)
    b.set_times(durations, dt, dt)
    return b


def cims_junction(T: float = 300.0) -> Junction:
    """SOT switching device.  Free: Ms 1.1 T, t 1.2 nm, Ku 3e5 J/m^3 along z
    taken as the effective perpendicular anisotropy (shape anisotropy folded
    in, so the free layer's demagnetisation tensor is zero), alpha 0.02,
    static SOT with tau_DL/j = 1e-12 T per A/m^2 (x TtoAm -> A/m) and
    FL/DL = 0.3, polarisation along y, starting at +z.  Anisotropy field
    2K/Ms = 545 kA/m; thermal stability K V / kT = 171 at 300 K, so no
    switching without current.  Pinned (the TMR reference): Ms 1.2 T, t 3 nm,
    Ku 1e5 J/m^3 along x, thin-film demagnetisation, alpha 0.02.
    S = (50 nm)^2 pi.  In-plane bias -20 kA/m along x (the current's axis)
    sets which way a damping-like torque switches the free layer; thermal
    field on both layers.  A damping-like field above about half the
    anisotropy field (j ~ 3.4e11 A/m^2) switches it to -z."""
    TtoAm = 795774.715459
    hdl = 1e-12 * TtoAm
    S = (50e-9) ** 2 * math.pi
    zero = [CVector(0, 0, 0), CVector(0, 0, 0), CVector(0, 0, 0)]
    free = Layer.createSOTLayer("free", mag=CVector(0.0, 0.0, 1.0), anis=CVector(0, 0, 1), Ms=1.1,
                                thickness=1.2e-9, cellSurface=S, demagTensor=zero, damping=0.02,
                                fieldLikeTorque=0.3 * hdl, dampingLikeTorque=hdl)
    free.setReferenceLayer(CVector(0, 1, 0))
    pinned = Layer("pinned", mag=CVector(1, 0, 0), anis=CVector(1, 0, 0), Ms=1.2, thickness=3e-9, cellSurface=S,
                   demagTensor=THIN, damping=0.02)
    j = Junction([free, pinned], 100, 200)
    j.setLayerAnisotropyDriver("free", constantDriver(3e5))
    j.setLayerAnisotropyDriver("pinned", constantDriver(1e5))
    j.setLayerTemperatureDriver("all", constantDriver(T))
    j.setLayerExternalFieldDriver("all", AxialDriver(-20e3, 0, 0))
    return j


def cims_pbit(n_currents: int = 256, n_seeds: int = 4096, total_time: float = 5e-9, dt: float = 1e-14,
              edge: float = 1e-9, plateau: float = 1.5e-9, precision: str = "fp32", seed: int = 1,
              j_min: float = 1e10, j_max: float = 1e12, T: float = 300.0) -> JunctionBatch:
    """Current pulses j_min..j_max A/m^2 (log-spaced, default 1e10..1e12) x
    seeds; final state only.  Each pulse is a trapezoid from t = 0: linear
    rise over `edge`, `plateau` at j, linear fall over `edge`, then zero
    current to total_time (default 1 + 1.5 + 1 ns, then 1.5 ns of
    relaxation).  Instance q = current (q // n_seeds), seed (q % n_seeds);
    a run switched when its final free-layer m_z < 0."""
    j = cims_junction(T)
    n = n_currents * n_seeds
    b = JunctionBatch.replicate(j, n, total_time, dt, total_time, SolverMode.EulerHeun, precision=precision,
                                log="none")
    jcur = np.repeat(np.logspace(math.log10(j_min), math.log10(j_max), n_currents), n_seeds)
    b.set_driver("free", "current", "type", np.full(n, 6, dtype=np.int32))  # trapezoid
    b.set_driver("free", "current", "amplitude", jcur)
    b.set_driver("free", "current", "t_start", np.zeros(n))
    b.set_driver("free", "current", "edge_time", np.full(n, edge))
    b.set_driver("free", "current", "steady_time", np.full(n, plateau))
    rng = np.random.default_rng(seed)
    b.set_seeds(rng.integers(0, 2**63, n, dtype=np.uint64))
    return b


def sto_junction() -> Junction:
    """Free: STT layer Ms 0.8 T, t 2 nm, in-plane Ku 2e3 J/m^3 along x,
    alpha 0.005, P 0.6, Lambda 1, polariser = the pinned layer below (TOP/
    BOTTOM reference); pinned: Ms 1.2 T, t 3 nm, Ku 1e5 along x; out-of-plane
    bias 200 kA/m; S 1e-16 m^2."""
    free = Layer.createSTTLayer("free", mag=CVector(1, 0.1, 0.1), anis=CVector(1, 0, 0), Ms=0.8, thickness=2e-9,
                                cellSurface=1e-16, demagTensor=THIN, damping=0.005,
                                SlonczewskiSpacerLayerParameter=1.0, beta=0.0, spinPolarisation=0.6)
    pinned = Layer("pinned", mag=CVector(1, 0, 0), anis=CVector(1, 0, 0), Ms=1.2, thickness=3e-9, cellSurface=1e-16,
                   demagTensor=THIN, damping=0.02)
    j = Junction([pinned, free], 100, 200)
    j.setLayerReferenceType("free", 3)  # BOTTOM: the pinned layer
    j.setLayerAnisotropyDriver("free", constantDriver(2e3))
    j.setLayerAnisotropyDriver("pinned", constantDriver(1e5))
    j.setLayerExternalFieldDriver("all", AxialDriver(0, 0, 200e3))
    return j


def sto(n_currents: int = 512, total_time: float = 100e-9, dt: float = 1e-13, wf: float = 1e-11,
        precision: str = "fp32", log: str = "full") -> JunctionBatch:
    """DC current sweep (A/m^2) of the STT oscillator, decimated log."""
    j = sto_junction()
    b = JunctionBatch.replicate(j, n_currents, total_time, dt, wf, SolverMode.RK4, precision=precision, log=log)
    I = np.linspace(-2e11, -2e10, n_currents)
    b.set_driver("free", "current", "constant", I)
    return b


def vsd_junction(H: float = 100e3, f: float = 5e9) -> Junction:
    """One free layer (Ms 1.0 T, t 1.5 nm, Ku 1e5 J/m^3 along z, alpha 0.02,
    S 1e-16 m^2) against a fixed reference along x (CLASSIC R, Rp 100 /
    Rap 200 Ohm); bias field H at 45 deg in the x-z plane; sine Oersted drive
    of 500 A/m at frequency f along y (the VSD_procedure excitation,
    cmtj/utils/procedures.py:280)."""
    free = Layer("free", mag=CVector(0.5, 0.0, 0.866), anis=CVector(0, 0, 1), Ms=1.0, thickness=1.5e-9,
                 cellSurface=1e-16, demagTensor=THIN, damping=0.02)
    free.setReferenceLayer(CVector(1, 0, 0))
    j = Junction([free], 100, 200)
    j.setLayerAnisotropyDriver("free", constantDriver(1e5))
    j.setLayerExternalFieldDriver("all", AxialDriver(H * math.sqrt(0.5), 0, H * math.sqrt(0.5)))
    j.setLayerOerstedFieldDriver("all", AxialDriver(NullDriver(), ScalarDriver.getSineDriver(0, 500, f, 0),
                                                    NullDriver()))
    return j


def chain_junction(n_layers: int, H: float = 200e3) -> Junction:
    """n_layers in-plane layers (Ms 1.2..1.5 T, t 1..2.5 nm, Ku 2e4..5e4 J/m^3
    along x, alpha 0.03) coupled by IEC J = -2e-4 J/m^2 between neighbours;
    in-plane bias H at 30 deg; Hoe_y step 1 kA/m over the first 10 ps; STRIP
    Rx0 100, Ry0 50, AMR 1, AHE 2 Ohm per layer."""
    layers = []
    for q in range(n_layers):
        layers.append(Layer(f"L{q}", mag=CVector(1, 0.1 * (q + 1), 0.02), anis=CVector(1, 0, 0), Ms=1.2 + 0.1 * q,
                            thickness=(1 + 0.5 * q) * 1e-9, cellSurface=1e-16, demagTensor=THIN, damping=0.03))
    z = [0.0] * n_layers
    j = Junction(layers, [100.0] * n_layers, [50.0] * n_layers, [1.0] * n_layers, [1.0] * n_layers, z, z,
                 [2.0] * n_layers)
    for q in range(n_layers):
        j.setLayerAnisotropyDriver(f"L{q}", constantDriver(2e4 + 1e4 * q))
    for q in range(n_layers - 1):
        j.setIECDriver(f"L{q}", f"L{q + 1}", constantDriver(-2e-4))
    j.setLayerExternalFieldDriver("all", AxialDriver(H * math.cos(math.pi / 6), H * math.sin(math.pi / 6), 0))
    j.setLayerOerstedFieldDriver("all", AxialDriver(NullDriver(), ScalarDriver.getStepDriver(0, 1e3, 0, 1e-11),
                                                    NullDriver()))
    return j


def serving_mix(n: int, seed: int = 0, t_min: float = 1e-9, t_max: float = 50e-9, dt_min: float = 1e-14,
                dt_max: float = 1e-13) -> list:
    """n request specs (kind, junction, totalTime, timeStep, writeFrequency,
    solverMode, log) of the mixed serving stream: PIMM (2 layers, RK4),
    VSD (1 layer, sine drive 1-20 GHz, RK4), CIMS (2 layers, static SOT at
    300 K, Euler-Heun, dt 1e-14) and 3-4 layer IEC chains (RK4); durations
    U(t_min, t_max), dt log-uniform in [dt_min, dt_max], logs decimated to
    1 ps.  Parameters are drawn in the ranges of the sweep workloads."""
    rng = np.random.default_rng(seed)
    kinds = rng.choice(["pimm", "vsd", "cims", "chain3", "chain4"], size=n, p=[0.35, 0.25, 0.2, 0.1, 0.1])
    out = []
    for q in range(n):
        k = str(kinds[q])
        T = float(rng.uniform(t_min, t_max))
        dt = float(np.exp(rng.uniform(np.log(dt_min), np.log(dt_max))))
        wf = max(1e-12, dt)
        if k == "pimm":
            j = pimm_junction()
            Hm = float(rng.uniform(0, 600e3))
            j.setLayerExternalFieldDriver("all", AxialDriver(Hm * math.cos(math.pi / 6), Hm * math.sin(math.pi / 6),
                                                             0))
            out.append((k, j, T, dt, wf, SolverMode.RK4, "mx"))
        elif k == "vsd":
            j = vsd_junction(float(rng.uniform(50e3, 400e3)), float(rng.uniform(1e9, 20e9)))
            out.append((k, j, T, dt, wf, SolverMode.RK4, "full"))
        elif k == "cims":
            j = cims_junction()
            jc = float(10 ** rng.uniform(10, 12))
            j.setLayerCurrentDriver("free", ScalarDriver.getTrapezoidDriver(0, jc, 0, T / 5, T / 5))
            j.setNoiseSeed(int(rng.integers(0, 2**62)))
            out.append((k, j, T, 1e-14, max(1e-12, 1e-14), SolverMode.EulerHeun, "m"))
        else:
            j = chain_junction(3 if k == "chain3" else 4, float(rng.uniform(50e3, 400e3)))
            out.append((k, j, T, dt, wf, SolverMode.RK4, "full"))
    return out


__all__ = ["r_h_loop", "pimm", "pimm_junction", "cims_pbit", "cims_junction", "sto", "sto_junction",
           "vsd_junction", "chain_junction", "serving_mix"]
