"""Junction configurations that exercise every term, driver kind, solver and
log channel of the hot path.  Each builder takes the Junction class to use
(GPU or oracle) and returns (junction, runSimulation kwargs).

Every case is well conditioned: a 1e-14 change of the initial state moves
the oracle's own trajectory by far less than the 1e-8 parity bar (checked by
tests/test_oracle.py), so a difference between two correct implementations
that round differently stays far below the bar.  (Weakly damped, strongly
coupled layers started on a saddle are chaotic; no bar tighter than O(1)
would hold for them between any two builds, so they are not parity cases.)"""
from __future__ import annotations

import math

from cmtj_amd import (AxialDriver, CVector, Layer, NullDriver, Reference, ScalarDriver, SolverMode,
                      constantDriver, gaussianImpulseDriver, posSineDriver, pulseDriver, sineDriver,
                      stepDriver, trapezoidDriver)

DEMAG = [CVector(0, 0, 0), CVector(0, 0, 0), CVector(0, 0, 1)]
SURF = 150e-9 * 150e-9 * math.pi


def rh_strip(J, H=200e3):
    """1 layer, STRIP Rx/Ry, uniaxial x, full demag diag: the R(H) loop point."""
    l = Layer("free", mag=CVector(1, 0, 0), anis=CVector(1, 0, 0), Ms=1.2, thickness=2e-9, cellSurface=1e-16,
              demagTensor=[CVector(0.00024, 0, 0), CVector(0, 0.0022, 0), CVector(0, 0, 0.99756)], damping=0.01)
    j = J([l], [100.0], [0.0], [1.0], [0.0], [0.0], [0.0], [0.0])
    j.setLayerAnisotropyDriver("free", constantDriver(5e3))
    j.setLayerExternalFieldDriver("all", AxialDriver(H, 0, 0))
    return j, dict(totalTime=1e-9, timeStep=1e-13, writeFrequency=1e-12)


def bilayer_iec(J, solver=SolverMode.RK4):
    """2 layers, bilinear + biquadratic IEC, step Oersted, CLASSIC R."""
    l1 = Layer("free", mag=CVector(0.8, 0.6, 0.1), anis=CVector(1, 0, 0), Ms=1.65, thickness=3e-9, cellSurface=SURF,
               demagTensor=DEMAG, damping=0.02)
    l2 = Layer("bottom", mag=CVector(1, 0.1, 0.1), anis=CVector(1, 0, 0), Ms=1.45, thickness=4e-9, cellSurface=SURF,
               demagTensor=DEMAG, damping=0.02)
    j = J([l1, l2], 100, 200)
    j.setLayerAnisotropyDriver("free", constantDriver(1.05e3))
    j.setLayerAnisotropyDriver("bottom", constantDriver(50e3))
    j.setIECDriver("free", "bottom", constantDriver(-1.78e-3))
    j.setQuadIECDriver("free", "bottom", constantDriver(-1.69e-4))
    j.setLayerExternalFieldDriver("all", AxialDriver(300e3, 10e3, 0))
    j.setLayerOerstedFieldDriver("all", AxialDriver(NullDriver(), stepDriver(0, 1e3, 0, 5e-12), NullDriver()))
    return j, dict(totalTime=2e-9, timeStep=1e-13, writeFrequency=1e-13, solverMode=solver)


def trilayer(J, solver=SolverMode.RK4):
    """3 layers with two IEC pairs (reference tests/conftest.py:183-237)."""
    Kdir = CVector(1, 0, 0)
    l1 = Layer("free", mag=CVector(0.6, 0.2, 0.7), anis=Kdir, Ms=1.65, thickness=3e-9, cellSurface=SURF,
               demagTensor=DEMAG, damping=0.05)
    l2 = Layer("middle", mag=CVector(0.5, -0.2, 0.8), anis=Kdir, Ms=1.45, thickness=4e-9, cellSurface=SURF,
               demagTensor=DEMAG, damping=0.05)
    l3 = Layer("bottom", mag=CVector(0.1, 0, 1.0), anis=CVector(0, 0, 1.0), Ms=0.7, thickness=1e-9,
               cellSurface=SURF, demagTensor=DEMAG, damping=0.05)
    j = J([l1, l2, l3])
    for lid, k in zip(["free", "middle", "bottom"], [1.05e3, 50e3, 1e4]):
        j.setLayerAnisotropyDriver(lid, constantDriver(k))
    j.setIECDriver("free", "middle", constantDriver(-1.78e-3))
    j.setQuadIECDriver("free", "middle", constantDriver(-1.69e-4))
    j.setIECDriver("middle", "bottom", constantDriver(3e-4))
    j.setQuadIECDriver("middle", "bottom", constantDriver(2e-6))
    j.setLayerExternalFieldDriver("all", AxialDriver(sineDriver(50e3, 20e3, 3e9, 0.3), constantDriver(0), constantDriver(100e3)))
    return j, dict(totalTime=2e-9, timeStep=1e-13, writeFrequency=1e-12, solverMode=solver)


def stt_drivers(J, alternative=False, ref=None):
    """STT with time-dependent current, anisotropy and field of every kind."""
    l = Layer.createSTTLayer("free", mag=CVector(0.1, 0.2, 1.0), anis=CVector(0, 0, 1.0), Ms=1.0, thickness=1.4e-9,
                             cellSurface=7e-10 * 7e-10, demagTensor=DEMAG, damping=0.05,
                             SlonczewskiSpacerLayerParameter=1.5, beta=0.3, spinPolarisation=0.6)
    l.setReferenceLayer(CVector(0, 1.0, 1.0))
    l.setKappa(0.9)
    if alternative:
        l.setAlternativeSTT(True)
    j = J([l], 100, 200)
    j.setLayerAnisotropyDriver("free", trapezoidDriver(300e3, 50e3, 2e-10, 3e-10, 4e-10))
    j.setLayerCurrentDriver("free", sineDriver(-6e9, 2e9, 2.5e9, 0.1))
    j.setLayerExternalFieldDriver("all", AxialDriver(pulseDriver(0, 5e3, 3e-10, 0.3),
                                                     posSineDriver(1e3, 4e3, 1.7e9, 0.5),
                                                     gaussianImpulseDriver(0, 10e3, 5e-10, 1e-10)))
    j.setLayerOerstedFieldDriver("all", AxialDriver(ScalarDriver.getHalfSineDriver(0, 2e3, 4e9, 0), NullDriver(), NullDriver()))
    return j, dict(totalTime=1.5e-9, timeStep=1e-13, writeFrequency=1e-13)


def stt_neighbour_reference(J, rtype=Reference.top):
    """STT on the free layer with the polariser taken from a neighbour."""
    l1 = Layer.createSTTLayer("free", mag=CVector(0.0, 0.3, 1.0), anis=CVector(1, 0, 0), Ms=0.8, thickness=2e-9,
                              cellSurface=1e-16, demagTensor=DEMAG, damping=0.005,
                              SlonczewskiSpacerLayerParameter=1.0, beta=0.0, spinPolarisation=0.6)
    l2 = Layer("ref", mag=CVector(1, 0, 0), anis=CVector(1, 0, 0), Ms=1.6, thickness=3e-9, cellSurface=1e-16,
               demagTensor=DEMAG, damping=0.02)
    layers = [l1, l2] if rtype == Reference.top else [l2, l1]
    j = J(layers, 100, 200)
    j.setLayerReferenceType("free", rtype)
    j.setLayerAnisotropyDriver("free", constantDriver(2e3))
    j.setLayerAnisotropyDriver("ref", constantDriver(1e4))
    j.setLayerCurrentDriver("free", constantDriver(5e10))
    j.setLayerExternalFieldDriver("all", AxialDriver(20e3, 0, 0))
    return j, dict(totalTime=1e-9, timeStep=1e-13, writeFrequency=1e-12)


def sot_dynamic(J, solver=SolverMode.RK4):
    """Dynamic SOT with primary and secondary torques on two layers."""
    a = Layer("free1", mag=CVector(0, 0.1, 1), anis=CVector(0, 0, 1), Ms=1, thickness=2e-9, cellSurface=1e-18,
              demagTensor=DEMAG, damping=0.01)
    b = Layer("free2", mag=CVector(1, 0, 0.2), anis=CVector(0, 0, 1), Ms=0.8, thickness=1e-9, cellSurface=1e-18,
              demagTensor=DEMAG, damping=0.02)
    j = J([a, b])
    j.setLayerAnisotropyDriver("all", constantDriver(3e5))
    j.setLayerFieldLikeTorqueDriver("free1", stepDriver(0, 2e3, 1e-10, 6e-10))
    j.setLayerDampingLikeTorqueDriver("free1", stepDriver(0, 5e3, 1e-10, 6e-10))
    j.setLayerSecondaryFieldLikeTorqueDriver("free1", constantDriver(300.0))
    j.setLayerSecondaryDampingLikeTorqueDriver("free1", constantDriver(400.0))
    j.setLayerPrimaryTorqueDrivers("free2", constantDriver(1e3), constantDriver(2e3))
    j.setLayerSecondaryTorqueDrivers("free2", constantDriver(3e2), constantDriver(4e2))
    j.setLayerReferenceLayer("free1", CVector(1, 0, 0))
    j.setLayerReferenceLayer("free2", CVector(1, 0, 0))
    j.getLayer("free1").setSecondaryReferenceLayer(CVector(0, 1, 0))
    j.getLayer("free2").setSecondaryReferenceLayer(CVector(0, 1, 0))
    j.setIECDriver("free1", "free2", constantDriver(-2e-4))
    return j, dict(totalTime=1e-9, timeStep=1e-13, writeFrequency=1e-12, solverMode=solver)


def sot_static(J):
    """Static SOT (createSOTLayer) with a step current: the CIMS path."""
    l = Layer.createSOTLayer("free", mag=CVector(0.001, 0.99, 0), anis=CVector(0, 0, 1.0), Ms=1.45,
                             thickness=1.7e-9, cellSurface=0.0, demagTensor=DEMAG, damping=0.0064,
                             dampingLikeTorque=0.000371 * 795774.715459 / 6e10,
                             fieldLikeTorque=-0.000433 * 795774.715459 / 6e10)
    j = J([l])
    j.setLayerAnisotropyDriver("free", constantDriver(0.3e6))
    j.setLayerReferenceLayer("free", CVector(0, 1, 0))
    j.setLayerCurrentDriver("all", stepDriver(0, 0.08e12, 1e-10, 1.2e-9))
    return j, dict(totalTime=2e-9, timeStep=5e-13, writeFrequency=5e-13)


def exotic_terms(J, solver=SolverMode.RK4):
    """IDMI, dipole tensors, off-diagonal demag, second-order anisotropy, Hdmi."""
    demag = [CVector(0.02, 0.001, 0.0005), CVector(0.001, 0.03, 0.0002), CVector(0.0005, 0.0002, 0.95)]
    dip = [CVector(0.01, 0.002, 0), CVector(0.002, 0.015, 0.001), CVector(0, 0.001, -0.02)]
    a = Layer("a", mag=CVector(0.1, 0.2, 1), anis=CVector(0, 0.2, 1), Ms=1.2, thickness=1.5e-9, cellSurface=1e-16,
              demagTensor=demag, damping=0.05)
    b = Layer("b", mag=CVector(1, 0.1, 0), anis=CVector(1, 0, 0), Ms=0.9, thickness=2e-9, cellSurface=1e-16,
              demagTensor=demag, damping=0.06)
    a.setTopDipoleTensor(dip)
    b.setBottomDipoleTensor(dip)
    j = J([a, b])
    j.setLayerAnisotropyDriver("a", constantDriver(4e5))
    j.setLayerAnisotropyDriver("b", constantDriver(2e4))
    j.setLayerSecondOrderAnisotropyDriver("a", constantDriver(-3e4))
    j.setIDMIDriver("a", "b", AxialDriver(1e-4, -2e-4, 5e-5))
    j.setIECDriver("a", "b", sineDriver(1e-4, 5e-5, 2e9, 0))
    j.setLayerHdmiDriver("all", AxialDriver(1e3, 0, -2e3))
    j.setLayerExternalFieldDriver("b", AxialDriver(100e3, 0, 0))
    return j, dict(totalTime=1e-9, timeStep=1e-13, writeFrequency=1e-12, solverMode=solver)


def energies(J):
    """calculateEnergies=True on an SOT/IEC bilayer."""
    j, kw = sot_dynamic(J)
    kw = dict(kw)
    kw["calculateEnergies"] = True
    return j, kw


def custom_driver(J):
    """A Python-callback anisotropy and current (tabulated at stage times)."""
    l = Layer.createSTTLayer("free", mag=CVector(0, 0.2, 1.0), anis=CVector(0, 0, 1.0), Ms=1.0, thickness=1.4e-9,
                             cellSurface=1e-16, demagTensor=DEMAG, damping=0.1, SlonczewskiSpacerLayerParameter=1.0,
                             beta=0.0, spinPolarisation=0.5)
    l.setReferenceLayer(CVector(0, 1, 0))
    j = J([l], 100, 200)
    j.setLayerAnisotropyDriver("free", ScalarDriver.getCustomDriver(lambda t: 3e5 + 1e5 * math.sin(2e9 * t)))
    j.setLayerCurrentDriver("free", ScalarDriver.getCustomDriver(lambda t: -5e9 * math.sqrt(t / 1e-9)))
    return j, dict(totalTime=5e-10, timeStep=1e-13, writeFrequency=1e-12, calculateEnergies=True)


def five_layer_chain(J, solver=SolverMode.RK4):
    """5 layers, IEC between every adjacent pair, STRIP R (8 lanes per
    instance in the one-layer-per-lane kernels)."""
    layers = []
    for q in range(5):
        layers.append(Layer(f"L{q}", mag=CVector(1, 0.1 * (q + 1), 0.05 * q), anis=CVector(1, 0, 0),
                            Ms=1.2 + 0.1 * q, thickness=(1 + 0.5 * q) * 1e-9, cellSurface=SURF,
                            demagTensor=DEMAG, damping=0.05))
    j = J(layers, [100.0 + q for q in range(5)], [50.0 + q for q in range(5)], [1.0] * 5, [1.0] * 5,
          [0.5] * 5, [0.5] * 5, [2.0] * 5)
    for q in range(5):
        j.setLayerAnisotropyDriver(f"L{q}", constantDriver(20e3 + 10e3 * q))
    for q in range(4):
        j.setIECDriver(f"L{q}", f"L{q + 1}", constantDriver(-2e-4 * (1 + 0.3 * q)))
    j.setLayerExternalFieldDriver("all", AxialDriver(200e3, 20e3, 5e3))
    j.setLayerOerstedFieldDriver("all", AxialDriver(NullDriver(), stepDriver(0, 2e3, 0, 5e-12), NullDriver()))
    return j, dict(totalTime=1e-9, timeStep=1e-13, writeFrequency=1e-12, solverMode=solver)


CASES = {
    "rh_strip": rh_strip,
    "bilayer_iec_rk4": bilayer_iec,
    "bilayer_iec_heun": lambda J: bilayer_iec(J, SolverMode.Heun),
    "bilayer_iec_eulerheun": lambda J: bilayer_iec(J, SolverMode.EulerHeun),
    "trilayer_rk4": trilayer,
    "trilayer_heun": lambda J: trilayer(J, SolverMode.Heun),
    "trilayer_eulerheun": lambda J: trilayer(J, SolverMode.EulerHeun),
    "stt_drivers": stt_drivers,
    "stt_alternative": lambda J: stt_drivers(J, alternative=True),
    "stt_ref_top": stt_neighbour_reference,
    "stt_ref_bottom": lambda J: stt_neighbour_reference(J, Reference.bottom),
    "sot_dynamic_rk4": sot_dynamic,
    "sot_dynamic_heun": lambda J: sot_dynamic(J, SolverMode.Heun),
    "sot_static": sot_static,
    "exotic_rk4": exotic_terms,
    "exotic_eulerheun": lambda J: exotic_terms(J, SolverMode.EulerHeun),
    "energies": energies,
    "custom_driver": custom_driver,
    "five_layer_chain": five_layer_chain,
}
