"""The cmtj-compatible API: names, defaults, operators, error types and
messages, log keys and run semantics (reference tests/test_drivers.py,
test_sanity.py, test_basic.py, test_constants.py, test_sot_torques.py).
Simulations here run through the CPU oracle (the descriptor packing and log
assembly under test are the product's own code)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import cmtj_amd as cmtj
from cmtj_amd import (AxialDriver, CVector, Junction, Layer, NullDriver, Reference, ScalarDriver, SolverMode,
                      constantDriver, sineDriver, stepDriver)
from oracle_backend import OracleJunction

DEMAG = [CVector(0, 0, 0), CVector(0, 0, 0), CVector(0, 0, 1)]


def _layer(id="free", **kw):
    args = dict(mag=CVector(0, 0, 1), anis=CVector(1, 0, 0), Ms=1.65, thickness=3e-9, cellSurface=1e-16,
                demagTensor=DEMAG, damping=0.005)
    args.update(kw)
    return Layer(id, **args)


# --- CVector and drivers (reference tests/test_drivers.py) -------------------

def test_cvector_operators():
    c1, c2, c3 = CVector(1.0, 2.0, 3.0), CVector(1.0, 2.0, 3.0), CVector(2, 4, 6)
    assert (c1[0], c1[1], c1[2]) == (1.0, 2.0, 3.0)
    assert c1 == c2 and c1 != c3
    assert (c1 + c2) == c3 and (c3 - c1) == c1
    assert 2 * c1 == c3 and c3 * 0.5 == c1
    c1 += CVector(1.0, 2.0, 3.0)
    assert c1 == c3
    c3 -= c2
    assert c3 == c2
    assert len(c1) == 3 and list(c1) == [2.0, 4.0, 6.0]
    assert str(CVector(1, 0.5, 1e-7)) == "[x:1, y:0.5, z:1e-07]"
    v = CVector(3, 4, 0)
    v.normalize()
    assert v.tolist() == [0.6, 0.8, 0.0]
    assert cmtj.c_dot(CVector(1, 2, 3), CVector(4, 5, 6)) == 32


def test_axial_definitions_and_aliases():
    vec = (1.0, 2.0, 3.0)
    assert AxialDriver(*vec).getCurrentAxialDrivers(0.0) == CVector(*vec)
    assert AxialDriver(CVector(*vec)).getCurrentAxialDrivers(0.0) == CVector(*vec)
    d = AxialDriver(constantDriver(1.0), constantDriver(2.0), constantDriver(3.0))
    assert d.getCurrentAxialDrivers(1e6) == CVector(1.0, 2.0, 3.0)
    d.applyMask([1, 0, 1])
    assert d.getCurrentAxialDrivers(0.0) == CVector(1.0, 0.0, 3.0)
    with pytest.raises(RuntimeError, match="binary"):
        d.applyMask([2, 0, 1])
    with pytest.raises(RuntimeError, match="3 axes"):
        AxialDriver([constantDriver(1.0)] * 2)


def test_driver_ops():
    d = sineDriver(10, 20, 1, 0)
    assert d.getCurrentScalarValue(1 / 4) == 30
    d *= 2
    assert d.getCurrentScalarValue(1 / 4) == 60
    d = sineDriver(10, 20, 1, 0)
    d += 2
    assert d.getCurrentScalarValue(1 / 4) == 34
    assert (sineDriver(10, 20, 1, 0) * 2).getCurrentScalarValue(1 / 4) == 60
    assert (sineDriver(10, 20, 1, 0) + 2).getCurrentScalarValue(1 / 4) == 34
    assert NullDriver().getCurrentScalarValue(5.0) == 0.0
    assert ScalarDriver().getCurrentScalarValue(5.0) == 0.0
    with pytest.raises(RuntimeError, match="Start time cannot be later"):
        stepDriver(0, 1, 2e-9, 1e-9)


def test_gaussian_step_factory_builds_an_impulse():
    # the reference's getGaussianStepDriver returns a gaussimpulse (core/drivers.hpp:241-245)
    d = ScalarDriver.getGaussianStepDriver(0, 1, 1e-9, 1e-10)
    assert d.update == cmtj.UpdateType.gaussimpulse
    assert d.getCurrentScalarValue(1e-9) == 1.0


def test_custom_driver():
    d = ScalarDriver.getCustomDriver(lambda t: math.sqrt(t))
    assert d.getCurrentScalarValue(1e-9) == math.sqrt(1e-9)

    def faulty(t):
        raise ValueError("Intentional exception")

    with pytest.raises(RuntimeError):
        ScalarDriver.getCustomDriver(faulty).getCurrentScalarValue(1e-9)

    def no_arg():
        return 1.0

    with pytest.raises(RuntimeError):
        ScalarDriver.getCustomDriver(no_arg)


def test_custom_driver_on_junction():
    # reference tests/test_drivers.py:126-137
    l = Layer("free", mag=CVector(0, 0, 1), anis=CVector(1, 0, 0), Ms=1.65, thickness=3e-9,
              cellSurface=150e-9 * 150e-9 * math.pi, demagTensor=DEMAG, damping=0.005)
    j = OracleJunction([l])
    j.setLayerAnisotropyDriver("all", ScalarDriver.getCustomDriver(lambda t: 12345))
    j.runSimulation(1e-9, 1e-13, 1e-13, calculateEnergies=True)
    assert all(x == 12345 for x in j.getLog()["free_K"])


# --- constants (reference tests/test_constants.py) ----------------------------

def test_constants_runtime_changes():
    PC = cmtj.constants.PhysicalConstants
    assert PC.gyromagnetic_ratio() == 220880.0
    assert PC.magnetic_permeability() == 12.566e-7
    assert PC.hbar() == 6.62607015e-34 / (2.0 * math.pi)
    new = dict(magnetic_permeability=15e-7, gyromagnetic_ratio=250000.0, TtoAm=800000.0, hbar=1.1e-34,
               elementary_charge=1.7e-19, boltzmann_constant=1.5e-23)
    PC.set_magnetic_permeability(new["magnetic_permeability"])
    PC.set_gyromagnetic_ratio(new["gyromagnetic_ratio"])
    PC.set_TtoAm(new["TtoAm"])
    PC.set_hbar(new["hbar"])
    PC.set_elementary_charge(new["elementary_charge"])
    PC.set_boltzmann_constant(new["boltzmann_constant"])
    for k, v in new.items():
        assert getattr(PC, k)() == v
    PC.resetToDefaults()
    assert PC.gyromagnetic_ratio() == 220880.0 and PC.TtoAm() == 795774.715459


def test_constants_reach_the_next_run():
    """A set_* between runs changes the next run (constants are read live)."""
    def final(gyro):
        cmtj.constants.PhysicalConstants.set_gyromagnetic_ratio(gyro)
        l = _layer(mag=CVector(1, 0, 0))
        j = OracleJunction([l])
        j.setLayerExternalFieldDriver("all", AxialDriver(0, 0, 1e5))
        j.runSimulation(1e-10, 1e-13, 1e-11)
        return j.getLayerMagnetisation("free")

    a, b = final(220880.0), final(2 * 220880.0)
    assert a != b


# --- construction errors (reference tests/test_sanity.py, test_basic.py) ------

def test_construction_errors():
    with pytest.raises(ValueError, match="Layers must have unique ids!"):
        Junction([_layer(), _layer()])
    with pytest.raises(ValueError, match="Passed a zero length Layer vector!"):
        Junction([])
    with pytest.raises(ValueError, match="must have a pinning"):
        Junction([_layer()], 100, 200)
    with pytest.raises(ValueError, match="only bilayers"):
        Junction([_layer("a"), _layer("b"), _layer("c")], 100, 200)
    with pytest.raises(ValueError, match="same size"):
        Junction([_layer()], [1, 2], [0], [0], [0], [0], [0], [0])
    with pytest.raises(RuntimeError, match="zero vector"):
        _layer(mag=CVector(0, 0, 0))
    with pytest.raises(RuntimeError, match="Anisotropy was set to a zero vector"):
        _layer(anis=CVector(0, 0, 0))
    j = Junction([_layer("a"), _layer("b")])
    with pytest.raises(RuntimeError, match="Failed to find a layer with a given id: c!"):
        j.setLayerAnisotropyDriver("c", constantDriver(1))
    with pytest.raises(RuntimeError, match="Failed to match the layer order"):
        j.setIECDriver("a", "x", constantDriver(1))


def test_sot_misuse_raises():
    # reference tests/test_sot_torques.py:158-181
    l = Layer.createSOTLayer(id="free", mag=CVector(0, 0, 1), anis=CVector(0, 0, 1), Ms=1, thickness=2e-9,
                             cellSurface=1e-18, demagTensor=[CVector(0, 0, 0)] * 3, damping=0.01)
    with pytest.raises(RuntimeError):
        l.setFieldLikeTorqueDriver(constantDriver(0.1))
    with pytest.raises(RuntimeError):
        l.setDampingLikeTorqueDriver(constantDriver(0.1))
    s = Layer.createSTTLayer(id="s", mag=CVector(0, 0, 1), anis=CVector(0, 0, 1), Ms=1, thickness=2e-9,
                             cellSurface=1e-18, demagTensor=DEMAG)
    with pytest.raises(RuntimeError, match="includeSTT"):
        s.setFieldLikeTorqueDriver(constantDriver(0.1))


def test_reference_layers():
    l = _layer()
    l.setReferenceLayer(CVector(1, 0, 0))
    assert l.getReferenceLayer() == CVector(1, 0, 0)
    l.setSecondaryReferenceLayer(CVector(0, 1, 0))
    assert l.getSecondaryReferenceLayer() == CVector(0, 1, 0)
    l2 = _layer()
    with pytest.raises(RuntimeError, match="Cannot set fixed polarisation layer to 0"):
        l2.setReferenceLayer(Reference.fixed)


def test_junction_copies_layers_and_getlayer_refers():
    l = _layer()
    j = Junction([l])
    l.setAnisotropyDriver(constantDriver(99.0))
    assert j.getLayer("free")._scalar["anisotropy"].constantValue == 0.0
    j.getLayer("free").setAnisotropyDriver(constantDriver(7.0))
    assert j.getLayer("free")._scalar["anisotropy"].constantValue == 7.0


def test_initial_magnetisation_kept_unnormalised_but_setter_normalises():
    l = _layer(mag=CVector(0, 0, 0.99))
    assert l.mag == CVector(0, 0, 0.99)
    j = Junction([l])
    j.setLayerMagnetisation("free", CVector(0, 3, 4))
    assert j.getLayerMagnetisation("free") == CVector(0, 0.6, 0.8)


# --- runs and logs ------------------------------------------------------------

def test_step_and_log_counts_follow_fp64_truncation():
    j = OracleJunction([_layer()])
    j.runSimulation(1e-9, 1e-11, 1e-11)  # reference tests/test_layer.cpp:36-56: 100 entries
    assert len(j.getLog()["time"]) == 100
    j.clearLog()
    # 100e-9 / 1e-13 truncates to 999999 steps; default writeFrequency gives W = 99
    jd, L, _, n, W = j._descriptor(100e-9, 1e-13, 1e-11, False, SolverMode.RK4)
    assert (n, W) == (999999, 99)
    assert int(jd["n_entries"]) == (999999 + 98) // 99


def test_a_run_shorter_than_one_step_logs_nothing():
    j = OracleJunction([_layer(mag=CVector(1, 0, 0))])
    j.runSimulation(0.5e-13, 1e-13, 1e-13)  # (unsigned)(0.5) = 0 steps
    assert j.getLog() == {}


def test_log_times_and_continuation():
    j = OracleJunction([_layer(mag=CVector(1, 0, 0))])
    j.setLayerExternalFieldDriver("all", AxialDriver(0, 0, 1e5))
    j.runSimulation(1e-11, 1e-13, 1e-12)
    t = j.getLog()["time"]
    assert t == [(k * 10) * 1e-13 for k in range(10)]  # entry k labelled (kW) dt
    m1 = j.getLayerMagnetisation("free")
    j.runSimulation(1e-11, 1e-13, 1e-12)  # t restarts at 0, m persists, log appends
    lg = j.getLog()
    assert lg["time"][10:] == t and len(lg["free_mx"]) == 20
    assert lg["free_mx"][10] != m1.x or lg["free_my"][10] != m1.y
    j.clearLog()
    assert j.getLog() == {}


def test_log_key_sets():
    # reference tests/test_basic.py:27-52 and :55-102, tests/test_sot_torques.py:35-82
    l = _layer(mag=CVector(0, 0, 1))
    l.setReferenceLayer(CVector(1, 0, 0))
    j = OracleJunction([l], 100, 200)
    j.setLayerAnisotropyDriver("free", constantDriver(1.05e3))
    j.setLayerExternalFieldDriver("all", AxialDriver(CVector(300e3, 10e3, 0)))
    j.runSimulation(10e-9, 1e-12, 1e-12)
    lg = j.getLog()
    assert set(lg) == {"free_mx", "free_my", "free_mz", "R", "time"}
    assert np.all(np.asarray(lg["R"]))

    a, b = _layer("a"), _layer("b")
    j2 = OracleJunction([a, b], 100, 200)
    j2.runSimulation(1e-11, 1e-13, 1e-12)
    assert "R_a_b" in j2.getLog()

    j3 = OracleJunction([_layer("a")], [100.0], [0.0], [1.0], [0.0], [0.0], [0.0], [0.0])
    j3.runSimulation(1e-11, 1e-13, 1e-12)
    lg3 = j3.getLog()
    assert {"Rx", "Ry", "Rz"} <= set(lg3)
    assert set(lg3["Ry"]) == {0.0}  # 1 / (1 / 0) with all-zero Ry terms
    assert set(lg3["Rz"]) == {0.0}

    l4 = Layer("free", mag=CVector(0, 0, 1), anis=CVector(0, 0, 1), Ms=0.5, thickness=1e-9, cellSurface=70e-9,
               demagTensor=[CVector(0, 0, 0)] * 3, damping=0.03)
    l4.setAnisotropyDriver(constantDriver(1e6))
    l4.setSecondOrderAnisotropyDriver(constantDriver(1e6))
    j4 = OracleJunction([l4])
    j4.runSimulation(1e-9, 1e-12, 1e-12, calculateEnergies=True)
    lg4 = j4.getLog()
    assert np.all(np.asarray(lg4["free_K2"]) == 1e6)
    keys = {"K", "K2", "Jbottom", "Jtop", "I"} | {f"{b}{c}" for b in ("Hext", "Hiec", "Hanis", "Hdemag", "Hth")
                                                  for c in "xyz"}
    assert {f"free_{k}" for k in keys} <= set(lg4)
    assert len(lg4["free_I"]) == len(lg4["time"])


def test_sot_energy_keys_and_doubled_current():
    l = Layer.createSOTLayer(id="free", mag=CVector(0, 0.1, 1), anis=CVector(0, 0, 1), Ms=1, thickness=2e-9,
                             cellSurface=1e-18, demagTensor=DEMAG, damping=0.01, fieldLikeTorque=1e-9,
                             dampingLikeTorque=2e-9)
    l.setReferenceLayer(CVector(0, 1, 0))
    j = OracleJunction([l])
    j.setLayerCurrentDriver("all", constantDriver(1e12))
    j.runSimulation(1e-11, 1e-13, 1e-12, calculateEnergies=True)
    lg = j.getLog()
    assert {f"free_H{b}{c}" for b in ("fl", "dl") for c in "xyz"} <= set(lg)
    assert len(lg["free_I"]) == 2 * len(lg["time"])  # pushed twice (core/junction.hpp:1593, 1611-1612)
    assert lg["free_I"][:2] == [1e12, 1e12]
    assert lg["free_Hfly"][0] == 1e-9 * 1e12 - 0.01 * 2e-9 * 1e12


def test_temperature_switches_to_euler_heun(capsys):
    l = _layer(cellSurface=1e-16)
    j = OracleJunction([l])
    j.setLayerTemperatureDriver("all", constantDriver(300))
    jd, *_ = j._descriptor(1e-11, 1e-13, 1e-12, False, SolverMode.RK4)
    assert int(jd["solver"]) == SolverMode.EulerHeun
    assert "Solver automatically changed to Euler Heun" in capsys.readouterr().out
    jd, *_ = j._descriptor(1e-11, 1e-13, 1e-12, False, SolverMode.Heun)
    assert int(jd["solver"]) == SolverMode.Heun


def test_run_errors():
    j = OracleJunction([_layer()])
    with pytest.raises(RuntimeError, match="The time step cannot be larger than write frequency!"):
        j.runSimulation(1e-9, 1e-12, 1e-13)
    j0 = OracleJunction([_layer(cellSurface=0.0)])
    j0.setLayerTemperatureDriver("all", constantDriver(300))
    with pytest.raises(RuntimeError, match="Cell surface cannot be 0"):
        j0.runSimulation(1e-11, 1e-13, 1e-12)
    with pytest.raises(NotImplementedError, match="1 to 4 layers"):
        OracleJunction([_layer(f"l{i}") for i in range(5)]).runSimulation(1e-11, 1e-13, 1e-12,
                                                                         solverMode=SolverMode.DormandPrince)
    jt = OracleJunction([_layer(thickness=0.0)])
    with pytest.raises(RuntimeError, match="divide vector by zero"):
        jt.runSimulation(1e-11, 1e-13, 1e-12)


def test_nan_magnetisation_raises_and_keeps_partial_log():
    j = OracleJunction([_layer(mag=CVector(1, 0, 0))])
    j.setLayerExternalFieldDriver("all", AxialDriver(stepDriver(0, float("inf"), 5e-12, 1e-9), constantDriver(0),
                                                     constantDriver(1e5)))
    with pytest.raises(RuntimeError, match="NAN magnetisation"):
        j.runSimulation(1e-11, 1e-13, 1e-13)
    lg = j.getLog()
    assert 0 < len(lg["time"]) < 100
    assert all(np.isfinite(lg["free_mx"]))
    assert math.isnan(j.getLayerMagnetisation("free").x)


def test_magnetoresistance_getter_and_save(tmp_path):
    l = _layer(mag=CVector(0, 0, 1))
    l.setReferenceLayer(CVector(1, 0, 0))
    j = OracleJunction([l], 100, 200)
    assert j.getMagnetoresistance() == [150.0]
    with pytest.raises(RuntimeError, match="not supported"):
        Junction([_layer()]).getMagnetoresistance()
    j.runSimulation(1e-12, 1e-13, 1e-13)
    f = tmp_path / "log.csv"
    j.saveLog(str(f))
    lines = f.read_text().splitlines()
    assert len(lines) == 1 + len(j.getLog()["time"])
    assert lines[0].split(";")[:-1] == list(j.getLog().keys())
    with pytest.raises(RuntimeError, match="filename may not be empty"):
        j.saveLog("")


def test_module_exports():
    for name in ("Junction", "Layer", "CVector", "ScalarDriver", "AxialDriver", "NullDriver", "constantDriver",
                 "pulseDriver", "sineDriver", "posSineDriver", "stepDriver", "trapezoidDriver",
                 "gaussianImpulseDriver", "gaussianStepDriver", "c_dot", "RK4", "Heun", "EulerHeun", "xaxis",
                 "constant", "sine"):
        assert hasattr(cmtj, name), name
    assert cmtj.RK4 == SolverMode.RK4 and cmtj.EulerHeun == 0
