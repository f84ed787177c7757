"""Coupled stacks on the CPU: the Python API's errors and logs (the
reference's tests/test_stack.py, run through the oracle), and the oracle's
stack runner pinned by what the mathematics fixes: uncoupled members evolve
exactly as the same junction alone, and every logged current and coupling
follows from the logged magnetisations of the previous step."""
from __future__ import annotations

from math import pi

import numpy as np
import pytest

import oracle
from cmtj_amd import CVector, Junction, Layer, ScalarDriver, SolverMode, _lib
from cmtj_amd.stack import ParallelStack, SeriesStack
from oracle_backend import OracleJunction, OracleParallelStack, OracleSeriesStack

DEMAG = [CVector(0, 0, 0), CVector(0, 0, 0), CVector(0, 0, 1)]
SURF = 150e-9 * 150e-9 * pi


def _single(J=Junction, K=1.05e3, Ms=1.65, mag=(0, 0, 1), anis=(1, 0, 0), stt=False):
    """tests/conftest.py single_layer_mtj_fictious of the reference (optionally
    an STT layer, so that the current moves m)."""
    if stt:
        l1 = Layer.createSTTLayer("free", mag=CVector(*mag), anis=CVector(*anis), Ms=Ms, thickness=3e-9,
                                  cellSurface=SURF, demagTensor=DEMAG, damping=0.005,
                                  SlonczewskiSpacerLayerParameter=1.0, beta=0.1, spinPolarisation=0.6)
    else:
        l1 = Layer("free", mag=CVector(*mag), anis=CVector(*anis), Ms=Ms, thickness=3e-9, cellSurface=SURF,
                   demagTensor=DEMAG, damping=0.005)
    l1.setReferenceLayer(CVector(1, 0, 0))
    j = J([l1], 100, 200)
    j.setLayerAnisotropyDriver("free", ScalarDriver.getConstantDriver(K))
    return j


def _two(J=Junction, stt=False, mag=(0, 0, 1)):
    """tests/conftest.py two_layer_mtj of the reference."""
    mk = Layer.createSTTLayer if stt else Layer
    kw = dict(SlonczewskiSpacerLayerParameter=1.0, beta=0.1, spinPolarisation=0.6) if stt else {}
    l1 = mk("free", mag=CVector(*mag), anis=CVector(1, 0, 0), Ms=1.65, thickness=3e-9, cellSurface=SURF,
            demagTensor=DEMAG, damping=0.005, **kw)
    l2 = Layer("bottom", mag=CVector(0, 0, 1), anis=CVector(1, 0, 0), Ms=1.45, thickness=4e-9, cellSurface=SURF,
               demagTensor=DEMAG, damping=0.005)
    if stt:
        l1.setReferenceLayer(CVector(1, 0, 0))
    j = J([l1, l2], 100, 200)
    j.setLayerAnisotropyDriver("free", ScalarDriver.getConstantDriver(1.05e3))
    j.setLayerAnisotropyDriver("bottom", ScalarDriver.getConstantDriver(50e3))
    j.setIECDriver("free", "bottom", ScalarDriver.getConstantDriver(-1.78e-3))
    return j


# --- API (reference tests/test_stack.py) ------------------------------------

def test_invalid_stack_indices():
    j = _single()
    with pytest.raises(RuntimeError, match="Asking for id of a non-existing junction!"):
        ParallelStack([j, j]).getLog(2)
    with pytest.raises(RuntimeError, match="Asking for id of a non-existing junction!"):
        SeriesStack([j, j]).getLog(10)
    with pytest.raises(TypeError):
        ParallelStack([j, j]).getLog(-1)
    with pytest.raises(TypeError):
        SeriesStack([j, j]).getLog(-10)


def test_invalid_stack():
    j = _single()
    for cls in (ParallelStack, SeriesStack):
        with pytest.raises(RuntimeError, match="must have at least 2 junctions"):
            cls([j])
    with pytest.raises(RuntimeError, match="must have at least 2 junctions"):
        ParallelStack([])
    strip = Junction([Layer("free", mag=CVector(0, 0, 1), anis=CVector(1, 0, 0), Ms=1.0, thickness=1e-9,
                            cellSurface=SURF, demagTensor=DEMAG)], [100.0], [0.0], [1.0], [0.0], [0.0], [0.0], [0.0])
    with pytest.raises(RuntimeError, match="non-classic magnetoresitance"):
        SeriesStack([strip, strip])


def test_configuration_errors():
    j = _single()
    s = SeriesStack([j, j, j])
    with pytest.raises(RuntimeError, match="size of stack size - 1"):
        s.setCouplingStrength([0.1, 0.2, 0.3])
    s.setCouplingStrength([0.1, 0.2])
    assert s.getCoupling(1) == 0.2
    with pytest.raises(RuntimeError, match="only supported for 2 layer stacks"):
        s.setDelayed(False)
    SeriesStack([_two(), _two()]).setDelayed(False)
    with pytest.raises(RuntimeError, match="time step cannot be larger"):
        s.runSimulation(1e-10, 1e-12, 1e-13)
    j2 = _single()
    j2.getLayer("free").setReferenceLayer(CVector(0, 1, 0))
    with pytest.raises(RuntimeError, match="Polarisation vectors are not equal"):
        OracleSeriesStack([j, j2]).runSimulation(1e-11, 1e-12, 1e-12)


@pytest.mark.parametrize("member", [_single, _two])
@pytest.mark.parametrize("cls", [OracleSeriesStack, OracleParallelStack])
def test_basic_stack_logs(member, cls):
    stack = cls([member(OracleJunction), member(OracleJunction)])
    for c in (0, 0.1, 0.5, 1):
        stack.setCouplingStrength(c)
        stack.runSimulation(2e-10, 1e-12, 1e-12)
    log = stack.getLog()
    assert set(log) == {"Resistance", "I_0", "I_1", "C_0_1", "time"}
    assert len(log["time"]) == 4 * 200
    stack.clearLogs()
    stack.runSimulation(2e-10, 1e-12, 1e-12)
    assert len(stack.getLog()["time"]) == 200 and len(stack.getLog(0)["time"]) == 200


@pytest.mark.parametrize("member", [_single, _two])
def test_series_resistance_exceeds_parallel(member):
    p = OracleParallelStack([member(OracleJunction), member(OracleJunction)])
    s = OracleSeriesStack([member(OracleJunction), member(OracleJunction)])
    p.runSimulation(2e-10, 1e-12, 1e-12)
    s.runSimulation(2e-10, 1e-12, 1e-12)
    pr, sr = p.getLog()["Resistance"], s.getLog()["Resistance"]
    assert all(r > 0 for r in pr) and all(r > 0 for r in sr)
    assert np.mean(sr) > np.mean(pr)


@pytest.mark.parametrize("cls", [OracleParallelStack, OracleSeriesStack])
def test_kcl_vs_non_kcl_current_behavior(cls):
    """The reference's test_kcl_vs_non_kcl_current_behavior."""
    j1 = _single(OracleJunction)
    j2 = _single(OracleJunction, K=1.05e2, Ms=1.35, mag=(0.3, 0.2, 0.9), anis=(1, 0.1, 0.3))
    non = cls([j1, j2], useKCL=False)
    kcl = cls([j1, j2], useKCL=True)
    for s in (non, kcl):
        s.setCouplingStrength(0.3)
        s.setCoupledCurrentDriver(ScalarDriver.getConstantDriver(2e12))
        s.runSimulation(0.1e-9, 1e-12, 1e-12)
    i_non = np.array(non.getLog()["I_0"]) / 1e12
    i_kcl = np.array(kcl.getLog()["I_0"]) / 1e12
    assert np.allclose(i_non, 2.0, rtol=1e-6)
    assert not np.allclose(np.mean(i_kcl), 2.0, rtol=1e-6)
    assert not np.allclose(np.mean(i_non), np.mean(i_kcl), rtol=1e-6)
    assert len(non.getLog()["time"]) == len(kcl.getLog()["time"]) and "I_1" in non.getLog()


# --- oracle pinning --------------------------------------------------------

@pytest.mark.parametrize("member", [_single, _two])
@pytest.mark.parametrize("kcl", [True, False])
def test_uncoupled_members_evolve_as_alone(member, kcl):
    """Coupling 0: the current is the driver's own value for every member, so
    each member's log equals the same junction run alone with that current
    driver (the stack's stepping and current override add nothing)."""
    # KCL holds the current of the step start through the step: a constant
    # driver; non-KCL evaluates the driver at the stage times: a sine
    drv = (ScalarDriver.getConstantDriver(5e11) if kcl else ScalarDriver.getSineDriver(5e11, 3e11, 4e9, 0.2))
    members = [member(OracleJunction, stt=True), member(OracleJunction, stt=True, mag=(0.2, 0.3, 0.9))]
    s = (OracleSeriesStack if kcl else OracleParallelStack)(members, useKCL=kcl, phaseOffset=0.0)
    s.setCouplingStrength(0.0)
    s.setCoupledCurrentDriver(drv)
    s.runSimulation(3e-10, 1e-13, 1e-12)
    for q, base in enumerate([member(OracleJunction, stt=True), member(OracleJunction, stt=True,
                                                                       mag=(0.2, 0.3, 0.9))]):
        base.setLayerCurrentDriver("all", drv)
        base.runSimulation(3e-10, 1e-13, 1e-12)
        a, b = base.getLog(), s.getLog(q)
        assert a.keys() == b.keys()
        for k in a:
            np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]))


def _m_start(log, ids, init):
    """m of the given layers at the start of each logged step (W = 1: the
    previous entry, the initial state for step 0)."""
    out = []
    for lid in ids:
        arr = np.stack([np.asarray(log[f"{lid}_m{c}"]) for c in "xyz"], axis=1)
        out.append(np.vstack([init[lid], arr[:-1]]))
    return out


@pytest.mark.parametrize("series", [True, False])
@pytest.mark.parametrize("kcl", [True, False])
@pytest.mark.parametrize("member", [_single, _two])
def test_logged_currents_follow_from_logged_magnetisations(series, kcl, member):
    """Every step's pair coupling and member currents, recomputed from the
    logged magnetisations of the step start (W = 1), match the stack log."""
    I0 = 1.5e12
    members = [member(OracleJunction, stt=True), member(OracleJunction, stt=True, mag=(0.4, 0.1, 0.9)),
               member(OracleJunction, stt=True, mag=(0.1, 0.5, 0.8))]
    cls = OracleSeriesStack if series else OracleParallelStack
    s = cls(members, useKCL=kcl)
    s.setCouplingStrength([0.2, -0.35])
    s.setCoupledCurrentDriver(ScalarDriver.getConstantDriver(I0))
    init = [{l.id: np.array(l.mag.tolist()) for l in j.layers} for j in s.junctionList]
    s.runSimulation(2e-10, 1e-13, 1e-13)
    two = s.isTwoLayerMemberStack()
    d = []
    for q, j in enumerate(s.junctionList):
        lg = j.getLog()
        ids = ["free", "bottom"] if two else ["free"]
        ms = _m_start(lg, ids, init[q])
        p = ms[1] if two else np.array([1.0, 0.0, 0.0])[None, :]
        d.append(ms[0][:, 0] * p[:, 0] + ms[0][:, 1] * p[:, 1] + ms[0][:, 2] * p[:, 2])
    c = [0.2, -0.35]
    C = [c[q] * (d[q] + d[q + 1] if series else d[q] - d[q + 1]) for q in range(2)]
    log = s.getLog()
    if kcl:
        total = I0 + C[0] * I0 + C[1] * I0
        for q in range(3):
            np.testing.assert_allclose(log[f"I_{q}"], total, rtol=1e-12)
        for q in range(2):
            np.testing.assert_allclose(log[f"C_{q}_{q + 1}"], C[q], rtol=1e-12, atol=1e-15)
    else:
        eff1 = 1 + C[0]
        eff2 = eff1 * (1 + C[1])
        np.testing.assert_allclose(log["I_0"], I0, rtol=1e-15)
        np.testing.assert_allclose(log["I_1"], I0 * eff1, rtol=1e-12)
        np.testing.assert_allclose(log["I_2"], I0 * eff2, rtol=1e-12)
        np.testing.assert_allclose(log["C_0_1"], eff1, rtol=1e-12)
        np.testing.assert_allclose(log["C_1_2"], eff2, rtol=1e-12)
    R = [np.asarray(s.getLog(q)["R" if not two else "R_free_bottom"]) for q in range(3)]
    want = R[0] + R[1] + R[2] if series else 1.0 / (1.0 / R[0] + 1.0 / R[1] + 1.0 / R[2])
    np.testing.assert_allclose(log["Resistance"], want, rtol=1e-14)


def test_current_moves_the_stt_members():
    """The stack current reaches the STT torque: with a large current the
    members' trajectories differ from a zero-current stack."""
    def run(I):
        s = OracleSeriesStack([_single(OracleJunction, stt=True), _single(OracleJunction, stt=True)])
        s.setCoupledCurrentDriver(ScalarDriver.getConstantDriver(I))
        s.setCouplingStrength(0.1)
        s.runSimulation(2e-10, 1e-13, 1e-12)
        return np.asarray(s.getLog(0)["free_mx"])
    assert np.max(np.abs(run(5e12) - run(0.0))) > 1e-4


def test_batched_stacks_equal_separate_runs_oracle():
    """run_stacks over several stacks gives each stack what its own
    runSimulation gives (the oracle backend runs them one by one)."""
    from cmtj_amd.stack import run_stacks

    def build(k):
        s = (OracleSeriesStack if k % 2 else OracleParallelStack)(
            [_single(stt=True, mag=(0.1 * k, 0.2, 1.0)), _single(stt=True), _single(stt=True, mag=(0.5, 0, 0.8))],
            useKCL=bool(k % 3))
        s.setCouplingStrength(0.1 * (k + 1))
        s.setCoupledCurrentDriver(ScalarDriver.getConstantDriver(1e12 * (k + 1)))
        return s

    batch = [build(k) for k in range(4)]
    alone = [build(k) for k in range(4)]
    for k in range(4):  # same noise seeds for the comparison
        for a, b in zip(batch[k].junctionList, alone[k].junctionList):
            b._seed = a._seed
    run_stacks(batch, 1e-10, 1e-13, 1e-12)
    for s in alone:
        s.runSimulation(1e-10, 1e-13, 1e-12)
    for a, b in zip(batch, alone):
        assert a.getLog().keys() == b.getLog().keys()
        for k in a.getLog():
            np.testing.assert_array_equal(np.asarray(a.getLog()[k]), np.asarray(b.getLog()[k]))
        for q in range(3):
            for k in a.getLog(q):
                np.testing.assert_array_equal(np.asarray(a.getLog(q)[k]), np.asarray(b.getLog(q)[k]))


@pytest.mark.parametrize("series", [True, False])
def test_non_delayed_currents_follow_from_logged_magnetisations(series):
    """setDelayed(False) (core/stack.hpp:334-343): member q's coupling to
    member q-1 reads q-1's magnetisations after this step (its log entry of
    the step, W = 1) and q's before it (q's previous entry).  Recomputed from
    the logs, every I_q and C_{q-1}_q matches; the delayed rule (both from
    the step start) does not."""
    I0 = 1.5e12
    members = [_two(OracleJunction, stt=True), _two(OracleJunction, stt=True, mag=(0.4, 0.1, 0.9)),
               _two(OracleJunction, stt=True, mag=(0.1, 0.5, 0.8))]
    cls = OracleSeriesStack if series else OracleParallelStack
    s = cls(members, useKCL=False)
    s.setDelayed(False)
    s.setCouplingStrength([0.2, -0.35])
    s.setCoupledCurrentDriver(ScalarDriver.getConstantDriver(I0))
    init = [{l.id: np.array(l.mag.tolist()) for l in j.layers} for j in s.junctionList]
    s.runSimulation(2e-10, 1e-13, 1e-13)
    after, before = [], []
    for q, j in enumerate(s.junctionList):
        lg = j.getLog()
        post = [np.stack([np.asarray(lg[f"{lid}_m{c}"]) for c in "xyz"], axis=1) for lid in ("free", "bottom")]
        pre = _m_start(lg, ["free", "bottom"], init[q])
        after.append(np.sum(post[0] * post[1], axis=1))
        before.append(np.sum(pre[0] * pre[1], axis=1))
    c = [0.2, -0.35]
    sgn = 1.0 if series else -1.0
    eff1 = 1 + c[0] * (after[0] + sgn * before[1])
    eff2 = eff1 * (1 + c[1] * (after[1] + sgn * before[2]))
    log = s.getLog()
    np.testing.assert_allclose(log["I_0"], I0, rtol=1e-15)
    np.testing.assert_allclose(log["I_1"], I0 * eff1, rtol=1e-12)
    np.testing.assert_allclose(log["I_2"], I0 * eff2, rtol=1e-12)
    np.testing.assert_allclose(log["C_0_1"], eff1, rtol=1e-12)
    np.testing.assert_allclose(log["C_1_2"], eff2, rtol=1e-12)
    delayed1 = 1 + c[0] * (before[0] + sgn * before[1])
    assert np.max(np.abs(delayed1 - eff1)) > 1e-6  # the two rules differ on this run


def test_non_delayed_needs_two_layer_members_in_the_engine():
    s = SeriesStack([_two(), _two()], useKCL=False)
    s.setDelayed(False)
    st, js, layers, *_ = s._pack(1e-11, 1e-13, 1e-12)
    assert int(st["delayed"]) == 0 and int(st["kcl"]) == 0 and int(st["two_layer"]) == 1


# --- callback drivers in stacks ------------------------------------------------

def _pulse(t):
    """A callback stack current: a smooth rise and a plateau (A/m^2)."""
    import math
    return 1.5e12 * (1 - math.exp(-t / 4e-11)) + 2e11


def test_kcl_callback_current_follows_the_callback():
    """KCL reads the stack driver once per step at the step's start
    (runSimulationKCL, core/stack.hpp:376): every logged I_q equals
    callback(t) (1 + sum of the pair couplings recomputed from the logged
    magnetisations), and a constant callback runs exactly as the constant
    driver."""
    members = [_single(OracleJunction, stt=True), _single(OracleJunction, stt=True, mag=(0.4, 0.1, 0.9)),
               _single(OracleJunction, stt=True, mag=(0.1, 0.5, 0.8))]
    s = OracleSeriesStack(members, useKCL=True)
    s.setCouplingStrength([0.2, -0.35])
    s.setCoupledCurrentDriver(ScalarDriver.getCustomDriver(_pulse))
    init = [{l.id: np.array(l.mag.tolist()) for l in j.layers} for j in s.junctionList]
    s.runSimulation(2e-10, 1e-13, 1e-13)
    d = []
    for q, j in enumerate(s.junctionList):
        ms = _m_start(j.getLog(), ["free"], init[q])
        d.append(ms[0][:, 0])  # pol = x
    C = [c * (d[q] + d[q + 1]) for q, c in enumerate([0.2, -0.35])]
    t = np.asarray(s.getLog()["time"])
    I0 = np.array([_pulse(x) for x in t])
    for q in range(3):
        np.testing.assert_allclose(s.getLog()[f"I_{q}"], I0 * (1 + C[0] + C[1]), rtol=1e-12)
    a = OracleSeriesStack([_single(OracleJunction, stt=True), _single(OracleJunction, stt=True, mag=(0.4, 0.1, 0.9))])
    b = OracleSeriesStack([_single(OracleJunction, stt=True), _single(OracleJunction, stt=True, mag=(0.4, 0.1, 0.9))])
    a.setCoupledCurrentDriver(ScalarDriver.getCustomDriver(lambda t: 7e11))
    b.setCoupledCurrentDriver(ScalarDriver.getConstantDriver(7e11))
    for x in (a, b):
        x.setCouplingStrength(0.3)
        x.runSimulation(1e-10, 1e-13, 1e-12)
    for k in a.getLog():
        np.testing.assert_array_equal(np.asarray(a.getLog()[k]), np.asarray(b.getLog()[k]))
    for q in range(2):
        for k in a.getLog(q):
            np.testing.assert_array_equal(np.asarray(a.getLog(q)[k]), np.asarray(b.getLog(q)[k]))


def test_non_kcl_callback_current_fails_as_in_the_reference():
    s = OracleSeriesStack([_single(OracleJunction), _single(OracleJunction)], useKCL=False)
    s.setCoupledCurrentDriver(ScalarDriver.getCustomDriver(_pulse))
    with pytest.raises(RuntimeError, match="Error in Python callback"):
        s.runSimulation(1e-11, 1e-13, 1e-12)


def test_member_callback_drivers_in_a_stack():
    """A member's own callback driver (an anisotropy that ramps) is sampled
    at its stage times as for a junction alone: with coupling 0 each member
    evolves as the same junction run alone."""
    def ramp(t):
        return 1.05e3 + 4e13 * t

    members = [_single(OracleJunction, stt=True), _single(OracleJunction, stt=True, mag=(0.2, 0.3, 0.9))]
    for j in members:
        j.setLayerAnisotropyDriver("free", ScalarDriver.getCustomDriver(ramp))
    s = OracleParallelStack(members, useKCL=False, phaseOffset=0.0)
    s.setCouplingStrength(0.0)
    drv = ScalarDriver.getSineDriver(5e11, 3e11, 4e9, 0.2)
    s.setCoupledCurrentDriver(drv)
    s.runSimulation(2e-10, 1e-13, 1e-12)
    for q, mag in enumerate([(0, 0, 1), (0.2, 0.3, 0.9)]):
        base = _single(OracleJunction, stt=True, mag=mag)
        base.setLayerAnisotropyDriver("free", ScalarDriver.getCustomDriver(ramp))
        base.setLayerCurrentDriver("all", drv)
        base.runSimulation(2e-10, 1e-13, 1e-12)
        a, b = base.getLog(), s.getLog(q)
        for k in a:
            np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]))
    # the ramp reaches the members: a constant anisotropy gives another path
    c = OracleParallelStack([_single(OracleJunction, stt=True), _single(OracleJunction, stt=True, mag=(0.2, 0.3, 0.9))],
                            useKCL=False, phaseOffset=0.0)
    c.setCouplingStrength(0.0)
    c.setCoupledCurrentDriver(drv)
    c.runSimulation(2e-10, 1e-13, 1e-12)
    assert np.max(np.abs(np.asarray(c.getLog(1)["free_mx"]) - np.asarray(s.getLog(1)["free_mx"]))) > 1e-6
