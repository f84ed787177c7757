"""Dormand-Prince on the GPU (k_dp / k_dp_ml: one adaptive lane per
instance of one or several layers) against the oracle: the same accepted steps, entry times and
magnetisations for an SOT-driven switch with second-order anisotropy, a sine
drive, an exact equilibrium (the step-size collapse of the reference's
controller) and a batch of mixed instances; at a tight tolerance, both on
the closed-form damped precession."""
from __future__ import annotations

import math

import numpy as np
import pytest

from cmtj_amd import (AdaptiveIntegrationParams, AxialDriver, CVector, Junction, Layer, SolverMode, constantDriver,
                      sineDriver, stepDriver)
from cmtj_amd.batch import JunctionBatch
from oracle_backend import OracleJunction

pytestmark = pytest.mark.gpu


def _case(name, J):
    demag = [CVector(0.017, 0, 0), CVector(0, 0.017, 0), CVector(0, 0, math.sqrt(1 - 2 * 0.017 ** 2))]
    if name == "precession":
        l = Layer("free", mag=CVector(math.sin(1.2), 0, math.cos(1.2)), anis=CVector(1, 0, 0), Ms=1.0,
                  thickness=1e-9, cellSurface=1e-16, demagTensor=[CVector(0, 0, 0)] * 3, damping=0.02)
        p = AdaptiveIntegrationParams()
        p.abs_tol = p.rel_tol = 1e-10
        l.setAdaptiveParams(p)
        j = J([l])
        j.setLayerExternalFieldDriver("all", AxialDriver(0, 0, 1e5))
        return j, (0.5e-9, 1e-13, 1e-12)
    if name == "sot_switch":  # the stability-diagram example of the reference's curated tests
        l = Layer("free", mag=CVector(0.1, 0.1, 0.99), anis=CVector(0, 0, 1), Ms=1.12, thickness=1e-9,
                  cellSurface=math.pi * (300e-9) ** 2, demagTensor=demag, damping=0.01)
        l.setReferenceLayer(CVector(0, 0, 1))
        l.setSecondOrderAnisotropyDriver(constantDriver(-30e3))
        l.setAnisotropyDriver(constantDriver(563e3))
        j = J([l])
        j.setLayerExternalFieldDriver("free", AxialDriver(2e4, 0, 1e5))
        j.setLayerFieldLikeTorqueDriver("free", stepDriver(0, 0.7 * 795774.715459 * 1e-3 * 0.04, 0, 1e-9))
        j.setLayerDampingLikeTorqueDriver("free", stepDriver(0, 0.7 * 795774.715459 * 16e-3 * 0.2, 0, 1e-9))
        return j, (2e-9, 1e-12, 1e-11)
    if name == "sine":
        l = Layer("free", mag=CVector(1, 0.1, 0), anis=CVector(1, 0, 0), Ms=1.65, thickness=3e-9,
                  cellSurface=1e-16, demagTensor=demag, damping=0.01)
        j = J([l])
        j.setLayerAnisotropyDriver("free", constantDriver(1e4))
        j.setLayerExternalFieldDriver("all", AxialDriver(1e5, 0, 0))
        j.setLayerOerstedFieldDriver("all", AxialDriver(constantDriver(0), sineDriver(0, 500, 7e9, 0),
                                                        constantDriver(0)))
        return j, (1e-9, 1e-13, 1e-12)
    if name == "equilibrium":  # zero torque: the controller shrinks the step to nothing and resets
        l = Layer("free", mag=CVector(0, 0, 1), anis=CVector(1, 0, 0), Ms=1.0, thickness=1e-9,
                  cellSurface=1e-16, demagTensor=[CVector(0, 0, 0)] * 3, damping=0.02)
        j = J([l])
        j.setLayerExternalFieldDriver("all", AxialDriver(0, 0, 1e5))
        return j, (3e-12, 1e-13, 1e-12)
    raise KeyError(name)


def _compare(a, b):
    la, lb = a.getLog(), b.getLog()
    assert la.keys() == lb.keys()
    assert len(la["time"]) == len(lb["time"])
    worst = 0.0
    for k in la:
        x, y = np.asarray(la[k]), np.asarray(lb[k])
        scale = float(np.max(np.abs(y))) if y.size else 1.0
        worst = max(worst, float(np.max(np.abs(x - y))) / max(scale, 1e-300) if y.size else 0.0)
    return worst


@pytest.mark.parametrize("name", ["sot_switch", "sine", "equilibrium"])
def test_dp_matches_oracle(name):
    g, args = _case(name, Junction)
    o, _ = _case(name, OracleJunction)
    g.runSimulation(*args, solverMode=SolverMode.DormandPrince)
    o.runSimulation(*args, solverMode=SolverMode.DormandPrince)
    worst = _compare(g, o)
    print(f"[dp] {name}: {len(g.getLog()['time'])} entries, max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8
    np.testing.assert_allclose(np.array(g.getLayerMagnetisation("free").tolist()),
                               np.array(o.getLayerMagnetisation("free").tolist()), atol=1e-8)


def test_dp_tight_tolerance_stays_on_the_exact_trajectory():
    """At tolerance 1e-10 the controller reads |m5 - m4| ~ 1e-10, so one-ulp
    differences between two builds move the proposed step by ~1e-7 relative
    and the two runs sample the same trajectory at slightly different times.
    Both stay on the closed-form damped precession, entry for entry."""
    from test_dp_cpu import _off_trajectory

    g, args = _case("precession", Junction)
    o, _ = _case("precession", OracleJunction)
    g.runSimulation(*args, solverMode=SolverMode.DormandPrince)
    o.runSimulation(*args, solverMode=SolverMode.DormandPrince)
    lg, lo = g.getLog(), o.getLog()
    assert len(lg["time"]) == len(lo["time"])
    for lg_ in (lg, lo):
        m = np.stack([lg_[f"free_m{c}"] for c in "xyz"], axis=1)
        off = float(_off_trajectory(m, 1.2, 1e5, 0.02)[0].max())
        print(f"[dp] precession: {len(lg_['time'])} entries, off the exact trajectory by {off:.3e} rad", flush=True)
        assert off < 1e-6
    np.testing.assert_allclose(lg["time"], lo["time"], rtol=1e-7)


def test_dp_batch_matches_oracle():
    """Instances of different lengths and dynamics in one launch, each lane
    with its own step history."""
    js, os_ = [], []
    for name in ("sot_switch", "sine", "equilibrium", "sine"):
        for J, out in ((Junction, js), (OracleJunction, os_)):
            j, _ = _case(name, J)
            out.append(j)
    T = [1e-9, 0.5e-9, 3e-12, 0.1e-9]
    b = JunctionBatch.from_junctions(js, T, 1e-13, 1e-12, SolverMode.DormandPrince, precision="fp64", log="full",
                                     log_dtype="f64")
    logs = run_and_split(b, js)
    for q, o in enumerate(os_):
        o.runSimulation(T[q], 1e-13, 1e-12, solverMode=SolverMode.DormandPrince)
        ol = o.getLog()
        assert len(ol["time"]) == len(logs[q]["time"]), q
        for k in ol:
            np.testing.assert_allclose(logs[q][k], ol[k], rtol=1e-8, atol=1e-12)


def run_and_split(b, js):
    from cmtj_amd.batch import run_batch  # noqa: F401  (same packing as run_batch)

    r = b.run()
    out = []
    for q in range(len(js)):
        out.append(r.instance_log(q))
    return out


# ---- several layers (k_dp_ml) ------------------------------------------------


def _ml_case(name, J):
    import parity_cases
    from test_dp_cpu import _pair

    if name == "shared_step":  # a torque-free first layer halves the shared step every attempt
        # tolerance 1e-6: at 1e-8 the error estimate |m5 - m4| keeps ~8 digits and one-ulp
        # differences between builds move each proposed step by ~1e-9 relative
        return _pair(J, 0.5, tol=1e-6), (0.5e-9, 2e-13, 2e-13)
    base = "bilayer_iec" if name.startswith("bilayer") else "trilayer"
    j, _ = getattr(parity_cases, base)(J)
    if name != "bilayer_rejects":  # controllers that keep the layers' steps within reach of each other
        for l in j.layers:
            p = AdaptiveIntegrationParams()
            p.abs_tol = p.rel_tol = 1e-6 if base == "bilayer_iec" else 1e-8
            p.safety_factor, p.max_factor = 0.3, 1.2
            l.setAdaptiveParams(p)
    return j, (1e-9, 1e-14, 1e-12)


def _oracle_sensitivity(name) -> float:
    """Relative response of the oracle's own log to a 1e-14 change of the
    last layer's initial m_y: what rounding differences between two correct
    builds can grow into.  Coupled layers with frozen neighbours and their
    own adaptive steps depend at first order on the step sequence, so this is
    far above 1e-14 (1.8e-8 for the bilayer, 5e-6 for the trilayer over
    1 ns)."""
    from cmtj_amd import CVector

    out = []
    for eps in (0.0, 1e-14):
        j, args = _ml_case(name, OracleJunction)
        l = j.layers[-1]
        l.mag = CVector(l.mag.x, l.mag.y + eps, l.mag.z)
        try:
            j.runSimulation(*args, solverMode=SolverMode.DormandPrince)
        except RuntimeError:
            pass
        out.append(j)
    return _compare(out[1], out[0])


@pytest.mark.parametrize("name", ["shared_step", "bilayer_iec", "trilayer", "bilayer_rejects"])
def test_multilayer_dp_matches_oracle(name):
    """Two and three coupled layers: the same accepted attempts and entry
    times as the oracle, and magnetisations and resistance within 1e-8, or,
    where the oracle's own trajectory is that sensitive, within 100 times
    its response to a 1e-14 change of the initial state; a run whose layer
    rejects a step ends with the same error and the same log up to it."""
    out = []
    for J in (Junction, OracleJunction):
        j, args = _ml_case(name, J)
        err = None
        try:
            j.runSimulation(*args, solverMode=SolverMode.DormandPrince)
        except RuntimeError as e:
            err = str(e)
        out.append((j, err))
    (g, eg), (o, eo) = out
    assert eg == eo
    worst = _compare(g, o)
    sens = _oracle_sensitivity(name)
    bar = max(1e-8, 100 * sens)
    print(f"[dp] {name}: {len(g.getLog()['time'])} entries ({'rejected' if eg else 'finished'}), "
          f"max relative difference {worst:.3e}, oracle sensitivity {sens:.3e}, bar {bar:.1e}", flush=True)
    assert worst <= bar
    assert (eg is not None) == (name == "bilayer_rejects")
    if name != "bilayer_rejects":
        assert len(g.getLog()["time"]) > (20 if name == "shared_step" else 500)
    for lid in g.getLayerIds():
        np.testing.assert_allclose(np.array(g.getLayerMagnetisation(lid).tolist()),
                                   np.array(o.getLayerMagnetisation(lid).tolist()), atol=bar)


def test_work_queue_equals_one_lane_per_instance(monkeypatch):
    """k_dp_q on a persistent grid of 64 lanes over 384 instances of very
    different lengths (every lane takes several instances from the queue)
    gives k_dp's results bit for bit."""
    monkeypatch.setenv("CMTJ_QUEUE_LANES", "64")  # read once, at the first queue launch of the process
    js, T = [], []
    for q in range(384):
        name = ("sot_switch", "sine", "equilibrium")[q % 3]
        j, args = _case(name, Junction)
        js.append(j)
        T.append(args[0] * (0.2 + 0.8 * ((q * 37) % 101) / 100.0))
    mk = lambda lanes: JunctionBatch.from_junctions(js, T, 1e-13, 1e-12, SolverMode.DormandPrince,  # noqa: E731
                                                     precision="fp64", log="full", log_dtype="f64").set_lanes(lanes)
    a, b = mk("auto"), mk("queue")
    assert "k_dp_q" in b.plan().info.kernel_name.decode()
    ra, rb = a.run(), b.run()
    sa, sb = ra.status(), rb.status()
    print(f"[dp] queue: {len(js)} instances, accepted steps {int(sa['steps_done'].min())}.."
          f"{int(sa['steps_done'].max())}", flush=True)
    assert sa.tobytes() == sb.tobytes()
    np.testing.assert_array_equal(ra.final_m.cpu().numpy(), rb.final_m.cpu().numpy())
    for i in range(len(js)):  # the written entries (adaptive runs fill their capacity in part)
        la, lb = ra.instance_log(i), rb.instance_log(i)
        for k in la:
            np.testing.assert_array_equal(la[k], lb[k])


# ---- the PID step controller (core/junction.hpp:953-961) ------------------------

PID = dict(kp=-0.15, ki=0.0, kd=0.01)
# an integral term makes the trajectory far more sensitive to rounding
# (ki > 0 shrinks the step until |m5 - m4| is rounding noise): the oracle's own
# response to a 1e-14 change of m0 sets that case's bar
PID_I = dict(kp=-0.15, ki=-0.0003, kd=0.01)


def _with_pid(j, tol=None, **gains):
    for l in j.layers:
        p = AdaptiveIntegrationParams()
        if tol is not None:
            p.abs_tol = p.rel_tol = tol
        p.use_pid_control = True
        for k, v in gains.items():
            setattr(p, k, v)
        l.setAdaptiveParams(p)
    return j


def _pid_state(j):
    return np.array([[l.adaptiveParams.integral_error, l.adaptiveParams.prev_error_ratio] for l in j.layers])


def _state_difference(a, b) -> float:
    """Largest relative difference of two controller states; an infinite
    entry (log of a zero error ratio) must be the same infinity."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    assert np.array_equal(np.isinf(a), np.isinf(b)) and np.array_equal(a[np.isinf(a)], b[np.isinf(b)])
    f = np.isfinite(a)
    return float(np.max(np.abs(a[f] - b[f]) / np.maximum(np.abs(b[f]), 1e-300))) if f.any() else 0.0


# The integral sums log(error ratio) over every attempt, including attempts
# whose error estimate |m5 - m4| is near the rounding level of m (a step
# clipped to the end of the run, a step just after a reset): there one
# rounding of m4 moves log(ratio) by up to ~1e-2, without changing the
# trajectory (the factor is clamped).  So the state is compared against 100
# times the oracle's own response to a 1e-14 change of the initial m (2e-7 to
# 2e-4 relative in the cases below), as the logs are.


def _pid_runs(name, gains, J, eps=0.0):
    j, args = _case(name, J)
    _with_pid(j, **gains)
    l = j.layers[0]
    l.mag = CVector(l.mag.x, l.mag.y + eps, l.mag.z)
    if not gains:
        args = (3e-12, 1e-13, 1e-13)
    states = []
    for _ in range(2):
        j.runSimulation(*args, solverMode=SolverMode.DormandPrince)
        states.append(_pid_state(j))
    return j, states


@pytest.mark.parametrize("name,gains", [("sot_switch", PID), ("sine", PID), ("sot_switch", PID_I), ("sine", {})])
def test_pid_dp_matches_oracle(name, gains):
    """One layer (k_dp) with the PID controller, twice in a row: the same
    entries and magnetisations as the oracle within 1e-8 (or 100 times the
    oracle's response to a 1e-14 change of the initial m), and the same
    controller state after each run (the second run starts from it).  The
    reference's default gains (kp = +0.2) shrink the step after every
    accepted attempt down to nothing and reset it, so that case runs 3 ps."""
    (g, sg), (o, so) = _pid_runs(name, gains, Junction), _pid_runs(name, gains, OracleJunction)
    p, sp = _pid_runs(name, gains, OracleJunction, 1e-14)
    bar = max(1e-8, 100 * _compare(p, o))
    worst = _compare(g, o)
    print(f"[dp] pid {name} {gains or 'default gains'}: {len(g.getLog()['time'])} entries, max relative "
          f"difference {worst:.3e}, bar {bar:.1e}, state {sg[-1].ravel()}", flush=True)
    assert worst <= bar
    assert len(g.getLog()["time"]) > 20
    for a, b, c in zip(sg, so, sp):
        d, sbar = _state_difference(a, b), max(1e-9, 100 * _state_difference(c, b))
        print(f"[dp]   state {a.ravel()} vs oracle {b.ravel()}: {d:.2e} (bar {sbar:.1e})", flush=True)
        assert d <= sbar
    assert not np.array_equal(sg[0], sg[1])


PAIR_PID = dict(kp=-0.15, ki=0.0, kd=0.01)


def _pid_pair(J, eps=0.0):
    """The torque-free layer first, both layers on the PID controller: the
    torque-free layer's error ratio is 0, log 0 = -inf makes its factor
    infinite or undefined, so it passes min_factor = 1/2 of the step on (its
    state goes to -inf); the precessing layer's controller sets the rest.
    (With ki > 0 the step here shrinks until |m5 - m4| is rounding noise.)"""
    import copy

    from test_dp_cpu import _pair

    j = _pair(J, 0.5, tol=1e-6)
    for l in j.layers:
        p = copy.copy(l.adaptiveParams)
        p.use_pid_control = True
        for k, v in PAIR_PID.items():
            setattr(p, k, v)
        l.setAdaptiveParams(p)
    l = j.getLayer("free")
    l.mag = CVector(l.mag.x, l.mag.y + eps, l.mag.z)
    j.runSimulation(1e-9, 2e-13, 2e-13, solverMode=SolverMode.DormandPrince)
    return j


def test_pid_multilayer_dp_matches_oracle():
    """Two layers (k_dp_ml), each with its own PID state, within 1e-8 or 100
    times the oracle's response to a 1e-14 change of the initial state."""
    g, o, s = _pid_pair(Junction), _pid_pair(OracleJunction), _pid_pair(OracleJunction, 1e-14)
    bar = max(1e-8, 100 * _compare(s, o))
    worst = _compare(g, o)
    sg, so = _pid_state(g), _pid_state(o)
    print(f"[dp] pid pair: {len(g.getLog()['time'])} entries, max relative difference {worst:.3e}, bar {bar:.1e}, "
          f"state {sg.ravel()}", flush=True)
    assert worst <= bar
    assert len(g.getLog()["time"]) > 50
    assert np.all(np.isneginf(sg[0])) and np.all(np.isneginf(so[0]))
    assert _state_difference(sg, so) <= max(1e-9, 100 * _state_difference(_pid_state(s), so))


def test_pid_batch_and_work_queue(monkeypatch):
    """A batch of PID instances: the PID state after the magnetisations in
    each instance's final-m block, against the oracle; the work-queue kernel
    gives k_dp's statuses (the same accepted steps and entries) and its
    magnetisations and logs within 1e-10.  (With the standard controller
    the two are bit-equal; with PID they were measured up to 5e-15 apart in
    m on 42 of 288 components: the two kernels' builds round some step of
    the controller apart, which this test does not isolate.)"""
    monkeypatch.setenv("CMTJ_QUEUE_LANES", "64")
    js, T, gains = [], [], []
    for q in range(96):
        name = ("sot_switch", "sine")[q % 2]
        j, args = _case(name, Junction)
        gains.append(dict(PID, kp=PID["kp"] * (1 + 0.2 * (q % 5) / 4)))
        _with_pid(j, **gains[-1])
        js.append(j)
        T.append(args[0] * (0.3 + 0.7 * ((q * 37) % 101) / 100.0))
    mk = lambda lanes: JunctionBatch.from_junctions(js, T, 1e-13, 1e-12, SolverMode.DormandPrince,  # noqa: E731
                                                     precision="fp64", log="full", log_dtype="f64").set_lanes(lanes)
    a, b = mk("auto"), mk("queue")
    assert "k_dp_q" in b.plan().info.kernel_name.decode()
    ra, rb = a.run(), b.run()
    fa, fb = ra.final_m.cpu().numpy(), rb.final_m.cpu().numpy()
    assert ra.status().tobytes() == rb.status().tobytes()
    np.testing.assert_allclose(fa[:, 0], fb[:, 0], rtol=0, atol=1e-10)  # m of the one layer
    for i in range(len(js)):
        la, lb = ra.instance_log(i), rb.instance_log(i)
        for k in la:
            np.testing.assert_allclose(la[k], lb[k], rtol=1e-10, atol=1e-10 if k != "time" else 0)
    worst = 0.0
    for q in range(0, len(js), 7):
        name = ("sot_switch", "sine")[q % 2]
        runs = []
        for eps in (0.0, 1e-14):
            o, _ = _case(name, OracleJunction)
            _with_pid(o, **gains[q])
            l = o.layers[0]
            l.mag = CVector(l.mag.x, l.mag.y + eps, l.mag.z)
            o.runSimulation(T[q], 1e-13, 1e-12, solverMode=SolverMode.DormandPrince)
            runs.append(o)
        o, p = runs
        so = _pid_state(o)[0]
        sbar = max(1e-9, 100 * _state_difference(_pid_state(p)[0], so))
        for f in (fa, fb):
            d = _state_difference(f[q].reshape(-1)[3:5], so)
            worst = max(worst, d / sbar)
            assert d <= sbar, (q, d, sbar)
        ol, gl = o.getLog(), ra.instance_log(q)
        assert len(ol["time"]) == len(gl["time"])
        lbar = max(1e-8, 100 * _compare(p, o))
        for k in ol:
            x, y = np.asarray(gl[k]), np.asarray(ol[k])
            d = float(np.max(np.abs(x - y))) / max(float(np.max(np.abs(y))), 1e-300) if y.size else 0.0
            worst = max(worst, d / lbar)
            assert d <= lbar, (q, k, d, lbar)
    print(f"[dp] pid batch: {len(js)} instances, worst difference {worst:.2f} of its bar", flush=True)
