"""Dormand-Prince (adaptive) runs through the oracle, pinned by what does not
come from the oracle: one step against scipy's Dormand-Prince 5(4) tableau
(scipy.integrate.RK45: A, B, C, E) on a textbook LLG right-hand side, and a
whole adaptive run against the closed-form damped precession."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.integrate import RK45

from cmtj_amd import AdaptiveIntegrationParams, AxialDriver, CVector, Junction, Layer, SolverMode, constantDriver
from oracle_backend import OracleJunction

GAMMA, MU0 = 220880.0, 12.566e-7


def _film(J, m0, H, K=0.0, demag_z=0.0, alpha=0.02, Ms=1.0, tol=None):
    demag = [CVector(0, 0, 0), CVector(0, 0, 0), CVector(0, 0, demag_z)]
    l = Layer("free", mag=CVector(*m0), anis=CVector(1, 0, 0), Ms=Ms, thickness=1e-9, cellSurface=1e-16,
              demagTensor=demag, damping=alpha)
    if tol is not None:
        p = AdaptiveIntegrationParams()
        p.abs_tol, p.rel_tol = tol
        l.setAdaptiveParams(p)
    j = J([l])
    j.setLayerExternalFieldDriver("all", AxialDriver(*[constantDriver(h) for h in H]))
    if K:
        j.setLayerAnisotropyDriver("free", constantDriver(K))
    return j


def _rhs(m, H, K, demag_z, alpha, Ms):
    """Textbook LLG: dm/dt = -gamma/(1+alpha^2) (m x Heff + alpha m x (m x Heff)),
    Heff = Hext + 2K (a.m)/Ms a - N m Ms/mu0 with a = x."""
    a = np.array([1.0, 0.0, 0.0])
    Heff = np.asarray(H, float) + 2 * K * (a @ m) / Ms * a - np.array([0, 0, demag_z * m[2]]) * (Ms / MU0)
    p = np.cross(m, Heff)
    return -GAMMA / (1 + alpha ** 2) * (p + alpha * np.cross(m, p))


def test_one_step_matches_scipy_tableau():
    m0 = np.array([0.3, 0.4, math.sqrt(1 - 0.25)])
    H, K, dz, alpha, Ms = (1e5, 2e5, 3e5), 4e4, 1.0, 0.02, 1.2
    h = 2e-13
    # tolerances so loose that the first attempt is accepted; the run ends
    # after one step because the clock then moves past totalTime
    j = _film(OracleJunction, m0, H, K=K, demag_z=dz, alpha=alpha, Ms=Ms, tol=(1.0, 1.0))
    j.runSimulation(h, h, h, solverMode=SolverMode.DormandPrince)
    log = j.getLog()
    f = lambda y: _rhs(y, H, K, dz, alpha, Ms)
    Ks = np.zeros((7, 3))
    Ks[0] = f(m0)
    for s in range(1, 6):
        Ks[s] = f(m0 + h * (RK45.A[s, :s] @ Ks[:s]))
    y5 = m0 + h * (RK45.B @ Ks[:6])
    Ks[6] = f(y5)
    err = np.linalg.norm(h * (Ks.T @ RK45.E))
    tol = 1.0 + 1.0 * np.linalg.norm(m0)
    factor = min(max(0.9 * (1.0 / (err / tol)) ** 0.2, 0.1), 5.0)
    want = y5 / np.linalg.norm(y5)
    got = np.array([log["free_mx"][-1], log["free_my"][-1], log["free_mz"][-1]])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)
    # the clock advances by the proposed next step (h factor)
    assert len(log["time"]) == 1
    np.testing.assert_allclose(log["time"][0], h * factor, rtol=1e-12)


def _off_trajectory(m, theta0, H, alpha):
    """Distance (rad) of each m from the exact damped precession about H z:
    tan(theta/2) = tan(theta0/2) exp(-alpha g H t), phi = g H t with
    g = gamma/(1+alpha^2).  The time an m belongs to is read from its theta,
    since the adaptive run's log labels are not that time: the reference
    advances its clock by the step proposed for the next step, not the step
    taken (junction.hpp:997 then :2029)."""
    g = GAMMA / (1 + alpha ** 2)
    th = np.arccos(np.clip(m[:, 2], -1, 1))
    t_true = np.log(np.tan(theta0 / 2) / np.tan(th / 2)) / (alpha * g * H)
    ph = np.arctan2(m[:, 1], m[:, 0])
    d = np.angle(np.exp(1j * (ph - g * H * t_true)))
    return np.abs(d), t_true


def test_adaptive_run_follows_damped_precession():
    theta0, H, alpha = 1.2, 1e5, 0.02
    j = _film(OracleJunction, (math.sin(theta0), 0.0, math.cos(theta0)), (0, 0, H), alpha=alpha,
              tol=(1e-10, 1e-10))
    j.runSimulation(0.5e-9, 1e-13, 1e-12, solverMode=SolverMode.DormandPrince)
    log = j.getLog()
    t = np.asarray(log["time"])
    m = np.stack([log["free_mx"], log["free_my"], log["free_mz"]], axis=1)
    off, t_true = _off_trajectory(m, theta0, H, alpha)
    assert float(off.max()) < 1e-6, float(off.max())
    np.testing.assert_allclose(np.linalg.norm(m, axis=1), 1.0, atol=1e-14)
    # the labels run ahead of the integrated time, never behind
    assert np.all(t >= t_true - 1e-18) and t_true[-1] > 0.4e-9
    # one entry per write interval crossed, times increasing
    k = np.floor(t / 1e-12)
    assert np.all(np.diff(t) > 0) and np.all(np.diff(k) >= 1)


def test_tighter_tolerance_stays_closer_to_the_exact_trajectory():
    off = {}
    for tol in (1e-5, 1e-10):
        j = _film(OracleJunction, (0.6, 0.0, 0.8), (0, 0, 1e5), tol=(tol, tol))
        j.runSimulation(0.2e-9, 1e-13, 1e-11, solverMode=SolverMode.DormandPrince)
        lg = j.getLog()
        m = np.stack([lg[f"free_m{c}"] for c in "xyz"], 1)
        off[tol] = float(_off_trajectory(m, math.acos(0.8), 1e5, 0.02)[0].max())
    assert off[1e-10] < off[1e-5] and off[1e-10] < 1e-6, off


def test_dp_restrictions():
    five = Junction([Layer(f"l{i}", mag=CVector(0, 0, 1), anis=CVector(1, 0, 0), Ms=1.0, thickness=1e-9,
                           cellSurface=1e-16, demagTensor=[CVector(0, 0, 0)] * 3) for i in range(5)])
    with pytest.raises(NotImplementedError, match="1 to 4 layers"):
        five.runSimulation(1e-11, 1e-13, 1e-12, solverMode=SolverMode.DormandPrince)
    with pytest.raises(NotImplementedError, match="energy"):
        _film(OracleJunction, (0, 0, 1), (0, 0, 1e5)).runSimulation(1e-11, 1e-13, 1e-12, calculateEnergies=True,
                                                                      solverMode=SolverMode.DormandPrince)


# ---- several layers (runMultiLayerSolver, core/junction.hpp:1682-1707) -------
#
# A layer held exactly along its field has zero torque, so its error ratio is
# 0, the standard controller's factor is infinite and falls back to
# min_factor (:967-969): that layer multiplies the shared step by min_factor
# on every attempt.  Put it first with min_factor c = 2^-1 or 2^2: the
# precessing second layer then steps with c H_n and proposes H_{n+1}; its own
# steps s_n = c H_n follow s_{n+1} = c s_n factor(s_n), which is a one-layer
# run with safety, min_factor and max_factor scaled by c and the first step
# c h0, while the two-layer clock adds H_{n+1} = s_{n+1} / c.  Scaling by a
# power of two is exact, so up to the first rejection the two-layer run is
# that one-layer run bit for bit, with its clock, total time and write
# interval divided by c.


def _pair(J, c, order=("fixed", "free"), tol=1e-8, fixed_min=None):
    fixed = Layer("fixed", mag=CVector(0, 0, 1), anis=CVector(1, 0, 0), Ms=1.2, thickness=2e-9, cellSurface=1e-16,
                  demagTensor=[CVector(0, 0, 0)] * 3, damping=0.02)
    free = Layer("free", mag=CVector(0.6, 0.0, 0.8), anis=CVector(1, 0, 0), Ms=1.0, thickness=1e-9,
                 cellSurface=1e-16, demagTensor=[CVector(0, 0, 0)] * 3, damping=0.02)
    pf = AdaptiveIntegrationParams()
    pf.min_factor = c if fixed_min is None else fixed_min
    pf.max_factor = max(5.0, c)
    fixed.setAdaptiveParams(pf)
    pv = AdaptiveIntegrationParams()
    pv.abs_tol = pv.rel_tol = tol
    free.setAdaptiveParams(pv)
    by = {"fixed": fixed, "free": free}
    j = J([by[order[0]], by[order[1]]])
    j.setLayerExternalFieldDriver("fixed", AxialDriver(0, 0, 2e5))
    j.setLayerExternalFieldDriver("free", AxialDriver(3e4, 0, 1e5))
    return j


def _single_equivalent(J, c, tol=1e-8):
    j = _film(J, (0.6, 0.0, 0.8), (3e4, 0, 1e5), tol=(tol, tol))
    p = j.getLayer("free").adaptiveParams
    p.safety_factor *= c
    p.min_factor *= c
    p.max_factor *= c
    j.getLayer("free").setAdaptiveParams(p)
    return j


def _two_vs_one(J2, c, T=0.5e-9, h0=2e-13, wf=2e-13, **kw):
    """(two-layer log, one-layer log, status message or None) of the pair."""
    two = _pair(J2, c, **kw)
    err = None
    try:
        two.runSimulation(T, h0, wf, solverMode=SolverMode.DormandPrince)
    except RuntimeError as e:
        err = str(e)
    one = _single_equivalent(OracleJunction, c)
    try:  # with c > 1 the one-layer controller grows the step after a rejection and never recovers
        one.runSimulation(T * c, h0 * c, wf * c, solverMode=SolverMode.DormandPrince)
    except RuntimeError:
        pass
    return two.getLog(), one.getLog(), err


def _equivalent(l2, l1, c) -> bool:
    n = len(l2["time"])
    if n == 0 or len(l1["time"]) < n:
        return False
    same_m = all(np.array_equal(np.asarray(l2[f"free_m{x}"]), np.asarray(l1[f"free_m{x}"])[:n]) for x in "xyz")
    same_t = np.array_equal(np.asarray(l2["time"]) * c, np.asarray(l1["time"])[:n])
    fixed = np.all(np.asarray(l2["fixed_mz"]) == 1.0)
    return bool(same_m and same_t and fixed)


@pytest.mark.parametrize("c", [0.5, 0.25])
def test_layers_share_one_step_size_in_order(c):
    l2, l1, err = _two_vs_one(OracleJunction, c)
    assert err is None, err
    assert len(l2["time"]) > 40 and len(l2["time"]) == len(l1["time"])
    assert _equivalent(l2, l1, c)
    # planted mistakes in the descriptor: the first layer passes the step on
    # unchanged, or the precessing layer steps first
    l2b, l1b, _ = _two_vs_one(OracleJunction, c, fixed_min=1.0)
    assert not _equivalent(l2b, l1b, c)
    l2c, l1c, _ = _two_vs_one(OracleJunction, c, order=("free", "fixed"))
    assert not _equivalent(l2c, l1c, c)


def test_a_rejected_layer_step_stops_the_run():
    # the first layer doubles the step every attempt until the second rejects
    l2, l1, err = _two_vs_one(OracleJunction, 2.0, T=1e-11, tol=1e-10, h0=1e-14, wf=1e-14)
    assert err is not None and "hopt = -1" in err
    n = len(l2["time"])
    print(f"{n} entries before the rejection")
    assert 2 < n <= len(l1["time"])
    assert _equivalent(l2, l1, 2.0)  # identical up to the rejection


# ---- the PID step controller (dormand_prince_step, core/junction.hpp:953-961)
#
# factor = exp(kp e + ki I + kd (e - prev)) with e = log(error ratio),
# I = integral + e, then prev = e; the state starts from the layer's
# integral_error and prev_error_ratio and stays in the layer after the run.
# A run whose first proposed step already carries the clock past totalTime
# takes exactly one attempt, and its one log entry is labelled with that
# proposal (the clock advances by the proposed step, :997 then :2029): the
# first attempt's error ratio r1 is the same under both controllers (same
# step h0 from the same m), so a standard-controller run tells r1 through
# h0 safety r1^(-1/5), and the PID run's entry, final integral and final
# prev then follow in closed form.


def _pid_film(params, field=(2e4, 0, 1e5)):
    j = _film(OracleJunction, (0.6, 0, 0.8), field, alpha=0.02)
    j.getLayer("free").setAdaptiveParams(params)
    return j


def _params(**kw):
    p = AdaptiveIntegrationParams()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


H0 = 2.0 ** -43  # 1.14e-13 s; sums of it are exact


def _first_error_log():
    """log(r1) of the first attempt, from a standard-controller run with the
    factor unclamped."""
    j = _pid_film(_params(max_factor=1e9))
    j.runSimulation(1.0001 * H0, H0, H0, solverMode=SolverMode.DormandPrince)
    t1 = j.getLog()["time"]
    assert len(t1) == 1 and t1[0] > 1.0001 * H0
    return -5.0 * math.log(t1[0] / (0.9 * H0))


@pytest.mark.parametrize("kp,ki,kd,I0,P0", [
    (-0.03, 0.0, 0.0, 0.0, 1.0),
    (-0.02, -0.01, 0.0, -3.0, 1.0),
    (-0.02, -0.01, 0.015, -3.0, -10.0),
    (0.0, 0.0, -0.02, 0.0, 1.0),  # the reference's initial prev_error_ratio 1.0 enters through kd
])
def test_pid_first_step_closed_form(kp, ki, kd, I0, P0):
    e1 = _first_error_log()
    assert e1 < -1  # an accepted first attempt
    x = kp * e1 + ki * (I0 + e1) + kd * (e1 - P0)
    assert math.log(1.0001) < x < math.log(5.0)  # unclamped, and one attempt ends the run
    p = _params(use_pid_control=True, kp=kp, ki=ki, kd=kd, integral_error=I0, prev_error_ratio=P0)
    j = _pid_film(p)
    j.runSimulation(1.0001 * H0, H0, H0, solverMode=SolverMode.DormandPrince)
    t = j.getLog()["time"]
    assert len(t) == 1
    np.testing.assert_allclose(t[0], H0 * math.exp(x), rtol=1e-11)
    ap = j.getLayer("free").adaptiveParams
    np.testing.assert_allclose(ap.integral_error, I0 + e1, rtol=1e-11)
    np.testing.assert_allclose(ap.prev_error_ratio, e1, rtol=1e-11)
    # the caller's parameters are the layer's own copy: untouched
    assert p.integral_error == I0 and p.prev_error_ratio == P0


def test_pid_integral_sums_every_attempt():
    """kp only: every attempt's factor exp(kp e_n) is the ratio of the step
    it proposes (the clock increment) to the step it took, so the log gives
    each e_n back; the layer's final integral is I0 + sum e_n and its prev
    the last e_n.  (kp = -1/5 would be the standard controller with safety 1,
    whose error ratios settle at the acceptance threshold.)"""
    kp, I0 = -0.01, 2.5
    j = _pid_film(_params(use_pid_control=True, kp=kp, ki=0.0, kd=0.0, integral_error=I0, max_factor=10.0))
    T = 60 * H0
    j.runSimulation(T, H0, H0, solverMode=SolverMode.DormandPrince)
    t = np.asarray(j.getLog()["time"])
    p = np.diff(np.concatenate([[0.0], t]))          # proposals = clock increments
    used = np.minimum(np.concatenate([[H0], p[:-1]]), T - np.concatenate([[0.0], t[:-1]]))
    e = np.log(p / used) / kp
    assert len(t) > 10 and np.all(e < 0)              # every attempt accepted, one entry each
    assert np.all(p / used < 10.0) and np.all(p / used > 0.1)  # no clamping
    ap = j.getLayer("free").adaptiveParams
    np.testing.assert_allclose(ap.integral_error, I0 + e.sum(), rtol=1e-9)
    np.testing.assert_allclose(ap.prev_error_ratio, e[-1], rtol=1e-9)


def test_pid_zero_gains_keep_the_first_step():
    """kp = ki = kd = 0: factor exp(0) = 1, every accepted step is h0, and the
    clock is the running sum of h0."""
    j = _pid_film(_params(use_pid_control=True, kp=0.0, ki=0.0, kd=0.0))
    n = 40
    j.runSimulation(n * H0, H0, H0, solverMode=SolverMode.DormandPrince)
    t = np.asarray(j.getLog()["time"])
    assert len(t) == n
    np.testing.assert_array_equal(t, H0 * np.arange(1, n + 1))
    ap = j.getLayer("free").adaptiveParams
    assert ap.prev_error_ratio < 0 and ap.integral_error < n * ap.prev_error_ratio * 0.5


def test_pid_state_carries_over_to_the_next_run():
    """Two runs in a row equal a second junction started from the first
    run's final m and controller state."""
    p = _params(use_pid_control=True, kp=-0.15, ki=-0.002, kd=0.01)
    a = _pid_film(p)
    a.runSimulation(0.1e-9, H0, 1e-11, solverMode=SolverMode.DormandPrince)
    ap = a.getLayer("free").adaptiveParams
    I1, P1 = ap.integral_error, ap.prev_error_ratio
    m1 = a.getLayer("free").mag
    n1 = len(a.getLog()["time"])
    a.runSimulation(0.1e-9, H0, 1e-11, solverMode=SolverMode.DormandPrince)
    b = _film(OracleJunction, (m1.x, m1.y, m1.z), (2e4, 0, 1e5), alpha=0.02)
    b.getLayer("free").setAdaptiveParams(_params(use_pid_control=True, kp=-0.15, ki=-0.002, kd=0.01,
                                                 integral_error=I1, prev_error_ratio=P1))
    b.runSimulation(0.1e-9, H0, 1e-11, solverMode=SolverMode.DormandPrince)
    la, lb = a.getLog(), b.getLog()
    for k in ("free_mx", "free_my", "free_mz"):
        np.testing.assert_array_equal(np.asarray(la[k])[n1:], lb[k])
    np.testing.assert_array_equal(np.asarray(la["time"])[n1:], np.asarray(lb["time"]) + 0.0)
    assert a.getLayer("free").adaptiveParams.integral_error == b.getLayer("free").adaptiveParams.integral_error
    # a plain second run from the default state differs
    c = _film(OracleJunction, (m1.x, m1.y, m1.z), (2e4, 0, 1e5), alpha=0.02)
    c.getLayer("free").setAdaptiveParams(_params(use_pid_control=True, kp=-0.15, ki=-0.002, kd=0.01))
    c.runSimulation(0.1e-9, H0, 1e-11, solverMode=SolverMode.DormandPrince)
    assert not np.array_equal(c.getLog()["time"], lb["time"])
