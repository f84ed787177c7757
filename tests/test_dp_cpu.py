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
    j = _film(OracleJunction, (0, 0, 1), (0, 0, 1e5))
    p = AdaptiveIntegrationParams()
    p.use_pid_control = True
    j.getLayer("free").setAdaptiveParams(p)
    with pytest.raises(NotImplementedError, match="PID"):
        j.runSimulation(1e-11, 1e-13, 1e-12, solverMode=SolverMode.DormandPrince)
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
