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
