"""Coupled stacks on the GPU (cmtj_run_stacks) against the oracle's
stack runner: series / parallel, KCL / non-KCL, one- and two-layer members,
RK4 and the thermal Euler-Heun path, and a NaN that stops the stack."""
from __future__ import annotations

import numpy as np
import pytest

from cmtj_amd import AxialDriver, ScalarDriver, _lib, constantDriver, stepDriver
from cmtj_amd.stack import ParallelStack, SeriesStack
from oracle_backend import OracleParallelStack, OracleSeriesStack
from test_stack_cpu import _single, _two

pytestmark = pytest.mark.gpu


def _members(member, n=3):
    mags = [(0, 0, 1), (0.4, 0.1, 0.9), (0.1, 0.5, 0.8), (0.7, 0.1, 0.3)]
    return [member(stt=True, mag=mags[q % len(mags)]) for q in range(n)]


def _compare(a, b, tol=1e-8):
    la, lb = a.getLog(), b.getLog()
    assert la.keys() == lb.keys()
    worst = 0.0
    for k in la:
        x, y = np.asarray(la[k]), np.asarray(lb[k])
        assert x.shape == y.shape, k
        scale = max(1.0, float(np.max(np.abs(y)))) if y.size else 1.0
        worst = max(worst, float(np.max(np.abs(x - y))) / scale if y.size else 0.0)
    for q in range(a.stackSize):
        ja, jb = a.getLog(q), b.getLog(q)
        assert ja.keys() == jb.keys()
        for k in ja:
            x, y = np.asarray(ja[k]), np.asarray(jb[k])
            scale = max(1.0, float(np.max(np.abs(y)))) if y.size else 1.0
            worst = max(worst, float(np.max(np.abs(x - y))) / scale if y.size else 0.0)
    return worst


@pytest.mark.parametrize("member", [_single, _two])
@pytest.mark.parametrize("kcl", [True, False])
@pytest.mark.parametrize("series", [True, False])
def test_stack_matches_oracle(member, kcl, series):
    js = _members(member)
    drv = ScalarDriver.getSineDriver(8e11, 5e11, 3e9, 0.4)
    gcls, ocls = (SeriesStack, OracleSeriesStack) if series else (ParallelStack, OracleParallelStack)
    stacks = [cls(js, useKCL=kcl, phaseOffset=0.3) for cls in (gcls, ocls)]
    for s in stacks:
        s.setCouplingStrength([0.25, -0.4])
        s.setCoupledCurrentDriver(drv)
        s.setExternalFieldDriver(AxialDriver(constantDriver(1e4), constantDriver(0), constantDriver(2e4)))
        s.runSimulation(0.5e-9, 1e-13, 1e-12)
    kern = _lib.last_kernels()  # the GPU stack ran first
    worst = _compare(*stacks)
    print(f"[stack] {member.__name__} kcl={kcl} series={series}: {kern}, max relative difference {worst:.3e}",
          flush=True)
    assert worst <= 1e-8
    assert len(kern) == 1 and kern[0].startswith("k_junction_cl<double")  # one lane per component


def test_stochastic_stack_matches_oracle_pathwise():
    """Members at 300 K step with Euler-Heun; the engine and the oracle share
    the Philox stream, so the paths agree."""
    js = _members(_single, n=4)
    for j in js:
        j.setLayerTemperatureDriver("all", constantDriver(300.0))
        j.setNoiseSeed(7)
    stacks = [cls(js) for cls in (SeriesStack, OracleSeriesStack)]
    for s in stacks:
        s.setCouplingStrength(0.3)
        s.setCoupledCurrentDriver(constantDriver(1e12))
        s.runSimulation(0.2e-9, 1e-13, 1e-12)
    worst = _compare(*stacks)
    print(f"[stack] Euler-Heun 300 K: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8


def test_nan_member_stops_the_stack():
    js = _members(_single, n=3)
    for q, j in enumerate(js):  # one structure for every member; member 1's step is infinite
        amp = float("inf") if q == 1 else 1e3
        j.setLayerExternalFieldDriver("all", AxialDriver(stepDriver(0, amp, 5e-12, 1e-9), constantDriver(0),
                                                         constantDriver(1e5)))
    g, o = SeriesStack(js), OracleSeriesStack(js)
    for s in (g, o):
        s.setCoupledCurrentDriver(constantDriver(1e12))
        with pytest.raises(RuntimeError, match="NAN magnetisation"):
            s.runSimulation(1e-10, 1e-13, 1e-13)
    assert 0 < len(g.getLog()["time"]) < 1000
    assert len(g.getLog()["time"]) == len(o.getLog()["time"])
    for q in range(3):
        assert len(g.getLog(q)["time"]) == len(g.getLog()["time"])


def test_batched_stacks_match_oracle():
    """Many 3-member stacks (4 lanes each, 16 per wave) in one launch: every
    stack as the oracle runs it alone; a NaN in one stack stops that stack
    only."""
    from cmtj_amd.stack import run_stacks

    def build(k, cls_s, cls_p, bad=False):
        js = _members(_single, n=3)
        for q, j in enumerate(js):
            amp = float("inf") if (bad and q == 1) else 1e3 * (k % 5)
            j.setLayerExternalFieldDriver("all", AxialDriver(stepDriver(0, amp, 5e-12, 1e-9), constantDriver(0),
                                                             constantDriver(1e5)))
        s = (cls_s if k % 2 else cls_p)(js, useKCL=bool(k % 3), phaseOffset=0.1 * k)
        s.setCouplingStrength([0.1 + 0.01 * k, -0.2])
        s.setCoupledCurrentDriver(ScalarDriver.getSineDriver(5e11, 2e11 * (k % 4), 2e9, 0.0))
        return s

    K = 40
    g = [build(k, SeriesStack, ParallelStack, bad=(k == 7)) for k in range(K)]
    o = [build(k, OracleSeriesStack, OracleParallelStack, bad=(k == 7)) for k in range(K)]
    for a, b in zip(g, o):
        for x, y in zip(a.junctionList, b.junctionList):
            y._seed = x._seed
    with pytest.raises(RuntimeError, match="NAN magnetisation"):
        run_stacks(g, 0.3e-9, 1e-13, 1e-12)
    worst = 0.0
    for k, b in enumerate(o):
        if k == 7:
            with pytest.raises(RuntimeError, match="NAN magnetisation"):
                b.runSimulation(0.3e-9, 1e-13, 1e-12)
        else:
            b.runSimulation(0.3e-9, 1e-13, 1e-12)
        worst = max(worst, _compare(g[k], b))
    print(f"[stack] {K} stacks in one launch: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8
    assert 0 < len(g[7].getLog()["time"]) < len(g[6].getLog()["time"])


@pytest.mark.parametrize("series", [True, False])
@pytest.mark.parametrize("n", [3, 5])
def test_non_delayed_stack_matches_oracle(series, n):
    """setDelayed(False): the members step one after another within a step,
    each coupled to its predecessor's new state (the wave runs one sub-step
    per member)."""
    js = _members(_two, n=n)
    drv = ScalarDriver.getSineDriver(8e11, 5e11, 3e9, 0.4)
    gcls, ocls = (SeriesStack, OracleSeriesStack) if series else (ParallelStack, OracleParallelStack)
    stacks = [cls(js, useKCL=False, phaseOffset=0.3) for cls in (gcls, ocls)]
    for s in stacks:
        s.setDelayed(False)
        s.setCouplingStrength([0.25, -0.4, 0.3, -0.2][: n - 1])
        s.setCoupledCurrentDriver(drv)
        s.setExternalFieldDriver(AxialDriver(constantDriver(1e4), constantDriver(0), constantDriver(2e4)))
        s.runSimulation(0.5e-9, 1e-13, 1e-12)
        if s is stacks[0]:
            kern = _lib.last_kernels()
            assert len(kern) == 1 and kern[0].startswith("k_junction<double")  # sub-steps: one lane per member
    worst = _compare(*stacks)
    print(f"[stack] non-delayed n={n} series={series}: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8


def test_mixed_delay_rules_in_one_launch():
    """Delayed and non-delayed stacks (and KCL ones) side by side in one wave:
    each as the oracle runs it alone."""
    from cmtj_amd.stack import run_stacks

    def build(k, cls_s, cls_p):
        js = _members(_two, n=3)
        for j in js:
            j.setLayerExternalFieldDriver("all", AxialDriver(constantDriver(1e3 * (k % 5)), constantDriver(0),
                                                             constantDriver(1e5)))
        s = (cls_s if k % 2 else cls_p)(js, useKCL=(k % 3 == 0), phaseOffset=0.1 * k)
        if k % 3 == 1:
            s.setDelayed(False)
        s.setCouplingStrength([0.1 + 0.01 * k, -0.2])
        s.setCoupledCurrentDriver(ScalarDriver.getSineDriver(5e11, 2e11 * (k % 4), 2e9, 0.0))
        return s

    K = 24
    g = [build(k, SeriesStack, ParallelStack) for k in range(K)]
    o = [build(k, OracleSeriesStack, OracleParallelStack) for k in range(K)]
    run_stacks(g, 0.3e-9, 1e-13, 1e-12)
    worst = 0.0
    for k, b in enumerate(o):
        b.runSimulation(0.3e-9, 1e-13, 1e-12)
        worst = max(worst, _compare(g[k], b))
    print(f"[stack] {K} stacks, three coupling rules, one launch: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8


@pytest.mark.parametrize("series", [True, False])
def test_callback_currents_and_drivers_match_oracle(series):
    """A KCL stack whose current is a Python callback (tabulated once per
    step) and whose members carry a callback anisotropy (tabulated at the
    RK4 stage times), against the oracle on the same tables."""
    from test_stack_cpu import _pulse

    def ramp(t):
        return 1.05e3 + 4e13 * t

    js = _members(_single, n=4)
    for j in js:
        j.setLayerAnisotropyDriver("free", ScalarDriver.getCustomDriver(ramp))
    gcls, ocls = (SeriesStack, OracleSeriesStack) if series else (ParallelStack, OracleParallelStack)
    stacks = [cls(js, useKCL=True) for cls in (gcls, ocls)]
    for s in stacks:
        s.setCouplingStrength([0.25, -0.4, 0.3])
        s.setCoupledCurrentDriver(ScalarDriver.getCustomDriver(_pulse))
        s.runSimulation(0.3e-9, 1e-13, 1e-12)
    worst = _compare(*stacks)
    print(f"[stack] callbacks series={series}: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8


def test_large_sweep_keeps_one_lane_per_member_and_matches_component_lanes():
    """A sweep larger than the chip's SIMDs at one wave each keeps one lane
    per member; the same stacks run few at a time take the component-lane
    kernel: both against the oracle."""
    from cmtj_amd.stack import run_stacks

    def build(k, cls_s):
        js = _members(_single, n=8)
        s = cls_s(js)
        s.setCouplingStrength(0.02 + 0.001 * (k % 17))
        s.setCoupledCurrentDriver(constantDriver(5e11 + 1e9 * (k % 5)))
        return s

    big = [build(k, SeriesStack) for k in range(2100)]  # 2100 x 32 lanes > 1024 waves
    run_stacks(big, 2e-11, 1e-13, 1e-12)
    kb = _lib.last_kernels()
    few = [build(k, SeriesStack) for k in range(3)]
    run_stacks(few, 2e-11, 1e-13, 1e-12)
    kf = _lib.last_kernels()
    print(f"[stack] 2100 stacks: {kb}; 3 stacks: {kf}", flush=True)
    assert kb[0].startswith("k_junction<double") and kf[0].startswith("k_junction_cl<double")
    worst = 0.0
    for k in (0, 1, 2, 999, 2099):
        o = build(k, OracleSeriesStack)
        o.runSimulation(2e-11, 1e-13, 1e-12)
        worst = max(worst, _compare(big[k], o))
        if k < 3:
            worst = max(worst, _compare(few[k], o))
    assert worst <= 1e-8
