"""Dipole-coupled groups on the GPU (cmtj_run_groups: one group per
workgroup, the group field built in LDS every step) against the oracle's
group runner: Kanao and Nomura coupling, one- and two-layer elements, the
thermal Euler-Heun path, a full 64-element group, many groups in one launch
and a NaN that stops its group only."""
from __future__ import annotations

import math

import numpy as np
import pytest

from cmtj_amd import AxialDriver, CVector, Junction, constantDriver, stepDriver
from cmtj_amd import reservoir
from cmtj_amd.reservoir import GroupInteraction, run_groups
from oracle_backend import OracleGroupInteraction
from test_group_cpu import POS, _elements

pytestmark = pytest.mark.gpu

_FN = {"kanao": reservoir.computeDipoleInteraction, "nomura": reservoir.computeDipoleInteractionNoumra,
       "null": reservoir.nullDipoleInteraction}


def _compare(a, b, n):
    worst = 0.0
    for q in range(n):
        la, lb = a.getLog(q), b.getLog(q)
        assert la.keys() == lb.keys()
        for k in la:
            x, y = np.asarray(la[k]), np.asarray(lb[k])
            assert x.shape == y.shape, k
            scale = max(1.0, float(np.max(np.abs(y)))) if y.size else 1.0
            worst = max(worst, float(np.max(np.abs(x - y))) / scale if y.size else 0.0)
    return worst


def _grid(n, pitch=55e-9):
    side = int(math.ceil(math.sqrt(n)))
    return [CVector((q % side) * pitch, (q // side) * pitch, 0.0) for q in range(n)]


@pytest.mark.parametrize("kind", ["kanao", "nomura", "null"])
@pytest.mark.parametrize("two", [False, True])
def test_group_matches_oracle(kind, two):
    js = _elements(Junction, 4, two)
    pos = [CVector(*p) for p in POS[:4]]
    g, o = GroupInteraction(pos, js), OracleGroupInteraction(pos, js)
    for x in (g, o):
        x.setInteractionFunction(_FN[kind])
        x.runSimulation(0.3e-9, 1e-13, 1e-12)
    worst = _compare(g, o, 4)
    print(f"[group] {kind} two_layer={two}: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8
    for q in range(4):
        np.testing.assert_allclose(np.array(g.junctionList[q].getLayerMagnetisation("free").tolist()),
                                   np.array(o.junctionList[q].getLayerMagnetisation("free").tolist()), atol=1e-8)


def test_thermal_group_matches_oracle_pathwise():
    """Elements at 300 K step with Euler-Heun; the engine and the oracle share
    the Philox stream, so the paths agree."""
    js = _elements(Junction, 3)
    for q, j in enumerate(js):
        j.setLayerTemperatureDriver("all", constantDriver(300.0))
        j.setNoiseSeed(11 + q)
    pos = [CVector(*p) for p in POS[:3]]
    g, o = GroupInteraction(pos, js), OracleGroupInteraction(pos, js)
    for x in (g, o):
        x.runSimulation(0.2e-9, 1e-13, 1e-12)
    worst = _compare(g, o, 3)
    print(f"[group] Euler-Heun 300 K: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8


def test_full_group_of_64():
    js = _elements(Junction, 64)
    pos = _grid(64)
    g, o = GroupInteraction(pos, js), OracleGroupInteraction(pos, js)
    for x in (g, o):
        x.runSimulation(0.1e-9, 1e-13, 1e-12)
    worst = _compare(g, o, 64)
    print(f"[group] 64 elements: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8


def test_nan_element_stops_its_group_only():
    """Many 5-element groups in one launch (one workgroup each); an element
    whose field steps to infinity ends its group's run at that step."""
    def build(k, J, G, bad=False):
        js = _elements(J, 5)
        for q, j in enumerate(js):
            amp = float("inf") if (bad and q == 2) else 1e3 * (k % 4)
            j.setLayerExternalFieldDriver("all", AxialDriver(stepDriver(0, amp, 5e-12, 1e-9), constantDriver(0),
                                                             constantDriver(5e3)))
        grp = G(_grid(5, pitch=50e-9 + 2e-9 * k), js)
        grp.setInteractionFunction(_FN["nomura" if k % 2 else "kanao"])
        return grp

    K = 24
    g = [build(k, Junction, GroupInteraction, bad=(k == 5)) for k in range(K)]
    o = [build(k, Junction, OracleGroupInteraction, bad=(k == 5)) for k in range(K)]
    with pytest.raises(RuntimeError, match="NAN magnetisation"):
        run_groups(g, 0.2e-9, 1e-13, 1e-12)
    worst = 0.0
    for k, b in enumerate(o):
        if k == 5:
            with pytest.raises(RuntimeError, match="NAN magnetisation"):
                b.runSimulation(0.2e-9, 1e-13, 1e-12)
        else:
            b.runSimulation(0.2e-9, 1e-13, 1e-12)
        worst = max(worst, _compare(g[k], b, 5))
    print(f"[group] {K} groups in one launch: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8
    assert 0 < len(g[5].getLog(0)["time"]) < len(g[4].getLog(0)["time"])


def test_single_element_group_equals_the_junction_alone():
    """One element: the sum has no source (its own inverse distance is 0), so
    the run is the junction's own run."""
    (j,) = _elements(Junction, 1, True)
    g = GroupInteraction([CVector(0, 0, 0)], [j])
    g.runSimulation(0.2e-9, 1e-13, 1e-12)
    j.runSimulation(0.2e-9, 1e-13, 1e-12)
    a, b = g.getLog(0), j.getLog()
    assert a.keys() == b.keys()
    for k in a:
        np.testing.assert_allclose(a[k], b[k], rtol=0, atol=1e-13)


def test_elements_of_different_structure_are_rejected():
    js = _elements(Junction, 2)
    js[1].setLayerOerstedFieldDriver("all", AxialDriver(stepDriver(0, 1e3, 1e-12, 2e-12), constantDriver(0),
                                                        constantDriver(0)))
    g = GroupInteraction([CVector(*p) for p in POS[:2]], js)
    with pytest.raises(RuntimeError, match="one structure"):
        g.runSimulation(1e-12, 1e-13, 1e-13)


@pytest.mark.parametrize("n", [9, 12, 33])
def test_four_wave_groups_match_oracle(n):
    """More than 8 elements in at most 256 groups: four waves per group, the
    sum over the sources split four ways (uneven shares included); two-layer
    elements, Nomura coupling, several groups in one launch."""
    def build(k, G):
        js = _elements(Junction, n, two=True)
        grp = G(_grid(n, pitch=70e-9 + 3e-9 * k), js)
        grp.setInteractionFunction(_FN["nomura" if k % 2 else "kanao"])
        return grp

    g = [build(k, GroupInteraction) for k in range(6)]
    o = [build(k, OracleGroupInteraction) for k in range(6)]
    run_groups(g, 0.1e-9, 1e-13, 1e-12)
    worst = 0.0
    for a, b in zip(g, o):
        b.runSimulation(0.1e-9, 1e-13, 1e-12)
        worst = max(worst, _compare(a, b, n))
    print(f"[group] {n} elements x 6 groups, four waves each: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8


def test_four_wave_group_nan_releases_every_wave():
    """A NaN in one element of a 12-element group (four waves) stops that
    group; the other waves of its workgroup leave with it and the other
    groups finish."""
    def build(k, G, bad):
        js = _elements(Junction, 12)
        for q, j in enumerate(js):
            amp = float("inf") if (bad and q == 7) else 1e3 * (k % 3)
            j.setLayerExternalFieldDriver("all", AxialDriver(stepDriver(0, amp, 5e-12, 1e-9), constantDriver(0),
                                                             constantDriver(5e3)))
        return G(_grid(12, pitch=60e-9), js)

    g = [build(k, GroupInteraction, k == 2) for k in range(4)]
    o = [build(k, OracleGroupInteraction, k == 2) for k in range(4)]
    with pytest.raises(RuntimeError, match="NAN magnetisation"):
        run_groups(g, 0.1e-9, 1e-13, 1e-12)
    for k, b in enumerate(o):
        if k == 2:
            with pytest.raises(RuntimeError, match="NAN magnetisation"):
                b.runSimulation(0.1e-9, 1e-13, 1e-12)
        else:
            b.runSimulation(0.1e-9, 1e-13, 1e-12)
        assert _compare(g[k], b, 12) <= 1e-8
    assert 0 < len(g[2].getLog(0)["time"]) < len(g[1].getLog(0)["time"])


def _biased_kanao(r1, r2, layer1, layer2):
    """A Python interaction: the Kanao field plus a constant bias (affine,
    not the module's function object)."""
    h = reservoir.computeDipoleInteraction(r1, r2, layer1, layer2)
    return CVector(h.x + 35.0, h.y - 20.0, h.z + 5.0)


@pytest.mark.parametrize("n", [4, 64])
def test_callable_interaction_matches_oracle(n):
    """A Python callable runs as per-pair affine terms read on the host; 64
    elements put the group on four waves (the helper waves sum their share
    of the sources from the same terms)."""
    js = _elements(Junction, n)
    pos = [CVector(*p) for p in POS[:4]] if n == 4 else _grid(n)
    g, o = GroupInteraction(pos, js), OracleGroupInteraction(pos, js)
    for x in (g, o):
        x.setInteractionFunction(_biased_kanao)
        x.runSimulation(0.1e-9, 1e-13, 1e-12)
    worst = _compare(g, o, n)
    print(f"[group] callable n={n}: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8


def test_callable_and_builtin_groups_in_one_launch():
    """Groups with a Python callable and groups with the built-in functions
    side by side in one launch; a callable computing the Kanao field runs as
    the built-in one within rounding."""
    from test_group_cpu import _my_kanao

    def build(k, G):
        js = _elements(Junction, 5)
        grp = G(_grid(5, pitch=50e-9 + 2e-9 * k), js)
        grp.setInteractionFunction([_my_kanao, _FN["kanao"], _biased_kanao, _FN["nomura"]][k % 4])
        return grp

    K = 12
    g = [build(k, GroupInteraction) for k in range(K)]
    o = [build(k, OracleGroupInteraction) for k in range(K)]
    run_groups(g, 0.2e-9, 1e-13, 1e-12)
    worst = 0.0
    for k, b in enumerate(o):
        b.runSimulation(0.2e-9, 1e-13, 1e-12)
        worst = max(worst, _compare(g[k], b, 5))
    print(f"[group] {K} groups, callables and built-ins: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8
    a = GroupInteraction(_grid(5), _elements(Junction, 5))
    b = GroupInteraction(_grid(5), _elements(Junction, 5))
    a.setInteractionFunction(_my_kanao)
    for x in (a, b):
        x.runSimulation(0.2e-9, 1e-13, 1e-12)
    d = _compare(a, b, 5)
    print(f"[group] Kanao as a callable vs built in: {d:.3e}", flush=True)
    assert d <= 1e-10
