"""Dipole-coupled groups (cmtj_amd.reservoir): the reference's own group tests
(tests/test_group_interaction.py) run through the oracle, the interaction
functions pinned by the textbook point-dipole field, and the oracle's group
runner pinned step by step against single-junction runs that carry the
textbook field as their reserved interaction field."""
from __future__ import annotations

import math

import numpy as np
import pytest

from cmtj_amd import AxialDriver, CVector, Junction, Layer, constantDriver
from cmtj_amd import reservoir
from cmtj_amd.constants import PhysicalConstants
from oracle_backend import OracleGroupInteraction, OracleJunction


def create_test_junction(pos=(0, 0, 0), Ms=1e6, J=Junction, mag=(0, 0, 1)):
    layer = Layer("free", mag=CVector(*mag), anis=CVector(0, 0, 1), Ms=Ms, thickness=2e-9,
                  cellSurface=np.pi * (20e-9) ** 2, demagTensor=[CVector(0, 0, 0)] * 3)
    return J([layer]), CVector(*pos)


# --- the reference's tests/test_group_interaction.py ------------------------

def test_group_interaction_initialization():
    j1, pos1 = create_test_junction((0, 0, 0))
    j2, pos2 = create_test_junction((100e-9, 0, 0))
    reservoir.GroupInteraction([pos1, pos2], [j1, j2])
    with pytest.raises(RuntimeError, match="same size"):
        reservoir.GroupInteraction([pos1], [j1, j2])
    with pytest.raises(RuntimeError, match="cannot be empty"):
        reservoir.GroupInteraction([], [])
    with pytest.raises(RuntimeError, match="unique"):
        reservoir.GroupInteraction([pos1, pos1], [j1, j2])


def test_dipole_interactions():
    j1, pos1 = create_test_junction((0, 0, 0))
    j2, pos2 = create_test_junction((100e-9, 0, 0))
    h = reservoir.nullDipoleInteraction(pos1, pos2, j1.getLayer("free"), j2.getLayer("free"))
    assert h.x == 0 and h.y == 0 and h.z == 0
    assert isinstance(reservoir.computeDipoleInteraction(pos1, pos2, j1.getLayer("free"), j2.getLayer("free")),
                      CVector)
    assert isinstance(reservoir.computeDipoleInteractionNoumra(pos1, pos2, j1.getLayer("free"),
                                                               j2.getLayer("free")), CVector)


def test_group_simulation():
    j1, pos1 = create_test_junction((0, 0, 0), J=OracleJunction)
    j2, pos2 = create_test_junction((100e-9, 0, 0), J=OracleJunction)
    group = OracleGroupInteraction([pos1, pos2], [j1, j2])
    with pytest.raises((RuntimeError, IndexError, TypeError)):
        group.getLog(-1)
    with pytest.raises((RuntimeError, IndexError)):
        group.getLog(2)
    group.runSimulation(1e-9, 1e-12)
    assert len(group.getLog(0)["time"]) > 0
    assert len(group.getLog(1)["time"]) > 0
    group.clearLogs()
    with pytest.raises(KeyError):
        group.getLog(0)["time"]
    with pytest.raises(KeyError):
        group.getLog(1)["time"]


def test_interaction_functions():
    j1, pos1 = create_test_junction((0, 0, 0))
    j2, pos2 = create_test_junction((100e-9, 0, 0))
    group = reservoir.GroupInteraction([pos1, pos2], [j1, j2])
    group.setInteractionFunction(reservoir.nullDipoleInteraction)
    group.setInteractionFunction(reservoir.computeDipoleInteractionNoumra)
    group.setInteractionFunction(reservoir.computeDipoleInteraction)


def test_invalid_log_access():
    j1, pos1 = create_test_junction((0, 0, 0))
    j2, pos2 = create_test_junction((100e-9, 0, 0))
    group = reservoir.GroupInteraction([pos1, pos2], [j1, j2])
    with pytest.raises(RuntimeError, match="non-existing"):
        group.getLog(2)


# --- interaction functions against the textbook dipole field ----------------

def _textbook(r, m, Ms, V):
    """Field of a point dipole of moment (Ms V / mu0) m at distance r:
    (1 / 4 pi) (3 (m . rhat) rhat - m) / |r|^3 (in the reference's units)."""
    mu0 = PhysicalConstants.magnetic_permeability()
    d = np.linalg.norm(r)
    rh = r / d
    return (Ms * V / mu0) * (3 * (m @ rh) * rh - m) / (4 * math.pi * d ** 3)


@pytest.mark.parametrize("seed", range(5))
def test_interactions_match_textbook_dipole(seed):
    rng = np.random.default_rng(seed)
    r1, r2 = rng.normal(size=3) * 50e-9, rng.normal(size=3) * 50e-9
    m = rng.normal(size=3)
    m /= np.linalg.norm(m)
    j1, _ = create_test_junction(Ms=0.7)
    j2, _ = create_test_junction(Ms=1.3, mag=tuple(m))
    l1, l2 = j1.getLayer("free"), j2.getLayer("free")
    want = _textbook(r1 - r2, m, l2.Ms, l2.thickness * l2.cellSurface)
    nom = np.array(reservoir.computeDipoleInteractionNoumra(CVector(*r1), CVector(*r2), l1, l2).tolist())
    kan = np.array(reservoir.computeDipoleInteraction(CVector(*r1), CVector(*r2), l1, l2).tolist())
    np.testing.assert_allclose(nom, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # Kanao's form carries no 1 / (4 pi)
    np.testing.assert_allclose(kan, 4 * math.pi * want, rtol=1e-12, atol=1e-12 * np.abs(kan).max())


def test_too_close_points():
    j1, _ = create_test_junction()
    l = j1.getLayer("free")
    with pytest.raises(RuntimeError, match="too close"):
        reservoir.computeDipoleInteraction(CVector(0, 0, 0), CVector(1e-11, 0, 0), l, l)
    reservoir.computeDipoleInteractionNoumra(CVector(0, 0, 0), CVector(1e-11, 0, 0), l, l)
    with pytest.raises(RuntimeError, match="too close"):
        reservoir.computeDipoleInteractionNoumra(CVector(0, 0, 0), CVector(1e-15, 0, 0), l, l)
    # a group whose points are too close fails at its first step
    js = [create_test_junction(J=OracleJunction)[0] for _ in range(2)]
    g = OracleGroupInteraction([CVector(0, 0, 0), CVector(1e-11, 0, 0)], js)
    with pytest.raises(RuntimeError, match="too close"):
        g.runSimulation(1e-12, 1e-13, 1e-13)


def test_non_affine_callables_are_refused():
    """The device applies f = T(r_i, r_j) m_j + b per pair: a callable that
    is not affine in the source's m, or that reads the target's m, is
    refused before the run."""
    js = [create_test_junction()[0] for _ in range(2)]
    for fn in (lambda r1, r2, a, b: CVector(b.mag.x * b.mag.x, 0, 0),
               lambda r1, r2, a, b: CVector(a.mag.x * 1e3, 0, 0)):
        g = reservoir.GroupInteraction([CVector(0, 0, 0), CVector(1e-7, 0, 0)], js)
        g.setInteractionFunction(fn)
        with pytest.raises(NotImplementedError, match="not affine"):
            g.runSimulation(1e-12, 1e-13, 1e-13)


# --- the oracle's group runner, step by step --------------------------------

def _elements(J, n=3, two=False):
    out = []
    mags = [(0.1, 0.2, 0.97), (0.6, -0.3, 0.74), (-0.2, 0.5, 0.84), (0.3, 0.3, 0.9)]
    for q in range(n):
        demag = [CVector(0.02, 0, 0), CVector(0, 0.03, 0), CVector(0, 0, 0.95)]
        free = Layer("free", mag=CVector(*mags[q % 4]), anis=CVector(0, 0, 1), Ms=1.1 + 0.05 * q, thickness=2e-9,
                     cellSurface=math.pi * (30e-9) ** 2, demagTensor=demag, damping=0.02)
        layers = [free]
        if two:
            layers.append(Layer("bottom", mag=CVector(0, 0.1, 0.99), anis=CVector(0, 0, 1), Ms=1.3,
                                thickness=3e-9, cellSurface=math.pi * (30e-9) ** 2, demagTensor=demag,
                                damping=0.03))
        j = J(layers)
        j.setLayerAnisotropyDriver("free", constantDriver(4e5))
        if two:
            j.setLayerAnisotropyDriver("bottom", constantDriver(6e5))
            j.setIECDriver("free", "bottom", constantDriver(-1e-4))
        j.setLayerExternalFieldDriver("all", AxialDriver(1e4 * q, 0, 5e3))
        out.append(j)
    return out


POS = [(0, 0, 0), (60e-9, 0, 0), (20e-9, 45e-9, 5e-9), (-30e-9, 10e-9, 0)]


def _my_kanao(r1, r2, layer1, layer2):
    """A Python interaction (not the module's function object): the Kanao
    field, so the run takes the per-pair tensor path."""
    return reservoir.computeDipoleInteraction(r1, r2, layer1, layer2)


@pytest.mark.parametrize("kind", ["kanao", "nomura", "callable"])
@pytest.mark.parametrize("two", [False, True])
def test_oracle_group_composes_single_steps(kind, two):
    """The group run equals, step for step, single-junction runs of one step
    each whose top layer carries, as its reserved interaction field, the
    textbook dipole field of the other elements at the step's start (Kanao:
    4 pi times it; a Python callable computing it: the same, through the
    per-pair tensors)."""
    n, dt, steps = 3, 1e-13, 12
    group = OracleGroupInteraction([CVector(*p) for p in POS[:n]], _elements(OracleJunction, n, two))
    if kind == "nomura":
        group.setInteractionFunction(reservoir.computeDipoleInteractionNoumra)
    elif kind == "callable":
        group.setInteractionFunction(_my_kanao)
    group.runSimulation(steps * dt, dt, dt)
    solo = _elements(OracleJunction, n, two)
    scale = 1.0 if kind == "nomura" else 4 * math.pi
    want = [[] for _ in range(n)]
    for _ in range(steps):
        ms = [np.array(j.getLayerMagnetisation("free").tolist()) for j in solo]
        for i, j in enumerate(solo):
            H = np.zeros(3)
            for k, other in enumerate(solo):
                if k == i:
                    continue
                lo = other.getLayer("free")
                H += scale * _textbook(np.array(POS[i], float) - np.array(POS[k], float), ms[k], lo.Ms,
                                       lo.thickness * lo.cellSurface)
            j.setLayerReservedInteractionField("free", AxialDriver(*H))
        for i, j in enumerate(solo):
            j.runSimulation(dt, dt, dt)
            want[i].append(np.array(j.getLayerMagnetisation("free").tolist()))
    for i in range(n):
        lg = group.getLog(i)
        got = np.stack([lg[f"free_m{c}"] for c in "xyz"], axis=1)
        assert got.shape == (steps, 3)
        np.testing.assert_allclose(got, np.stack(want[i]), rtol=0, atol=1e-12)
        np.testing.assert_allclose(lg["time"], np.arange(steps) * dt)


def test_oracle_null_interaction_equals_alone():
    js = _elements(OracleJunction, 3, True)
    group = OracleGroupInteraction([CVector(*p) for p in POS[:3]], js)
    group.setInteractionFunction(reservoir.nullDipoleInteraction)
    group.runSimulation(2e-11, 1e-13, 1e-12)
    for q, j in enumerate(js):  # the group holds copies; the originals run alone
        j.runSimulation(2e-11, 1e-13, 1e-12)
        a, b = group.getLog(q), j.getLog()
        assert a.keys() == b.keys()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k])


def test_group_coupling_changes_the_dynamics():
    """Kanao coupling at 60 nm moves the trajectories away from the
    uncoupled ones; the two interaction kinds differ by 4 pi in strength, so
    Kanao's departure is the larger one."""
    def run(fn):
        g = OracleGroupInteraction([CVector(*p) for p in POS[:3]], _elements(OracleJunction, 3))
        g.setInteractionFunction(fn)
        g.runSimulation(5e-11, 1e-13, 1e-12)
        return np.array(g.getLog(0)["free_mx"])

    null, kan, nom = (run(f) for f in (reservoir.nullDipoleInteraction, reservoir.computeDipoleInteraction,
                                       reservoir.computeDipoleInteractionNoumra))
    d_kan, d_nom = np.abs(kan - null).max(), np.abs(nom - null).max()
    assert d_nom > 1e-6 and d_kan > 3 * d_nom


def test_group_copies_junctions_and_default_write_frequency():
    js = _elements(OracleJunction, 2)
    g = OracleGroupInteraction([CVector(*p) for p in POS[:2]], js)
    g.runSimulation(1e-10, 1e-13)  # writeFrequency defaults to 1e-11
    # (unsigned)(1e-11 / 1e-13) = 99 and (unsigned)(1e-10 / 1e-13) in FP64
    W, n = int(1e-11 / 1e-13), int(1e-10 / 1e-13)
    assert W == 99 and len(g.getLog(0)["time"]) == -(-n // W) == 11
    assert js[0].getLog() == {}


def test_mixed_solvers_are_rejected():
    js = _elements(OracleJunction, 2)
    js[1].setLayerTemperatureDriver("all", constantDriver(300.0))
    g = OracleGroupInteraction([CVector(*p) for p in POS[:2]], js)
    with pytest.raises(RuntimeError, match="different solver modes"):
        g.runSimulation(1e-12, 1e-13, 1e-13)


def test_callable_tensors_are_the_textbook_dipole_tensor():
    """The per-pair terms read off a Python Kanao callable are
    (Ms_j V_j / mu0) (3 r r^T / |r|^5 - I / |r|^3), r = r_i - r_j, with
    b = 0, for every ordered pair."""
    js = _elements(OracleJunction, 4)
    g = OracleGroupInteraction([CVector(*p) for p in POS[:4]], js)
    g.setInteractionFunction(_my_kanao)
    pt = g._pair_terms()
    mu0 = PhysicalConstants.magnetic_permeability()
    for i in range(4):
        assert np.all(pt[i, :, i] == 0)
        for j in range(4):
            if i == j:
                continue
            r = np.array(POS[i], float) - np.array(POS[j], float)
            lj = js[j].getLayer("free")
            d = np.linalg.norm(r)
            want = (lj.Ms / mu0) * lj.thickness * lj.cellSurface * (3 * np.outer(r, r) / d ** 5 - np.eye(3) / d ** 3)
            np.testing.assert_allclose(pt[j, :9, i].reshape(3, 3), want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
            assert np.all(pt[j, 9:, i] == 0)


def test_callable_constant_field_is_an_offset():
    """A callable that returns a constant vector c is all offset (T = 0): each
    element gets (n - 1) c, as a junction alone whose reserved field is
    (n - 1) c -- bit for bit."""
    c = (120.0, -40.0, 15.0)
    js = _elements(OracleJunction, 3, True)
    g = OracleGroupInteraction([CVector(*p) for p in POS[:3]], js)
    g.setInteractionFunction(lambda r1, r2, a, b: CVector(*c))
    g.runSimulation(3e-11, 1e-13, 1e-12)
    for q, j in enumerate(_elements(OracleJunction, 3, True)):
        j.setLayerReservedInteractionField("free", AxialDriver(*(2 * x for x in c)))
        j.runSimulation(3e-11, 1e-13, 1e-12)
        a, b = g.getLog(q), j.getLog()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k])
