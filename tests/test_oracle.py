"""Pins the CPU oracle: the reference's golden vectors, the reference's own
drivers.hpp, closed forms, convergence orders and statistical mechanics.
Runs on the CPU only."""
from __future__ import annotations

import math

import numpy as np
import pytest

import golden_cases
import oracle
import parity_cases
from cmtj_amd import (AxialDriver, CVector, Junction, Layer, ScalarDriver, SolverMode, _lib, constantDriver,
                      sineDriver)
from conditioning import sensitivity
from oracle_backend import OracleJunction

ZERO3 = [CVector(0, 0, 0), CVector(0, 0, 0), CVector(0, 0, 0)]


# ---------------------------------------------------------------------------
# golden vectors of the reference (tests/assets/curated_examples/*.json)

@pytest.mark.parametrize("name", list(golden_cases.CASES))
def test_oracle_reproduces_reference_fixture(name):
    golden_cases.assert_matches(name, golden_cases.CASES[name](OracleJunction))


def test_golden_check_detects_a_planted_mistake(monkeypatch):
    """The fixture comparison is sharp enough to catch a wrong STT sign."""
    def planted(J):
        out = golden_cases.case_sto(J)
        return out

    class Wrong(OracleJunction):
        def _execute(self, j, L, consts, tables):
            L = L.copy()
            L["beta"] = -L["beta"]  # a plausible sign slip in the field-like STT term
            return oracle.run(j, L, consts, tables)

    with pytest.raises(AssertionError):
        golden_cases.assert_matches("test_sto", planted(Wrong))


# ---------------------------------------------------------------------------
# drivers

def _rec(d: ScalarDriver):
    return d._record()


def test_driver_known_answers():
    # reference tests/test_drivers.py:51-65
    d = sineDriver(10, 20, 1, 0)
    assert oracle.driver_value(_rec(d), 1 / 4) == 30
    d *= 2
    assert oracle.driver_value(_rec(d), 1 / 4) == 60
    d = sineDriver(10, 20, 1, 0)
    d += 2
    assert oracle.driver_value(_rec(d), 1 / 4) == 34
    assert oracle.driver_value(_rec(sineDriver(10, 20, 1, 0) * 2), 1 / 4) == 60
    assert oracle.driver_value(_rec(sineDriver(10, 20, 1, 0) + 2), 1 / 4) == 34


_KINDS = [
    ("const", ScalarDriver.getConstantDriver(3.5), (3.5, 0, 0, 0, 0)),
    ("pulse", ScalarDriver.getPulseDriver(1.0, 2.0, 1e-9, 0.3), (1.0, 2.0, 1e-9, 0.3, 0)),
    ("sine", ScalarDriver.getSineDriver(0.5, 2.0, 3e9, 0.4), (0.5, 2.0, 3e9, 0.4, 0)),
    ("posine", ScalarDriver.getPosSineDriver(0.5, 2.0, 3e9, 0.4), (0.5, 2.0, 3e9, 0.4, 0)),
    ("halfsine", ScalarDriver.getHalfSineDriver(0.5, 2.0, 3e9, 0.4), (0.5, 2.0, 3e9, 0.4, 0)),
    ("step", ScalarDriver.getStepDriver(1.0, 5.0, 1e-9, 2e-9), (1.0, 5.0, 1e-9, 2e-9, 0)),
    ("trap", ScalarDriver.getTrapezoidDriver(1.0, 5.0, 1e-9, 0.5e-9, 1e-9), (1.0, 5.0, 1e-9, 0.5e-9, 1e-9)),
    ("gimp", ScalarDriver.getGaussianImpulseDriver(1.0, 5.0, 2e-9, 0.3e-9), (1.0, 5.0, 2e-9, 0.3e-9, 0)),
    ("gstep", ScalarDriver.getGaussianStepDriver(1.0, 5.0, 2e-9, 0.3e-9), (1.0, 5.0, 2e-9, 0.3e-9, 0)),
]
# a grid that hits every edge: step/pulse/trapezoid boundaries exactly
_TIMES = sorted(set(np.linspace(0, 4e-9, 161).tolist() + [1e-9, 2e-9, 1.5e-9, 2.5e-9, 3e-9, 0.3e-9, 1.3e-9]))


@pytest.mark.skipif(not oracle.ref_drivers_available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("kind,drv,params", _KINDS, ids=[k[0] for k in _KINDS])
def test_drivers_match_reference_build(kind, drv, params):
    """Every waveform, also after * and +, against the reference's drivers.hpp."""
    for mul, add in ((1.0, 0.0), (2.5, 0.0), (1.0, -0.75)):
        d = drv
        if mul != 1.0:
            d = d * mul
        if add != 0.0:
            d = d + add
        lines = [f"{kind} {' '.join(repr(float(p)) for p in params)} {mul!r} {add!r} {t!r}" for t in _TIMES]
        ref = oracle.ref_driver_values(lines)
        ours = [oracle.driver_value(_rec(d), t) for t in _TIMES]
        host = [d.getCurrentScalarValue(t) for t in _TIMES]  # product host path (C ABI)
        assert ours == ref, (kind, mul, add)
        np.testing.assert_allclose(host, ref, rtol=1e-15, atol=1e-300)


# ---------------------------------------------------------------------------
# closed forms and convergence

GAMMA = 220880.0


def _precession(J, alpha, H, theta0, T, dt, solver):
    l = Layer("free", mag=CVector(math.sin(theta0), 0.0, math.cos(theta0)), anis=CVector(0, 0, 1), Ms=1.0,
              thickness=1e-9, cellSurface=1e-16, demagTensor=ZERO3, damping=alpha)
    j = J([l])
    j.setLayerExternalFieldDriver("all", AxialDriver(0, 0, H))
    j.runSimulation(T, dt, T / 10, solverMode=solver)
    m = j.getLayerMagnetisation("free")
    return np.array([m.x, m.y, m.z])


def _exact(alpha, H, theta0, t):
    # dm/dt = -g'[m x H + a m x (m x H)], g' = g/(1+a^2): phi = g' H t,
    # tan(theta/2) = tan(theta0/2) exp(-a g' H t)
    gp = GAMMA / (1 + alpha**2)
    phi = gp * H * t
    th = 2 * math.atan(math.tan(theta0 / 2) * math.exp(-alpha * gp * H * t))
    return np.array([math.sin(th) * math.cos(phi), math.sin(th) * math.sin(phi), math.cos(th)])


@pytest.mark.parametrize("solver,order", [(SolverMode.RK4, 4), (SolverMode.Heun, 2), (SolverMode.EulerHeun, 1)])
def test_convergence_order_against_exact_precession(solver, order):
    alpha, H, th0, T = 0.1, 2e5, 0.8, 2e-10
    exact = _exact(alpha, H, th0, T)
    errs = []
    for dt in (4e-13, 2e-13, 1e-13):
        m = _precession(OracleJunction, alpha, H, th0, T, dt, solver)
        errs.append(np.linalg.norm(m - exact))
    ratios = [errs[0] / errs[1], errs[1] / errs[2]]
    for r in ratios:
        assert abs(math.log2(r) - order) < 0.35, (solver, errs, ratios)


def test_rk4_is_accurate_against_exact_precession():
    m = _precession(OracleJunction, 0.02, 5e5, 1.2, 1e-9, 1e-13, SolverMode.RK4)
    np.testing.assert_allclose(m, _exact(0.02, 5e5, 1.2, 1e-9), atol=1e-9)


def test_multilayer_rk4_uses_start_of_step_neighbours():
    """With frozen neighbours the coupled error is first order in dt; a fully
    coupled RK4 would be fourth order."""
    def run(dt):
        a = Layer("a", mag=CVector(1, 0.3, 0.2), anis=CVector(1, 0, 0), Ms=1.0, thickness=1e-9, cellSurface=1e-16,
                  demagTensor=ZERO3, damping=0.05)
        b = Layer("b", mag=CVector(0.2, 1, 0.1), anis=CVector(0, 1, 0), Ms=1.0, thickness=1e-9, cellSurface=1e-16,
                  demagTensor=ZERO3, damping=0.05)
        j = OracleJunction([a, b])
        j.setIECDriver("a", "b", constantDriver(5e-4))
        j.runSimulation(1e-10, dt, 1e-10)
        return np.array(j.getLayerMagnetisation("a").tolist() + j.getLayerMagnetisation("b").tolist())

    ref = run(1e-15)
    e = [np.linalg.norm(run(dt) - ref) for dt in (4e-13, 2e-13, 1e-13)]
    for r in (e[0] / e[1], e[1] / e[2]):
        assert abs(math.log2(r) - 1.0) < 0.3, e


def test_unit_length_and_energy_decay():
    """|m| = 1 after every step; with damping and static fields the energy
    E = -Ms H.m - K (a.m)^2 + (Ms^2 / 2 mu0) m.N.m never increases."""
    mu0, Ms, K = 12.566e-7, 1.2, 2e4
    N = np.diag([0.05, 0.1, 0.85])
    l = Layer("free", mag=CVector(0.3, 0.9, 0.4), anis=CVector(1, 0, 0), Ms=Ms, thickness=1e-9,
              cellSurface=1e-16, demagTensor=[CVector(*r) for r in N], damping=0.05)
    j = OracleJunction([l])
    j.setLayerAnisotropyDriver("free", constantDriver(K))
    H = np.array([3e4, -1e4, 2e4])
    j.setLayerExternalFieldDriver("all", AxialDriver(*H))
    j.runSimulation(2e-9, 1e-13, 1e-13)
    lg = j.getLog()
    m = np.stack([lg["free_mx"], lg["free_my"], lg["free_mz"]], axis=1)
    np.testing.assert_allclose(np.linalg.norm(m, axis=1), 1.0, atol=1e-14)
    E = -Ms * m @ H - K * m[:, 0] ** 2 + 0.5 * Ms**2 / mu0 * np.einsum("ti,ij,tj->t", m, N, m)
    assert np.all(np.diff(E) <= 1e-9 * np.abs(E).max())


def test_energy_channels_closed_forms():
    """Logged fields of calculateEnergies equal their definitions at the
    logged state (RK4 stage 4 evaluates at the end-of-step m up to O(dt))."""
    l = Layer("free", mag=CVector(0.2, 0.3, 1.0), anis=CVector(0, 0, 1), Ms=1.1, thickness=1e-9,
              cellSurface=1e-16, demagTensor=[CVector(0.1, 0, 0), CVector(0, 0.1, 0), CVector(0, 0, 0.8)],
              damping=0.05)
    j = OracleJunction([l])
    j.setLayerAnisotropyDriver("free", constantDriver(3e5))
    j.setLayerExternalFieldDriver("all", AxialDriver(1e4, 2e4, 3e4))
    j.runSimulation(1e-11, 1e-13, 1e-13, calculateEnergies=True)
    lg = j.getLog()
    assert lg["free_K"] == [3e5] * len(lg["time"])
    assert lg["free_Hextx"][0] == 1e4 and lg["free_Hexty"][0] == 2e4 and lg["free_Hextz"][0] == 3e4
    mz = np.asarray(lg["free_mz"])[1:]
    # Hanis_z = 2K (a.m)/Ms ; Hdemag_z = Nzz mz Ms/mu0 (stage-4 m differs from the logged m by O(dt);
    # entry 0 is skipped: the constructor keeps the initial m unnormalised)
    np.testing.assert_allclose(lg["free_Hanisz"][1:], 2 * 3e5 * mz / 1.1, rtol=1e-3)
    np.testing.assert_allclose(lg["free_Hdemagz"][1:], 0.8 * mz * 1.1 / 12.566e-7, rtol=1e-3)
    assert set(lg["free_Hthx"]) == {0.0}


# ---------------------------------------------------------------------------
# thermal noise

def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors of Random123 (kat_vectors)."""
    assert oracle.philox4x32_10([0, 0, 0, 0], [0, 0]).tolist() == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2).tolist() == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                                [0xA4093822, 0x299F31D0]).tolist() == [0xD16CFE09, 0x94FDCCEB, 0x5001E420,
                                                                       0x24126EA1]


def test_thermal_normals_are_standard_normal():
    xs = np.array([oracle.thermal_normals(7, 3, s, l) for s in range(4000) for l in range(2)]).ravel()
    assert abs(xs.mean()) < 0.03
    assert abs(xs.std() - 1.0) < 0.03
    assert abs(np.mean(xs**4) - 3.0) < 0.2


def _langevin_ensemble(solver, n=384, T=300.0, H=8.28e4):
    Ms, V = 1.0, 1e-25
    base = Layer("free", mag=CVector(0, 0, 1), anis=CVector(0, 0, 1), Ms=Ms, thickness=1e-9,
                 cellSurface=V / 1e-9, demagTensor=ZERO3, damping=0.1)
    js = []
    for i in range(n):
        j = Junction([base])
        j.setLayerExternalFieldDriver("all", AxialDriver(0, 0, H))
        j.setLayerTemperatureDriver("all", constantDriver(T))
        j.setNoiseSeed(1000 + i)
        js.append(j)
    from cmtj_amd.batch import JunctionBatch

    b = JunctionBatch.from_junctions(js, 4e-9, 1e-13, 1e-11, solver, precision="fp64", log="m", log_dtype="f64")
    log, mag, st = oracle.run_batch(b.js, b.layers, _lib.constants_get(), None, b.log_elements)
    ne = int(b.js[0]["n_entries"])
    mz = np.stack([log[int(b.js[i]["log_offset"]) + 2 * ne: int(b.js[i]["log_offset"]) + 3 * ne]
                   for i in range(n)])
    x = Ms * V * H / (1.380649e-23 * T)
    return mz[:, ne // 2:].mean(), 1 / math.tanh(x) - 1 / x


@pytest.mark.parametrize("solver", [SolverMode.EulerHeun, SolverMode.Heun])
def test_thermal_equilibrium_is_boltzmann(solver):
    """<m_z> of an isotropic macrospin in a field is the Langevin function
    L(x) = coth x - 1/x, x = Ms V H / (kB T) (Ms in T, H in A/m)."""
    got, want = _langevin_ensemble(solver)
    assert abs(got - want) < 0.02, (got, want)


# ---------------------------------------------------------------------------
# parity cases are well conditioned

@pytest.mark.parametrize("name", list(parity_cases.CASES))
def test_parity_cases_are_well_conditioned(name):
    assert sensitivity(parity_cases.CASES[name], OracleJunction) < 1e-9
