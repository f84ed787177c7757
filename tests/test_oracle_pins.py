"""Pins the oracle's coupling, torque and resistance branches with checks
whose answers the mathematics fixes, independently of the oracle's code:

* one-sided coupling to a torque-free neighbour (zero field, so its m never
  moves) turns each interlayer term into a drive a single layer already has:
  bilinear IEC and IDMI into a constant field, biquadratic IEC into a
  uniaxial anisotropy along the neighbour's m, a dipole tensor into a
  constant field scaled by this layer's Ms (core/junction.hpp:583-585,
  654-701);
* field-like SOT equals an external field along the polariser; damping-like
  SOT in zero field has a closed-form trajectory; the secondary torque pair
  obeys the same forms with its own polariser (:770-812);
* a full (off-diagonal) demagnetisation tensor makes the dynamics covariant
  under rotations of the whole problem; a non-symmetric tensor is applied
  row by row (:54-62);
* second-order anisotropy on the field axis reduces the LLG to a scalar ODE
  in the polar angle, integrated here by scipy (:645-652);
* STRIP resistances equal the utils' parallel-resistor model recomputed from
  the logged magnetisations, including the 1/(1/0) = 0 case (:1835-1857);
* deterministic coupled Heun takes its corrector neighbours from the
  predictor, which makes it second order on a coupled system; a scheme with
  frozen neighbours is first order (:1756-1815).

Every check is paired with a planted mistake, applied to the descriptor the
oracle receives, that must make it fail.  Runs on the CPU only."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from cmtj_amd import AxialDriver, CVector, Layer, SolverMode, _lib, constantDriver
from oracle_backend import OracleJunction

ZERO3 = [CVector(0, 0, 0), CVector(0, 0, 0), CVector(0, 0, 0)]
GAMMA = 220880.0
MU0 = 12.566e-7
MS_A, T_A = 1.1, 1.3e-9
MS_B, T_B = 1.6, 2.5e-9
M_B = np.array([0.3, -0.5, 0.81]) / np.linalg.norm([0.3, -0.5, 0.81])
RUN = dict(totalTime=1e-9, timeStep=1e-13, writeFrequency=1e-11)


def planted(fn):
    """An oracle junction whose descriptor is altered by fn(j, L) first."""

    class Planted(OracleJunction):
        def _execute(self, j, L, consts, tables):
            j, L = j.copy(), L.copy()
            fn(j, L)
            return oracle.run(j, L, consts, tables)

    return Planted


def _layer_a(name="a", anis=(0, 0, 1)):
    return Layer(name, mag=CVector(0.6, 0.2, 0.77), anis=CVector(*anis), Ms=MS_A, thickness=T_A,
                 cellSurface=1e-16, demagTensor=ZERO3, damping=0.05)


def _neighbour():
    b = Layer("b", mag=CVector(1, 0, 0), anis=CVector(0, 0, 1), Ms=MS_B, thickness=T_B, cellSurface=1e-16,
              demagTensor=ZERO3, damping=0.05)
    b.setMagnetisation(CVector(*M_B))
    return b


def _drive_a(j, layer="a"):
    # a field that keeps layer a precessing; the neighbour gets none
    j.setLayerExternalFieldDriver(layer, AxialDriver(2e4, -1e4, 3e4))


def _m(log, lid):
    return np.stack([log[f"{lid}_mx"], log[f"{lid}_my"], log[f"{lid}_mz"]], axis=1)


def _coupled_vs_single(J, couple, equivalent, side="top"):
    """max|m_a(coupled) - m_a(single layer with the equivalent drive)| and the
    neighbour's largest excursion from its initial m."""
    a = _layer_a()
    couple(a, side)
    j2 = J([a, _neighbour()] if side == "top" else [_neighbour(), a])
    _drive_a(j2)
    j2.runSimulation(**RUN)
    l2 = j2.getLog()
    s = _layer_a()
    equivalent(s)
    j1 = J([s])
    _drive_a(j1)
    j1.runSimulation(**RUN)
    l1 = j1.getLog()
    moved = float(np.max(np.abs(_m(l2, "b") - M_B)))
    return float(np.max(np.abs(_m(l2, "a") - _m(l1, "a")))), moved


# ---------------------------------------------------------------------------
# interlayer terms against a frozen neighbour

J_IEC, J2_IEC = 3e-4, -2e-4
D_VEC = np.array([1.2e-4, -0.7e-4, 0.4e-4])
N_DIP = np.array([[0.02, 0.005, -0.003], [0.001, 0.03, 0.004], [-0.006, 0.002, -0.05]])  # not symmetric


def _iec(a, side):
    (a.setIECDriverTop if side == "top" else a.setIECDriverBottom)(constantDriver(J_IEC))


def _iec_equiv(s):
    h = J_IEC * M_B / (MS_A * T_A)
    s.setReservedInteractionField(AxialDriver(*h))


def _quad(a, side):
    (a.setQuadIECDriverTop if side == "top" else a.setQuadIECDriverBottom)(constantDriver(J2_IEC))


def _quad_equiv(s):
    # 2 J2 (m_b.m) m_b / (Ms t) is the uniaxial field 2 K (u.m) u / Ms with u = m_b, K = J2 / t
    s.anis = CVector(*M_B)
    s.setAnisotropyDriver(constantDriver(J2_IEC / T_A))


def _idmi(a, side):
    (a.setIDMIDriverTop if side == "top" else a.setIDMIDriverBottom)(AxialDriver(*D_VEC))


def _idmi_equiv(s):
    # the reference's interfacial DMI field: -(D x m_neighbour) / (Ms t)
    s.setReservedInteractionField(AxialDriver(*(-np.cross(D_VEC, M_B) / (MS_A * T_A))))


def _dipole(a, side):
    rows = [CVector(*r) for r in N_DIP]
    (a.setTopDipoleTensor if side == "top" else a.setBottomDipoleTensor)(rows)


def _dipole_equiv(s):
    # H = -(N m_neighbour) Ms_this / mu0: rows of N dot m, this layer's Ms
    s.setReservedInteractionField(AxialDriver(*(-(N_DIP @ M_B) * MS_A / MU0)))


def _swap_top_bottom(j, L):
    for a, b in ((_lib.SLOT_J_TOP, _lib.SLOT_J_BOTTOM), (_lib.SLOT_J2_TOP, _lib.SLOT_J2_BOTTOM)):
        L["drv"][:, [a, b]] = L["drv"][:, [b, a]]
    for k in range(3):
        a, b = _lib.SLOT_IDMI_TOP + k, _lib.SLOT_IDMI_BOTTOM + k
        L["drv"][:, [a, b]] = L["drv"][:, [b, a]]
    L["dipole_top"], L["dipole_bottom"] = L["dipole_bottom"].copy(), L["dipole_top"].copy()


def _halve_quad(j, L):
    for s in (_lib.SLOT_J2_TOP, _lib.SLOT_J2_BOTTOM):
        L["drv"][:, s]["constant"] *= 0.5  # a dropped factor 2 in 2 J2 (m_c.m)


def _flip_idmi(j, L):
    for s in range(_lib.SLOT_IDMI_TOP, _lib.SLOT_IDMI_BOTTOM + 3):
        L["drv"][:, s]["constant"] *= -1.0  # D x m_c written as m_c x D


def _transpose_dipole(j, L):
    for f in ("dipole_top", "dipole_bottom"):
        L[f] = L[f].reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1, 9)


def _neighbour_ms_for_dipole(j, L):
    # the field of the neighbour's moment taken with the neighbour's Ms; layer a
    # (row with a dipole tensor) has no other Ms-dependent term in this check
    rows = np.flatnonzero(np.abs(L["dipole_top"]).sum(1) + np.abs(L["dipole_bottom"]).sum(1))
    L["Ms"][rows] = MS_B


TERMS = {
    "iec": (_iec, _iec_equiv, [_swap_top_bottom]),
    "biquadratic_iec": (_quad, _quad_equiv, [_swap_top_bottom, _halve_quad]),
    "idmi": (_idmi, _idmi_equiv, [_swap_top_bottom, _flip_idmi]),
    "dipole": (_dipole, _dipole_equiv, [_swap_top_bottom, _transpose_dipole, _neighbour_ms_for_dipole]),
}


@pytest.mark.parametrize("side", ["top", "bottom"])
@pytest.mark.parametrize("term", list(TERMS))
def test_one_sided_coupling_equals_single_layer_drive(term, side):
    couple, equiv, plants = TERMS[term]
    d, moved = _coupled_vs_single(OracleJunction, couple, equiv, side)
    assert moved < 1e-14  # the neighbour is torque-free
    assert d < 1e-11, (term, side, d)
    for p in plants:
        dp, _ = _coupled_vs_single(planted(p), couple, equiv, side)
        assert dp > 1e-4, (term, side, p.__name__, dp)


# ---------------------------------------------------------------------------
# spin-orbit torques

P_SOT = np.array([0.2, 0.9, -0.4])  # not unit length: the reference uses it as given


def _sot_layer(damping=0.02, K=4e4):
    l = Layer("a", mag=CVector(0.5, -0.3, 0.81), anis=CVector(0.2, 0.1, 1.0), Ms=1.0, thickness=1e-9,
              cellSurface=1e-16, demagTensor=[CVector(0.02, 0, 0), CVector(0, 0.03, 0), CVector(0, 0, 0.9)],
              damping=damping)
    return l, K


def _fl_vs_field(J, secondary=False):
    """Dynamic field-like SOT of strength Hfl along p equals an external
    field Hfl p (damping-like zero)."""
    Hfl, H0 = 3.5e4, np.array([1e4, 2e4, -5e3])
    l, K = _sot_layer()
    if secondary:
        l.setSecondaryReferenceLayer(CVector(*P_SOT))
        l.setSecondaryTorqueDrivers(constantDriver(Hfl), constantDriver(0.0))
    else:
        l.setReferenceLayer(CVector(*P_SOT))
        l.setPrimaryTorqueDrivers(constantDriver(Hfl), constantDriver(0.0))
    j = J([l])
    j.setLayerAnisotropyDriver("a", constantDriver(K))
    j.setLayerExternalFieldDriver("a", AxialDriver(*H0))
    j.runSimulation(**RUN)
    l2, _ = _sot_layer()
    e = J([l2])
    e.setLayerAnisotropyDriver("a", constantDriver(K))
    e.setLayerExternalFieldDriver("a", AxialDriver(*(H0 + Hfl * P_SOT)))
    e.runSimulation(**RUN)
    return float(np.max(np.abs(_m(j.getLog(), "a") - _m(e.getLog(), "a"))))


def _dl_closed_form(J, secondary=False, alpha=0.02):
    """Damping-like SOT alone, p = z, no field: theta' = -g' Hdl sin(theta)
    and phi' = -alpha g' Hdl, i.e. tan(theta/2) = tan(theta0/2) exp(-g' Hdl t)."""
    Hdl, th0, ph0 = 1e4, 2.0, 0.3
    l = Layer("a", mag=CVector(math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)),
              anis=CVector(0, 0, 1), Ms=1.0, thickness=1e-9, cellSurface=1e-16, demagTensor=ZERO3, damping=alpha)
    if secondary:
        l.setSecondaryReferenceLayer(CVector(0, 0, 1))
        l.setSecondaryTorqueDrivers(constantDriver(0.0), constantDriver(Hdl))
    else:
        l.setReferenceLayer(CVector(0, 0, 1))
        l.setPrimaryTorqueDrivers(constantDriver(0.0), constantDriver(Hdl))
    j = J([l])
    j.runSimulation(**RUN)
    lg = j.getLog()
    t = np.asarray(lg["time"]) + 1e-13  # an entry labelled i dt holds m after step i, at (i + 1) dt
    gp = GAMMA / (1 + alpha**2)
    th = 2 * np.arctan(math.tan(th0 / 2) * np.exp(-gp * Hdl * t))
    ph = ph0 - alpha * gp * Hdl * t
    want = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    return float(np.max(np.abs(_m(lg, "a") - want)))


def _swap_fl_dl(j, L):
    L["drv"][:, [_lib.SLOT_FL, _lib.SLOT_DL]] = L["drv"][:, [_lib.SLOT_DL, _lib.SLOT_FL]]
    L["drv"][:, [_lib.SLOT_FL2, _lib.SLOT_DL2]] = L["drv"][:, [_lib.SLOT_DL2, _lib.SLOT_FL2]]


def _flip_dl(j, L):
    for s in (_lib.SLOT_DL, _lib.SLOT_DL2):
        L["drv"][:, s]["constant"] *= -1.0


def _swap_polarisers(j, L):
    L["ref"], L["ref2"] = L["ref2"].copy(), L["ref"].copy()


def _secondary_reads_primary_drivers(j, L):
    L["drv"][:, _lib.SLOT_FL2] = L["drv"][:, _lib.SLOT_FL]
    L["drv"][:, _lib.SLOT_DL2] = L["drv"][:, _lib.SLOT_DL]


@pytest.mark.parametrize("secondary", [False, True], ids=["primary", "secondary"])
def test_field_like_sot_is_a_field_along_the_polariser(secondary):
    assert _fl_vs_field(OracleJunction, secondary) < 1e-11
    plants = [_swap_fl_dl, _swap_polarisers] + ([_secondary_reads_primary_drivers] if secondary else [])
    for p in plants:
        assert _fl_vs_field(planted(p), secondary) > 1e-4, p.__name__


@pytest.mark.parametrize("alpha", [0.0, 0.3])
@pytest.mark.parametrize("secondary", [False, True], ids=["primary", "secondary"])
def test_damping_like_sot_closed_form(secondary, alpha):
    assert _dl_closed_form(OracleJunction, secondary, alpha) < 1e-9
    plants = [_swap_fl_dl, _flip_dl, _swap_polarisers]
    if secondary:
        plants.append(_secondary_reads_primary_drivers)
    for p in plants:
        assert _dl_closed_form(planted(p), secondary, alpha) > 1e-3, p.__name__


# ---------------------------------------------------------------------------
# demagnetisation tensor with off-diagonal entries

def _rotation(seed=4):
    q, r = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def _demag_run(J, R, N):
    Nr = R @ N @ R.T
    m0 = R @ np.array([0.3, 0.9, 0.4])
    l = Layer("a", mag=CVector(*m0), anis=CVector(*(R @ np.array([1.0, 0.2, 0.1]))), Ms=1.2, thickness=1e-9,
              cellSurface=1e-16, demagTensor=[CVector(*row) for row in Nr], damping=0.05)
    j = J([l])
    j.setLayerAnisotropyDriver("a", constantDriver(2e4))
    j.setLayerExternalFieldDriver("a", AxialDriver(*(R @ np.array([3e4, -1e4, 2e4]))))
    j.runSimulation(**RUN)
    return _m(j.getLog(), "a")


def _covariance_error(J, N):
    R = _rotation()
    plain = _demag_run(J, np.eye(3), N)
    rotated = _demag_run(J, R, N)
    return float(np.max(np.abs(rotated - plain @ R.T)))


def _drop_off_diagonal(j, L):
    L["demag"] = L["demag"] * np.eye(3).ravel()[None, :]


def _scramble_off_diagonal(j, L):
    d = L["demag"].copy()
    L["demag"][:, [1, 2]] = d[:, [2, 1]]  # N_xy and N_xz exchanged


@pytest.mark.parametrize("full", [False, True], ids=["diagonal_frame", "full_frame"])
def test_demag_tensor_is_rotation_covariant(full):
    N = np.diag([0.05, 0.1, 0.85])
    if full:  # already off-diagonal in the unrotated frame
        Q = _rotation(9)
        N = Q @ N @ Q.T
    assert _covariance_error(OracleJunction, N) < 1e-11
    for p in (_drop_off_diagonal, _scramble_off_diagonal):
        assert _covariance_error(planted(p), N) > 1e-4, p.__name__


def test_energy_with_full_demag_tensor_never_increases():
    Ms, K = 1.2, 2e4
    Q = _rotation(9)
    N = Q @ np.diag([0.05, 0.1, 0.85]) @ Q.T
    l = Layer("a", mag=CVector(0.3, 0.9, 0.4), anis=CVector(1, 0, 0), Ms=Ms, thickness=1e-9, cellSurface=1e-16,
              demagTensor=[CVector(*r) for r in N], damping=0.05)
    j = OracleJunction([l])
    j.setLayerAnisotropyDriver("a", constantDriver(K))
    H = np.array([3e4, -1e4, 2e4])
    j.setLayerExternalFieldDriver("a", AxialDriver(*H))
    j.runSimulation(2e-9, 1e-13, 1e-13)
    m = _m(j.getLog(), "a")
    E = -Ms * m @ H - K * m[:, 0] ** 2 + 0.5 * Ms**2 / MU0 * np.einsum("ti,ij,tj->t", m, N, m)
    assert np.all(np.diff(E) <= 1e-9 * np.abs(E).max())
    assert E[-1] < E[0] - 1e3  # it relaxes


# ---------------------------------------------------------------------------
# second-order anisotropy: exact reduction to a scalar ODE

def _k2_run(J, alpha=0.05, Ms=1.0, K1=2e4, K2=5e4, H0=-3e4, th0=0.9, ph0=0.2):
    l = Layer("a", mag=CVector(math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)),
              anis=CVector(0, 0, 1), Ms=Ms, thickness=1e-9, cellSurface=1e-16, demagTensor=ZERO3, damping=alpha)
    j = J([l])
    j.setLayerAnisotropyDriver("a", constantDriver(K1))
    j.setLayerSecondOrderAnisotropyDriver("a", constantDriver(K2))
    j.setLayerExternalFieldDriver("a", AxialDriver(0, 0, H0))
    j.runSimulation(2e-9, 1e-13, 1e-11)
    lg = j.getLog()
    m = _m(lg, "a")
    t = np.asarray(lg["time"]) + 1e-13

    # on the axis, H = Hz(theta) z with Hz = H0 + 2 K1 cos/Ms + 4 K2 cos^3/Ms, and
    # dm/dt = -g'[m x H + a m x (m x H)] gives theta' = -a g' Hz sin(theta), phi' = g' Hz
    from scipy.integrate import solve_ivp

    gp = GAMMA / (1 + alpha**2)

    def rhs(_, y):
        c = math.cos(y[0])
        hz = H0 + 2 * K1 * c / Ms + 4 * K2 * c**3 / Ms
        return [-alpha * gp * hz * math.sin(y[0]), gp * hz]

    sol = solve_ivp(rhs, (0, t[-1]), [th0, ph0], method="DOP853", t_eval=t, rtol=1e-13, atol=1e-15)
    th, ph = sol.y
    want = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    return float(np.max(np.abs(m - want)))


def _k2_coefficient_2(j, L):
    L["drv"][:, _lib.SLOT_K2]["constant"] *= 0.5  # 2 K2 instead of 4 K2


def _k2_sign(j, L):
    L["drv"][:, _lib.SLOT_K2]["constant"] *= -1.0


def test_second_order_anisotropy_reduces_to_the_polar_ode():
    assert _k2_run(OracleJunction) < 1e-8
    for p in (_k2_coefficient_2, _k2_sign):
        assert _k2_run(planted(p)) > 1e-3, p.__name__


# ---------------------------------------------------------------------------
# STRIP resistance recomputed from the logged magnetisations

STRIP = dict(Rx0=[100.0, 80.0, 120.0], Ry0=[3.0, 2.0, 5.0], AMR_X=[1.5, 0.8, 2.0], AMR_Y=[0.7, 1.1, 0.4],
             SMR_X=[-0.4, 0.3, 0.6], SMR_Y=[0.2, -0.5, 0.9], AHE=[2.5, -1.0, 4.0])


def _strip_junction(J, params):
    ls = []
    for q, m in enumerate([(0.3, 0.8, 0.5), (-0.6, 0.2, 0.7), (0.1, -0.9, 0.4)]):
        l = Layer(f"l{q}", mag=CVector(*m), anis=CVector(0, 0, 1), Ms=1.0 + 0.2 * q, thickness=1e-9,
                  cellSurface=1e-16, demagTensor=ZERO3, damping=0.03)
        ls.append(l)
    j = J(ls, params["Rx0"], params["Ry0"], params["AMR_X"], params["AMR_Y"], params["SMR_X"], params["SMR_Y"],
          params["AHE"])
    j.setLayerExternalFieldDriver("all", AxialDriver(5e4, -2e4, 3e4))
    j.setIECDriver("l0", "l1", constantDriver(2e-4))
    j.runSimulation(5e-10, 1e-13, 1e-12)
    return j.getLog()


def _strip_error(J, params=STRIP):
    from cmtj_amd.utils.resistance import calculate_resistance_parallel

    lg = _strip_junction(J, params)
    m = np.stack([_m(lg, f"l{q}").T for q in range(3)])  # [layer, 3, T]
    ones = [1.0] * 3
    rx, _ = calculate_resistance_parallel(params["Rx0"], params["Ry0"], params["AMR_X"], params["AHE"],
                                          params["SMR_X"], m, l=ones, w=ones)
    _, ry = calculate_resistance_parallel(params["Rx0"], params["Ry0"], params["AMR_Y"], params["AHE"],
                                          params["SMR_Y"], m, l=ones, w=ones)
    assert set(lg["Rz"]) == {0.0}
    return max(float(np.max(np.abs(np.asarray(lg["Rx"]) - rx))), float(np.max(np.abs(np.asarray(lg["Ry"]) - ry))))


def _swap_amr_smr_x(j, L):
    j["AMR_X"], j["SMR_X"] = j["SMR_X"].copy(), j["AMR_X"].copy()


def _ahe_without_half(j, L):
    j["AHE"] = j["AHE"] * 2.0


def test_strip_resistance_matches_the_parallel_resistor_model():
    assert _strip_error(OracleJunction) < 1e-11
    for p in (_swap_amr_smr_x, _ahe_without_half):
        assert _strip_error(planted(p)) > 1e-4, p.__name__


def test_strip_transverse_resistance_of_an_open_circuit_is_zero():
    """All-zero transverse coefficients: 1/(sum of 1/0) = 1/inf = 0."""
    params = dict(STRIP, Ry0=[0.0] * 3, AMR_Y=[0.0] * 3, SMR_Y=[0.0] * 3, AHE=[0.0] * 3)
    lg = _strip_junction(OracleJunction, params)
    assert set(lg["Ry"]) == {0.0} and all(np.isfinite(lg["Rx"]))


# ---------------------------------------------------------------------------
# coupled Heun: corrector neighbours from the predictor (second order)

def _coupled_final(J, dt, solver):
    a = Layer("a", mag=CVector(1, 0.3, 0.2), anis=CVector(1, 0, 0), Ms=1.0, thickness=1e-9, cellSurface=1e-16,
              demagTensor=ZERO3, damping=0.05)
    b = Layer("b", mag=CVector(0.2, 1, 0.1), anis=CVector(0, 1, 0), Ms=1.0, thickness=1e-9, cellSurface=1e-16,
              demagTensor=ZERO3, damping=0.05)
    for l in (a, b):  # the constructor keeps m unnormalised, which would add an O(dt) first step
        l.setMagnetisation(l.mag)
    j = J([a, b])
    j.setIECDriver("a", "b", constantDriver(5e-4))
    j.setLayerAnisotropyDriver("all", constantDriver(1e4))
    # powers of two: totalTime / timeStep is exact, so every run ends at the same time
    T = 2.0**-33
    j.runSimulation(T, dt, T, solverMode=solver)
    return np.array(j.getLayerMagnetisation("a").tolist() + j.getLayerMagnetisation("b").tolist())


def _order(J, solver):
    ref = _coupled_final(J, 2.0**-49, solver)
    e = [np.linalg.norm(_coupled_final(J, dt, solver) - ref) for dt in (2.0**-41, 2.0**-42, 2.0**-43)]
    return [math.log2(e[0] / e[1]), math.log2(e[1] / e[2])]


def _frozen_neighbours(j, L):
    j["solver"] = int(SolverMode.RK4)  # the RK4 runner's rule: neighbours held at the step start


def test_coupled_heun_is_second_order():
    orders = _order(OracleJunction, SolverMode.Heun)
    assert all(abs(o - 2.0) < 0.3 for o in orders), orders
    frozen = _order(planted(_frozen_neighbours), SolverMode.Heun)
    assert frozen[-1] < 1.4, frozen  # heading to first order: the check above fails
