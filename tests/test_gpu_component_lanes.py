"""The one-component-per-lane kernels (anchored binary32 RK4, and the
binary64 variant) against the oracle (binary64): workloads and multilayer
cases of the lean feature set,
the magnetoresistance channels, the NaN semantics of a multilayer step and
the automatic lane choice."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import parity_cases
from cmtj_amd import (AxialDriver, CVector, Junction, Layer, NullDriver, Reference, SolverMode, _lib,
                      constantDriver, sineDriver, stepDriver, workloads)
from cmtj_amd.batch import JunctionBatch, _channel_names

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-4   # max |dm| of the FP32-bulk mode against FP64


def _chain(n_layers, mr="strip"):
    """n layers coupled by bilinear IEC, constant and step Oersted drives;
    mr: "strip", "classic" (bilayers) or "none"."""
    S, D = parity_cases.SURF, parity_cases.DEMAG
    layers = [Layer(f"L{q}", mag=CVector(1, 0.1 * (q + 1), 0.05 * q), anis=CVector(1, 0, 0), Ms=1.2 + 0.1 * q,
                    thickness=(1 + 0.5 * q) * 1e-9, cellSurface=S, demagTensor=D, damping=0.05)
              for q in range(n_layers)]
    if mr == "strip":
        j = Junction(layers, [100.0 + q for q in range(n_layers)], [50.0 + q for q in range(n_layers)],
                     [1.0] * n_layers, [1.0] * n_layers, [0.5] * n_layers, [0.5] * n_layers, [2.0] * n_layers)
    elif mr == "classic":
        j = Junction(layers, 100.0, 200.0)
    else:
        j = Junction(layers)
    for q in range(n_layers):
        j.setLayerAnisotropyDriver(f"L{q}", constantDriver(20e3 + 10e3 * q))
    for q in range(n_layers - 1):
        j.setIECDriver(f"L{q}", f"L{q + 1}", constantDriver(-2e-4 * (1 + 0.3 * q)))
    j.setLayerExternalFieldDriver("all", AxialDriver(200e3, 20e3, 5e3))
    j.setLayerOerstedFieldDriver("all", AxialDriver(NullDriver(), stepDriver(0, 2e3, 0, 5e-12), NullDriver()))
    return j


def _trilayer_sine():
    """Three layers, bilinear IEC, a sine and a step drive on different axes
    (the generic lean kernel: drive kinds resolved at run time)."""
    j = _chain(3, mr="none")
    j.setLayerExternalFieldDriver("all", AxialDriver(sineDriver(50e3, 20e3, 3e9, 0.3), constantDriver(0),
                                                     constantDriver(100e3)))
    return j


def _batch(name, precision="fp32"):
    if name == "pimm":
        return workloads.pimm(n_points=16, t_min=0.2e-9, t_max=0.5e-9, precision=precision, log="full",
                              log_dtype="f64")
    if name == "pimm_delayed":
        b = workloads.pimm(n_points=16, t_min=0.2e-9, t_max=0.5e-9, precision=precision, log="mx", log_dtype="f64")
        n = len(b.js)
        b.set_driver("all", "oersted", "t_start", np.full((n, 3), 5.05e-12))
        b.set_driver("all", "oersted", "t_stop", np.full((n, 3), 2.0e-11))
        return b
    if name == "r_h_loop":
        b = workloads.r_h_loop(n_fields=11, total_time=0.5e-9, precision=precision, log="full")
        b.log_dtype = _lib.LOG_F64
        return b
    if name == "vsd":
        b = JunctionBatch.replicate(workloads.vsd_junction(), 8, 1e-9, 1e-13, 1e-12, SolverMode.RK4,
                                    precision=precision, log="full", log_dtype="f64")
        f = np.zeros((8, 3))
        f[:, 1] = np.linspace(1e9, 20e9, 8)
        b.set_driver("free", "oersted", "frequency", f)
        return b
    if name == "rh_strip":
        j, kw = parity_cases.rh_strip(Junction)
        return JunctionBatch.from_junctions([j] * 5, 0.5e-9, 1e-13, 1e-12, SolverMode.RK4, precision=precision,
                                            log="full", log_dtype="f64")
    if name == "chain2_strip":
        return JunctionBatch.from_junctions([_chain(2)] * 3, 0.5e-9, 1e-13, 1e-12, SolverMode.RK4,
                                            precision=precision, log="full", log_dtype="f64")
    if name == "chain4_strip":
        return JunctionBatch.from_junctions([_chain(4)] * 3, 0.5e-9, 1e-13, 1e-12, SolverMode.RK4,
                                            precision=precision, log="full", log_dtype="f64")
    if name == "trilayer_sine":
        return JunctionBatch.from_junctions([_trilayer_sine()] * 3, 1e-9, 1e-13, 3e-13, SolverMode.RK4,
                                            precision=precision, log="full", log_dtype="f64")
    raise KeyError(name)


def _channels(b, log, i):
    j = b.js[i]
    names = _channel_names(b.metas[b.meta_index[i]], int(j["log_mask"]))
    ne, off = int(j["n_entries"]), int(j["log_offset"])
    return {k: np.asarray(log[off + c * ne: off + (c + 1) * ne], dtype=np.float64) for c, k in enumerate(names)}


@pytest.mark.parametrize("name", ["pimm", "pimm_delayed", "r_h_loop", "vsd", "rh_strip", "chain2_strip",
                                  "chain4_strip", "trilayer_sine"])
def test_component_lanes_match_oracle(name):
    b = _batch(name).set_lanes("per_component")
    kname = b.plan().info.kernel_name.decode()
    print(f"[cl] {name}: {kname}", flush=True)
    assert "k_junction_cl" in kname
    r = b.run()
    olog, omag, ost = oracle.run_batch(b.js, b.layers, _lib.constants_get(), b.tables, b.log_elements)
    st = r.status()
    assert (st["code"] == 0).all()
    assert (st["steps_done"] == ost["steps_done"]).all() and (st["entries_done"] == ost["entries_done"]).all()
    glog = r.log.cpu().numpy()
    wm, wr = 0.0, 0.0
    for i in range(len(b.js)):
        g, o = _channels(b, glog, i), _channels(b, olog, i)
        assert g.keys() == o.keys()
        for k in o:
            d = float(np.max(np.abs(g[k] - o[k]))) if len(o[k]) else 0.0
            if "_m" in k:
                wm = max(wm, d)
            else:  # resistance channels, relative
                wr = max(wr, d / max(1.0, float(np.max(np.abs(o[k])))))
    nl = int(b.js[0]["n_layers"])
    fm = float(np.max(np.abs(r.final_m.cpu().numpy()[:, :nl] - omag[:, :nl])))
    print(f"[cl] {name}: max|dm| logs {wm:.3e}, final {fm:.3e}, R rel {wr:.3e}", flush=True)
    assert wm <= FP32_TOL and fm <= FP32_TOL and wr <= FP32_TOL


@pytest.mark.parametrize("name", ["pimm", "pimm_delayed", "r_h_loop", "vsd", "rh_strip", "chain2_strip",
                                  "chain4_strip", "trilayer_sine"])
def test_fp64_component_lanes_match_oracle(name):
    """The binary64 variant (every RK4 stage in binary64): the drop-in's
    tolerance against the oracle."""
    b = _batch(name, precision="fp64").set_lanes("per_component")
    kname = b.plan().info.kernel_name.decode()
    print(f"[cl64] {name}: {kname}", flush=True)
    assert kname.startswith("k_junction_cl<double")
    r = b.run()
    olog, omag, ost = oracle.run_batch(b.js, b.layers, _lib.constants_get(), b.tables, b.log_elements)
    st = r.status()
    assert (st["code"] == 0).all()
    assert (st["steps_done"] == ost["steps_done"]).all() and (st["entries_done"] == ost["entries_done"]).all()
    glog = r.log.cpu().numpy()
    worst = 0.0
    for i in range(len(b.js)):
        g, o = _channels(b, glog, i), _channels(b, olog, i)
        assert g.keys() == o.keys()
        for k in o:
            if len(o[k]):
                worst = max(worst, float(np.max(np.abs(g[k] - o[k]))) / max(1.0, float(np.max(np.abs(o[k])))))
    nl = int(b.js[0]["n_layers"])
    fm = float(np.max(np.abs(r.final_m.cpu().numpy()[:, :nl] - omag[:, :nl])))
    print(f"[cl64] {name}: max relative difference {worst:.3e}, final |dm| {fm:.3e}", flush=True)
    assert worst <= 1e-8 and fm <= 1e-8


def _stt_constant(alternative):
    """STT at a constant current against an unnormalised fixed polariser
    (kappa, beta, alternative form), anisotropy and constant field."""
    l = Layer.createSTTLayer("free", mag=CVector(0.1, 0.2, 1.0), anis=CVector(0, 0, 1.0), Ms=1.0, thickness=1.4e-9,
                             cellSurface=7e-10 * 7e-10, demagTensor=parity_cases.DEMAG, damping=0.05,
                             SlonczewskiSpacerLayerParameter=1.5, beta=0.3, spinPolarisation=0.6)
    l.setReferenceLayer(CVector(0, 1.0, 1.0))
    l.setKappa(0.9)
    l.setAlternativeSTT(alternative)
    j = Junction([l], 100, 200)
    j.setLayerAnisotropyDriver("free", constantDriver(300e3))
    j.setLayerCurrentDriver("free", constantDriver(-6e9))
    j.setLayerExternalFieldDriver("all", AxialDriver(5e3, 1e3, 2e3))
    return j


def _sot_constant(dynamic):
    """SOT at a constant current (static coefficients) or with constant
    field-like / damping-like drivers (dynamic), two layers with IEC."""
    kw = dict(mag=CVector(0.001, 0.99, 0.1), anis=CVector(0, 0, 1.0), Ms=1.45, thickness=1.7e-9, cellSurface=1e-16,
              demagTensor=parity_cases.DEMAG, damping=0.0064)
    if dynamic:  # the reference refuses dynamic drivers on a static SOT layer
        a = Layer("free", **kw)
    else:
        a = Layer.createSOTLayer("free", dampingLikeTorque=0.000371 * 795774.715459 / 6e10,
                                 fieldLikeTorque=-0.000433 * 795774.715459 / 6e10, **kw)
    b = Layer("pinned", mag=CVector(0, 0, 1), anis=CVector(0, 0, 1), Ms=1.2, thickness=2e-9, cellSurface=1e-16,
              demagTensor=parity_cases.DEMAG, damping=0.02)
    j = Junction([a, b], 100, 200)
    j.setLayerAnisotropyDriver("free", constantDriver(0.3e6))
    j.setLayerAnisotropyDriver("pinned", constantDriver(1e6))
    j.setIECDriver("free", "pinned", constantDriver(-1e-4))
    j.setLayerReferenceLayer("free", CVector(0, 1, 0))
    if dynamic:
        j.setLayerPrimaryTorqueDrivers("free", constantDriver(-2e3), constantDriver(5e3))
    else:
        j.setLayerCurrentDriver("free", constantDriver(0.08e12))
    return j


def _exotic(dynamic):
    """Full demag tensor, second-order anisotropy and biquadratic IEC on two
    layers (constant, or time-dependent second-order and biquadratic terms)."""
    N = [CVector(0.02, 0.01, 0.0), CVector(0.01, 0.03, 0.005), CVector(0.0, 0.005, 0.95)]
    a = Layer("free", mag=CVector(0.3, 0.2, 0.9), anis=CVector(0.1, 0, 1.0), Ms=1.1, thickness=1.2e-9,
              cellSurface=1e-16, demagTensor=N, damping=0.02)
    b = Layer("ref", mag=CVector(0.9, 0.1, 0.2), anis=CVector(1, 0, 0), Ms=1.4, thickness=2e-9, cellSurface=1e-16,
              demagTensor=N, damping=0.03)
    j = Junction([a, b], 100, 200)
    j.setLayerAnisotropyDriver("free", constantDriver(4e5))
    j.setLayerAnisotropyDriver("ref", constantDriver(1e5))
    j.setIECDriver("free", "ref", constantDriver(-1e-4))
    if dynamic:
        j.setLayerSecondOrderAnisotropyDriver("free", sineDriver(-2e4, 1e4, 0.8e9, 0.1))
        j.setQuadIECDriver("free", "ref", sineDriver(-3e-5, 1e-5, 1.1e9, 0.0))
    else:
        j.setLayerSecondOrderAnisotropyDriver("free", constantDriver(-2e4))
        j.setQuadIECDriver("free", "ref", constantDriver(-3e-5))
    j.setLayerExternalFieldDriver("all", AxialDriver(2e4, 1e4, 5e4))
    return j


def _iec_dynamic():
    j = _chain(2, mr="classic")
    j.setIECDriver("L0", "L1", sineDriver(-2e-4, 5e-5, 1.3e9, 0.2))
    j.setLayerAnisotropyDriver("L1", sineDriver(30e3, 10e3, 0.7e9, 0.0))
    return j


@pytest.mark.parametrize("name", ["sto", "stt_fixed", "stt_alternative", "stt_ref_top", "stt_ref_bottom",
                                  "sot_static", "sot_dynamic", "stt_drivers", "stt_drivers_alternative",
                                  "sot_step_current", "iec_dynamic", "exotic", "exotic_dynamic", "bilayer_iec2",
                                  "exotic_terms", "sot_secondary", "callback_drivers"])
def test_fp64_component_lanes_spin_transfer_torque(name):
    """Spin torques and time-dependent scalar drives in the binary64
    component-lane kernels: STT with fixed, top and bottom polarisers and
    both eta forms, SOT with static and dynamic coefficients, sine and step
    currents, trapezoid and sine anisotropy, a sine IEC coupling, against the
    oracle at the drop-in's tolerance."""
    if name == "sto":
        b = workloads.sto(n_currents=8, total_time=0.5e-9, precision="fp64", log="full")
        b.log_dtype = _lib.LOG_F64
    else:
        if name.startswith("stt_ref"):
            j, _ = parity_cases.stt_neighbour_reference(Junction, Reference.top if name.endswith("top")
                                                        else Reference.bottom)
        elif name in ("sot_static", "sot_dynamic"):
            j = _sot_constant(name == "sot_dynamic")
        elif name.startswith("stt_drivers"):
            j, _ = parity_cases.stt_drivers(Junction, alternative=name.endswith("alternative"))
        elif name == "sot_step_current":
            j, _ = parity_cases.sot_static(Junction)
        elif name == "iec_dynamic":
            j = _iec_dynamic()
        elif name.startswith("exotic"):
            j = _exotic(name.endswith("dynamic"))
        elif name == "bilayer_iec2":
            j, _ = parity_cases.bilayer_iec(Junction)
        elif name == "exotic_terms":
            j, _ = parity_cases.exotic_terms(Junction)
        elif name == "sot_secondary":
            j, _ = parity_cases.sot_dynamic(Junction)
        elif name == "callback_drivers":
            j, _ = parity_cases.custom_driver(Junction)
        else:
            j = _stt_constant(name == "stt_alternative")
        b = JunctionBatch.from_junctions([j] * 3, 1e-9, 1e-13, 1e-12, SolverMode.RK4, precision="fp64", log="full",
                                         log_dtype="f64")
    b.set_lanes("per_component")
    kname = b.plan().info.kernel_name.decode()
    print(f"[cl64-stt] {name}: {kname}", flush=True)
    assert kname.startswith("k_junction_cl<double")
    r = b.run()
    olog, omag, ost = oracle.run_batch(b.js, b.layers, _lib.constants_get(), b.tables, b.log_elements)
    assert (r.status()["code"] == 0).all()
    glog = r.log.cpu().numpy()
    worst = 0.0
    for i in range(len(b.js)):
        g, o = _channels(b, glog, i), _channels(b, olog, i)
        for k in o:
            if len(o[k]):
                worst = max(worst, float(np.max(np.abs(g[k] - o[k]))) / max(1.0, float(np.max(np.abs(o[k])))))
    print(f"[cl64-stt] {name}: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8


@pytest.mark.parametrize("solver", [SolverMode.EulerHeun, SolverMode.Heun])
@pytest.mark.parametrize("name", ["sot_step_300K", "chain2_300K", "sto_300K", "stt_sine_T_ramp", "stt_ref_top_300K",
                                  "chain2"])
def test_fp64_component_lanes_stochastic_solvers(name, solver):
    """Stochastic runs (Euler-Heun, Heun) in the binary64 component-lane
    kernels: the engine and the oracle share the Philox stream, so the paths
    agree sample by sample; a constant temperature on one layer or all, a
    temperature that changes in time, a polariser taken from the (predicted)
    neighbour, and a deterministic run."""
    if name == "sot_step_300K":
        j, _ = parity_cases.sot_static(Junction)
        j.getLayer("free").cellSurface = 1e-16
        j.getLayer("free").cellVolume = 1e-16 * j.getLayer("free").thickness
        j.setLayerTemperatureDriver("all", constantDriver(300.0))
    elif name == "chain2_300K":
        j = _chain(2, mr="classic")
        j.setLayerTemperatureDriver("L1", constantDriver(300.0))
    elif name == "sto_300K":
        j = workloads.sto_junction()
        j.setLayerCurrentDriver("free", constantDriver(-1e11))
        j.setLayerTemperatureDriver("all", constantDriver(300.0))
    elif name == "stt_ref_top_300K":
        j, _ = parity_cases.stt_neighbour_reference(Junction, Reference.top)
        j.setLayerTemperatureDriver("all", constantDriver(300.0))
    elif name == "chain2":
        j = _chain(2, mr="classic")
    else:
        j, _ = parity_cases.stt_drivers(Junction)
        j.setLayerTemperatureDriver("all", sineDriver(200.0, 100.0, 1e9, 0.0))
    j.setNoiseSeed(1234)
    b = JunctionBatch.from_junctions([j] * 3, 0.5e-9, 1e-13, 1e-12, solver, precision="fp64", log="full",
                                     log_dtype="f64")
    b.set_lanes("per_component")
    kname = b.plan().info.kernel_name.decode()
    print(f"[cl64-eh] {name} {solver.name}: {kname}", flush=True)
    assert kname.startswith("k_junction_cl<double") and f",{solver.name}," in kname
    r = b.run()
    olog, omag, ost = oracle.run_batch(b.js, b.layers, _lib.constants_get(), b.tables, b.log_elements)
    glog = r.log.cpu().numpy()
    worst = 0.0
    for i in range(len(b.js)):
        g, o = _channels(b, glog, i), _channels(b, olog, i)
        for k in o:
            if len(o[k]):
                worst = max(worst, float(np.max(np.abs(g[k] - o[k]))) / max(1.0, float(np.max(np.abs(o[k])))))
    print(f"[cl64-eh] {name} {solver.name}: max relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-8


@pytest.mark.parametrize("bad_layer", [0, 1])
def test_component_lanes_nan_commits_like_the_reference(bad_layer):
    """A NaN in layer b stops the run: layers up to and including it take
    the step's value, later layers keep the previous one."""
    js = []
    for q in range(3):
        j = _chain(2, mr="classic")
        if q == 1:
            j.setLayerExternalFieldDriver(f"L{bad_layer}", AxialDriver(stepDriver(0, float("inf"), 5e-12, 1e-9),
                                                                       constantDriver(0), constantDriver(1e5)))
        js.append(j)
    b = JunctionBatch.from_junctions(js, 1e-10, 1e-13, 1e-13, SolverMode.RK4, precision="fp32", log="full",
                                     log_dtype="f64").set_lanes("per_component")
    assert "k_junction_cl" in b.plan().info.kernel_name.decode()
    r = b.run()
    st = r.status()
    olog, omag, ost = oracle.run_batch(b.js, b.layers, _lib.constants_get(), b.tables, b.log_elements)
    for i in range(3):
        for f in ("code", "steps_done", "entries_done", "nan_layer"):
            assert int(st[i][f]) == int(ost[i][f]), (i, f)
    assert int(st[1]["code"]) == _lib.ERR_NAN and int(st[1]["nan_layer"]) == bad_layer
    gm = r.final_m.cpu().numpy()
    assert np.array_equal(np.isnan(gm[:, :2]), np.isnan(omag[:, :2]))
    ok = np.isfinite(omag[:, :2])
    assert float(np.max(np.abs(gm[:, :2][ok] - omag[:, :2][ok]))) <= FP32_TOL


def test_pimm_log_layout_kernel_matches_oracle():
    """The PIMM sweep's kernel with its log layout compiled in (m_x of both
    layers and the classic resistance as binary32 entries) against the
    oracle: the same entries as the general layout, at the binary32 mode's
    tolerance (R relative, as the log entries are binary32)."""
    b = workloads.pimm(n_points=16, t_min=0.2e-9, t_max=0.5e-9, precision="fp32", log="mx", log_dtype="f32")
    kname = b.plan().info.kernel_name.decode()
    print(f"[cl] pimm log layout: {kname}", flush=True)
    assert "log m_x+R f32" in kname
    r = b.run()
    olog, omag, ost = oracle.run_batch(b.js, b.layers, _lib.constants_get(), b.tables, b.log_elements)
    st = r.status()
    assert (st["entries_done"] == ost["entries_done"]).all()
    glog = r.log.cpu().numpy().astype(np.float64)
    wm = wr = 0.0
    for i in range(len(b.js)):
        g, o = _channels(b, glog, i), _channels(b, olog, i)
        assert g.keys() == o.keys() and len(o) == 3
        for k in o:
            d = float(np.max(np.abs(g[k] - o[k])))
            if "_m" in k:
                wm = max(wm, d)
            else:
                wr = max(wr, d / float(np.max(np.abs(o[k]))))
    print(f"[cl] pimm log layout: max|dm| {wm:.3e}, R rel {wr:.3e}", flush=True)
    assert wm <= FP32_TOL and wr <= FP32_TOL


def test_component_lanes_chosen_automatically_on_small_batches():
    small = workloads.pimm(n_points=64, t_min=1e-11, t_max=2e-11, precision="fp32", log="mx")
    assert "k_junction_cl" in small.plan().info.kernel_name.decode()
    assert int(small.plan().info.lanes_per_instance) == 8
    big = workloads.pimm(n_points=16384, t_min=1e-11, t_max=2e-11, precision="fp32", log="mx")
    assert "k_junction_ll" in big.plan().info.kernel_name.decode()
    fp64 = workloads.pimm(n_points=64, t_min=1e-11, t_max=2e-11, precision="fp64", log="mx")
    assert fp64.plan().info.kernel_name.decode().startswith("k_junction_cl<double")


def test_component_lanes_sliced_run_equals_uninterrupted():
    """Resumable slices of the component-lane kernels are bit-identical to an
    uninterrupted run."""
    js = [_chain(2), _chain(4), _trilayer_sine(), parity_cases.rh_strip(Junction)[0]]
    mk = lambda: JunctionBatch.from_junctions(js, [0.5e-9, 0.4e-9, 0.3e-9, 0.6e-9], 1e-13, 1e-12, SolverMode.RK4,
                                              precision="fp32", log="full").set_lanes("per_component")
    a = mk()
    kname = a.plan().info.kernel_name.decode()
    assert kname.startswith("k_junction_cl") and "k_junction_ll" not in kname and "k_junction<" not in kname
    ra = a.run()
    rb = mk().run_sliced(slice_steps=777, n_streams=3)
    la, lb = ra.log.cpu().numpy(), rb.log.cpu().numpy()
    print(f"[cl slices] max log difference {np.max(np.abs(la - lb)):.3e}", flush=True)
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(ra.final_m.cpu().numpy(), rb.final_m.cpu().numpy())
    sa, sb = ra.status(), rb.status()
    assert (sa["steps_done"] == sb["steps_done"]).all() and (sa["entries_done"] == sb["entries_done"]).all()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("w", [1, 3])
def test_component_lanes_nan_in_step_passes(precision, w):
    """The step loop runs in wave-uniform passes and a NaN ends a pass: the
    lanes of the instance it stops keep the reference's state, the other
    lanes take that step again.  Eight bilayers share one wave, with uneven
    lengths; NaNs come at different steps and layers, one at the first step
    of a pass (the step where another instance's run ends) and one at an
    instance's last step.  Every instance's status, final m and log entries
    (every step, w = 1, and every third step) match the oracle."""
    dt = 1e-13
    n_steps = [400, 350, 420, 200, 380, 410, 300, 390]
    nan_at = {1: (0, 50), 2: (1, 51), 5: (0, 200), 6: (1, 299)}  # instance: (layer, step)
    js = []
    for q in range(8):
        j = _chain(2, mr="classic")
        lay, s = nan_at.get(q, (0, 10**6))
        amp = float("inf") if q in nan_at else 0.0
        # the same driver kinds on every instance (one structure group, one
        # wave); the x field turns infinite a quarter step into step s, so
        # the step's midpoint stages see it
        j.setLayerExternalFieldDriver(f"L{lay}", AxialDriver(stepDriver(200e3, amp, (s + 0.25) * dt, 1e-6),
                                                             constantDriver(20e3), constantDriver(5e3)))
        j.setLayerExternalFieldDriver(f"L{1 - lay}", AxialDriver(stepDriver(200e3, 0.0, 1.0, 2.0),
                                                                 constantDriver(20e3), constantDriver(5e3)))
        js.append(j)
    T = np.array(n_steps, dtype=np.float64) * dt
    b = JunctionBatch.from_junctions(js, T, dt, w * dt, SolverMode.RK4, precision=precision, log="full",
                                     log_dtype="f64").set_lanes("per_component")
    info = b.plan().info
    print(f"[cl] nan passes {precision} w={w}: {info.kernel_name.decode()}, launches {info.n_kernels}", flush=True)
    assert "k_junction_cl" in info.kernel_name.decode() and info.n_kernels == 1
    r = b.run()
    st = r.status()
    olog, omag, ost = oracle.run_batch(b.js, b.layers, _lib.constants_get(), b.tables, b.log_elements)
    tol = FP32_TOL if precision == "fp32" else 1e-8
    glog = r.log.cpu().numpy()
    gm = r.final_m.cpu().numpy()
    for i in range(8):
        for f in ("code", "steps_done", "entries_done", "nan_layer"):
            assert int(st[i][f]) == int(ost[i][f]), (i, f, int(st[i][f]), int(ost[i][f]))
        if i in nan_at:
            assert int(st[i]["code"]) == _lib.ERR_NAN and int(st[i]["nan_layer"]) == nan_at[i][0]
            assert int(st[i]["steps_done"]) == nan_at[i][1]
        else:
            # (the step count truncates totalTime / timeStep in binary64, as
            # the reference does: 380 * dt / dt gives 379)
            assert int(st[i]["code"]) == _lib.OK and n_steps[i] - 1 <= int(st[i]["steps_done"]) <= n_steps[i]
        ne = int(ost[i]["entries_done"])
        g, o = _channels(b, glog, i), _channels(b, olog, i)
        for k in o:
            d = float(np.max(np.abs(g[k][:ne] - o[k][:ne]) / np.maximum(1.0, np.abs(o[k][:ne]))))
            assert d <= tol, (i, k, d)
    assert np.array_equal(np.isnan(gm[:, :2]), np.isnan(omag[:, :2]))
    ok = np.isfinite(omag[:, :2])
    assert float(np.max(np.abs(gm[:, :2][ok] - omag[:, :2][ok]))) <= tol
