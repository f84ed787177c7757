"""Batched GPU runs: FP64 batches against the oracle, the FP32-bulk mode
against FP64, uneven/ragged batches, per-instance failure isolation, and
the thermal path (same-seed trajectories and Boltzmann statistics)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
import parity_cases
from cmtj_amd import (AxialDriver, CVector, Junction, Layer, SolverMode, _lib, constantDriver, stepDriver,
                      workloads)
from cmtj_amd.batch import JunctionBatch, _channel_names, run_batch
from oracle_backend import OracleJunction

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-4   # max |dm| of the FP32-bulk mode against FP64


def _oracle_blocks(b):
    log, mag, st = oracle.run_batch(b.js, b.layers, _lib.constants_get(), b.tables, b.log_elements)
    return log, mag, st


def _m_channels(b, log, i):
    j = b.js[i]
    names = _channel_names(b.metas[b.meta_index[i]], int(j["log_mask"]))
    ne, off = int(j["n_entries"]), int(j["log_offset"])
    return {k: np.asarray(log[off + c * ne: off + (c + 1) * ne], dtype=np.float64) for c, k in enumerate(names)}


@pytest.mark.parametrize("lanes", ["one", "per_layer"])
def test_heterogeneous_fp64_batch_matches_oracle(lanes):
    names = list(parity_cases.CASES)
    names.remove("energies")  # energy channels are a drop-in feature
    js, kws = [], []
    for n in names:
        j, kw = parity_cases.CASES[n](Junction)
        j.setNoiseSeed(77)
        js.append(j)
        kws.append(kw)
    b = JunctionBatch.from_junctions(js, [k["totalTime"] for k in kws], [k["timeStep"] for k in kws],
                                     [k["writeFrequency"] for k in kws],
                                     [k.get("solverMode", SolverMode.RK4) for k in kws],
                                     precision="fp64", log="full", log_dtype="f64").set_lanes(lanes)
    print(f"[batch] {lanes}: {len(js)} instances, {b.plan().info.n_kernels} groups: "
          f"{b.plan().info.kernel_name.decode()}", flush=True)
    r = b.run()
    glog = r.log.cpu().numpy()
    olog, omag, ost = _oracle_blocks(b)
    st = r.status()
    for i, n in enumerate(names):
        assert int(st[i]["code"]) == 0, n
        g, o = _m_channels(b, glog, i), _m_channels(b, olog, i)
        for k in o:
            d = float(np.max(np.abs(g[k] - o[k]))) if o[k].size else 0.0
            tol = 1e-8 if k.endswith(("_mx", "_my", "_mz")) else 1e-8 * max(1.0, float(np.max(np.abs(o[k]))))
            assert d <= tol, (n, k, d)
    np.testing.assert_allclose(r.final_m.cpu().numpy(), omag, atol=1e-8)


def _vsd_sweep(p, n=64):
    b = JunctionBatch.replicate(workloads.vsd_junction(), n, 2e-9, 1e-13, 1e-12, SolverMode.RK4, precision=p,
                                log="full")
    f = np.zeros((n, 3))
    f[:, 1] = np.linspace(1e9, 20e9, n)
    b.set_driver("free", "oersted", "frequency", f)
    return b


@pytest.mark.parametrize("lanes", ["one", "per_layer"])
@pytest.mark.parametrize("builder", ["r_h_loop", "pimm", "sto", "vsd"])
def test_fp32_bulk_within_1e4_of_fp64(builder, lanes):
    if builder == "vsd":
        mk = _vsd_sweep
    elif builder == "r_h_loop":
        mk = lambda p: workloads.r_h_loop(n_fields=101, total_time=2e-9, precision=p, log="full")
    elif builder == "pimm":
        mk = lambda p: workloads.pimm(n_points=256, t_min=1e-9, t_max=3e-9, precision=p, log="mx",
                                      log_dtype="f64")
    else:
        mk = lambda p: workloads.sto(n_currents=128, total_time=5e-9, precision=p, log="full")
    b64, b32 = mk("fp64").set_lanes(lanes), mk("fp32").set_lanes(lanes)
    b32.log_dtype = _lib.LOG_F64
    b32._plan = None
    r64, r32 = b64.run(), b32.run()
    worst = 0.0
    for i in range(len(b64.js)):
        a, c = _m_channels(b64, r64.log.cpu().numpy(), i), _m_channels(b32, r32.log.cpu().numpy(), i)
        for k in a:
            if "_m" in k:
                worst = max(worst, float(np.max(np.abs(a[k] - c[k]))))
    fm = float((r64.final_m - r32.final_m).abs().max())
    print(f"[fp32] {builder} {lanes}: max|dm| over logs = {worst:.3e}, final = {fm:.3e}", flush=True)
    assert worst <= FP32_TOL and fm <= FP32_TOL


def _workload(name):
    if name == "pimm":
        return workloads.pimm(n_points=16, t_min=0.2e-9, t_max=0.5e-9, precision="fp64", log="mx", log_dtype="f64")
    if name == "pimm_delayed":  # the Oersted step switches on and off mid-run
        b = workloads.pimm(n_points=16, t_min=0.2e-9, t_max=0.5e-9, precision="fp64", log="mx", log_dtype="f64")
        n = len(b.js)
        b.set_driver("all", "oersted", "t_start", np.full((n, 3), 5.05e-12))
        b.set_driver("all", "oersted", "t_stop", np.full((n, 3), 2.0e-11))
        return b
    if name == "r_h_loop":
        return workloads.r_h_loop(n_fields=11, total_time=0.5e-9, precision="fp64")
    if name == "vsd":
        return _vsd_sweep("fp64", n=8)
    if name == "sto":
        return workloads.sto(n_currents=8, total_time=0.5e-9, precision="fp64")
    return workloads.cims_pbit(n_currents=4, n_seeds=4, total_time=0.1e-9, dt=1e-14, edge=0.02e-9,
                               plateau=0.03e-9, precision="fp64")


@pytest.mark.parametrize("lanes", ["one", "per_layer"])
@pytest.mark.parametrize("name", ["pimm", "pimm_delayed", "r_h_loop", "vsd", "sto", "cims"])
def test_workload_kernels_match_oracle(name, lanes):
    """The kernels specialised to the workloads' exact feature sets against
    the oracle (the CIMS ensemble pathwise: same Philox stream)."""
    b = _workload(name).set_lanes(lanes)
    r = b.run()
    print(f"[workload] {name} {lanes}: {b.plan().info.kernel_name.decode()}", flush=True)
    olog, omag, ost = _oracle_blocks(b)
    st = r.status()
    assert (st["code"] == 0).all() and (st["steps_done"] == ost["steps_done"]).all()
    glog = r.log.cpu().numpy()
    worst = 0.0
    for i in range(len(b.js)):
        g, o = _m_channels(b, glog, i), _m_channels(b, olog, i)
        for k in o:
            if "_m" in k:
                worst = max(worst, float(np.max(np.abs(g[k] - o[k]))))
    nl = int(b.js[0]["n_layers"])
    worst = max(worst, float(np.max(np.abs(r.final_m.cpu().numpy()[:, :nl] - omag[:, :nl]))))
    print(f"[workload] {name} {lanes}: max|dm| vs oracle {worst:.3e}", flush=True)
    assert worst <= 1e-8


def test_uneven_ragged_batch_and_failure_isolation():
    """Different layer counts, durations, time steps and solvers in one batch;
    one instance blows up (NaN) without touching the others."""
    js = []
    for n in ("rh_strip", "trilayer_rk4", "sot_static", "bilayer_iec_heun"):
        j, _ = parity_cases.CASES[n](Junction)
        js.append(j)
    bad = Layer("free", mag=CVector(1, 0, 0), anis=CVector(1, 0, 0), Ms=1.0, thickness=1e-9, cellSurface=1e-16,
                demagTensor=parity_cases.DEMAG, damping=0.01)
    jb = Junction([bad])
    jb.setLayerExternalFieldDriver("all", AxialDriver(stepDriver(0, float("inf"), 5e-12, 1e-9), constantDriver(0),
                                                      constantDriver(1e5)))
    js.append(jb)
    T = [1e-10, 3e-10, 2e-10, 1.5e-10, 1e-11]
    dt = [1e-13, 1e-13, 5e-13, 1e-13, 1e-13]
    sm = [SolverMode.RK4, SolverMode.RK4, SolverMode.RK4, SolverMode.Heun, SolverMode.RK4]
    b = JunctionBatch.from_junctions(js, T, dt, dt, sm, precision="fp64", log="full", log_dtype="f64")
    r = b.run()
    st = r.status()
    olog, omag, ost = _oracle_blocks(b)
    for i in range(len(js)):
        assert int(st[i]["code"]) == int(ost[i]["code"])
        assert int(st[i]["steps_done"]) == int(ost[i]["steps_done"])
        assert int(st[i]["entries_done"]) == int(ost[i]["entries_done"])
    assert int(st[4]["code"]) == _lib.ERR_NAN and int(st[4]["nan_layer"]) == 0
    for i in range(4):
        g, o = _m_channels(b, r.log.cpu().numpy(), i), _m_channels(b, olog, i)
        for k in o:
            if "_m" in k:
                assert float(np.max(np.abs(g[k] - o[k]))) <= 1e-8


@pytest.mark.parametrize("lanes", ["one", "per_layer"])
@pytest.mark.parametrize("bad_layer", [0, 1])
def test_multilayer_nan_commits_like_the_reference(bad_layer, lanes):
    """A NaN in layer b of a bilayer stops the run; the layers up to and
    including the first NaN layer take the step's value, later ones keep the
    previous one (runMultiLayerSolver updates layers in turn)."""
    js = []
    for q in range(3):
        j, _ = parity_cases.bilayer_iec(Junction, SolverMode.RK4)
        if q == 1:
            name = j.getLayerIds()[bad_layer]
            j.setLayerExternalFieldDriver(name, AxialDriver(stepDriver(0, float("inf"), 5e-12, 1e-9),
                                                            constantDriver(0), constantDriver(1e5)))
        js.append(j)
    b = JunctionBatch.from_junctions(js, 1e-10, 1e-13, 1e-13, SolverMode.RK4, precision="fp64", log="full",
                                     log_dtype="f64").set_lanes(lanes)
    r = b.run()
    st = r.status()
    olog, omag, ost = _oracle_blocks(b)
    for i in range(3):
        for f in ("code", "steps_done", "entries_done", "nan_layer"):
            assert int(st[i][f]) == int(ost[i][f]), (i, f)
    assert int(st[1]["code"]) == _lib.ERR_NAN and int(st[1]["nan_layer"]) == bad_layer
    gm = r.final_m.cpu().numpy()
    np.testing.assert_allclose(gm[:, :2], omag[:, :2], atol=1e-8, equal_nan=True)
    assert np.isnan(gm[1, bad_layer]).any()
    if bad_layer == 0:
        assert np.isfinite(gm[1, 1]).all()  # the later layer keeps its previous value
    for i in (0, 2):
        g, o = _m_channels(b, r.log.cpu().numpy(), i), _m_channels(b, olog, i)
        for k in o:
            if "_m" in k:
                assert float(np.max(np.abs(g[k] - o[k]))) <= 1e-8


def test_lane_mappings_agree():
    """One lane per instance and one lane per layer run the same arithmetic
    per layer (FMA contraction may still differ between the two kernels)."""
    for n in ("bilayer_iec_rk4", "trilayer_rk4", "bilayer_iec_heun", "trilayer_eulerheun", "stt_ref_top",
              "sot_dynamic_heun", "five_layer_chain"):
        outs = []
        for lanes in ("one", "per_layer"):
            j, kw = parity_cases.CASES[n](Junction)
            j.setNoiseSeed(5)
            b = JunctionBatch.from_junctions([j] * 3, kw["totalTime"], kw["timeStep"], kw["writeFrequency"],
                                             kw.get("solverMode", SolverMode.RK4), precision="fp64", log="full",
                                             log_dtype="f64").set_lanes(lanes)
            r = b.run()
            outs.append((r.log.cpu().numpy(), r.final_m.cpu().numpy(), b.plan().info.lanes_per_instance))
        print(f"[lanes] {n}: lanes/instance {outs[0][2]} vs {outs[1][2]}", flush=True)
        assert outs[0][2] == 1 and outs[1][2] > 1
        d = max(float(np.max(np.abs(outs[0][1] - outs[1][1]))), float(np.max(np.abs(outs[0][0] - outs[1][0]))))
        print(f"[lanes] {n}: max difference {d:.3e}", flush=True)
        np.testing.assert_allclose(outs[0][1], outs[1][1], rtol=0, atol=1e-12)
        np.testing.assert_allclose(outs[0][0], outs[1][0], rtol=1e-12, atol=1e-12)


def test_drop_in_nan_raises_like_the_reference():
    j = Junction([Layer("free", mag=CVector(1, 0, 0), anis=CVector(1, 0, 0), Ms=1.0, thickness=1e-9,
                        cellSurface=1e-16, demagTensor=parity_cases.DEMAG, damping=0.01)])
    j.setLayerExternalFieldDriver("all", AxialDriver(stepDriver(0, float("inf"), 5e-12, 1e-9), constantDriver(0),
                                                     constantDriver(1e5)))
    with pytest.raises(RuntimeError, match="NAN magnetisation"):
        j.runSimulation(1e-11, 1e-13, 1e-13)
    assert 0 < len(j.getLog()["time"]) < 100


def test_run_batch_equals_sequential_runs():
    a = [parity_cases.stt_drivers(Junction)[0] for _ in range(3)]
    s = [parity_cases.stt_drivers(Junction)[0] for _ in range(3)]
    logs = run_batch(a, 3e-10, 1e-13, 1e-12, precision="fp64")
    for j, lg in zip(s, logs):
        j.runSimulation(3e-10, 1e-13, 1e-12)
        ref = j.getLog()
        for k in ref:
            np.testing.assert_allclose(lg[k], ref[k], rtol=0, atol=1e-12)
    for ja, js_ in zip(a, s):
        assert max(abs(p - q) for p, q in zip(ja.getLayerMagnetisation("free").tolist(),
                                              js_.getLayerMagnetisation("free").tolist())) <= 1e-12


def test_stochastic_fp64_trajectory_matches_oracle_with_same_seed():
    """Same Philox stream on both sides: sLLGS trajectories agree pathwise."""
    for solver in (SolverMode.EulerHeun, SolverMode.Heun):
        jg, kw = parity_cases.bilayer_iec(Junction, solver)
        jo, _ = parity_cases.bilayer_iec(OracleJunction, solver)
        for j in (jg, jo):
            j.setLayerTemperatureDriver("all", constantDriver(300.0))
            j.setNoiseSeed(2024)
        kw = dict(kw, totalTime=5e-10)
        jg.runSimulation(**kw)
        jo.runSimulation(**kw)
        lg, lo = jg.getLog(), jo.getLog()
        d = max(float(np.max(np.abs(np.asarray(lg[k]) - np.asarray(lo[k])))) for k in lo if "_m" in k)
        print(f"[thermal] {solver.name}: pathwise max|dm| = {d:.3e}", flush=True)
        assert d <= 1e-8


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_thermal_equilibrium_on_gpu(precision):
    """<m_z> = coth x - 1/x for an isotropic macrospin (Euler-Heun ensemble)."""
    Ms, V, H, T = 1.0, 1e-25, 8.28e4, 300.0
    l = Layer("free", mag=CVector(0, 0, 1), anis=CVector(0, 0, 1), Ms=Ms, thickness=1e-9, cellSurface=V / 1e-9,
              demagTensor=[CVector(0, 0, 0)] * 3, damping=0.1)
    j = Junction([l])
    j.setLayerExternalFieldDriver("all", AxialDriver(0, 0, H))
    j.setLayerTemperatureDriver("all", constantDriver(T))
    n = 4096
    b = JunctionBatch.replicate(j, n, 4e-9, 1e-13, 1e-11, SolverMode.EulerHeun, precision=precision, log="m",
                                log_dtype="f64")
    b.set_seeds(np.arange(n, dtype=np.uint64) * 7919 + 11)
    r = b.run()
    ne = int(b.js[0]["n_entries"])
    mz = r.log.view(n, 3, ne)[:, 2, ne // 2:].mean().item()
    x = Ms * V * H / (1.380649e-23 * T)
    want = 1 / math.tanh(x) - 1 / x
    print(f"[thermal] {precision}: <mz> = {mz:.4f}, Langevin = {want:.4f}", flush=True)
    assert abs(mz - want) < 0.01


def test_zero_and_short_runs():
    """totalTime below one step: no step, no entry, m unchanged (the
    reference's truncation); a write frequency longer than the run logs only
    the first step."""
    j, kw = parity_cases.CASES["bilayer_iec_rk4"](Junction)
    m0 = [j.getLayerMagnetisation(l).tolist() for l in j.getLayerIds()]
    j.runSimulation(0.5e-13, 1e-13, 1e-13)
    assert j.getLog() == {}
    assert [j.getLayerMagnetisation(l).tolist() for l in j.getLayerIds()] == m0
    js = [parity_cases.CASES["bilayer_iec_rk4"](Junction)[0] for _ in range(3)]
    b = JunctionBatch.from_junctions(js, [0.0, 5e-13, 3e-12], 1e-13, [1e-13, 1e-11, 1e-12], SolverMode.RK4,
                                     precision="fp64", log="full", log_dtype="f64")
    r = b.run()
    st = r.status()
    assert list(st["steps_done"]) == [0, 5, 30] and list(st["entries_done"]) == [0, 1, 3]
    olog, omag, ost = _oracle_blocks(b)
    assert list(ost["steps_done"]) == [0, 5, 30] and list(ost["entries_done"]) == [0, 1, 3]
    np.testing.assert_allclose(r.final_m.cpu().numpy()[:, :2], omag[:, :2], atol=1e-12)


def test_fp32_plans_report_binary64_groups():
    """An FP32 batch says which of its groups run binary64 kernels: the STT
    oscillator on one lane per component has no binary32 kernel, the PIMM
    sweep does."""
    sto = workloads.sto(n_currents=8, total_time=0.05e-9, precision="fp32")
    pimm = workloads.pimm(n_points=8, t_min=0.05e-9, t_max=0.05e-9, precision="fp32")
    assert sto.fp64_groups == 1 and "double" in sto.plan().info.kernel_name.decode()
    assert pimm.fp64_groups == 0
    assert sto.set_lanes("per_layer").fp64_groups == 0
