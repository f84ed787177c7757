"""Benchmark: junction-steps/s of the FP32-bulk batched engine on the PIMM
sweep (4096 two-layer MTJ points, field 0..600 kA/m at 30 deg in-plane,
per-point duration U(2, 20) ns at dt = 1e-13 s, Hoe_y step excitation,
m_x of both layers and R logged every step), next to the FP64 drop-in's
per-instance launch, the FP32-vs-FP64 error, a roofline figure and a CPU
baseline (the FP64 oracle on host threads, bounded sample).

One "step" = one pass of every instance of the sweep (one plan execute).
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# The serving leg runs one stream per structure group; with 4 hardware queues
# (HIP's default, and what some environments export) a fifth group queues
# behind another group's slices.  HIP reads the count once, at its first call,
# which here comes before the serving leg: raise it now.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

FLOP_PER_STEP = 666.0  # algorithmic FLOP per junction-step of this 2-layer RK4 workload
PEAK_FP32 = 157.3      # MI355X published FP32 vector peak, TFLOP/s
PEAK_FP64 = 78.6


def _pimm(precision, n, log_dtype=None):
    from cmtj_amd import workloads

    return workloads.pimm(n_points=n, seed=0, precision=precision, log="mx", log_dtype=log_dtype)


COPY_BYTES = 1 << 30


def _pmc_pass(b) -> int:
    """Two sweep launches (the first loads the code object) and a plain
    device copy of COPY_BYTES, whose FETCH_SIZE / WRITE_SIZE calibrate the
    counters against a known byte count."""
    import torch

    b.execute()
    b.execute()
    x = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    y.copy_(x)
    torch.cuda.synchronize()
    return 0


def _traffic():
    """HBM bytes per sweep launch from the committed counter summary (two
    rocprofv3 --pmc runs of `bench.py --pmc-pass`, FETCH_SIZE then
    WRITE_SIZE, each scaled by the copy kernel's reading), or None."""
    f = ROOT / "profiles" / "pimm_fp32_traffic.json"
    if not f.exists():
        return None
    return json.loads(f.read_text())


def _serving(qps: float, seconds: float, seed: int = 0, lanes: str = "per_component") -> dict:
    """Mixed request stream (PIMM / VSD / CIMS / 3-4 layer chains, 1-50 ns,
    dt 1e-14..1e-13 s) with Poisson arrivals at `qps` for `seconds`, served
    by the continuous-batching loop in FP32; latency = submit to result."""
    import numpy as np

    from cmtj_amd import workloads
    from cmtj_amd.serving import Server

    rng = np.random.default_rng(seed + 1)
    mix = workloads.serving_mix(int(qps * seconds * 1.5) + 16, seed=seed)
    futs, kinds = [], []
    with Server(precision="fp32", slice_ms=20.0, n_streams=8, log_elements=1 << 28, lanes=lanes) as srv:
        t0 = time.perf_counter()
        nxt = t0
        q = 0
        while q < len(mix):
            now = time.perf_counter()
            if now - t0 >= seconds:
                break
            if now >= nxt:
                kind, j, T, dt, wf, sm, log = mix[q]
                futs.append(srv.submit(j, T, dt, wf, sm, log=log))
                kinds.append(kind)
                q += 1
                nxt += rng.exponential(1.0 / qps)
            else:
                time.sleep(min(nxt - now, 0.005))
        res = [f.result() for f in futs]
        ticks = srv.ticks
        host_s = dict(srv.host_s)
    lat = np.array([r.latency for r in res])
    steps = sum(r.steps for r in res)
    wall = max(r.t_done for r in res) - t0
    iso = np.array([_isolated_s(mix[q], lanes) for q in range(len(res))])
    slow = lat / iso
    by_kind = {}
    for k in sorted(set(kinds)):
        sel = np.array([kk == k for kk in kinds])
        by_kind[k] = {"n": int(sel.sum()), "p50": float(np.percentile(slow[sel], 50)),
                      "p99": float(np.percentile(slow[sel], 99))}
    return {"qps": qps, "arrival_window_s": seconds, "requests": len(res),
            "mix": {k: kinds.count(k) for k in sorted(set(kinds))},
            "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": float(lat.max())},
            "slowdown": {"p50": float(np.percentile(slow, 50)), "p99": float(np.percentile(slow, 99)),
                         "by_kind": by_kind,
                         "definition": "latency / the request's kernel time alone (its own junction, "
                                       "lane mapping and log, up to 20000 steps timed with HIP events, "
                                       "scaled to its step count)"},
            "junction_steps": int(steps), "junction_steps_per_s": steps / wall, "wall_s": wall,
            "slices": ticks, "slice_ms_target": 20.0, "host_s": host_s, "precision": "fp32", "lanes": lanes}


def _isolated_s(spec, lanes: str, cap: int = 20000) -> float:
    """Kernel time of one served request run alone: its own junction, lane
    mapping and log, truncated to `cap` steps, timed with HIP events on the
    launch stream and scaled to the request's full step count."""
    import torch

    from cmtj_amd.batch import JunctionBatch

    kind, j, T, dt, wf, sm, log = spec
    n_full = int(T / dt)
    Tc = min(T, cap * dt)
    b = JunctionBatch.from_junctions([j], Tc, dt, max(dt, min(wf, Tc)), sm, precision="fp32",
                                     log=log).set_lanes(lanes)
    b.plan()
    b._buffers("cuda")
    stream = torch.cuda.current_stream()
    b.execute(stream)  # warm-up
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    b.execute(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    n_c = max(1, int(b.js[0]["n_steps"]))
    return e0.elapsed_time(e1) * 1e-3 * n_full / n_c


def _timed_run(b, reps: int = 1):
    """Warm-up launch on a tiny slice of the batch, then `reps` timed
    launches of the whole batch (HIP events on the launch stream)."""
    import numpy as np
    import torch

    b.subset(np.arange(min(64, len(b.js)))).run()
    b.plan()  # descriptors uploaded before the timed region
    b._buffers("cuda")
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        r = b.execute(stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    return r, float(np.mean(ms))


def _bench_rh() -> dict:
    """R(H) loop, 101 fields x 10 ns, 1 layer, RK4 dt 1e-13, W = 10: the FP64
    drop-in run field by field (as the reference's scripts do), the same
    sweep as one FP64 and one FP32 launch, and the oracle on the host."""
    import numpy as np

    import oracle
    from cmtj_amd import AxialDriver, _lib, workloads

    b64 = workloads.r_h_loop(precision="fp64")
    r64, ms64 = _timed_run(b64)
    b32 = workloads.r_h_loop(precision="fp32")
    b32.log_dtype = _lib.LOG_F64
    r32, ms32 = _timed_run(b32)
    steps = b64.total_steps
    # drop-in: a few fields, one runSimulation each
    from cmtj_amd import CVector, Junction, Layer, constantDriver

    base = b64.js[0]
    t = []
    for H in (-800e3, -400e3, 0.0, 400e3, 800e3):
        l = Layer("free", mag=CVector(1, 0, 0), anis=CVector(1, 0, 0), Ms=1.2, thickness=2e-9, cellSurface=1e-16,
                  demagTensor=[CVector(0.00024, 0, 0), CVector(0, 0.0022, 0), CVector(0, 0, 0.99756)], damping=0.01)
        jj = Junction([l], [100.0], [0.0], [1.0], [0.0], [0.0], [0.0], [0.0])
        jj.setLayerAnisotropyDriver("free", constantDriver(5e3))
        jj.setLayerExternalFieldDriver("all", AxialDriver(H, 0, 0))
        t0 = time.perf_counter()
        jj.runSimulation(10e-9, 1e-13, 1e-12)
        t.append(time.perf_counter() - t0)
    plain_rate = int(base["n_steps"]) / float(np.median(t))
    t0 = time.perf_counter()
    olog, omag, ost = oracle.run_batch(b64.js, b64.layers, _lib.constants_get(), None, b64.log_elements,
                                       n_threads=_host_cpus()[0])
    cpu_s = time.perf_counter() - t0
    glog = r64.log.cpu().numpy()
    dm64 = float(np.max(np.abs(r64.final_m.cpu().numpy()[:, 0] - omag[:, 0])))
    rx = r64.channel("Rx")[:, -1].cpu().numpy()
    dm32 = float((r32.log.double() - r64.log).abs().max())
    return {"workload": "r_h_loop_101_fields", "junction_steps": steps,
            "fp64_batched_junction_steps_per_s": steps / (ms64 * 1e-3), "fp64_batched_ms": ms64,
            "fp32_batched_junction_steps_per_s": steps / (ms32 * 1e-3), "fp32_batched_ms": ms32,
            "fp64_plain_junction_steps_per_s": plain_rate,
            "cpu_oracle_junction_steps_per_s": steps / cpu_s, "cpu_threads": _host_cpus()[0],
            "max_dm_fp64_vs_oracle_final": dm64, "max_dlog_fp32_vs_fp64": dm32,
            "log_max_abs_difference_fp64_vs_oracle": float(np.max(np.abs(glog - olog))),
            "final_Rx_first_mid_last": [float(rx[0]), float(rx[50]), float(rx[-1])]}


def _bench_stack(n_members: int = 8, total_time: float = 10e-9) -> dict:
    """Series stack of STT junctions sharing one current (KCL, coupling
    0.05): one stack run on the GPU (one wave, member j on lane j) against
    the oracle's stack runner on one host core; RK4 dt 1e-13, W = 10."""
    import numpy as np

    import oracle
    from cmtj_amd import CVector, Junction, Layer, _lib, constantDriver
    from cmtj_amd.stack import SeriesStack

    demag = [CVector(0, 0, 0), CVector(0, 0, 0), CVector(0, 0, 1)]
    members = []
    for q in range(n_members):
        l = Layer.createSTTLayer("free", mag=CVector(0.1 * q, 0.2, 1.0), anis=CVector(1, 0, 0), Ms=1.65 - 0.02 * q,
                                 thickness=3e-9, cellSurface=7e-14, demagTensor=demag, damping=0.005,
                                 SlonczewskiSpacerLayerParameter=1.0, beta=0.1, spinPolarisation=0.6)
        l.setReferenceLayer(CVector(1, 0, 0))
        j = Junction([l], 100, 200)
        j.setLayerAnisotropyDriver("free", constantDriver(1.05e3))
        members.append(j)

    class _OracleStack(SeriesStack):
        def _execute(self, stack, js, layers, consts, tables=None):
            return oracle.run_stack(stack, js, layers, consts, tables)

    out = {}
    for name, cls in (("gpu", SeriesStack), ("cpu_oracle", _OracleStack)):
        s = cls(members)
        s.setCouplingStrength(0.05)
        s.setCoupledCurrentDriver(constantDriver(5e11))
        s.runSimulation(1e-11, 1e-13, 1e-12)  # warm-up (code objects, buffers)
        s.clearLogs()
        t0 = time.perf_counter()
        s.runSimulation(total_time, 1e-13, 1e-12)
        out[name] = (time.perf_counter() - t0, s)
    steps = n_members * int(total_time / 1e-13)
    # a sweep: n_stacks stacks (coupling and current varied) in one launch
    from cmtj_amd.stack import run_stacks

    n_stacks = 4096
    sweep = []
    for k in range(n_stacks):
        s = SeriesStack(members)
        s.setCouplingStrength(0.01 + 0.1 * (k % 64) / 64)
        s.setCoupledCurrentDriver(constantDriver(1e11 * (1 + k // 64)))
        sweep.append(s)
    run_stacks(sweep[:64], 1e-11, 1e-13, 1e-12)  # warm-up
    t0 = time.perf_counter()
    run_stacks(sweep, 1e-9, 1e-13, 1e-12)
    sweep_s = time.perf_counter() - t0
    sweep_steps = n_stacks * n_members * int(1e-9 / 1e-13)
    g, o = out["gpu"][1], out["cpu_oracle"][1]
    dR = float(np.max(np.abs(np.asarray(g.getLog()["Resistance"]) - np.asarray(o.getLog()["Resistance"]))))
    dm = max(float(np.max(np.abs(np.asarray(g.getLog(q)["free_mx"]) - np.asarray(o.getLog(q)["free_mx"]))))
             for q in range(n_members))
    return {"workload": f"series_stack_{n_members}_stt_members_kcl", "junction_steps": steps,
            "gpu_s": out["gpu"][0], "gpu_junction_steps_per_s": steps / out["gpu"][0],
            "cpu_oracle_s": out["cpu_oracle"][0], "cpu_oracle_junction_steps_per_s": steps / out["cpu_oracle"][0],
            "cpu_threads": 1, "max_dm_gpu_vs_oracle": dm, "max_dresistance_gpu_vs_oracle": dR,
            "conditioning": "a 1e-14 change of one member's initial m grows to 3.9e-6 in the oracle itself by 10 ns",
            "sweep": {"stacks": n_stacks, "members": n_members, "total_time": 1e-9, "junction_steps": sweep_steps,
                      "s": sweep_s, "junction_steps_per_s": sweep_steps / sweep_s},
            "what": "host-to-host Stack.runSimulation / run_stacks (packing, upload, kernel, log copy back)"}


def _bench_group(side: int = 8, total_time: float = 2e-9, n_sweep: int = 1024) -> dict:
    """Dipole-coupled reservoir: a side x side square array (pitch 60 nm) of
    perpendicular single-layer junctions with Kanao point-dipole coupling,
    each in its own in-plane field; RK4 dt 1e-13, one entry per ps.  One
    group on the GPU (one workgroup) against the oracle's group runner on one
    host core (a sample of the same run), then n_sweep groups (an input-field
    sweep, as a reservoir's readout training would run) in one launch."""
    import math

    import numpy as np

    import oracle
    from cmtj_amd import AxialDriver, CVector, Junction, Layer, constantDriver
    from cmtj_amd.reservoir import GroupInteraction, run_groups

    n = side * side
    demag = [CVector(0.04, 0, 0), CVector(0, 0.04, 0), CVector(0, 0, 0.92)]
    pos = [CVector((q % side) * 60e-9, (q // side) * 60e-9, 0.0) for q in range(n)]

    def elements(h0):
        js = []
        for q in range(n):
            th = 0.2 + 0.01 * q
            l = Layer("free", mag=CVector(math.sin(th), 0.0, math.cos(th)), anis=CVector(0, 0, 1), Ms=1.2,
                      thickness=1.5e-9, cellSurface=math.pi * (25e-9) ** 2, demagTensor=demag, damping=0.01)
            j = Junction([l])
            j.setLayerAnisotropyDriver("free", constantDriver(8e5))
            j.setLayerExternalFieldDriver("all", AxialDriver(h0 + 500.0 * q, 0.0, 0.0))
            js.append(j)
        return js

    class _OracleGroup(GroupInteraction):
        def _execute(self, g, js, layers, consts, pair_terms=None):
            return oracle.run_group(g, js, layers, consts, pair_terms)

    base = elements(2e4)
    out = {}
    sample = total_time / 20
    for name, cls, T in (("gpu", GroupInteraction, total_time), ("cpu_oracle", _OracleGroup, sample)):
        g = cls(pos, base)
        g.runSimulation(1e-12, 1e-13, 1e-12)  # warm-up (code objects, buffers)
        g = cls(pos, base)
        t0 = time.perf_counter()
        g.runSimulation(T, 1e-13, 1e-12)
        out[name] = (time.perf_counter() - t0, g, n * int(T / 1e-13))
    # the GPU run against the oracle over the sample's span
    ref = out["cpu_oracle"][1]
    gpu_g = out["gpu"][1]
    k = len(ref.getLog(0)["time"])
    dm = max(float(np.max(np.abs(np.asarray(gpu_g.getLog(q)["free_mx"][:k]) -
                                 np.asarray(ref.getLog(q)["free_mx"])))) for q in range(n))
    groups = []
    for k_ in range(n_sweep):
        grp = GroupInteraction(pos, elements(1e4 + 40.0 * k_))
        groups.append(grp)
    run_groups(groups[:4], 1e-12, 1e-13, 1e-12)  # warm-up
    for grp in groups[:4]:
        grp.clearLogs()
    sweep_T = 1e-9
    t0 = time.perf_counter()
    run_groups(groups, sweep_T, 1e-13, 1e-12)
    sweep_s = time.perf_counter() - t0
    sweep_steps = n_sweep * n * int(sweep_T / 1e-13)
    # the engine call alone on the packed sweep (upload, kernel, log copy back)
    from cmtj_amd import _lib

    packs = [grp._pack(sweep_T, 1e-13, 1e-12) for grp in groups]
    recs = np.stack([p_[0] for p_ in packs])
    js_all, rows_all, rows, off = [], [], 0, 0
    for p_ in packs:
        jq = p_[1].copy()
        jq["layer_offset"] += rows
        jq["log_offset"] += off
        rows += len(p_[2])
        off += p_[6]
        js_all.append(jq)
        rows_all.append(p_[2])
    js_all, rows_all = np.concatenate(js_all), np.concatenate(rows_all)
    consts = _lib.constants_get()
    t0 = time.perf_counter()
    _lib.run_groups(recs, js_all, rows_all, consts)
    engine_s = time.perf_counter() - t0
    g_s, o_s = out["gpu"][0], out["cpu_oracle"][0]
    return {"workload": f"dipole_reservoir_{side}x{side}_kanao", "elements": n, "gpu_total_time": total_time,
            "gpu_junction_steps": out["gpu"][2], "gpu_s": g_s, "gpu_junction_steps_per_s": out["gpu"][2] / g_s,
            "cpu_oracle_total_time": sample, "cpu_oracle_junction_steps": out["cpu_oracle"][2],
            "cpu_oracle_s": o_s, "cpu_oracle_junction_steps_per_s": out["cpu_oracle"][2] / o_s, "cpu_threads": 1,
            "max_dm_gpu_vs_oracle": dm,
            "sweep": {"groups": n_sweep, "elements": n, "total_time": sweep_T, "junction_steps": sweep_steps,
                      "s": sweep_s, "junction_steps_per_s": sweep_steps / sweep_s, "engine_s": engine_s,
                      "engine_junction_steps_per_s": sweep_steps / engine_s},
            "what": "host-to-host GroupInteraction.runSimulation / run_groups (packing, upload, kernel, log copy "
                    "back); engine_s: cmtj_run_groups on the packed sweep"}


def _bench_sweep_forms(n_points: int = 65536) -> dict:
    """The batching forms on an uneven PIMM sweep larger than the chip
    (65,536 points, durations U(2, 20) ns, FP32, m_x and R logged every 10
    steps): the default launch, then
    the autotuner's measured candidates and its pick, each timed with HIP
    events over one whole pass."""
    import tempfile

    from cmtj_amd import autotune

    import numpy as np

    os.environ.setdefault("CMTJ_AMD_TUNE_CACHE", os.path.join(tempfile.mkdtemp(), "tune.json"))
    b = _pimm("fp32", n_points)
    # the sweep's durations, logged every 10 steps (every step would be 86 GB)
    b.set_times(np.random.default_rng(0).uniform(2e-9, 20e-9, n_points), 1e-13, 1e-12)
    b.set_lanes("auto")
    _, ms_auto = _timed_run(b)
    name_auto = b.plan().info.kernel_name.decode()
    pick = autotune.tune(b, force=True, verbose=True)
    b.set_lanes(pick["lanes"])
    b.plan()
    if pick["graph"] and b.plan().info.n_kernels > 1:
        b.capture()
    _, ms_pick = _timed_run(b)
    return {"workload": f"pimm_{n_points}_uneven", "junction_steps": b.total_steps,
            "auto": {"kernel": name_auto, "ms": ms_auto, "junction_steps_per_s": b.total_steps / (ms_auto * 1e-3)},
            "autotune": {"lanes": pick["lanes"], "graph": pick["graph"], "kernel": pick["kernels"], "ms": ms_pick,
                         "junction_steps_per_s": b.total_steps / (ms_pick * 1e-3),
                         "trials": [{k: t[k] for k in ("lanes", "graph", "ms_probe", "kernels")}
                                    for t in pick["trials"]]}}


def _bench_dp(n_fields: int = 64, n_currents: int = 64, total_time: float = 3e-9) -> dict:
    """Dormand-Prince stability-diagram sweep: single-layer SOT junctions
    (perpendicular, second-order anisotropy, 1 ns field- and damping-like
    torque pulses) over fields x current densities, default step control,
    initial step 1e-12 s, one entry per 10 ps; the whole grid in one launch
    against the oracle on the host (a sample, 16 threads)."""
    import math

    import numpy as np

    import oracle
    from cmtj_amd import AxialDriver, CVector, Junction, Layer, SolverMode, _lib, constantDriver, stepDriver
    from cmtj_amd.batch import JunctionBatch

    p = 0.017
    demag = [CVector(p, 0, 0), CVector(0, p, 0), CVector(0, 0, math.sqrt(1 - 2 * p * p))]
    TtoAm = 795774.715459
    js = []
    for Hz in np.linspace(-4e5, 4e5, n_fields):
        for jden in np.linspace(0.0, 0.3, n_currents):
            l = Layer("free", mag=CVector(0.1, 0.1, 0.99), anis=CVector(0, 0, 1), Ms=1.12, thickness=1e-9,
                      cellSurface=math.pi * (300e-9) ** 2, demagTensor=demag, damping=0.01)
            l.setReferenceLayer(CVector(0, 0, 1))
            l.setSecondOrderAnisotropyDriver(constantDriver(-30e3))
            l.setAnisotropyDriver(constantDriver(563e3))
            j = Junction([l])
            j.setLayerExternalFieldDriver("free", AxialDriver(0.1 * abs(Hz) + 1e3, 0, float(Hz)))
            j.setLayerFieldLikeTorqueDriver("free", stepDriver(0, 0.7 * TtoAm * 1e-3 * jden ** 2, 0, 1e-9))
            j.setLayerDampingLikeTorqueDriver("free", stepDriver(0, 0.7 * TtoAm * 16e-3 * jden, 0, 1e-9))
            js.append(j)
    b = JunctionBatch.from_junctions(js, total_time, 1e-12, 1e-11, SolverMode.DormandPrince, precision="fp64",
                                     log="full", log_dtype="f64")
    r, ms = _timed_run(b)
    st = r.status()
    steps = int(st["steps_done"].astype(np.int64).sum())
    idx = np.arange(0, len(js), max(1, len(js) // 256))
    sub = b.subset(idx)
    t0 = time.perf_counter()
    olog, omag, ost = oracle.run_batch(sub.js, sub.layers, _lib.constants_get(), None, sub.log_elements,
                                       n_threads=_host_cpus()[0])
    cpu_s = time.perf_counter() - t0
    cpu_steps = int(ost["steps_done"].astype(np.int64).sum())
    same = bool(np.all(ost["steps_done"] == st["steps_done"][idx]) and np.all(ost["entries_done"] ==
                                                                               st["entries_done"][idx]))
    dm = float(np.max(np.abs(r.final_m.cpu().numpy()[idx, 0] - omag[:, 0])))
    return {"workload": f"dormand_prince_sot_stability_{n_fields}x{n_currents}", "instances": len(js),
            "accepted_steps": steps, "gpu_ms": ms, "gpu_accepted_steps_per_s": steps / (ms * 1e-3),
            "cpu_oracle_accepted_steps_per_s": cpu_steps / cpu_s, "cpu_threads": _host_cpus()[0],
            "cpu_sample": f"{len(idx)} instances", "same_step_and_entry_counts_on_sample": same,
            "max_dm_final_gpu_vs_oracle_sample": dm}


def _bench_cims(n_currents: int = 256, n_seeds: int = 4096) -> dict:
    """CIMS / p-bit ensemble: 256 currents x 4096 seeds, 2 layers, static SOT
    at 300 K, Euler-Heun dt 1e-14 s, 5 ns; switching probability per current
    in FP32, an FP64 sample with the same noise, and the oracle on a smaller
    sample (pathwise: the same Philox streams)."""
    import numpy as np

    import oracle
    from cmtj_amd import _lib, analysis, workloads

    b = workloads.cims_pbit(n_currents=n_currents, n_seeds=n_seeds, precision="fp32")
    r, ms = _timed_run(b)
    P, se = analysis.switching_probability(r.final_m[:, 0, 2], n_currents)
    ns64 = 64
    idx64 = (np.arange(n_currents)[:, None] * n_seeds + np.arange(ns64)[None, :]).ravel()
    b64 = b.subset(idx64, precision="fp64")
    r64, ms64 = _timed_run(b64)
    s32 = (r.final_m[idx64, 0, 2] < 0).cpu().numpy()
    s64 = (r64.final_m[:, 0, 2] < 0).cpu().numpy()
    dm = float((r.final_m[idx64] - r64.final_m).abs().max())
    idxo = (np.arange(0, n_currents, 8)[:, None] * n_seeds + np.arange(2)[None, :]).ravel()
    bo = b.subset(idxo, precision="fp64")
    t0 = time.perf_counter()
    _, omag, _ = oracle.run_batch(bo.js, bo.layers, _lib.constants_get(), None, 0,
                                  n_threads=_host_cpus()[0])
    cpu_s = time.perf_counter() - t0
    ro = b64.subset(np.searchsorted(idx64, idxo)).run()  # the GPU FP64 of the oracle's sample
    dmo = float(np.max(np.abs(ro.final_m.cpu().numpy()[:, :2] - omag[:, :2])))
    Pn = P.cpu().numpy()
    jcur = np.logspace(10, 12, n_currents)
    k50 = int(np.argmax(Pn >= 0.5)) if np.any(Pn >= 0.5) else -1
    return {"workload": "cims_pbit_256x4096", "instances": len(b.js), "junction_steps": b.total_steps,
            "j50_A_per_m2": float(jcur[k50]) if k50 >= 0 else None,
            "currents_with_0_lt_P_lt_1": int(np.sum((Pn > 0) & (Pn < 1))),
            "fp32_junction_steps_per_s": b.total_steps / (ms * 1e-3), "fp32_ms": ms,
            "fp64_sample_junction_steps_per_s": b64.total_steps / (ms64 * 1e-3),
            "switching_probability_first_mid_last": [float(Pn[0]), float(Pn[n_currents // 2]), float(Pn[-1])],
            "switching_stderr_max": float(se.max()),
            "fp32_vs_fp64_same_outcome_fraction": float(np.mean(s32 == s64)),
            "fp32_vs_fp64_max_dm_final": dm, "fp64_sample": f"{ns64} seeds x {n_currents} currents",
            "fp64_gpu_vs_oracle_max_dm_final": dmo, "oracle_sample": f"{len(idxo)} instances",
            "cpu_oracle_junction_steps_per_s": bo.total_steps / cpu_s}


def _bench_sto(n_currents: int = 512) -> dict:
    """Spin-torque oscillator precision study: 512 DC currents x 1e6 RK4 steps
    (100 ns at 1e-13 s), FP32 and FP64; oscillation frequency and linewidth
    from the FFT of the free layer's m_x (logged every 99 steps), max|dm|,
    and the oracle on a sample."""
    import numpy as np

    import oracle
    from cmtj_amd import _lib, analysis, workloads

    # the binary32 leg on one lane per layer: left to the automatic choice,
    # an FP32 batch this small runs the binary64 component-lane kernel
    b32 = workloads.sto(n_currents=n_currents, precision="fp32", log="full").set_lanes("per_layer")
    b32.log_dtype = _lib.LOG_F64
    b64 = workloads.sto(n_currents=n_currents, precision="fp64", log="full")
    r32, ms32 = _timed_run(b32)
    r64, ms64 = _timed_run(b64)
    # the same FP32 request with the lane mapping left to the engine
    b32a = workloads.sto(n_currents=n_currents, precision="fp32", log="full")
    _, ms32a = _timed_run(b32a)
    auto_info = b32a.plan().info
    del b32a
    W = int(b64.js[0]["write_every"])
    dtl = W * float(b64.js[0]["dt"])
    f32, w32, _ = analysis.spectrum_peaks(r32.channel("free_mx"), dtl)
    f64, w64, _ = analysis.spectrum_peaks(r64.channel("free_mx"), dtl)
    dm = float((r32.log - r64.log).abs().max())
    dm_m = max(float((r32.channel(c) - r64.channel(c)).abs().max()) for c in ("free_mx", "free_my", "free_mz"))
    osc = (f64 > 0.2e9)
    idx = np.arange(0, n_currents, n_currents // 8)
    bo = b64.subset(idx)
    t0 = time.perf_counter()
    olog, omag, _ = oracle.run_batch(bo.js, bo.layers, _lib.constants_get(), None, bo.log_elements,
                                     n_threads=_host_cpus()[0])
    cpu_s = time.perf_counter() - t0
    ro = bo.run()
    dmo = float(np.max(np.abs(ro.log.cpu().numpy() - olog)))
    df = (f32 - f64).abs()
    dw = (w32 - w64).abs()
    return {"workload": "sto_512_currents_1e6_steps", "junction_steps": b64.total_steps,
            "fp32_junction_steps_per_s": b32.total_steps / (ms32 * 1e-3), "fp32_ms": ms32,
            "fp32_kernel": b32.plan().info.kernel_name.decode(),
            "fp32_auto": {"junction_steps_per_s": b32.total_steps / (ms32a * 1e-3), "ms": ms32a,
                          "kernel": auto_info.kernel_name.decode(), "fp64_groups": int(auto_info.fp64_groups)},
            "fp64_junction_steps_per_s": b64.total_steps / (ms64 * 1e-3), "fp64_ms": ms64,
            "oscillating_points": int(osc.sum()),
            "frequency_GHz_first_mid_last": [float(f64[0]) / 1e9, float(f64[n_currents // 2]) / 1e9,
                                             float(f64[-1]) / 1e9],
            "fp32_vs_fp64_max_dfreq_Hz": float(df.max()), "fp32_vs_fp64_max_dlinewidth_Hz": float(dw.max()),
            "frequency_resolution_Hz": 1.0 / (0.8 * r64.channel("free_mx").shape[1] * dtl),
            "fp32_vs_fp64_max_dm": dm_m, "fp32_vs_fp64_max_dlog": dm,
            "fp64_gpu_vs_oracle_max_dlog": dmo, "oracle_sample": f"{len(idx)} currents",
            "cpu_oracle_junction_steps_per_s": bo.total_steps / cpu_s}


def _host_cpus():
    """(threads to use, nproc, CPUs this process may run on).  The box gives a
    GPU's process a share of 16 host CPUs, whatever nproc shows."""
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = nproc
    return min(16, avail), nproc, avail


def _cpu_baseline(b, args):
    """FP64 oracle (-O3 -march=native build of oracle.c, no FP contraction) on
    a stratified sample of the sweep at full durations: all the process's host
    CPUs (at most the box's 16-CPU share), then one core on a smaller sample."""
    import tempfile

    import numpy as np

    import oracle
    from cmtj_amd import _lib

    ncores, nproc, avail = _host_cpus()
    try:
        lib_path, build = oracle.build_native(tempfile.mkdtemp(prefix="oracle_native_")), "-O3 -march=native"
    except Exception as e:  # no compiler: the shipped generic build
        lib_path, build = None, f"-O3 generic (native build failed: {type(e).__name__})"

    def sample(n_pts):
        idx = np.linspace(0, len(b.js) - 1, n_pts).astype(int)
        bs = _pimm("fp64", args.points, log_dtype="f64")
        js = bs.js[idx].copy()
        layers = np.concatenate([bs.layers[int(js[q]["layer_offset"]): int(js[q]["layer_offset"]) + 2]
                                 for q in range(len(idx))])
        js["layer_offset"] = np.arange(len(idx), dtype=np.uint32) * 2
        sizes = 3 * js["n_entries"].astype(np.uint64)
        js["log_offset"] = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.uint64)
        return js, layers, int(sizes.sum())

    def timed(n_pts, threads):
        js, layers, nlog = sample(n_pts)
        t0 = time.perf_counter()
        oracle.run_batch(js, layers, _lib.constants_get(), None, nlog, n_threads=threads, lib_path=lib_path)
        s = time.perf_counter() - t0
        return int(js["n_steps"].astype(np.int64).sum()), s

    steps, cs = timed(args.cpu_sample, ncores)
    n1 = max(8, args.cpu_sample // 32)
    steps1, cs1 = timed(n1, 1)
    cpu = {"value": steps / cs, "unit": "junction-steps/s", "cores": ncores, "kind": "port",
           "host_nproc": nproc, "host_cpus_available": avail, "build": build,
           "single_core": steps1 / cs1,
           "sample": f"{args.cpu_sample} of {len(b.js)} sweep points (stratified), full durations, "
                     f"{steps:.3e} junction-steps, {cs:.1f} s on {ncores} threads; single core: {n1} points, "
                     f"{steps1:.3e} junction-steps, {cs1:.1f} s"}
    print(f"[bench] cpu oracle: {cpu['value']:.4e} junction-steps/s on {ncores} threads "
          f"(nproc {nproc}, {avail} available), single core {cpu['single_core']:.4e}", file=sys.stderr, flush=True)
    return cpu


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=4096)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=2048)
    ap.add_argument("--lanes", default="auto", choices=["auto", "one", "per_layer", "per_component"])
    ap.add_argument("--serving-lanes", default="per_component", choices=["one", "per_layer", "per_component"])
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-qps", type=float, default=20.0)
    ap.add_argument("--serving-seconds", type=float, default=10.0)
    ap.add_argument("--workload", default="pimm", choices=["pimm", "rh", "cims", "sto", "serving", "stack", "dp", "group",
                                                            "forms"],
                    help="pimm: the headline sweep (default); rh, cims, sto: the other sweep workloads, "
                         "one JSON line each")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="run the sweep kernel twice and a 1 GiB device copy, print nothing (for counter runs)")
    ap.add_argument("--pmc-fixed-ns", type=float, default=0.0,
                    help="with --pmc-pass: give every sweep point this duration (ns)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(instances shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from cmtj_amd import _lib, workloads
    from cmtj_amd.batch import JunctionBatch

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.workload != "pimm":
        fn = {"rh": _bench_rh, "cims": _bench_cims, "sto": _bench_sto, "stack": _bench_stack, "dp": _bench_dp,
              "group": _bench_group, "forms": _bench_sweep_forms,
              "serving": lambda: _serving(args.serving_qps, args.serving_seconds, lanes=args.serving_lanes)}[args.workload]
        print(json.dumps(fn()))
        return 0

    # ---- FP32 batched, inputs resident in HBM -------------------------------
    if args.pmc_pass and args.pmc_fixed_ns:
        # every point the same duration: counters divide evenly by waves x steps
        from cmtj_amd import workloads

        T = args.pmc_fixed_ns * 1e-9
        b = workloads.pimm(n_points=args.points, seed=0, t_min=T, t_max=T, precision="fp32", log="mx")
        b.set_lanes(args.lanes)
    else:
        b = _pimm("fp32", args.points).set_lanes(args.lanes)
    plan = b.plan()
    if args.pmc_pass:
        return _pmc_pass(b)
    info = plan.info
    stream = torch.cuda.current_stream()
    print(f"[bench] pimm: {len(b.js)} instances, {b.total_steps:.3e} junction-steps/pass, "
          f"max {info.max_steps} steps, kernel {info.kernel_name.decode()}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        b.execute(stream)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        r32 = b.execute(stream)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [s.elapsed_time(e) for s, e in zip(starts, ends)]  # ms, kernel time on the launch stream
    ms_step = wall * 1e3 / args.steps
    value = b.total_steps * args.steps / wall
    kern_ms = float(np.mean(per))
    achieved = FLOP_PER_STEP * b.total_steps / (kern_ms * 1e-3) / 1e12
    print(f"[bench] fp32 batched: {value:.4e} junction-steps/s, {ms_step:.1f} ms/pass", file=sys.stderr, flush=True)

    # ---- FP64 drop-in, one instance per launch -------------------------------
    from cmtj_amd import AxialDriver

    j = workloads.pimm_junction()
    j.setLayerExternalFieldDriver("all", AxialDriver(300e3 * math.cos(math.pi / 6), 300e3 * math.sin(math.pi / 6), 0))
    j.runSimulation(2e-10, 1e-13, 1e-13)  # warm
    plain_T = 2e-9
    t1 = time.perf_counter()
    j.runSimulation(plain_T, 1e-13, 1e-13)
    plain_s = time.perf_counter() - t1
    plain_rate = int(plain_T / 1e-13) / plain_s
    print(f"[bench] fp64 plain per-instance: {plain_rate:.4e} junction-steps/s", file=sys.stderr, flush=True)

    # ---- FP32 vs FP64 on the whole sweep --------------------------------------
    max_dm = None
    if not args.no_accuracy:
        b64 = _pimm("fp64", args.points, log_dtype="f64")
        r64 = b64.run()
        r32 = b.execute(stream)
        torch.cuda.synchronize()
        nch = 3
        ent = torch.as_tensor(b.js["n_entries"].astype(np.int64), device=dev)
        reps = ent.repeat_interleave(nch)
        chan = torch.arange(nch, device=dev, dtype=torch.int8).repeat(len(b.js)).repeat_interleave(reps)
        m_sel = chan < 2  # m_x of the two layers
        d = (r32.log.double() - r64.log)[m_sel].abs().max().item()
        dfin = (r32.final_m - r64.final_m).abs().max().item()
        max_dm = max(d, dfin)
        del r64, b64, chan, m_sel, reps
        torch.cuda.empty_cache()
        print(f"[bench] fp32 vs fp64: max|dm| = {max_dm:.3e}", file=sys.stderr, flush=True)

    # ---- CPU baseline: FP64 oracle, bounded stratified sample ----------------
    cpu = None
    if not args.no_cpu:
        cpu = _cpu_baseline(b, args)

    serving = None
    if not args.no_serving:
        serving = _serving(args.serving_qps, args.serving_seconds, lanes=args.serving_lanes)
        print(f"[bench] serving: {serving['requests']} requests at {args.serving_qps} QPS, "
              f"p50 {serving['latency_s']['p50']:.3f} s, p99 {serving['latency_s']['p99']:.3f} s, "
              f"{serving['junction_steps_per_s']:.3e} junction-steps/s", file=sys.stderr, flush=True)

    tr = _traffic()
    out = {
        "metric": "junction-steps/s FP32 batched vs FP64 plain at max|dm|<=1e-4; p50/p99 latency",
        "value": value,
        "unit": "junction-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (anchored RK4: stages 2-4 in fp32; k1, state, clock in fp64)",
        "data": "synthetic",
        "config": {
            "workload": "pimm_2layer_mtj_sweep",
            "instances": len(b.js),
            "layers": 2,
            "solver": "RK4",
            "dt": 1e-13,
            "durations_ns": [2, 20],
            "junction_steps_per_pass": b.total_steps,
            "log": "m_x (2 layers) + R every step, float32",
            "lanes_per_instance": int(info.lanes_per_instance),
            "kernel": info.kernel_name.decode(),
        },
        "fp64_plain_junction_steps_per_s": plain_rate,
        "fp32_batched_over_fp64_plain": value / plain_rate,
        "max_dm_fp32_vs_fp64": max_dm,
        "latency_ms": {"p50": float(np.percentile(per, 50)), "p99": float(np.percentile(per, 99)),
                       "what": "one full sweep pass (kernel time on the launch stream)"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "traffic": None if tr is None else tr["bytes_per_launch"],
                     "flop_per_junction_step": FLOP_PER_STEP, "kernel_ms": kern_ms,
                     "traffic_algorithmic": 12.0 * b.total_steps},
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
