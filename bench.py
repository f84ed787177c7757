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


def _serving(qps: float, seconds: float, seed: int = 0) -> dict:
    """Mixed request stream (PIMM / VSD / CIMS / 3-4 layer chains, 1-50 ns,
    dt 1e-14..1e-13 s) with Poisson arrivals at `qps` for `seconds`, served
    by the continuous-batching loop in FP32; latency = submit to result."""
    import numpy as np

    from cmtj_amd import workloads
    from cmtj_amd.serving import Server

    rng = np.random.default_rng(seed + 1)
    mix = workloads.serving_mix(int(qps * seconds * 1.5) + 16, seed=seed)
    futs, kinds = [], []
    with Server(precision="fp32", slice_ms=20.0, n_streams=8, log_elements=1 << 28) as srv:
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
    return {"qps": qps, "arrival_window_s": seconds, "requests": len(res),
            "mix": {k: kinds.count(k) for k in sorted(set(kinds))},
            "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": float(lat.max())},
            "junction_steps": int(steps), "junction_steps_per_s": steps / wall, "wall_s": wall,
            "slices": ticks, "slice_ms_target": 20.0, "host_s": host_s, "precision": "fp32", "lanes": "per_layer"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=4096)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=2048)
    ap.add_argument("--lanes", default="auto", choices=["auto", "one", "per_layer"])
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-qps", type=float, default=20.0)
    ap.add_argument("--serving-seconds", type=float, default=10.0)
    ap.add_argument("--pmc-pass", action="store_true",
                    help="run the sweep kernel twice and a 1 GiB device copy, print nothing (for counter runs)")
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

    # ---- FP32 batched, inputs resident in HBM -------------------------------
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
        import oracle

        ncores = min(16, os.cpu_count() or 1)
        idx = np.linspace(0, len(b.js) - 1, args.cpu_sample).astype(int)
        bs = _pimm("fp64", args.points, log_dtype="f64")
        js = bs.js[idx].copy()
        layers = np.concatenate([bs.layers[int(js[q]["layer_offset"]): int(js[q]["layer_offset"]) + 2]
                                 for q in range(len(idx))])
        js["layer_offset"] = np.arange(len(idx), dtype=np.uint32) * 2
        sizes = 3 * js["n_entries"].astype(np.uint64)
        js["log_offset"] = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.uint64)
        t2 = time.perf_counter()
        oracle.run_batch(js, layers, _lib.constants_get(), None, int(sizes.sum()), n_threads=ncores)
        cs = time.perf_counter() - t2
        csteps = int(js["n_steps"].astype(np.int64).sum())
        cpu = {"value": csteps / cs, "unit": "junction-steps/s", "cores": ncores, "kind": "port",
               "sample": f"{len(idx)} of {len(b.js)} sweep points (every {len(b.js) // len(idx)}th), "
                         f"full durations, {csteps:.3e} junction-steps, {cs:.1f} s"}
        print(f"[bench] cpu oracle: {cpu['value']:.4e} junction-steps/s on {ncores} threads", file=sys.stderr,
              flush=True)

    serving = None
    if not args.no_serving:
        serving = _serving(args.serving_qps, args.serving_seconds)
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
