"""Launch-configuration autotuner for batches.

How fast a batch runs depends on how its instances map onto lanes (all layers
of an instance in one lane, one lane per layer, or one lane per vector
component) and, for a batch of several structure groups, on whether the
groups' launches follow one another on a stream or run concurrently from one
HIP graph (cmtj_plan_graph_create).  Which wins depends on the batch size
against the chip and on the spread of the instances' step counts: a sweep
too small to fill the chip is bound by its longest instance's step latency,
a large one by throughput.  The tuner measures instead of guessing: every
candidate runs a probe copy of the batch whose step counts are all scaled by
`scale` (the same shape of the length distribution, so the same balance of
critical path and throughput), timed with HIP events on the launch stream;
the fastest candidate is kept in a JSON cache keyed by the batch's structure
signature and size class, so later batches of the same kind reuse the pick
without measuring.

Replaces the reference's choice of a process-pool size for sweeps
(cmtj/utils/parallel.py:47,84), which has no per-batch tuning.
"""
from __future__ import annotations

import hashlib
import json
import math
import os
from pathlib import Path

import numpy as np

from . import _lib

LANES = ("auto", "per_component", "per_layer", "one", "queue")


def cache_path() -> Path:
    p = os.environ.get("CMTJ_AMD_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "cmtj_amd" / "autotune.json"


def signature(b) -> str:
    """Structure groups of the batch (their 64-bit keys), arithmetic, log
    dtype, log2 of the instance count (rounded) and the longest-to-shortest
    step ratio in factors of 4: batches that share it share a best
    configuration."""
    keys = set()
    for i in range(len(b.js)):
        off, nl = int(b.js[i]["layer_offset"]), int(b.js[i]["n_layers"])
        keys.add(int(_lib.group_key(b.js[i: i + 1], b.layers[off: off + nl])))
    steps = b.js["n_steps"].astype(np.float64)
    spread = float(steps.max() / max(steps.min(), 1.0)) if len(steps) else 1.0
    parts = [sorted(keys), int(b.precision), int(b.log_dtype), int(round(math.log2(max(len(b.js), 1)))),
             int(math.log2(max(spread, 1.0)) // 2)]
    return hashlib.sha1(json.dumps(parts).encode()).hexdigest()[:16]


def load_cache() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def save_pick(sig: str, pick: dict) -> None:
    c = load_cache()
    c[sig] = pick
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(c, indent=1, sort_keys=True))
        os.replace(tmp, p)
    except OSError:
        pass  # a read-only home: the pick still applies to this process


def _probe(b, scale: float):
    """A copy of the batch with every instance's step count scaled (at least
    one step), same noise streams and log layout rules."""
    q = b.subset(np.arange(len(b.js)))
    if np.any(q.js["solver"] == _lib.DORMAND_PRINCE) or scale >= 1.0:
        return q
    dt = q.js["dt"].astype(np.float64)
    steps = np.maximum(1, np.floor(q.js["n_steps"].astype(np.float64) * scale))
    wf = q.js["write_every"].astype(np.float64) * dt
    q.set_times(steps * dt * (1 + 1e-12), dt, wf * (1 + 1e-12))
    return q


def _time(q, lanes: str, graph: bool, reps: int = 2) -> tuple[float, str]:
    import torch

    q.set_lanes(lanes)
    p = q.plan()
    name = p.info.kernel_name.decode()
    stream = torch.cuda.current_stream()
    if graph:
        q.capture()
        run = lambda: q.replay(stream)  # noqa: E731
    else:
        run = lambda: q.execute(stream)  # noqa: E731
    run()  # warm-up: code objects loaded, caches filled
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        run()
        e1.record(stream)
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best, name


def tune(b, scale: float = 0.05, persist: bool = True, force: bool = False, verbose: bool = False) -> dict:
    """The fastest launch configuration of batch b: {"lanes", "graph",
    "ms_probe", "kernels", "trials"}.  Uses the cache unless `force`."""
    sig = signature(b)
    if not force:
        hit = load_cache().get(sig)
        if hit:
            return dict(hit, cached=True)
    q = _probe(b, scale)
    trials, seen = [], set()
    for lanes in LANES:
        for graph in (False, True):
            try:
                q.set_lanes(lanes)
                p = q.plan()
            except RuntimeError:
                continue
            if graph and p.info.n_kernels < 2:
                continue  # one launch: nothing for a graph to overlap
            name = p.info.kernel_name.decode()
            if (name, graph) in seen:
                continue  # another lane request chose the same kernels
            seen.add((name, graph))
            ms, name = _time(q, lanes, graph)
            trials.append({"lanes": lanes, "graph": graph, "ms_probe": ms, "kernels": name})
            if verbose:
                print(f"[autotune] {lanes:13s} graph={graph!s:5s} {ms:9.3f} ms  {name}", flush=True)
    if not trials:
        raise RuntimeError("autotune: no launch configuration covers this batch")
    best = min(trials, key=lambda t: t["ms_probe"])
    pick = {"lanes": best["lanes"], "graph": best["graph"], "ms_probe": best["ms_probe"],
            "kernels": best["kernels"], "scale": scale, "trials": trials}
    if persist:
        save_pick(sig, pick)
    return dict(pick, cached=False)


def run(b, pick: dict | None = None, stream=None):
    """Runs batch b with its tuned configuration (tuning first if no pick is
    given and none is cached); returns the BatchResult."""
    import torch

    pick = pick or tune(b)
    b.set_lanes(pick["lanes"])
    if pick["graph"] and b.plan().info.n_kernels > 1:
        b.capture()
        r = b.replay(stream)
    else:
        r = b.execute(stream)
    torch.cuda.synchronize()
    return r


__all__ = ["tune", "run", "signature", "cache_path", "load_cache", "save_pick"]
