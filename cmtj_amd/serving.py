"""Serving: a request loop with continuous batching over the HIP engine.

The reference answers one ``runSimulation`` per call, on the caller's thread.
Here a ``Server`` owns the GPU and a background loop:

  * ``submit(junction, totalTime, ...)`` packs the request on the caller's
    thread and returns a ``concurrent.futures.Future``;
  * the loop admits queued requests while their log blocks fit the device
    log arena (admission by HBM log-buffer footprint) and slots are free;
  * every active request advances by a slice of at most ``slice_steps`` steps
    per tick (``cmtj_plan_execute_slice``); its magnetisation, step and log
    position stay resident on the device between ticks;
  * requests that finish are retired at the end of their tick (log copied to
    the host, arena block freed) and newly arrived ones join the next tick,
    so work admitted at different times overlaps on the device;
  * groups of different structure are launched on several streams.

A sliced run takes the same steps with the same step times and noise
counters as an uninterrupted one, so a request's result does not depend on
what else is in flight (with the lane mapping fixed, as the server does).
"""
from __future__ import annotations

import bisect
import os
import threading
import time
import warnings
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .batch import _LANES, _LOG_MODES, JunctionBatch, _channel_names, _Meta
from .core import SolverMode


_cat = _lib.cat_records


def _raise_hw_queues(torch, want: int) -> bool:
    """HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues
    (4 unless set), read once when HIP initialises; kernels of two streams
    that share a queue run one after the other, however idle the chip is.
    The server runs one stream per structure group, so it raises the count
    to `want` (at most 32) while HIP is not yet initialised in this process,
    and warns when that is too late.  Returns whether the count now holds
    `want` queues."""
    want = min(int(want), 32)
    try:
        have = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
    except ValueError:
        have = 4
    if have >= want:
        return True
    if torch.cuda.is_initialized():
        warnings.warn(f"GPU_MAX_HW_QUEUES is {have} and HIP is already initialised: groups beyond {have} share "
                      f"hardware queues and wait for each other's slices (set GPU_MAX_HW_QUEUES={want} before "
                      f"the first GPU call)", RuntimeWarning, stacklevel=3)
        return False
    os.environ["GPU_MAX_HW_QUEUES"] = str(want)
    return True


class LogArena:
    """First-fit allocator of element ranges in a fixed device log buffer."""

    def __init__(self, capacity: int) -> None:
        self.capacity = int(capacity)
        self._free = [(0, self.capacity)]  # sorted (offset, size)
        self.used = 0

    def alloc(self, size: int):
        size = max(int(size), 1)
        for q, (off, sz) in enumerate(self._free):
            if sz >= size:
                if sz == size:
                    del self._free[q]
                else:
                    self._free[q] = (off + size, sz - size)
                self.used += size
                return off
        return None

    def free(self, off: int, size: int) -> None:
        size = max(int(size), 1)
        q = bisect.bisect_left(self._free, (off, 0))
        self._free.insert(q, (off, size))
        # merge with the neighbours
        if q + 1 < len(self._free) and off + size == self._free[q + 1][0]:
            self._free[q] = (off, size + self._free[q + 1][1])
            del self._free[q + 1]
        if q > 0 and self._free[q - 1][0] + self._free[q - 1][1] == self._free[q][0]:
            self._free[q - 1] = (self._free[q - 1][0], self._free[q - 1][1] + self._free[q][1])
            del self._free[q]
        self.used -= size


@dataclass
class Result:
    """Outcome of one request: the getLog-style dict, final magnetisations
    [layers, 3] and the request's timeline (host clock, seconds)."""
    log: dict
    final_m: np.ndarray
    steps: int
    t_submit: float
    t_admit: float
    t_done: float

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


@dataclass
class _Req:
    jd: np.ndarray
    layers: np.ndarray
    tables: np.ndarray | None
    meta: _Meta
    n_elems: int
    key: int
    future: Future
    t_submit: float
    t_admit: float = 0.0
    log_off: int = -1
    state: np.ndarray | None = None  # initial state row until the request joins its group's plan


class _Group:
    """Requests of one structure signature: one plan, one state array and
    one stream; advances independently of the other groups."""

    def __init__(self, key: int, stream, slice_steps: int) -> None:
        self.key = key
        self.stream = stream
        self.slice_steps = slice_steps
        self.active: list = []
        self.incoming: list = []
        self.plan = None
        # host-mapped buffers (cmtj_host_alloc): descriptors, slice state and
        # status live in host memory that the kernels read and write directly
        self.harena = self.hstate = self.hstatus = None
        self.n = 0
        self.nsteps = np.zeros(0, dtype=np.int64)
        self.index: dict = {}
        self.changed = False
        self.busy = False
        self.ev0 = self.ev1 = None
        self.slices = 0


class Server:
    """Continuous-batching request loop (see module docstring).

    precision: "fp32" or "fp64"; log_elements: capacity of the device log
    arena (elements of the log dtype); max_active: slots; slice_ms: target
    device time of one slice (each group adapts its step count to it);
    lanes: fixed lane mapping ("per_component", "per_layer" or "one"; a
    group's kernel then depends on its structure only, never on how many
    requests share it, so a result is the same alone and in a mix);
    n_streams: streams created up front (more are created when more groups
    are active); hw_queues: hardware queues asked of HIP through
    GPU_MAX_HW_QUEUES when the server starts before HIP is initialised
    (None leaves the environment alone).  With HIP's default of 4, a fifth
    concurrent group shares a queue with another and both take twice as
    long."""

    def __init__(self, precision: str = "fp32", log_elements: int = 1 << 28, max_active: int = 65536,
                 slice_ms: float = 20.0, slice_steps: int = 5000, lanes: str = "per_component", n_streams: int = 8,
                 log_dtype=None, device: str = "cuda", hw_queues: int | None = 16) -> None:
        if lanes not in ("per_component", "per_layer", "one"):
            raise ValueError("a server fixes the lane mapping: 'per_component', 'per_layer' or 'one'")
        self.precision = _lib.FP32 if precision in ("fp32", _lib.FP32) else _lib.FP64
        if log_dtype is None:
            log_dtype = "f32" if self.precision == _lib.FP32 else "f64"
        self.log_dtype = _lib.LOG_F32 if log_dtype in ("f32", _lib.LOG_F32) else _lib.LOG_F64
        self.lanes = _LANES[lanes]
        self.arena = LogArena(log_elements)
        self.max_active = int(max_active)
        self.slice_ms = float(slice_ms)
        self.slice_steps = int(slice_steps)  # first slice of a new group
        self.n_streams = max(1, int(n_streams))
        self.hw_queues = None if hw_queues is None else max(1, int(hw_queues))
        self.device = device
        self._pending: deque = deque()
        self._groups: dict = {}
        self._n_active = 0
        self._cv = threading.Condition()
        self._stop = False
        self._thread = None
        self.ticks = 0
        self.steps_done = 0
        self.host_s = {"admit": 0.0, "rebuild": 0.0, "retire": 0.0, "launch": 0.0, "complete": 0.0}
        self._finishing: list = []  # (event, pinned log copies, offsets, requests) of retired requests
        self._hpool: list = []      # free host-mapped buffers
        self.error = None

    # --- client side ------------------------------------------------------------
    def submit(self, junction, totalTime: float, timeStep: float = 1e-13, writeFrequency: float = 1e-11,
               solverMode=SolverMode.RK4, log: str = "full") -> Future:
        """Queues one runSimulation-equivalent request; the future resolves to
        a Result (or raises RuntimeError("NAN magnetisation"))."""
        jd, L, tabs, meta = JunctionBatch._pack(junction, totalTime, timeStep, writeFrequency, solverMode)
        jd = np.array(jd, dtype=_lib.JUNCTION_DTYPE).reshape(1)
        if int(jd["solver"][0]) == _lib.DORMAND_PRINCE:
            raise NotImplementedError("Dormand-Prince runs are not served in slices (their step count is adaptive)")
        mask = _LOG_MODES[log] if isinstance(log, str) else int(log)
        if mask & _lib.LOG_ENERGY:
            raise ValueError("energy channels are a drop-in (Junction.runSimulation) feature")
        jd["log_mask"] = mask
        nch = _lib.n_channels(int(jd["n_layers"][0]), int(jd["mr_mode"][0]), mask)
        n_elems = nch * int(jd["n_entries"][0])
        if n_elems > self.arena.capacity:
            raise ValueError(f"request log ({n_elems} elements) exceeds the server's log arena "
                             f"({self.arena.capacity} elements)")
        key = _lib.group_key(jd, L)
        st = np.zeros(1, dtype=_lib.STATE_DTYPE)
        st["nan_layer"] = -1
        st["m"][0, : L.size] = L["mag"]
        fut: Future = Future()
        r = _Req(jd, L.copy(), tabs, meta, n_elems, key, fut, time.perf_counter(), state=st)
        with self._cv:
            if self._stop:
                raise RuntimeError("server is stopped")
            self._pending.append(r)
            self._cv.notify()
        return fut

    def start(self) -> "Server":
        if self._thread is None:
            import torch

            if self.hw_queues is not None:
                _raise_hw_queues(torch, self.hw_queues)
            d = torch.device(self.device)
            self.device = torch.device(d.type, d.index if d.index is not None else torch.cuda.current_device())
            self._thread = threading.Thread(target=self._loop, name="cmtj-server", daemon=True)
            self._thread.start()
        return self

    def stop(self) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify()
        if self._thread is not None:
            self._thread.join()
            self._thread = None

    def __enter__(self) -> "Server":
        return self.start()

    def __exit__(self, *exc) -> None:
        self.stop()

    # --- the loop ------------------------------------------------------------------
    def _admit(self, torch, streams) -> None:
        """FIFO admission while the log arena and the slots allow (no
        overtaking: a request that does not fit holds the ones behind it)."""
        while self._pending and self._n_active < self.max_active:
            r = self._pending[0]
            off = self.arena.alloc(r.n_elems)
            if off is None:
                break
            self._pending.popleft()
            r.log_off = off
            r.t_admit = time.perf_counter()
            g = self._groups.get(r.key)
            if g is None:
                # a stream of its own: a group syncs its stream when it
                # rebuilds or retires and must not wait for another group's slice
                st = streams.pop() if streams else torch.cuda.Stream(device=self.device)
                g = _Group(r.key, st, self.slice_steps)
                self._groups[r.key] = g
            g.incoming.append(r)
            self._n_active += 1

    def _hget(self, nbytes: int, old=None):
        """A host-mapped buffer of at least nbytes (old is kept if it is big
        enough, else returned to the pool)."""
        if old is not None and old.nbytes >= nbytes:
            return old
        if old is not None:
            self._hpool.append(old)
        fit = [b for b in self._hpool if b.nbytes >= nbytes]
        if fit:
            b = min(fit, key=lambda x: x.nbytes)
            self._hpool.remove(b)
            return b
        return _lib.HostBuffer(max(1 << max(int(nbytes) - 1, 1).bit_length(), 1 << 16))

    def _rebuild(self, torch, g: _Group) -> None:
        """Plan + state of the group's active set, written in place into the
        group's host-mapped buffers.  Carried-over requests keep their state
        row, new ones start from their initial row.  Runs while the group's
        stream is idle (no launch of the group reads the buffers)."""
        g.active = g.active + g.incoming
        g.incoming = []
        n = len(g.active)
        if n == 0:
            for b in (g.harena, g.hstate, g.hstatus):
                if b is not None:
                    self._hpool.append(b)
            g.plan = g.harena = g.hstate = g.hstatus = None
            g.index, g.n = {}, 0
            return
        js = _cat([r.jd for r in g.active], _lib.JUNCTION_DTYPE)
        nls = js["n_layers"].astype(np.int64)
        js["layer_offset"] = np.concatenate([[0], np.cumsum(nls)[:-1]]).astype(np.uint32)
        js["log_offset"] = np.fromiter((r.log_off for r in g.active), dtype=np.uint64, count=n)
        g.nsteps = js["n_steps"].astype(np.int64)
        layers = _cat([r.layers for r in g.active], _lib.LAYER_DTYPE)
        tabs, tab = [], 0
        for i, r in enumerate(g.active):  # only requests with callback drivers carry tables
            if r.tables is not None:
                a0 = int(js["layer_offset"][i])
                d = layers["drv"][a0: a0 + int(nls[i])]
                d["table_index"][d["type"] == _lib.TABLE] += tab
                tabs.append(r.tables)
                tab += r.tables.size
        tables = np.concatenate(tabs) if tabs else None
        nbytes = _lib.plan_arena_size(js, layers, 0 if tables is None else tables.size, self.precision,
                                      self.log_dtype, self.lanes)
        sz = _lib.STATE_DTYPE.itemsize
        # state rows: gather the carried-over rows (a copy) before the buffer is rewritten
        kept = [(i, g.index[id(r)]) for i, r in enumerate(g.active) if r.state is None]
        old_rows = None
        if kept:
            dst, src = np.asarray(kept, dtype=np.int64).T
            old_rows = g.hstate.array[: g.n * sz].reshape(g.n, sz)[src]
        g.hstate = self._hget(n * sz, g.hstate)
        hv = g.hstate.array[: n * sz].reshape(n, sz)
        if kept:
            hv[dst] = old_rows
        for i, r in enumerate(g.active):
            if r.state is not None:
                hv[i] = r.state.view(np.uint8).reshape(sz)
                r.state = None
        g.hstatus = self._hget(n * _lib.STATUS_DTYPE.itemsize, g.hstatus)
        g.hstatus.array[: n * _lib.STATUS_DTYPE.itemsize] = 0
        g.harena = self._hget(nbytes, g.harena)
        g.plan = _lib.Plan(js, layers, _lib.constants_get(), tables, self.precision, self.log_dtype, self.lanes,
                           arena=(g.harena.host, g.harena.nbytes, g.stream.cuda_stream, True))
        g.n = n
        g.index = {id(r): i for i, r in enumerate(g.active)}

    def _retire(self, torch, g: _Group, log) -> None:
        """Called once the group's slice (and its status/state copies) are
        done: takes finished requests out of the group and starts the copies
        of their logs; _complete hands out the results when those land."""
        sh = g.hstatus.array[: g.n * _lib.STATUS_DTYPE.itemsize].view(_lib.STATUS_DTYPE)
        done = np.flatnonzero((sh["code"] != _lib.OK) | (sh["steps_done"].astype(np.int64) >= g.nsteps))
        if done.size == 0:
            return
        done = done.tolist()
        states = g.hstate.array[: g.n * _lib.STATE_DTYPE.itemsize].view(_lib.STATE_DTYPE)
        reqs = [(g.active[i], sh[i].copy(), states[i]["m"].copy()) for i in done]
        offs = np.concatenate([[0], np.cumsum([r.n_elems for r, _, _ in reqs])]).astype(np.int64)
        buf = torch.empty(int(offs[-1]), dtype=log.dtype, pin_memory=True)
        with torch.cuda.stream(g.stream):
            for q, (r, _, _) in enumerate(reqs):
                buf[int(offs[q]): int(offs[q + 1])].copy_(log[r.log_off: r.log_off + r.n_elems], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(g.stream)
        self._finishing.append((ev, buf, offs, reqs))
        gone = set(done)
        g.active = [r for i, r in enumerate(g.active) if i not in gone]
        g.changed = True

    def _complete(self) -> bool:
        """Results of retired requests whose log copies have landed; their
        log blocks return to the arena only now."""
        ready, waiting = [], []
        for f in self._finishing:
            (ready if f[0].query() else waiting).append(f)
        if not ready:
            return False
        self._finishing = waiting
        t = time.perf_counter()
        freed = []
        for ev, buf, offs, reqs in ready:
            allv = buf.numpy().astype(np.float64)
            for q, (r, st, m) in enumerate(reqs):
                j = r.jd[0]
                names = _channel_names(r.meta, int(j["log_mask"]))
                ne = int(j["n_entries"])
                blk = allv[int(offs[q]): int(offs[q]) + len(names) * ne].reshape(len(names), ne)
                out = {k: blk[c] for c, k in enumerate(names)}
                if r.meta.mr_mode == _lib.MR_STRIP and int(j["log_mask"]) & _lib.LOG_MR:
                    out["Rz"] = np.zeros(ne)
                out["time"] = (np.arange(ne, dtype=np.float64) * int(j["write_every"])) * float(j["dt"])
                nl = int(j["n_layers"])
                if int(st["code"]) == _lib.ERR_NAN:
                    k = int(st["entries_done"])
                    e = RuntimeError("NAN magnetisation")
                    e.partial_log = {key: v[:k] for key, v in out.items()}
                    r.future.set_exception(e)
                else:
                    self.steps_done += int(st["steps_done"])
                    r.future.set_result(Result(out, m[:nl].copy(), int(st["steps_done"]), r.t_submit, r.t_admit, t))
                freed.append(r)
        with self._cv:
            for r in freed:
                self.arena.free(r.log_off, r.n_elems)
            self._n_active -= len(freed)
        return True

    def _launch(self, torch, g: _Group, log) -> None:
        """One slice on the group's stream; status and state land in the
        group's host-mapped buffers."""
        if g.ev1 is None:
            g.ev0 = torch.cuda.Event(enable_timing=True)
            g.ev1 = torch.cuda.Event(enable_timing=True)
        g.ev0.record(g.stream)
        g.plan.execute_slice(log.data_ptr(), log.numel(), None, g.hstatus.device, g.hstate.device, g.slice_steps,
                             [g.stream.cuda_stream])
        g.ev1.record(g.stream)
        g.busy = True
        g.slices += 1
        self.ticks += 1

    def _loop(self) -> None:
        import torch

        try:
            torch.cuda.set_device(self.device)
            dtype = torch.float32 if self.log_dtype == _lib.LOG_F32 else torch.float64
            log = torch.empty(self.arena.capacity, dtype=dtype, device=self.device)
            torch.cuda.synchronize()
            streams = [torch.cuda.Stream(device=self.device) for _ in range(self.n_streams)]
            clk = time.perf_counter
            while True:
                t0 = clk()
                with self._cv:
                    self._admit(torch, streams)
                    idle = not self._groups and not self._pending
                    if idle and not self._finishing:
                        if self._stop:
                            return
                        self._cv.wait(timeout=0.05)
                        continue
                self.host_s["admit"] += clk() - t0
                t0 = clk()
                progressed = self._complete()
                self.host_s["complete"] += clk() - t0
                for key in list(self._groups):
                    g = self._groups[key]
                    if g.busy:
                        if not g.ev1.query():
                            continue
                        g.busy = False
                        ms = g.ev0.elapsed_time(g.ev1)
                        if ms > 0:  # steer the next slice towards slice_ms of device time
                            want = g.slice_steps * self.slice_ms / ms
                            g.slice_steps = int(min(max(want, 0.5 * g.slice_steps, 500), 2.0 * g.slice_steps,
                                                    2_000_000))
                        t0 = clk()
                        self._retire(torch, g, log)
                        self.host_s["retire"] += clk() - t0
                        progressed = True
                    if g.incoming or g.changed:
                        t0 = clk()
                        self._rebuild(torch, g)
                        g.changed = False
                        self.host_s["rebuild"] += clk() - t0
                    if g.active:
                        t0 = clk()
                        self._launch(torch, g, log)
                        self.host_s["launch"] += clk() - t0
                        progressed = True
                    elif not g.incoming:
                        streams.append(g.stream)  # back to the pool
                        del self._groups[key]
                if not progressed:
                    time.sleep(0.0005)
        except BaseException as e:  # fail every open request rather than hang its caller
            self.error = e
            with self._cv:
                for _, _, _, reqs in self._finishing:
                    for r, _, _ in reqs:
                        if not r.future.done():
                            r.future.set_exception(e)
                self._finishing = []
                for g in self._groups.values():
                    for r in g.active + g.incoming:
                        if not r.future.done():
                            r.future.set_exception(e)
                for r in self._pending:
                    if not r.future.done():
                        r.future.set_exception(e)
                self._groups, self._pending = {}, deque()
        finally:
            # host-mapped buffers go back to the driver once no launch can read them
            try:
                torch.cuda.synchronize()
            except Exception:
                pass
            for g in self._groups.values():
                for b in (g.harena, g.hstate, g.hstatus):
                    if b is not None:
                        self._hpool.append(b)
            for b in self._hpool:
                b.close()
            self._hpool = []


__all__ = ["Server", "Result", "LogArena"]
