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
import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .batch import _LANES, _LOG_MODES, JunctionBatch, _channel_names, _Meta
from .core import SolverMode


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
    future: Future
    t_submit: float
    t_admit: float = 0.0
    log_off: int = -1
    state: np.ndarray | None = None  # initial state row
    slot_rows: int = 0
    extra: dict = field(default_factory=dict)


class Server:
    """Continuous-batching request loop (see module docstring).

    precision: "fp32" or "fp64"; log_elements: capacity of the device log
    arena (elements of the log dtype); max_active: slots; slice_steps: steps
    per tick; lanes: fixed lane mapping ("per_layer" or "one"); n_streams:
    streams the groups of a tick are spread over."""

    def __init__(self, precision: str = "fp32", log_elements: int = 1 << 28, max_active: int = 65536,
                 slice_steps: int = 20000, lanes: str = "per_layer", n_streams: int = 4, log_dtype=None,
                 device: str = "cuda") -> None:
        if lanes not in ("per_layer", "one"):
            raise ValueError("a server fixes the lane mapping: 'per_layer' or 'one'")
        self.precision = _lib.FP32 if precision in ("fp32", _lib.FP32) else _lib.FP64
        if log_dtype is None:
            log_dtype = "f32" if self.precision == _lib.FP32 else "f64"
        self.log_dtype = _lib.LOG_F32 if log_dtype in ("f32", _lib.LOG_F32) else _lib.LOG_F64
        self.lanes = _LANES[lanes]
        self.arena = LogArena(log_elements)
        self.max_active = int(max_active)
        self.slice_steps = int(slice_steps)
        self.n_streams = max(1, int(n_streams))
        self.device = device
        self._pending: deque = deque()
        self._active: list = []
        self._cv = threading.Condition()
        self._stop = False
        self._thread = None
        self.ticks = 0
        self.steps_done = 0
        self.error = None

    # --- client side ------------------------------------------------------------
    def submit(self, junction, totalTime: float, timeStep: float = 1e-13, writeFrequency: float = 1e-11,
               solverMode=SolverMode.RK4, log: str = "full") -> Future:
        """Queues one runSimulation-equivalent request; the future resolves to
        a Result (or raises RuntimeError("NAN magnetisation"))."""
        jd, L, tabs, meta = JunctionBatch._pack(junction, totalTime, timeStep, writeFrequency, solverMode)
        jd = np.array(jd, dtype=_lib.JUNCTION_DTYPE).reshape(1)
        mask = _LOG_MODES[log] if isinstance(log, str) else int(log)
        if mask & _lib.LOG_ENERGY:
            raise ValueError("energy channels are a drop-in (Junction.runSimulation) feature")
        jd["log_mask"] = mask
        nch = _lib.n_channels(int(jd["n_layers"][0]), int(jd["mr_mode"][0]), mask)
        n_elems = nch * int(jd["n_entries"][0])
        if n_elems > self.arena.capacity:
            raise ValueError(f"request log ({n_elems} elements) exceeds the server's log arena "
                             f"({self.arena.capacity} elements)")
        st = np.zeros(1, dtype=_lib.STATE_DTYPE)
        st["nan_layer"] = -1
        st["m"][0, : L.size] = L["mag"]
        fut: Future = Future()
        r = _Req(jd, L.copy(), tabs, meta, n_elems, fut, time.perf_counter(), state=st)
        with self._cv:
            if self._stop:
                raise RuntimeError("server is stopped")
            self._pending.append(r)
            self._cv.notify()
        return fut

    def start(self) -> "Server":
        if self._thread is None:
            import torch

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
    def _admit(self) -> bool:
        changed = False
        while self._pending and len(self._active) < self.max_active:
            r = self._pending[0]
            off = self.arena.alloc(r.n_elems)
            if off is None:
                break  # wait for memory to free up (FIFO: no overtaking)
            self._pending.popleft()
            r.log_off = off
            r.t_admit = time.perf_counter()
            self._active.append(r)
            changed = True
        return changed

    def _rebuild(self, torch, old_state, old_index):
        """Plan + state of the current active set; carried-over requests keep
        their device state rows."""
        n = len(self._active)
        js = np.concatenate([r.jd for r in self._active])
        layers, tabs = [], []
        row = tab = 0
        for i, r in enumerate(self._active):
            L = r.layers.copy()
            if r.tables is not None:
                for q in range(L.size):
                    d = L[q]["drv"]
                    sel = d["type"] == _lib.TABLE
                    d["table_index"][sel] += tab
                tabs.append(r.tables)
                tab += r.tables.size
            js["layer_offset"][i] = row
            js["log_offset"][i] = r.log_off
            row += L.size
            layers.append(L)
        layers = np.concatenate(layers)
        tables = np.concatenate(tabs) if tabs else None
        nbytes = _lib.plan_arena_size(js, layers, 0 if tables is None else tables.size, self.precision,
                                      self.log_dtype, self.lanes)
        arena = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        plan = _lib.Plan(js, layers, _lib.constants_get(), tables, self.precision, self.log_dtype, self.lanes,
                         arena=(arena.data_ptr(), nbytes, torch.cuda.current_stream().cuda_stream))
        sz = _lib.STATE_DTYPE.itemsize
        state = torch.empty((n, sz), dtype=torch.uint8, device=self.device)
        fresh = [i for i, r in enumerate(self._active) if r.state is not None]
        kept = [i for i, r in enumerate(self._active) if r.state is None]
        if kept:
            src = torch.as_tensor([old_index[id(self._active[i])] for i in kept], device=self.device)
            state[torch.as_tensor(kept, device=self.device)] = old_state.index_select(0, src)
        if fresh:
            rows = np.concatenate([self._active[i].state for i in fresh]).view(np.uint8).reshape(len(fresh), sz)
            state[torch.as_tensor(fresh, device=self.device)] = torch.as_tensor(rows).to(self.device)
            for i in fresh:
                self._active[i].state = None
        index = {id(r): i for i, r in enumerate(self._active)}
        status = torch.zeros((n, _lib.STATUS_DTYPE.itemsize), dtype=torch.uint8, device=self.device)
        return plan, arena, state, index, status

    def _retire(self, torch, log, state, status_host):
        keep, done = [], []
        for i, r in enumerate(self._active):
            st = status_host[i]
            if int(st["code"]) != _lib.OK or int(st["steps_done"]) >= int(r.jd["n_steps"][0]):
                done.append((i, r, st))
            else:
                keep.append(r)
        if not done:
            return False
        states = state[torch.as_tensor([i for i, _, _ in done], device=self.device)].cpu().numpy()
        states = states.view(_lib.STATE_DTYPE).reshape(-1)
        for q, (i, r, st) in enumerate(done):
            j = r.jd[0]
            names = _channel_names(r.meta, int(j["log_mask"]))
            ne = int(j["n_entries"])
            blk = log[r.log_off: r.log_off + len(names) * ne].double().cpu().numpy().reshape(len(names), ne)
            self.arena.free(r.log_off, r.n_elems)
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
                r.future.set_result(Result(out, states[q]["m"][:nl].copy(), int(st["steps_done"]), r.t_submit,
                                           r.t_admit, time.perf_counter()))
        self._active = keep
        return True

    def _loop(self) -> None:
        import torch

        try:
            torch.cuda.set_device(self.device)
            dtype = torch.float32 if self.log_dtype == _lib.LOG_F32 else torch.float64
            log = torch.empty(self.arena.capacity, dtype=dtype, device=self.device)
            streams = [torch.cuda.Stream(device=self.device) for _ in range(self.n_streams)]
            sptr = [s.cuda_stream for s in streams]
            plan = arena = state = status = None
            index: dict = {}
            changed = False
            while True:
                with self._cv:
                    changed |= self._admit()
                    if not self._active:
                        if self._stop and not self._pending:
                            return
                        self._cv.wait(timeout=0.05)
                        continue
                if changed:
                    torch.cuda.synchronize()
                    plan, arena, state, index, status = self._rebuild(torch, state, index)
                    changed = False
                plan.execute_slice(log.data_ptr(), log.numel(), None, status.data_ptr(), state.data_ptr(),
                                   self.slice_steps, sptr)
                torch.cuda.synchronize()
                self.ticks += 1
                sh = status.cpu().numpy().view(_lib.STATUS_DTYPE).reshape(-1)
                with self._cv:
                    changed |= self._retire(torch, log, state, sh)
        except BaseException as e:  # fail every open request rather than hang its caller
            self.error = e
            with self._cv:
                for r in list(self._active) + list(self._pending):
                    if not r.future.done():
                        r.future.set_exception(e)
                self._active, self._pending = [], deque()


__all__ = ["Server", "Result", "LogArena"]
