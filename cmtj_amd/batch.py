"""Batched runs: many (uneven) junction instances in one device-resident plan.

The reference runs sweep points one ``runSimulation`` at a time, in a Python
loop (cmtj/utils/procedures.py:141, 280) or a process pool
(cmtj/utils/parallel.py:47, 84).  Here a sweep becomes one plan: descriptors
are uploaded to HBM once, instances are grouped by structure and ordered by
cost, and ``execute`` launches the kernels with no allocation.  Logs land in a
device tensor laid out [instance][channel][entry].

Two ways to build a batch:
  * ``JunctionBatch.from_junctions(list, ...)``: arbitrary junctions (layer
    counts, durations, time steps and solvers may all differ);
  * ``JunctionBatch.replicate(junction, n, ...)`` followed by vectorised
    per-instance overrides (fields, driver parameters, magnetisations,
    durations, seeds) -- the form for large sweeps.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass

import numpy as np

from . import _lib
from .core import Junction, SolverMode, _AXIAL_SLOTS, _SCALAR_SLOTS

_LOG_MODES = {
    "full": _lib.LOG_FULL,                       # m (x, y, z) of every layer + resistance
    "mx": _lib.LOG_MX | _lib.LOG_MR,             # m_x of every layer + resistance (PIMM spectra)
    "m": _lib.LOG_MX | _lib.LOG_MYZ,             # m only
    "r": _lib.LOG_MR,                            # resistance only (spin-diode voltages)
    "none": 0,                                   # final state only
}


@dataclass
class _Meta:
    layer_ids: list
    mr_mode: int
    rtag: str


def _channel_names(meta: _Meta, mask: int) -> list:
    names = []
    for lid in meta.layer_ids:
        if mask & _lib.LOG_MX:
            names.append(f"{lid}_mx")
        if mask & _lib.LOG_MYZ:
            names += [f"{lid}_my", f"{lid}_mz"]
    if mask & _lib.LOG_MR:
        if meta.mr_mode == _lib.MR_CLASSIC:
            names.append(meta.rtag)
        elif meta.mr_mode == _lib.MR_STRIP:
            names += ["Rx", "Ry"]
    if mask & _lib.LOG_TIME:  # Dormand-Prince: the adaptive entry times
        names.append("time")
    return names


class BatchResult:
    """Outputs of one execute: device log, final magnetisations, status."""

    def __init__(self, batch: "JunctionBatch", log, mag, status) -> None:
        self.batch = batch
        self.log = log          # torch tensor [log_elements]
        self.final_m = mag      # torch tensor [n, MAX_LAYERS, 3] (float64)
        self.status_dev = status

    def status(self) -> np.ndarray:
        import torch

        raw = self.status_dev.cpu().numpy().view(np.uint8)
        return raw.view(_lib.STATUS_DTYPE).reshape(-1)

    def instance_log(self, i: int) -> dict:
        """Log of instance i as the dict getLog returns (numpy arrays)."""
        b = self.batch
        j = b.js[i]
        meta = b.metas[b.meta_index[i]]
        names = _channel_names(meta, int(j["log_mask"]))
        ne = int(j["n_entries"])
        off = int(j["log_offset"])
        block = self.log[off: off + len(names) * ne].double().cpu().numpy().reshape(len(names), ne)
        dp = int(j["log_mask"]) & _lib.LOG_TIME
        if dp:  # adaptive runs fill their log capacity only in part
            block = block[:, : int(self.status()[i]["entries_done"])]
        out = {k: block[c] for c, k in enumerate(names)}
        if meta.mr_mode == _lib.MR_STRIP and int(j["log_mask"]) & _lib.LOG_MR:
            out["Rz"] = np.zeros(block.shape[1])
        if not dp:
            out["time"] = (np.arange(ne, dtype=np.float64) * int(j["write_every"])) * float(j["dt"])
        return out

    def channel(self, name: str):
        """[n, entries] view of one channel when every instance logs the same
        number of entries (a uniform sweep); otherwise raises."""
        b = self.batch
        ne = b.js["n_entries"]
        if not np.all(ne == ne[0]):
            raise ValueError("instances log different numbers of entries")
        meta = b.metas[0]
        names = _channel_names(meta, int(b.js[0]["log_mask"]))
        c = names.index(name)
        nch = len(names)
        n = len(b.js)
        return self.log[: n * nch * int(ne[0])].view(n, nch, int(ne[0]))[:, c, :]


_LANES = {"auto": 0, "one": 1, "per_layer": 2, "per_component": 3, "queue": 4}


class JunctionBatch:
    """A packed batch of junction instances (see module docstring)."""

    def __init__(self, js, layers, tables, metas, meta_index, precision="fp32", log="full", log_dtype=None):
        self.js = js
        self.layers = layers
        self.tables = tables
        self.metas = metas
        self.meta_index = np.asarray(meta_index, dtype=np.int32)
        self.precision = _lib.FP32 if precision in ("fp32", _lib.FP32) else _lib.FP64
        if log_dtype is None:
            log_dtype = "f32" if self.precision == _lib.FP32 else "f64"
        self.log_dtype = _lib.LOG_F32 if log_dtype in ("f32", _lib.LOG_F32) else _lib.LOG_F64
        self.set_log(log)
        self.lanes = "auto"
        self._plan = None
        self._bufs = None

    # --- construction ---------------------------------------------------------
    @staticmethod
    def _pack(j: Junction, totalTime, timeStep, writeFrequency, solverMode):
        """Packs one run of `j`.  Every packing takes the junction's next
        thermal-noise stream, as every runSimulation call does: a junction
        packed twice (submitted twice, or listed twice in a batch) runs two
        independent noise paths."""
        jd, L, tabs, _, _ = j._descriptor(totalTime, timeStep, writeFrequency, False, solverMode)
        jd["stream"] = j._runs
        j._runs += 1
        meta = _Meta(j.getLayerIds(), j.MR_mode, j.Rtag if j.layerNo > 1 else "R")
        return jd, L, tabs, meta

    @classmethod
    def from_junctions(cls, junctions, totalTime, timeStep=1e-13, writeFrequency=1e-11,
                       solverMode=SolverMode.RK4, precision="fp32", log="full", log_dtype=None):
        n = len(junctions)

        def per(v):
            return list(v) if np.ndim(v) else [v] * n

        T, dt, wf, sm = per(totalTime), per(timeStep), per(writeFrequency), per(solverMode)
        js = np.zeros(n, dtype=_lib.JUNCTION_DTYPE)
        rows, tabs, metas, midx = [], [], [], []
        row = tab = 0
        for i, jn in enumerate(junctions):
            jd, L, tb, meta = cls._pack(jn, T[i], dt[i], wf[i], sm[i])
            if tb is not None:
                for q in range(L.size):
                    d = L[q]["drv"]
                    sel = d["type"] == _lib.TABLE
                    d["table_index"][sel] += tab
                tabs.append(tb)
                tab += tb.size
            jd["layer_offset"] = row
            row += L.size
            js[i] = jd
            rows.append(L)
            metas.append(meta)
            midx.append(i)
        layers = np.concatenate(rows)
        tables = np.concatenate(tabs) if tabs else None
        return cls(js, layers, tables, metas, midx, precision, log, log_dtype)

    @classmethod
    def replicate(cls, junction: Junction, n: int, totalTime, timeStep=1e-13, writeFrequency=1e-11,
                  solverMode=SolverMode.RK4, precision="fp32", log="full", log_dtype=None):
        jd, L, tb, meta = cls._pack(junction, np.max(totalTime), timeStep, writeFrequency, solverMode)
        if tb is not None:
            raise ValueError("callback drivers cannot be replicated; use from_junctions")
        nl = L.size
        # byte-level copies: numpy's structured-dtype repeat is element-wise and slow
        js = np.tile(np.ascontiguousarray(jd.reshape(1)).view(np.uint8), n).view(_lib.JUNCTION_DTYPE)
        layers = np.tile(np.ascontiguousarray(L).view(np.uint8), n).view(_lib.LAYER_DTYPE)
        js["layer_offset"] = np.arange(n, dtype=np.uint32) * nl
        # instance i runs the junction's stream base + i; the junction moves past them
        base = int(jd["stream"])
        js["stream"] = base + np.arange(n, dtype=np.uint64)
        junction._runs = base + n
        b = cls(js, layers, None, [meta], np.zeros(n, dtype=np.int32), precision, log, log_dtype)
        b.set_times(totalTime, timeStep, writeFrequency)
        return b

    # --- per-instance overrides (vectorised) ----------------------------------
    def _layer_rows(self, layer_id: str) -> np.ndarray:
        """Row indices [n, k] of the named layer(s) in every instance."""
        out = []
        for i in range(len(self.js)):
            meta = self.metas[self.meta_index[i]]
            base = int(self.js[i]["layer_offset"])
            ids = range(len(meta.layer_ids)) if layer_id == "all" else [meta.layer_ids.index(layer_id)]
            out.append([base + q for q in ids])
        return np.asarray(out)

    def _rows(self, layer_id):
        if len(self.metas) == 1:
            meta = self.metas[0]
            ids = range(len(meta.layer_ids)) if layer_id == "all" else [meta.layer_ids.index(layer_id)]
            base = self.js["layer_offset"].astype(np.int64)[:, None]
            return base + np.asarray(list(ids))[None, :]
        return self._layer_rows(layer_id)

    def set_driver(self, layer_id: str, slot: str, param: str, values) -> None:
        """Sets one parameter of a driver slot per instance.

        slot: a scalar slot (anisotropy, current, temperature, ...) with values
        of shape [n], or an axial slot (external, oersted, ...) with values of
        shape [n, 3].  param: constant, amplitude, frequency, phase, period,
        cycle, t_start, t_stop, edge_time, steady_time, or type."""
        rows = self._rows(layer_id)
        v = np.asarray(values)
        drv = self.layers["drv"]
        if slot in _SCALAR_SLOTS:
            s = _SCALAR_SLOTS[slot]
            for q in range(rows.shape[1]):
                drv[param][rows[:, q], s] = v
        elif slot in _AXIAL_SLOTS:
            s = _AXIAL_SLOTS[slot]
            v = v.reshape(len(self.js), 3)
            for q in range(rows.shape[1]):
                for a in range(3):
                    drv[param][rows[:, q], s + a] = v[:, a]
        else:
            raise KeyError(slot)
        self._plan = None

    def set_external_field(self, layer_id: str, H) -> None:
        """Constant external field [n, 3] (A/m)."""
        self.set_driver(layer_id, "external", "type", np.zeros((len(self.js), 3), dtype=np.int32))
        self.set_driver(layer_id, "external", "constant", H)

    def set_magnetisation(self, layer_id: str, m) -> None:
        """Initial magnetisation [n, 3], normalised like setMagnetisation."""
        m = np.asarray(m, dtype=np.float64).reshape(len(self.js), 3)
        nrm = np.linalg.norm(m, axis=1, keepdims=True)
        if np.any(nrm == 0):
            raise RuntimeError("Initial magnetisation was set to a zero vector!")
        rows = self._rows(layer_id)
        for q in range(rows.shape[1]):
            self.layers["mag"][rows[:, q]] = m / nrm
        self._plan = None

    def set_seeds(self, seeds) -> None:
        self.js["seed"] = np.asarray(seeds, dtype=np.uint64)
        self._plan = None

    def set_times(self, totalTime, timeStep=None, writeFrequency=None) -> None:
        """Per-instance durations / steps with the reference's truncations
        (core/junction.hpp:1956-1959)."""
        if np.any(self.js["solver"] == _lib.DORMAND_PRINCE):
            raise NotImplementedError("Dormand-Prince batches take their times from from_junctions")
        n = len(self.js)
        dt = np.broadcast_to(np.asarray(self.js["dt"] if timeStep is None else timeStep, dtype=np.float64), (n,))
        T = np.broadcast_to(np.asarray(totalTime, dtype=np.float64), (n,))
        if writeFrequency is None:
            wf = self.js["write_every"].astype(np.float64) * dt
            W = self.js["write_every"].astype(np.uint32)
        else:
            wf = np.broadcast_to(np.asarray(writeFrequency, dtype=np.float64), (n,))
            if np.any(dt > wf):
                raise RuntimeError("The time step cannot be larger than write frequency!")
            W = np.maximum(1, np.trunc(wf / dt)).astype(np.uint32)
        steps = np.maximum(0, np.trunc(T / dt)).astype(np.uint32)
        self.js["dt"] = dt
        self.js["n_steps"] = steps
        self.js["write_every"] = W
        self.js["n_entries"] = (steps.astype(np.uint64) + W - 1) // W
        self._layout()
        self._plan = None

    def set_log(self, log) -> None:
        mask = _LOG_MODES[log] if isinstance(log, str) else int(log)
        if mask & _lib.LOG_ENERGY:
            raise ValueError("energy channels are a drop-in (Junction.runSimulation) feature")
        # Dormand-Prince instances log their adaptive entry times as well
        dp = self.js["solver"] == _lib.DORMAND_PRINCE
        self.js["log_mask"] = np.where(dp & (mask != 0), mask | _lib.LOG_TIME, mask)
        self._layout()
        self._plan = None

    def _layout(self) -> None:
        nch = np.array([_lib.n_channels(int(self.js[i]["n_layers"]), int(self.js[i]["mr_mode"]),
                                        int(self.js[i]["log_mask"]))
                        for i in range(len(self.js))], dtype=np.uint64) \
            if len(self.metas) > 1 or len(np.unique(self.js["log_mask"])) > 1 else \
            np.full(len(self.js), _lib.n_channels(int(self.js[0]["n_layers"]), int(self.js[0]["mr_mode"]),
                                                  int(self.js[0]["log_mask"])), dtype=np.uint64)
        sizes = nch * self.js["n_entries"].astype(np.uint64)
        offs = np.zeros(len(self.js), dtype=np.uint64)
        if len(sizes) > 1:
            offs[1:] = np.cumsum(sizes[:-1])
        self.js["log_offset"] = offs
        self.log_elements = int(sizes.sum())
        self.total_steps = int(self.js["n_steps"].astype(np.uint64).sum())

    def subset(self, idx, precision=None, log=None, log_dtype=None) -> "JunctionBatch":
        """A batch of the instances idx (same descriptors, seeds and noise
        streams), e.g. to re-run a sample in another precision."""
        idx = np.asarray(idx, dtype=np.int64)
        js = self.js[idx].copy()
        rows, off = [], 0
        for q, i in enumerate(idx):
            a, nl = int(self.js[i]["layer_offset"]), int(self.js[i]["n_layers"])
            rows.append(self.layers[a: a + nl])
            js["layer_offset"][q] = off
            off += nl
        b = JunctionBatch(js, np.concatenate(rows), self.tables, self.metas, self.meta_index[idx],
                          precision=self.precision if precision is None else precision,
                          log=int(self.js[0]["log_mask"]) if log is None else log, log_dtype=log_dtype)
        b.lanes = self.lanes
        return b

    def set_lanes(self, lanes: str) -> "JunctionBatch":
        """Lane mapping of the instances: "one" (all layers of an instance in
        one GPU lane), "per_layer" (one lane per layer), "per_component" (one
        lane per vector component of each layer: binary32 RK4 groups of the
        lean feature set; binary64 groups of every feature but energy
        channels), "auto" (per component where such a kernel covers
        the group and the batch leaves SIMDs idle, else per layer for
        multilayer instances) or "queue" (single-layer Dormand-Prince groups
        on a persistent grid whose lanes take instances from a device work
        queue; other groups as "auto")."""
        if lanes not in _LANES:
            raise ValueError(f"lanes must be one of {sorted(_LANES)}")
        self.lanes = lanes
        self._plan = None
        return self

    # --- execution --------------------------------------------------------------
    def plan(self) -> _lib.Plan:
        if self._plan is None:
            consts = _lib.constants_get()
            self._plan = _lib.Plan(self.js, self.layers, consts, self.tables, self.precision, self.log_dtype,
                                   _LANES[self.lanes])
            self._bufs = None
        return self._plan

    def _buffers(self, device):
        import torch

        if self._bufs is None:
            dtype = torch.float32 if self.log_dtype == _lib.LOG_F32 else torch.float64
            n = len(self.js)
            log = torch.empty(max(self.log_elements, 1), dtype=dtype, device=device)
            mag = torch.zeros((n, _lib.MAX_LAYERS, 3), dtype=torch.float64, device=device)
            st = torch.zeros((n, _lib.STATUS_DTYPE.itemsize), dtype=torch.uint8, device=device)
            self._bufs = (log, mag, st)
        return self._bufs

    def execute(self, stream=None, device="cuda") -> BatchResult:
        """Launches every instance once (asynchronous on `stream`, default the
        current torch stream)."""
        import torch

        p = self.plan()
        log, mag, st = self._buffers(device)
        if stream is None:
            stream = torch.cuda.current_stream()
        sp = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        p.execute(log.data_ptr() if self.log_elements else None, log.numel(), mag.data_ptr(), st.data_ptr(), sp)
        return BatchResult(self, log, mag, st)

    @property
    def fp64_groups(self) -> int:
        """Structure groups of an FP32 batch that run binary64 kernels (spin
        torques or scalar drives on one lane per component, where no
        binary32 kernel covers them): their arithmetic is the drop-in's."""
        return int(self.plan().info.fp64_groups)

    def capture(self, device="cuda") -> "JunctionBatch":
        """Captures one execute of the batch, bound to its device buffers, as
        a HIP graph (cmtj_plan_graph_create): one kernel node per structure
        group, no edges between them.  replay() then runs it."""
        p = self.plan()
        log, mag, st = self._buffers(device)
        p.graph_create(log.data_ptr() if self.log_elements else None, log.numel(), mag.data_ptr(), st.data_ptr())
        self._graph_bufs = self._bufs
        return self

    def replay(self, stream=None) -> BatchResult:
        """Launches the captured graph (asynchronous on `stream`, default the
        current torch stream); the same results as execute()."""
        import torch

        if getattr(self, "_graph_bufs", None) is None or self._bufs is not self._graph_bufs:
            raise RuntimeError("capture() the batch first")
        if stream is None:
            stream = torch.cuda.current_stream()
        sp = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        self.plan().graph_launch(sp)
        return BatchResult(self, *self._graph_bufs)

    def run_sliced(self, slice_steps: int, n_streams: int = 1, device="cuda") -> BatchResult:
        """Runs every instance to the end in slices of at most slice_steps
        steps (cmtj_plan_execute_slice), state kept on the device between
        slices; the result equals run()'s."""
        import torch

        p = self.plan()
        log, mag, st = self._buffers(device)
        init = np.zeros(len(self.js), dtype=_lib.STATE_DTYPE)
        init["nan_layer"] = -1
        for i in range(len(self.js)):
            off, nl = int(self.js[i]["layer_offset"]), int(self.js[i]["n_layers"])
            init["m"][i, :nl] = self.layers["mag"][off: off + nl]
        state = torch.as_tensor(init.view(np.uint8).reshape(len(self.js), -1)).to(device)
        streams = [torch.cuda.Stream(device=device) for _ in range(max(1, n_streams))]
        need = self.js["n_steps"].astype(np.int64)
        while True:
            torch.cuda.synchronize()
            p.execute_slice(log.data_ptr() if self.log_elements else None, log.numel(), mag.data_ptr(),
                            st.data_ptr(), state.data_ptr(), slice_steps, [s.cuda_stream for s in streams])
            torch.cuda.synchronize()
            sh = st.cpu().numpy().view(_lib.STATUS_DTYPE).reshape(-1)
            if np.all((sh["code"] != _lib.OK) | (sh["steps_done"].astype(np.int64) >= need)):
                break
        return BatchResult(self, log, mag, st)

    def run(self, device="cuda") -> BatchResult:
        """execute + synchronize + raise on a NaN instance."""
        import torch

        r = self.execute(device=device)
        torch.cuda.synchronize()
        return r


def run_batch(junctions, totalTime, timeStep=1e-13, writeFrequency=1e-11, solverMode=SolverMode.RK4,
              precision="fp64", update=True) -> list:
    """Runs each junction once, as its own runSimulation would, in one launch.

    Returns the per-junction log dicts (lists, as getLog) and, with
    update=True, appends them to each junction's log and takes over the final
    magnetisations -- the loop ``for j in junctions: j.runSimulation(...)``
    in one call."""
    import torch

    b = JunctionBatch.from_junctions(junctions, totalTime, timeStep, writeFrequency, solverMode,
                                     precision=precision, log="full", log_dtype="f64")
    r = b.run()
    mags = r.final_m.cpu().numpy()
    st = r.status()
    out = []
    for i, j in enumerate(junctions):
        lg = r.instance_log(i)
        out.append({k: v.tolist() for k, v in lg.items()})
        if update:
            from .core import CVector

            for k, v in lg.items():
                j._append(k, v)
            j.logLength += int(st[i]["entries_done"])
            for q, l in enumerate(j.layers):
                l.mag = CVector(*mags[i, q])
    bad = [i for i in range(len(junctions)) if int(st[i]["code"]) == _lib.ERR_NAN]
    if bad:
        raise RuntimeError(f"NAN magnetisation (instances {bad[:8]})")
    return out


__all__ = ["JunctionBatch", "BatchResult", "run_batch"]
