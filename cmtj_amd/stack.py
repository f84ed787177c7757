"""Electrically coupled stacks of junctions: ``SeriesStack`` and
``ParallelStack`` with the reference's Python API (python/cmtj.cpp:361-426,
core/stack.hpp).  A stack runs on the GPU as one wave, member j on lane j,
the members' current recomputed every step from their magnetisations
(``cmtj_run_stacks``); logs, errors and defaults follow the reference.
"""
from __future__ import annotations

import copy

import numpy as np

from . import _lib
from .core import CVector, NullDriver, ScalarDriver, SolverMode, UpdateType


class _Stack:
    """Stack<T> (core/stack.hpp:16-411)."""

    _series = True

    def __init__(self, junctionList, topId_a: str = "free", bottomId: str = "bottom", phaseOffset: float = 0.0,
                 useKCL: bool = True) -> None:
        junctions = list(junctionList)
        # Stack::Stack, :128-152
        if len(junctions) < 2:
            raise RuntimeError("Stack must have at least 2 junctions!")
        if any(j.MR_mode != _lib.MR_CLASSIC for j in junctions):
            raise RuntimeError("Junction has a non-classic magnetoresitance mode!"
                               " Define the junction with Rp and Rap resistance values.")
        # the reference copies the junctions into the stack
        self.junctionList = [copy.deepcopy(j) for j in junctions]
        self.stackSize = len(self.junctionList)
        self.topId = topId_a
        self.bottomId = bottomId
        self.phaseOffset = float(phaseOffset)
        self.useKCL = bool(useKCL)
        self.couplingStrength = [0.0]
        self.delayed = True
        self.currentDriver = ScalarDriver.getConstantDriver(0.0)
        self.currentDriverSet = False
        self._log: dict[str, list] = {}

    # --- configuration (:154-258) -------------------------------------------
    def getCoupling(self, order: int) -> float:
        if not self.couplingStrength:
            raise RuntimeError("Coupling strength is not set!")
        if len(self.couplingStrength) == 1:
            return self.couplingStrength[0]
        return self.couplingStrength[order]

    def setDelayed(self, delayed: bool) -> None:
        if not delayed and not self.isTwoLayerMemberStack():
            raise RuntimeError("Non delayed coupling is only supported for 2 layer stacks!")
        self.delayed = bool(delayed)

    def setMagnetisation(self, junction: int, layerId: str, mag: CVector) -> None:
        self.junctionList[junction].setLayerMagnetisation(layerId, mag)

    def getMagnetisation(self, junction: int, layerId: str) -> CVector:
        return self.junctionList[junction].getLayerMagnetisation(layerId)

    def getJunction(self, junctionId: int):
        if junctionId < 0 or junctionId >= len(self.junctionList):
            raise IndexError("junction index out of range")
        return self.junctionList[junctionId]

    def setJunctionAnisotropyDriver(self, junctionId: int, layerId: str, k: ScalarDriver) -> None:
        self.junctionList[junctionId].setLayerAnisotropyDriver(layerId, k)

    def setOerstedFieldDriver(self, driver) -> None:
        for j in self.junctionList:
            j.setLayerOerstedFieldDriver("all", driver)

    def setExternalFieldDriver(self, driver) -> None:
        for j in self.junctionList:
            j.setLayerExternalFieldDriver("all", driver)

    def resetCoupledCurrentDriver(self) -> None:
        self.currentDriver = NullDriver()
        for j in self.junctionList:
            j.setLayerCurrentDriver("all", self.currentDriver)
        self.currentDriverSet = False

    def setCoupledCurrentDriver(self, driver: ScalarDriver) -> None:
        self.currentDriver = driver
        for j in self.junctionList:
            j.setLayerCurrentDriver("all", driver)
        self.currentDriverSet = True

    def setCouplingStrength(self, coupling) -> None:
        if np.ndim(coupling) == 0:
            self.couplingStrength = [float(coupling)]
            return
        vals = [float(c) for c in coupling]
        if len(vals) != self.stackSize - 1:
            raise RuntimeError("Coupling strength vector must have size of stack size - 1!")
        self.couplingStrength = vals

    def getPolarisationVector(self) -> CVector:
        probe = self.junctionList[0].getLayer(self.topId).referenceLayer
        for j in self.junctionList[1:]:
            if probe != j.getLayer(self.topId).referenceLayer:
                raise RuntimeError("Polarisation vectors are not equal in stack")
        if not probe.length():
            raise RuntimeError("Polarisation is not set!")
        return probe

    def isTwoLayerMemberStack(self) -> bool:
        for j in self.junctionList:
            if j.layerNo >= 3:
                raise RuntimeError("At least one junction has more than 2 layers! It's not supported now.")
            if j.layerNo != 2:
                return False
        return True

    # --- logs (:160-258) -----------------------------------------------------
    def clearLogs(self) -> None:
        for j in self.junctionList:
            j.clearLog()
        self._log = {}

    def getLog(self, id=None) -> dict:
        if id is None:
            return {k: (v[0] if len(v) == 1 else np.concatenate(v)).tolist() for k, v in self._log.items()}
        if isinstance(id, bool) or not isinstance(id, (int, np.integer)) or id < 0:
            raise TypeError("getLog(): the junction id is an unsigned integer")
        if id < len(self.junctionList):
            return self.junctionList[id].getLog()
        raise RuntimeError("Asking for id of a non-existing junction!")

    def saveLogs(self, fileSave: str) -> None:
        if fileSave == "":
            print("Ignoring file save to an empty filename")
            return
        log = self.getLog()
        with open(fileSave, "w") as f:
            f.write("".join(f"{k};" for k in log) + "\n")
            for i in range(len(log.get("time", []))):
                f.write("".join(f"{v[i]};" for v in log.values()) + "\n")

    # --- run -------------------------------------------------------------------
    def _execute(self, stack, js, layers, consts, tables=None):
        return _lib.run_stacks(stack, js, layers, consts, tables)

    def _pack(self, totalTime: float, timeStep: float, writeFrequency: float):
        """Descriptors of one run: (stack record, member records with layer
        and log offsets from 0, layer rows, channels per member, steps, W,
        log size, callback tables or None)."""
        callback = self.currentDriver.update == UpdateType.custom
        if callback and not self.useKCL:
            # runSimulationNonKCL scales the driver with ScalarDriver::operator*
            # (core/drivers.hpp:313-325), which builds a driver without the
            # callback: the reference then fails in the callback call
            raise RuntimeError("Error in Python callback: non-KCL coupling scales the stack current, "
                               "which drops a callback driver")
        # initializeSimulation, :36-62 (the time-step check first)
        packed = [j._descriptor(totalTime, timeStep, writeFrequency, False, SolverMode.RK4)
                  for j in self.junctionList]
        solvers = {int(p[0]["solver"]) for p in packed}
        if len(solvers) > 1:
            raise RuntimeError("Junctions have different solver modes!"
                               " Set the same solver mode for all junctions explicitly."
                               " Do not mix stochastic and deterministic solvers!")
        # the members' callback samples in one table, each member's indices
        # moved past the previous members' samples
        tabs, toff = [], 0
        for p in packed:
            if p[2] is not None and len(p[2]):
                drv = p[1]["drv"]
                sel = drv["type"] == _lib.TABLE
                drv["table_index"][sel] += toff
                tabs.append(np.asarray(p[2], dtype=np.float64))
                toff += len(p[2])
        two = self.isTwoLayerMemberStack()
        n_steps, W = packed[0][3], packed[0][4]
        n = self.stackSize
        if n > _lib.MAX_STACK:
            raise NotImplementedError(f"stacks of more than {_lib.MAX_STACK} junctions")
        st = np.zeros((), dtype=_lib.STACK_DTYPE)
        st["series"] = 1 if self._series else 0
        st["kcl"] = 1 if self.useKCL else 0
        st["delayed"] = 1 if self.delayed else 0
        st["n_junctions"] = n
        st["two_layer"] = 1 if two else 0
        st["phase_offset"] = self.phaseOffset
        if not two and n_steps > 0:
            st["pol"] = self.getPolarisationVector().tolist()
        if callback:
            # KCL reads the stack driver once per step, at the step's start
            # time (runSimulationKCL, :376): one sample per step
            from .core import _call_custom

            t = np.arange(n_steps, dtype=np.float64) * timeStep
            vals = np.fromiter((_call_custom(self.currentDriver._callback, float(x)) for x in t), dtype=np.float64,
                               count=n_steps)
            rec = np.zeros((), dtype=_lib.DRIVER_DTYPE)
            rec["type"] = _lib.TABLE
            rec["table_index"] = toff
            st["current"] = rec
            tabs.append(vals)
            toff += n_steps
        else:
            st["current"] = self.currentDriver._record()
        for q in range(n - 1):
            st["coupling"][q] = self.getCoupling(q)
        for q, j in enumerate(self.junctionList):
            if n_steps > 0:
                ids = j.getLayerIds()
                st["top"][q] = ids.index(j.getLayer(self.topId).id)
                if two:
                    st["bottom"][q] = ids.index(j.getLayer(self.bottomId).id)
        js = _lib.cat_records([p[0] for p in packed], _lib.JUNCTION_DTYPE)
        layers, nchs = [], []
        rows = off = 0
        for q, (jd, L, _t, _ns, _w) in enumerate(packed):
            js["layer_offset"][q] = rows
            js["log_offset"][q] = off
            rows += len(L)
            layers.append(L)
            nch = _lib.n_channels(int(jd["n_layers"]), int(jd["mr_mode"]), int(jd["log_mask"]))
            nchs.append(nch)
            off += (nch + 2) * int(jd["n_entries"])
        return (st, js, _lib.cat_records(layers, _lib.LAYER_DTYPE), nchs, n_steps, W, off,
                (np.concatenate(tabs) if tabs else None))

    def _finish(self, js, log, mag, status, nchs, W: int, timeStep: float) -> None:
        """Members take over their logs and final state; the stack log."""
        ne_done = int(status["entries_done"].min())
        for q, j in enumerate(self.junctionList):
            ne = int(js[q]["n_entries"])
            o = int(js[q]["log_offset"])
            blk = log[o:o + (nchs[q] + 2) * ne].reshape(nchs[q] + 2, ne)
            stq = status[q].copy()
            stq["entries_done"] = ne_done
            j._runs += 1
            j._apply(js[q], blk[: nchs[q]], mag[q, : j.layerNo], stq, timeStep, W)
        self._log_stack(js, log, nchs, ne_done, W, timeStep)
        if (status["code"] == _lib.ERR_NAN).any():
            raise RuntimeError("NAN magnetisation")
        if (status["code"] != _lib.OK).any():
            raise RuntimeError(f"engine status {int(status['code'].max())}")

    def runSimulation(self, totalTime: float, timeStep: float = 1e-13, writeFrequency: float = 1e-11) -> None:
        """Stack<T>::runSimulation (:296-300): KCL or non-KCL coupling, every
        member stepping with RK4 (Euler-Heun when stochastic)."""
        st, js, layers, nchs, n_steps, W, _, tables = self._pack(totalTime, timeStep, writeFrequency)
        if n_steps == 0:
            return
        log, mag, status = self._execute(st, js, layers, _lib.constants_get(), tables)
        self._finish(js, log, mag, status, nchs, W, timeStep)

    def _stack_resistance(self, R: list) -> np.ndarray:
        raise NotImplementedError

    def _log_stack(self, js, log, nchs, ne: int, W: int, dt: float) -> None:
        """logStackData (:206-217) for every logged step: stack resistance
        from the members' logged resistances, then I_j, C_j_{j+1}, time."""
        if ne == 0:
            return
        n = len(nchs)
        blocks = []
        for q in range(n):
            ne_q = int(js[q]["n_entries"])
            o = int(js[q]["log_offset"])
            blocks.append(log[o:o + (nchs[q] + 2) * ne_q].reshape(nchs[q] + 2, ne_q)[:, :ne])
        # the classic resistance follows the members' m channels
        R = [blocks[q][3 * int(js[q]["n_layers"])] for q in range(n)]

        def put(k, v):
            self._log.setdefault(k, []).append(np.asarray(v, dtype=np.float64))

        put("Resistance", self._stack_resistance(R))
        for q in range(n):
            put(f"I_{q}", blocks[q][nchs[q]])
            if q < n - 1:
                put(f"C_{q}_{q + 1}", blocks[q][nchs[q] + 1])
        put("time", (np.arange(ne, dtype=np.float64) * W) * dt)


class SeriesStack(_Stack):
    """SeriesStack<T> (core/stack.hpp:414-441): resistances add; pair
    coupling coupling * (m_j . p_j + m_{j+1} . p_{j+1}); phase offset grows
    with the member index."""

    _series = True

    def _stack_resistance(self, R: list) -> np.ndarray:
        acc = np.zeros_like(R[0])
        for r in R:
            acc = acc + r
        return acc


class ParallelStack(_Stack):
    """ParallelStack<T> (core/stack.hpp:443-480): conductances add; pair
    coupling coupling * (m_j . p_j - m_{j+1} . p_{j+1}); one phase offset for
    every member."""

    _series = False

    def _stack_resistance(self, R: list) -> np.ndarray:
        inv = np.zeros_like(R[0])
        for r in R:
            inv = inv + 1.0 / r
        return 1.0 / inv


def run_stacks(stacks, totalTime: float, timeStep: float = 1e-13, writeFrequency: float = 1e-11) -> None:
    """Runs many stacks in one launch (no reference counterpart: a parameter
    sweep over stacks, each stack as by its own runSimulation).  Every stack
    has the same member count and every member the same structure; logs and
    final states go to each stack as runSimulation leaves them.  Raises
    RuntimeError("NAN magnetisation") after the run if any stack hit a NaN."""
    stacks = list(stacks)
    if not stacks:
        return
    parts = [s._pack(totalTime, timeStep, writeFrequency) for s in stacks]
    if len({p[0]["n_junctions"].item() for p in parts}) != 1:
        raise ValueError("every stack of a batch has the same member count")
    if parts[0][4] == 0:
        return
    recs = _lib.cat_records([p[0] for p in parts], _lib.STACK_DTYPE)
    js, layers, tabs = [], [], []
    rows = off = toff = 0
    for k, (st, jq, L, _n, _s, _w, size, tab) in enumerate(parts):
        jq = jq.copy()
        jq["layer_offset"] += rows
        jq["log_offset"] += off
        js.append(jq)
        if tab is not None and len(tab):  # this stack's table indices follow the previous stacks' samples
            drv = L["drv"]
            drv["table_index"][drv["type"] == _lib.TABLE] += toff
            if int(recs[k]["current"]["type"]) == _lib.TABLE:
                recs[k]["current"]["table_index"] += toff
            tabs.append(tab)
            toff += len(tab)
        layers.append(L)
        rows += len(L)
        off += size
    js = _lib.cat_records(js, _lib.JUNCTION_DTYPE)
    log, mag, status = stacks[0]._execute(recs, js, _lib.cat_records(layers, _lib.LAYER_DTYPE), _lib.constants_get(),
                                          np.concatenate(tabs) if tabs else None)
    failed = False
    for k, (s, p) in enumerate(zip(stacks, parts)):
        n = s.stackSize
        sl = slice(k * n, (k + 1) * n)
        try:
            s._finish(js[sl], log, mag[sl], status[sl], p[3], p[5], timeStep)
        except RuntimeError:
            failed = True
    if failed:
        raise RuntimeError("NAN magnetisation")


__all__ = ["SeriesStack", "ParallelStack", "run_stacks"]
