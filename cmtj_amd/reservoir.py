"""Dipole-coupled groups of junctions: ``GroupInteraction`` and the
interaction functions of the reference's ``cmtj.reservoir`` module
(python/cmtj.cpp:428-446, core/reservoir.hpp:1-230).

A group runs on the GPU as one workgroup, element j on lane j: every step
the elements' top-layer magnetisations go through LDS, each lane sums the
dipole field of the others on its own top layer and takes its step
(``cmtj_run_groups``).  Logs, errors and defaults follow the reference.
The interaction functions themselves are also callable from Python (one
host evaluation, as in the reference's bindings).
"""
from __future__ import annotations

import copy
import math

import numpy as np

from . import _lib
from .core import CVector, SolverMode


def _vec(v) -> np.ndarray:
    return np.array([v.x, v.y, v.z] if isinstance(v, CVector) else list(v), dtype=np.float64)


def _length(r: np.ndarray) -> float:
    # CVector::length, core/cvector.hpp:179-181
    return math.sqrt(r[0] ** 2 + r[1] ** 2 + r[2] ** 2)


def nullDipoleInteraction(r1, r2, layer1, layer2) -> CVector:
    """nullDipoleInteraction (core/reservoir.hpp:63-68): no coupling."""
    return CVector(0.0, 0.0, 0.0)


def computeDipoleInteraction(r1, r2, layer1, layer2) -> CVector:
    """computeDipoleInteraction (core/reservoir.hpp:79-96, Kanao et al. 2019):
    the point-dipole field at r1 of layer2's moment at r2,
    (Ms2 / mu0) V2 (3 (m2 . r) r / |r|^5 - m2 / |r|^3), r = r1 - r2."""
    from .constants import PhysicalConstants

    rij = _vec(r1) - _vec(r2)
    if _length(rij) < 1e-10:
        raise RuntimeError("Points are too close for stable dipole calculation")
    r3, r5 = _length(rij) ** 3, _length(rij) ** 5
    m1 = _vec(layer2.mag)
    pref = (layer2.Ms / PhysicalConstants.magnetic_permeability()) * (layer2.thickness * layer2.cellSurface)
    h = pref * (3 * float(m1 @ rij) * rij / r5 - m1 / r3)
    return CVector(*h.tolist())


def computeDipoleInteractionNoumra(r1, r2, layer1, layer2) -> CVector:
    """computeDipoleInteractionNoumra (core/reservoir.hpp:106-119, Nomura et
    al. 2019): T m2 (Ms2 V2 / mu0) with the dipole tensor
    T = (r r^T - |r|^2 / 3) 3 / (4 pi |r|^5) (getDipoleTensorFromRelPositions,
    :36-56)."""
    from .constants import PhysicalConstants

    rij = _vec(r1) - _vec(r2)
    if _length(rij) < 1e-14:
        raise RuntimeError("Points are too close for stable dipole calculation")
    r_mag = _length(rij) ** 2
    mult = 3 / (4 * math.pi * _length(rij) ** 5)
    T = (np.outer(rij, rij) - np.eye(3) * (r_mag / 3)) * mult
    m1 = _vec(layer2.mag)
    h = (T @ m1) * ((layer2.Ms * layer2.thickness * layer2.cellSurface) / PhysicalConstants.magnetic_permeability())
    return CVector(*h.tolist())


_KINDS = {
    nullDipoleInteraction: _lib.INTERACTION_NULL,
    computeDipoleInteraction: _lib.INTERACTION_KANAO,
    computeDipoleInteractionNoumra: _lib.INTERACTION_NOMURA,
}


class GroupInteraction:
    """GroupInteraction (core/reservoir.hpp:121-230)."""

    def __init__(self, coordinateMatrix, junctionList, topId: str = "free") -> None:
        coords = [CVector(*_vec(c).tolist()) for c in coordinateMatrix]
        junctions = list(junctionList)
        # GroupInteraction::GroupInteraction, :150-175
        if len(coords) != len(junctions):
            raise RuntimeError("Coordinate matrix and junction list must have the same size!")
        if not coords or not junctions:
            raise RuntimeError("Coordinate matrix and junction list cannot be empty!")
        for i in range(len(coords)):
            for j in range(i + 1, len(coords)):
                if coords[i] == coords[j]:
                    raise RuntimeError("Coordinate vectors must be unique!")
        self.topId = str(topId)
        self.coordinateMatrix = coords
        # the reference copies the junctions into the group
        self.junctionList = [copy.deepcopy(j) for j in junctions]
        self.noElements = len(self.junctionList)
        self.interactionFunc = computeDipoleInteraction

    def setInteractionFunction(self, interactionFunc) -> None:
        self.interactionFunc = interactionFunc

    def clearLogs(self) -> None:
        for j in self.junctionList:
            j.clearLog()

    def getLog(self, id) -> dict:
        if isinstance(id, bool) or not isinstance(id, (int, np.integer)) or id < 0:
            raise TypeError("getLog(): the junction id is an unsigned integer")
        if id < len(self.junctionList):
            return self.junctionList[id].getLog()
        raise RuntimeError("Asking for id of a non-existing junction!")

    # --- run -------------------------------------------------------------------
    def _kind(self) -> int:
        # the module's three functions run on the device from the positions;
        # any other callable is tabulated (_pair_terms)
        return _KINDS.get(self.interactionFunc, _lib.INTERACTION_TENSOR)

    def _pair_terms(self) -> np.ndarray:
        """The interaction callable as per-pair affine terms: for every
        ordered pair (i, j), f(r_i, r_j, layer_i, layer_j) = T_ij m_j + b_ij,
        read off calls with m_j = +-x, +-y, +-z (T's columns are the odd
        parts, b the even part) on copies of the top layers, and checked at
        random unit m_j and m_i: a callable that is not affine in m_j, or
        that reads m_i, raises NotImplementedError (the device cannot call
        Python; the reference calls the function per pair and step,
        core/reservoir.hpp:135-142).  Returns [n][12][n] ([j][k][i]: T_ij
        row-major for k < 9, b_ij for k >= 9), zero on the diagonal."""
        f = self.interactionFunc
        n = self.noElements
        tops = [copy.deepcopy(j.getLayer(self.topId)) for j in self.junctionList]
        rng = np.random.default_rng(0x5EED)
        out = np.zeros((n, 12, n), dtype=np.float64)
        eye = np.eye(3)
        for i in range(n):
            for j in range(n):
                if i == j:
                    continue
                li, lj = tops[i], tops[j]

                def call(mj, mi=None):
                    lj.mag = CVector(*[float(x) for x in mj])
                    if mi is not None:
                        li.mag = CVector(*[float(x) for x in mi])
                    return _vec(f(self.coordinateMatrix[i], self.coordinateMatrix[j], li, lj))

                fp = [call(eye[k]) for k in range(3)]
                fm = [call(-eye[k]) for k in range(3)]
                T = np.stack([(fp[k] - fm[k]) * 0.5 for k in range(3)], axis=1)
                b = sum((fp[k] + fm[k]) * 0.5 for k in range(3)) / 3.0
                for _ in range(2):
                    u = rng.normal(size=3)
                    w = rng.normal(size=3)
                    u /= np.linalg.norm(u)
                    w /= np.linalg.norm(w)
                    got = call(u, w)
                    want = T @ u + b
                    scale = float(np.abs(T).sum() + np.abs(b).sum() + np.abs(got).sum())
                    if not np.all(np.abs(got - want) <= 1e-12 * scale + 1e-300):
                        raise NotImplementedError(
                            "the interaction function is not affine in the source layer's magnetisation "
                            "(or reads the target's): the GPU engine applies f = T(r_i, r_j) m_j + b per pair")
                out[j, :9, i] = T.ravel()
                out[j, 9:, i] = b
        return out

    def _pack(self, totalTime: float, timeStep: float, writeFrequency: float):
        """Descriptors of one run: (group record, element records with layer
        and log offsets from 0, layer rows, channels per element, steps, W)."""
        kind = self._kind()
        if self.noElements > _lib.MAX_GROUP:
            raise NotImplementedError(f"groups of more than {_lib.MAX_GROUP} junctions")
        # runSimulation, :181-212: every element with RK4 (Euler-Heun when
        # stochastic); the time-step check first
        packed = [j._descriptor(totalTime, timeStep, writeFrequency, False, SolverMode.RK4)
                  for j in self.junctionList]
        solvers = {int(p[0]["solver"]) for p in packed}
        if len(solvers) > 1:
            raise RuntimeError("Junctions have different solver modes!"
                               " Set the same solver mode for all junctions explicitly."
                               " Do not mix stochastic and deterministic solvers!")
        if any(p[2] is not None and len(p[2]) for p in packed):
            raise NotImplementedError("callback drivers on group elements are not provided by the GPU engine")
        n_steps, W = packed[0][3], packed[0][4]
        n = self.noElements
        g = np.zeros((), dtype=_lib.GROUP_DTYPE)
        g["n_elements"] = n
        g["interaction"] = kind
        for q, (j, c) in enumerate(zip(self.junctionList, self.coordinateMatrix)):
            if n_steps > 0:
                g["top"][q] = j.getLayerIds().index(j.getLayer(self.topId).id)
            g["pos"][q] = [c.x, c.y, c.z]
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
            off += nch * int(jd["n_entries"])
        pt = self._pair_terms() if (kind == _lib.INTERACTION_TENSOR and n_steps > 0) else None
        return g, js, _lib.cat_records(layers, _lib.LAYER_DTYPE), nchs, n_steps, W, off, pt

    def _execute(self, g, js, layers, consts, pair_terms=None):
        return _lib.run_groups(g, js, layers, consts, pair_terms)

    def _finish(self, js, log, mag, status, nchs, W: int, timeStep: float) -> None:
        """Elements take over their logs and final state."""
        ne_done = int(status["entries_done"].min())
        for q, j in enumerate(self.junctionList):
            ne = int(js[q]["n_entries"])
            o = int(js[q]["log_offset"])
            blk = log[o:o + nchs[q] * ne].reshape(nchs[q], ne)
            stq = status[q].copy()
            stq["entries_done"] = ne_done
            j._runs += 1
            j._apply(js[q], blk, mag[q, : j.layerNo], stq, timeStep, W)
        if (status["code"] == _lib.ERR_NAN).any():
            raise RuntimeError("NAN magnetisation")
        if (status["code"] != _lib.OK).any():
            raise RuntimeError(f"engine status {int(status['code'].max())}")

    def runSimulation(self, totalTime: float, timeStep: float = 1e-13, writeFrequency: float = 1e-11) -> None:
        """GroupInteraction::runSimulation (:181-221): each step the group
        field from the top layers at the step's start, then one step of every
        element; every element logs each write_every steps."""
        g, js, layers, nchs, n_steps, W, _, pt = self._pack(totalTime, timeStep, writeFrequency)
        if n_steps == 0:
            return
        log, mag, status = self._execute(g, js, layers, _lib.constants_get(), pt)
        self._finish(js, log, mag, status, nchs, W, timeStep)


def run_groups(groups, totalTime: float, timeStep: float = 1e-13, writeFrequency: float = 1e-11) -> None:
    """Runs many GroupInteraction objects of one element count and structure
    in one launch (group g on workgroup g): each ends as if its own
    runSimulation had run.  Replaces a host loop over groups (e.g. a
    reservoir's input sweep); no single-call counterpart in the reference."""
    groups = list(groups)
    if not groups:
        return
    packs = [g._pack(totalTime, timeStep, writeFrequency) for g in groups]
    n_steps, W = packs[0][4], packs[0][5]
    if n_steps == 0:
        return
    recs = _lib.cat_records([p[0] for p in packs], _lib.GROUP_DTYPE)
    n = groups[0].noElements
    pts = None
    if any(p[7] is not None for p in packs):  # every group's slot of the pair terms, zeros for the others
        pts = np.zeros((len(packs), n, 12, n), dtype=np.float64)
        for k, p in enumerate(packs):
            if p[7] is not None:
                pts[k] = p[7]
    js_all, layers_all = [], []
    rows = off = 0
    for g_, js, layers, nchs, _ns, _w, size, _pt in packs:
        js = js.copy()
        js["layer_offset"] += rows
        js["log_offset"] += off
        rows += len(layers)
        off += size
        js_all.append(js)
        layers_all.append(layers)
    js_all = _lib.cat_records(js_all, _lib.JUNCTION_DTYPE)
    log, mag, status = groups[0]._execute(recs, js_all, _lib.cat_records(layers_all, _lib.LAYER_DTYPE),
                                          _lib.constants_get(), pts)
    nan = False
    base = 0
    for grp, (g_, js, layers, nchs, _ns, _w, size, _pt) in zip(groups, packs):
        n = grp.noElements
        sl = slice(base, base + n)
        jsq = js_all[sl]
        try:
            grp._finish(jsq, log, mag[sl], status[sl], nchs, W, timeStep)
        except RuntimeError as e:
            if "NAN" not in str(e):
                raise
            nan = True
        base += n
    if nan:
        raise RuntimeError("NAN magnetisation")
