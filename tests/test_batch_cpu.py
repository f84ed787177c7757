"""Batch packing (the product's host code) checked on the CPU: a packed
batch run through the oracle equals each junction run alone."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import parity_cases
from cmtj_amd import SolverMode, _lib, workloads
from cmtj_amd.batch import JunctionBatch, _channel_names
from oracle_backend import OracleJunction


def _blocks(b, log, i):
    j = b.js[i]
    meta = b.metas[b.meta_index[i]]
    names = _channel_names(meta, int(j["log_mask"]))
    ne = int(j["n_entries"])
    off = int(j["log_offset"])
    return {k: log[off + c * ne: off + (c + 1) * ne] for c, k in enumerate(names)}


def test_heterogeneous_batch_equals_individual_runs():
    names = ["rh_strip", "bilayer_iec_rk4", "trilayer_heun", "stt_drivers", "sot_static", "exotic_eulerheun"]
    js, kws = [], []
    for n in names:
        j, kw = parity_cases.CASES[n](OracleJunction)
        js.append(j)
        kws.append(kw)
    b = JunctionBatch.from_junctions(js, [k["totalTime"] for k in kws], [k["timeStep"] for k in kws],
                                     [k["writeFrequency"] for k in kws],
                                     [k.get("solverMode", SolverMode.RK4) for k in kws], precision="fp64",
                                     log="full", log_dtype="f64")
    assert b.log_elements == sum(
        _lib.n_channels(int(j["n_layers"]), int(j["mr_mode"]), int(j["log_mask"])) * int(j["n_entries"])
        for j in b.js)
    log, mag, st = oracle.run_batch(b.js, b.layers, _lib.constants_get(), b.tables, b.log_elements)
    for i, (j, kw) in enumerate(zip(js, kws)):
        j.runSimulation(**kw)
        alone = j.getLog()
        got = _blocks(b, log, i)
        for k, v in got.items():
            np.testing.assert_array_equal(v, np.asarray(alone[k]), err_msg=f"{names[i]}:{k}")
        for q, l in enumerate(j.layers):
            np.testing.assert_array_equal(mag[i, q], l.mag.tolist())
        assert int(st[i]["code"]) == 0


def test_replicate_overrides():
    b = workloads.r_h_loop(n_fields=5, total_time=1e-11)
    H = b.layers["drv"][:, _lib.SLOT_HEXT]["constant"]
    np.testing.assert_array_equal(H, np.linspace(-800e3, 800e3, 5))
    # the reference's truncations: 1e-11 / 1e-13 = 99.99999999999999 -> 99 steps
    assert int(1e-11 / 1e-13) == 99
    assert np.all(b.js["n_steps"] == 99) and np.all(b.js["write_every"] == 10)
    assert list(b.js["log_offset"]) == [i * 5 * 10 for i in range(5)]  # 3 m + Rx + Ry channels

    p = workloads.pimm(n_points=16, seed=3)
    steps = p.js["n_steps"]
    assert steps.min() >= 20000 and steps.max() <= 200000 and len(set(steps.tolist())) == 16
    assert p.total_steps == int(steps.sum())
    offs = p.js["log_offset"].astype(np.int64)
    np.testing.assert_array_equal(np.diff(offs), 3 * steps[:-1].astype(np.int64))  # mx x2 + R per step


def test_replicated_sweep_equals_individual_runs():
    b = workloads.r_h_loop(n_fields=4, total_time=2e-10, precision="fp64")
    log, mag, st = oracle.run_batch(b.js, b.layers, _lib.constants_get(), None, b.log_elements)
    from cmtj_amd import AxialDriver, CVector, Layer, constantDriver

    for i, H in enumerate(np.linspace(-800e3, 800e3, 4)):
        l = Layer("free", mag=CVector(1, 0, 0), anis=CVector(1, 0, 0), Ms=1.2, thickness=2e-9, cellSurface=1e-16,
                  demagTensor=[CVector(0.00024, 0, 0), CVector(0, 0.0022, 0), CVector(0, 0, 0.99756)],
                  damping=0.01)
        j = OracleJunction([l], [100.0], [0.0], [1.0], [0.0], [0.0], [0.0], [0.0])
        j.setLayerAnisotropyDriver("free", constantDriver(5e3))
        j.setLayerExternalFieldDriver("all", AxialDriver(H, 0, 0))
        j.runSimulation(2e-10, 1e-13, 1e-12)
        got = _blocks(b, log, i)
        np.testing.assert_array_equal(got["Rx"], np.asarray(j.getLog()["Rx"]))
        np.testing.assert_array_equal(got["free_mx"], np.asarray(j.getLog()["free_mx"]))


def test_set_times_rejects_write_frequency_below_step():
    b = workloads.r_h_loop(n_fields=2, total_time=1e-11)
    with pytest.raises(RuntimeError, match="write frequency"):
        b.set_times(1e-11, 1e-12, 1e-13)
