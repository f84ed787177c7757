"""Host logic of the serving loop: the log-arena allocator and request
validation (no GPU needed)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from cmtj_amd import workloads
from cmtj_amd.serving import LogArena, Server


def test_arena_first_fit_and_merge():
    a = LogArena(100)
    x = a.alloc(30)
    y = a.alloc(30)
    z = a.alloc(30)
    assert (x, y, z) == (0, 30, 60)
    assert a.alloc(20) is None  # 10 left
    a.free(y, 30)
    assert a.alloc(25) == 30    # first fit into the hole
    a.free(x, 30)
    a.free(30, 25)
    a.free(z, 30)
    assert a.used == 0 and a._free == [(0, 100)]  # everything merged back


def test_arena_random_alloc_free_keeps_blocks_disjoint():
    rng = np.random.default_rng(3)
    a = LogArena(10_000)
    live = {}
    for _ in range(2000):
        if live and rng.random() < 0.45:
            off = list(live)[int(rng.integers(len(live)))]
            a.free(off, live.pop(off))
        else:
            sz = int(rng.integers(1, 400))
            off = a.alloc(sz)
            if off is not None:
                live[off] = sz
        spans = sorted((o, o + s) for o, s in live.items())
        assert all(spans[q][1] <= spans[q + 1][0] for q in range(len(spans) - 1))
        assert all(0 <= o and e <= a.capacity for o, e in spans)
        assert a.used == sum(live.values())


def test_submit_rejects_a_log_larger_than_the_arena():
    s = Server(log_elements=1000)  # not started: no device work
    j = workloads.pimm_junction()
    with pytest.raises(ValueError, match="exceeds"):
        s.submit(j, 1e-9, 1e-13, 1e-13, log="mx")  # 3 x 10000 elements


def test_serving_mix_covers_the_request_kinds():
    mix = workloads.serving_mix(200, seed=2)
    kinds = {m[0] for m in mix}
    assert kinds == {"pimm", "vsd", "cims", "chain3", "chain4"}
    layers = {m[1].layerNo for m in mix}
    assert layers == {1, 2, 3, 4}
    for k, j, T, dt, wf, sm, log in mix:
        assert 1e-9 <= T <= 50e-9 and 1e-14 <= dt <= 1e-13 + 1e-27 and wf >= dt


def test_group_key_follows_the_structure():
    from cmtj_amd import _lib
    from cmtj_amd.batch import JunctionBatch

    def key(j, T=1e-9, dt=1e-13):
        jd, L, _, _ = JunctionBatch._pack(j, T, dt, 1e-12, workloads.SolverMode.RK4)
        return _lib.group_key(jd, L)

    a, b = workloads.pimm_junction(), workloads.pimm_junction()
    b.setLayerExternalFieldDriver("all", workloads.AxialDriver(1e5, 0, 0))  # numbers only
    assert key(a) == key(b) == key(a, T=2e-9, dt=2e-13)
    assert key(a) != key(workloads.chain_junction(3)) != key(workloads.chain_junction(4))
    assert key(workloads.vsd_junction()) != key(a)  # sine drive, one layer


def test_each_submission_takes_its_own_noise_stream():
    """Two submissions of one 300 K junction run two independent thermal
    paths, as two runSimulation calls do; so do repeated entries of a batch."""
    from cmtj_amd import constantDriver
    from cmtj_amd.batch import JunctionBatch

    j = workloads.pimm_junction()
    j.setLayerTemperatureDriver("all", constantDriver(300.0))
    j.setNoiseSeed(42)
    s = Server(log_elements=1 << 20)  # not started: requests stay queued, no device work
    s.submit(j, 1e-10, 1e-13, 1e-13, log="mx")
    s.submit(j, 1e-10, 1e-13, 1e-13, log="mx")
    pairs = [(int(r.jd["seed"][0]), int(r.jd["stream"][0])) for r in s._pending]
    assert pairs[0] != pairs[1] and pairs[0][0] == pairs[1][0] == 42
    b = JunctionBatch.from_junctions([j, j, j], 1e-10, 1e-13, 1e-13)
    assert len(set(b.js["stream"].tolist())) == 3
    assert min(b.js["stream"].tolist()) > max(p[1] for p in pairs)
    r = JunctionBatch.replicate(j, 4, 1e-10, 1e-13, 1e-13)
    assert r.js["stream"].tolist() == list(range(5, 9)) and j._runs == 9


class _FakeCuda:
    def __init__(self, initialised: bool) -> None:
        self._init = initialised

    def is_initialized(self) -> bool:
        return self._init


class _FakeTorch:
    def __init__(self, initialised: bool) -> None:
        self.cuda = _FakeCuda(initialised)


def test_server_raises_hardware_queues_before_hip_starts(monkeypatch):
    """Streams beyond GPU_MAX_HW_QUEUES share hardware queues (their kernels
    serialise), so the server raises the count while HIP is uninitialised,
    keeps a larger one, caps at 32 and only warns once HIP has started."""
    from cmtj_amd.serving import _raise_hw_queues

    monkeypatch.setenv("GPU_MAX_HW_QUEUES", "4")
    assert _raise_hw_queues(_FakeTorch(False), 16)
    assert os.environ["GPU_MAX_HW_QUEUES"] == "16"
    monkeypatch.setenv("GPU_MAX_HW_QUEUES", "24")
    assert _raise_hw_queues(_FakeTorch(False), 16)
    assert os.environ["GPU_MAX_HW_QUEUES"] == "24"
    monkeypatch.delenv("GPU_MAX_HW_QUEUES")
    assert _raise_hw_queues(_FakeTorch(False), 100)
    assert os.environ["GPU_MAX_HW_QUEUES"] == "32"
    monkeypatch.setenv("GPU_MAX_HW_QUEUES", "4")
    with pytest.warns(RuntimeWarning, match="share hardware queues"):
        assert not _raise_hw_queues(_FakeTorch(True), 16)
    assert os.environ["GPU_MAX_HW_QUEUES"] == "4"
