"""The autotuner's host logic: a batch's signature depends on its structure
groups and size class (not on parameter values), and picks persist in the
JSON cache named by CMTJ_AMD_TUNE_CACHE."""
from __future__ import annotations

import numpy as np

from cmtj_amd import autotune, workloads


def test_signature_follows_structure_and_size_class():
    a = workloads.pimm(n_points=64, precision="fp32")
    b = workloads.pimm(n_points=64, seed=9, precision="fp32")   # other durations, same shape
    c = workloads.pimm(n_points=1024, precision="fp32")         # another size class
    d = workloads.pimm(n_points=64, precision="fp64")           # other arithmetic
    e = workloads.sto(n_currents=64, precision="fp32")          # other structure
    sa = autotune.signature(a)
    assert sa == autotune.signature(b)
    assert len({sa, autotune.signature(c), autotune.signature(d), autotune.signature(e)}) == 4


def test_probe_scales_step_counts_and_keeps_streams():
    b = workloads.pimm(n_points=32, precision="fp32")
    q = autotune._probe(b, 0.05)
    ratio = q.js["n_steps"].astype(float) / b.js["n_steps"].astype(float)
    assert np.all(np.abs(ratio - 0.05) < 1e-4)
    assert np.array_equal(q.js["seed"], b.js["seed"]) and np.array_equal(q.js["stream"], b.js["stream"])
    assert np.array_equal(q.js["write_every"], b.js["write_every"])


def test_picks_persist(tmp_path, monkeypatch):
    monkeypatch.setenv("CMTJ_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    assert autotune.load_cache() == {}
    autotune.save_pick("abc", {"lanes": "per_layer", "graph": False})
    autotune.save_pick("def", {"lanes": "one", "graph": True})
    c = autotune.load_cache()
    assert c["abc"]["lanes"] == "per_layer" and c["def"]["graph"] is True
