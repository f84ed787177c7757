"""The autotuner on the GPU: it measures every launch configuration that
covers a mixed batch, persists the fastest, reuses it without measuring, and
the tuned run gives the default run's results (lane mappings differ only in
the association of sums: within 1e-10 for binary64)."""
from __future__ import annotations

import numpy as np
import pytest

from cmtj_amd import autotune
from test_gpu_graph import _mixed

pytestmark = pytest.mark.gpu


def test_tune_persists_and_reproduces_the_default_run(tmp_path, monkeypatch):
    monkeypatch.setenv("CMTJ_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    b = _mixed("fp64")
    pick = autotune.tune(b, scale=0.2, verbose=True)
    assert not pick["cached"] and len(pick["trials"]) >= 2
    assert any(t["graph"] for t in pick["trials"])
    again = autotune.tune(b)
    assert again["cached"] and again["lanes"] == pick["lanes"] and again["graph"] == pick["graph"]
    ref = _mixed("fp64")
    r0 = ref.run()
    r1 = autotune.run(b, pick)
    worst = 0.0
    for i in range(len(b.js)):
        l0, l1 = r0.instance_log(i), r1.instance_log(i)
        for k in l0:
            worst = max(worst, float(np.max(np.abs(l0[k] - l1[k]))) if l0[k].size else 0.0)
    print(f"[autotune] pick {pick['lanes']} graph={pick['graph']}: max difference to the default run {worst:.3e}",
          flush=True)
    assert worst <= 1e-10
