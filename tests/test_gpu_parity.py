"""GPU FP64 drop-in vs the CPU oracle on the same junctions: identical log
keys and lengths, max|dm| <= 1e-8 over every logged magnetisation, and the
resistance and field channels within 1e-8 relative.  Then the reference's own
golden fixtures reproduced by the GPU path."""
from __future__ import annotations

import numpy as np
import pytest

import golden_cases
import parity_cases
from cmtj_amd import Junction
from oracle_backend import OracleJunction

pytestmark = pytest.mark.gpu

M_TOL = 1e-8  # max |dm| of the FP64 drop-in against the oracle


def _compare(name):
    jg, kw = parity_cases.CASES[name](Junction)
    jo, kwo = parity_cases.CASES[name](OracleJunction)
    for j in (jg, jo):
        j.setNoiseSeed(1234)
    jg.runSimulation(**kw)
    jo.runSimulation(**kwo)
    lg, lo = jg.getLog(), jo.getLog()
    assert sorted(lg) == sorted(lo), set(lg) ^ set(lo)
    worst = 0.0
    for k in lo:
        a, b = np.asarray(lg[k]), np.asarray(lo[k])
        assert a.shape == b.shape, (k, a.shape, b.shape)
        if k.endswith(("_mx", "_my", "_mz")):
            d = float(np.max(np.abs(a - b))) if a.size else 0.0
            worst = max(worst, d)
            assert d <= M_TOL, (name, k, d)
        elif k == "time":
            assert np.array_equal(a, b)
        else:
            scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
            d = float(np.max(np.abs(a - b))) if a.size else 0.0
            assert d <= 1e-8 * scale, (name, k, d, scale)
    for l in jo.layers:
        ma, mb = jg.getLayerMagnetisation(l.id), jo.getLayerMagnetisation(l.id)
        assert max(abs(ma.x - mb.x), abs(ma.y - mb.y), abs(ma.z - mb.z)) <= M_TOL
    print(f"[parity] {name}: max|dm| = {worst:.3e}", flush=True)


@pytest.mark.parametrize("name", list(parity_cases.CASES))
def test_fp64_matches_oracle(name):
    _compare(name)


@pytest.mark.parametrize("name", list(golden_cases.CASES))
def test_golden_fixture_on_gpu(name):
    golden_cases.assert_matches(name, golden_cases.CASES[name](Junction))
