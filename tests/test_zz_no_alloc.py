"""cmtj_run_noalloc, the allocation-free variant of the drop-in cmtj_run:
same results bit for bit, on two calls in a row (the second smaller than the
first, so stale scratch contents of the first would show), on the default
stream and on a caller stream; oversize runs are rejected before any launch."""
from __future__ import annotations

import numpy as np
import pytest

import parity_cases
from cmtj_amd import Junction, SolverMode, _lib

pytestmark = pytest.mark.gpu


def _packed(builder, **over):
    j, kw = builder(Junction)
    kw = dict(kw, **over)
    desc = j._descriptor(kw["totalTime"], kw.get("timeStep", 1e-13), kw.get("writeFrequency", 1e-11), False,
                         kw.get("solverMode", SolverMode.RK4))
    jd, L, tables = desc[0], desc[1], desc[2]
    return jd, L, tables


def _same(a, b):
    la, ma, sa = a
    lb, mb, sb = b
    assert la.shape == lb.shape and ma.shape == mb.shape
    assert np.array_equal(la, lb), float(np.max(np.abs(la - lb)))
    assert np.array_equal(ma, mb)
    assert sa.tobytes() == sb.tobytes()


def test_noalloc_equals_cmtj_run_on_two_calls_in_a_row():
    import torch

    consts = _lib.constants_get()
    big = _packed(parity_cases.bilayer_iec, totalTime=2e-9)           # 20000 entries x 7 channels
    small = _packed(parity_cases.rh_strip, totalTime=3e-10)           # 300 entries, one layer
    for case in (big, small):
        print(f"noalloc case: layers={int(case[0]['n_layers'])} entries={int(case[0]['n_entries'])}", flush=True)
        want = _lib.run_junction(case[0], case[1], consts, case[2])
        got = _lib.run_junction(case[0], case[1], consts, case[2], noalloc=True)
        _same(want, got)
        assert int(got[2]["code"]) == _lib.OK and int(got[2]["entries_done"]) == int(case[0]["n_entries"])
    # a caller stream, after the smaller run left its contents in the scratch
    s = torch.cuda.Stream()
    got = _lib.run_junction(big[0], big[1], consts, big[2], noalloc=True, stream=s.cuda_stream)
    _same(_lib.run_junction(big[0], big[1], consts, big[2]), got)


def test_noalloc_rejects_runs_beyond_its_scratch():
    consts = _lib.constants_get()
    jd, L, tables = _packed(parity_cases.bilayer_iec, totalTime=2e-9)
    jd = jd.copy()
    # 7 channels x 2e6 entries x 8 B = 112 MB > 64 MiB; rejected on the host before any launch
    jd["n_steps"] = 2_000_000
    jd["write_every"] = 1
    jd["n_entries"] = 2_000_000
    with pytest.raises(RuntimeError, match="device scratch"):
        _lib.run_junction(jd, L, consts, tables, noalloc=True)
