"""The C ABI: the library loads without a GPU, exports every entry point
declared in include/*.h, and its structs match the host binding."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from cmtj_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def _declared() -> set:
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(cmtj_[a-z_0-9]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    declared = _declared()
    assert declared, "no declarations found"
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(_lib.EXPORTS)


def test_struct_layout_matches_binding():
    _lib.lib()  # raises ImportError on any mismatch
    assert _lib.LAYER_DTYPE.itemsize == 3144
    assert _lib.JUNCTION_DTYPE.itemsize == 544
    assert _lib.STACK_DTYPE.itemsize == 1168
    assert _lib.GROUP_DTYPE.itemsize == 1800


def test_host_entry_points_need_no_gpu():
    c = _lib.constants_get()
    assert float(c["gyro"]) == 220880.0
    rec = _lib.ScalarDriverRecord() if hasattr(_lib, "ScalarDriverRecord") else None
    import numpy as np

    r = np.zeros((), dtype=_lib.DRIVER_DTYPE)
    r["type"] = _lib.STEP
    r["amplitude"] = 2.0
    r["t_start"] = 1.0
    r["t_stop"] = 2.0
    assert _lib.driver_eval(r, 1.0) == 2.0 and _lib.driver_eval(r, 2.5) == 0.0
    r["type"] = _lib.TABLE
    try:
        _lib.driver_eval(r, 0.0)
        raise AssertionError("tabulated drivers must be rejected")
    except RuntimeError as e:
        assert "tabulated" in str(e)


def test_plan_arena_size_runs_without_a_gpu():
    """Sizing a plan validates and groups the batch on the host only."""
    from cmtj_amd import workloads

    b = workloads.pimm(n_points=64, t_min=1e-10, t_max=2e-10)
    n1 = _lib.plan_arena_size(b.js, b.layers, 0, b.precision, b.log_dtype)
    b2 = workloads.pimm(n_points=128, t_min=1e-10, t_max=2e-10)
    n2 = _lib.plan_arena_size(b2.js, b2.layers, 0, b2.precision, b2.log_dtype)
    assert n1 >= 64 * _lib.JUNCTION_DTYPE.itemsize + 128 * _lib.LAYER_DTYPE.itemsize
    assert n2 > n1
    bad = b.js.copy()
    bad["n_layers"][3] = 0
    with pytest.raises(RuntimeError):
        _lib.plan_arena_size(bad, b.layers, 0, b.precision, b.log_dtype)


def test_plan_glue_passes_arguments(monkeypatch):
    """Python side of the plan calls, against a stand-in library object."""
    calls = []

    class Fake:
        def __getattr__(self, name):
            def f(*a):
                calls.append((name, a))
                return 0
            return f

    p = _lib.Plan.__new__(_lib.Plan)
    p._h = 1234
    monkeypatch.setattr(_lib, "lib", lambda: Fake())
    p.execute(1, 2, 3, 4, 5)
    p.execute_slice(1, 2, 3, 4, 6, 100, [7, 8])
    assert calls[0][0] == "cmtj_plan_execute" and calls[0][1] == (1234, 1, 2, 3, 4, 5)
    name, a = calls[1]
    assert name == "cmtj_plan_execute_slice" and a[:7] == (1234, 1, 2, 3, 4, 6, 100) and a[8] == 2
    assert list(a[7]) == [7, 8]
    p._h = None
