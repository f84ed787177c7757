"""CPU FP64 oracle of the (s)LLGS path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
this package.  It wraps liboracle.so (oracle.c, a plain-C restatement of the
reference algorithm) and the optional _ref/ref_drivers harness built from the
reference's own core/drivers.hpp.

Pinning (what fixes each oracle function independently of its own formula):
  * driver waveforms: reference known answers (tests/test_drivers.py:51-65)
    and the compiled reference drivers.hpp (_ref/ref_drivers) on a time grid;
  * field/torque/RK4 on one layer: the exact damped-precession trajectory in a
    uniform field and RK4's 4th-order convergence; test_sto.json,
    test_harmonic_hall.json, test_cims_hysteresis.json golden fixtures;
  * multi-layer stepping: test_fmr_basic.json, test_ahe_loops.json fixtures and
    the first-order convergence that start-of-step neighbours imply;
  * Heun / Euler-Heun: convergence orders 2 / 1 against the exact trajectory;
  * thermal field: Philox4x32-10 known-answer vectors (Random123) and the
    Boltzmann (Langevin-function) equilibrium of an ensemble;
  * energy channels: key sets of tests/test_basic.py and closed forms of the
    logged fields;
  * coupled stacks: zero coupling equals each member alone, logged currents
    and couplings recomputed from the logged magnetisations;
  * Dormand-Prince: scipy's RK45 tableau on a textbook LLG right-hand side
    and the closed-form damped precession;
  * dipole-coupled groups: single-junction runs carrying the textbook
    point-dipole field as their reserved field, step by step.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_DRIVERS = HERE / "_ref" / "ref_drivers"

_lib = None


def _bind(path):
    from cmtj_amd import _lib as abi  # struct dtypes only

    lib = ctypes.CDLL(str(path))
    P = ctypes.c_void_p
    lib.oracle_run.argtypes = [P, P, P, P, ctypes.c_size_t, P, P, P]
    lib.oracle_run_batch.argtypes = [P, ctypes.c_size_t, P, P, P, ctypes.c_size_t, P, P, P, ctypes.c_int]
    lib.oracle_driver_value.argtypes = [P, ctypes.c_double, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, P]
    lib.oracle_driver_value.restype = ctypes.c_double
    lib.oracle_philox4x32_10.argtypes = [P, P, P]
    lib.oracle_thermal_normals.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                           ctypes.c_uint32, P]
    lib.oracle_n_channels.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.oracle_run_stack.argtypes = [P, P, P, P, P, P, P, P]
    lib.oracle_run_group.argtypes = [P, P, P, P, P, P, P, P]
    return (lib, abi)


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} missing: run `make -C oracle`")
        _lib = _bind(LIB)
    return _lib


_native = {}


def build_native(out_dir) -> Path:
    """Builds the oracle tuned for this machine (-march=native, still without
    FP contraction or fast-math) into out_dir, for CPU-baseline timing only."""
    out = Path(out_dir) / "liboracle.so"
    subprocess.run(["make", "-s", "-C", str(HERE), "native", f"NATIVE_OUT={out}"], check=True,
                   capture_output=True)
    return out


def _p(a):
    return None if a is None else a.ctypes.data


def run(j, layers, consts, tables=None, pid_out=None):
    """One junction; returns (log [ch, entries], mag [L, 3], status) like
    cmtj_run; pid_out ([L, 2]) receives the final PID controller state of a
    Dormand-Prince run whose layers use it."""
    lib, abi = _load()
    j = np.ascontiguousarray(j, dtype=abi.JUNCTION_DTYPE).reshape(())
    layers = np.ascontiguousarray(layers, dtype=abi.LAYER_DTYPE)
    nch = lib.oracle_n_channels(int(j["n_layers"]), int(j["mr_mode"]), int(j["log_mask"]))
    nl = int(j["n_layers"])
    log = np.zeros((nch, int(j["n_entries"])), dtype=np.float64)
    out = np.zeros(5 * nl, dtype=np.float64)
    mag = out[: 3 * nl].reshape(nl, 3)
    st = np.zeros((), dtype=abi.STATUS_DTYPE)
    tabs = None if tables is None else np.ascontiguousarray(tables, dtype=np.float64)
    lib.oracle_run(_p(j), _p(layers), _p(np.ascontiguousarray(consts)), _p(tabs),
                   0 if tabs is None else tabs.size, _p(log), _p(out), _p(st))
    if pid_out is not None and int(j["solver"]) == abi.DORMAND_PRINCE and np.any(layers["dp_use_pid"][:nl]):
        pid_out[...] = out[3 * nl:].reshape(nl, 2)
    return log, mag, st


def run_batch(js, layers, consts, tables=None, log_elements=0, n_threads=None, lib_path=None):
    """Many junctions over host threads; log blocks at each log_offset (FP64).
    lib_path: another build of oracle.c (build_native) to run instead."""
    if lib_path is None:
        lib, abi = _load()
    else:
        if str(lib_path) not in _native:
            _native[str(lib_path)] = _bind(lib_path)
        lib, abi = _native[str(lib_path)]
    js = np.ascontiguousarray(js, dtype=abi.JUNCTION_DTYPE)
    layers = np.ascontiguousarray(layers, dtype=abi.LAYER_DTYPE)
    log = np.zeros(max(int(log_elements), 1), dtype=np.float64)
    mag = np.zeros((js.size, abi.MAX_LAYERS, 3), dtype=np.float64)
    st = np.zeros(js.size, dtype=abi.STATUS_DTYPE)
    tabs = None if tables is None else np.ascontiguousarray(tables, dtype=np.float64)
    n_threads = n_threads or os.cpu_count() or 1
    lib.oracle_run_batch(_p(js), js.size, _p(layers), _p(np.ascontiguousarray(consts)), _p(tabs),
                         0 if tabs is None else tabs.size, _p(log), _p(mag), _p(st), int(n_threads))
    return log, mag, st


def run_stack(stacks, js, layers, consts, tables=None):
    """Coupled stacks, same layout as cmtj_run_stacks (js[s * members + j]:
    member j of stack s), one after another: returns (log, mag [n, 8, 3],
    status [n])."""
    lib, abi = _load()
    stacks = np.ascontiguousarray(stacks, dtype=abi.STACK_DTYPE).reshape(-1)
    js = np.ascontiguousarray(js, dtype=abi.JUNCTION_DTYPE)
    layers = np.ascontiguousarray(layers, dtype=abi.LAYER_DTYPE)
    need = 0
    for q in range(js.size):
        nch = lib.oracle_n_channels(int(js[q]["n_layers"]), int(js[q]["mr_mode"]), int(js[q]["log_mask"])) + 2
        need = max(need, int(js[q]["log_offset"]) + nch * int(js[q]["n_entries"]))
    log = np.zeros(max(need, 1), dtype=np.float64)
    mag = np.zeros((js.size, abi.MAX_LAYERS, 3), dtype=np.float64)
    st = np.zeros(js.size, dtype=abi.STATUS_DTYPE)
    consts = np.ascontiguousarray(consts)
    tabs = None if tables is None or len(tables) == 0 else np.ascontiguousarray(tables, dtype=np.float64)
    n = int(stacks[0]["n_junctions"]) if stacks.size else 0
    for s in range(stacks.size):
        sub = js[s * n:(s + 1) * n]
        rc = lib.oracle_run_stack(_p(stacks[s:s + 1]), _p(sub), _p(layers), _p(consts), _p(tabs), _p(log),
                                  _p(mag[s * n:(s + 1) * n]), _p(st[s * n:(s + 1) * n]))
        if rc:
            raise ValueError("oracle_run_stack: bad stack descriptor")
    return log[:need], mag, st


def run_group(groups, js, layers, consts, pair_terms=None):
    """Dipole-coupled groups, same layout as cmtj_run_groups (js[g * elements
    + e]: element e of group g; pair_terms [groups][n][12][n] for tensor
    interactions), one after another: returns (log, mag [n, 8, 3], status
    [n])."""
    lib, abi = _load()
    groups = np.ascontiguousarray(groups, dtype=abi.GROUP_DTYPE).reshape(-1)
    js = np.ascontiguousarray(js, dtype=abi.JUNCTION_DTYPE)
    layers = np.ascontiguousarray(layers, dtype=abi.LAYER_DTYPE)
    need = 0
    for q in range(js.size):
        nch = lib.oracle_n_channels(int(js[q]["n_layers"]), int(js[q]["mr_mode"]), int(js[q]["log_mask"]))
        need = max(need, int(js[q]["log_offset"]) + nch * int(js[q]["n_entries"]))
    log = np.zeros(max(need, 1), dtype=np.float64)
    mag = np.zeros((js.size, abi.MAX_LAYERS, 3), dtype=np.float64)
    st = np.zeros(js.size, dtype=abi.STATUS_DTYPE)
    consts = np.ascontiguousarray(consts)
    n = int(groups[0]["n_elements"]) if groups.size else 0
    pt = None if pair_terms is None else np.ascontiguousarray(pair_terms, dtype=np.float64).reshape(-1)
    for g in range(groups.size):
        ptg = None if pt is None else pt[g * 12 * n * n:(g + 1) * 12 * n * n]
        rc = lib.oracle_run_group(_p(groups[g:g + 1]), _p(js[g * n:(g + 1) * n]), _p(layers), _p(consts), _p(ptg),
                                  _p(log), _p(mag[g * n:(g + 1) * n]), _p(st[g * n:(g + 1) * n]))
        if rc == 2:
            raise RuntimeError("Points are too close for stable dipole calculation")
        if rc:
            raise ValueError("oracle_run_group: bad group descriptor")
    return log[:need], mag, st


def driver_value(rec, t: float) -> float:
    lib, abi = _load()
    rec = np.ascontiguousarray(rec, dtype=abi.DRIVER_DTYPE).reshape(())
    return lib.oracle_driver_value(_p(rec), float(t), 0, 0, 1, None)


def philox4x32_10(ctr, key):
    lib, _ = _load()
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib.oracle_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def thermal_normals(seed: int, stream: int, step: int, layer: int):
    lib, _ = _load()
    xi = np.zeros(3, dtype=np.float64)
    lib.oracle_thermal_normals(seed, stream, step, layer, _p(xi))
    return xi


def ref_drivers_available() -> bool:
    return REF_DRIVERS.exists()


def ref_driver_values(lines: list[str]) -> list[float]:
    """Evaluates reference-built drivers (see ref_drivers_harness.cpp)."""
    out = subprocess.run([str(REF_DRIVERS)], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True)
    return [float.fromhex(x) for x in out.stdout.split()]
