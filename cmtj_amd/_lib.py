"""ctypes binding of libcmtj_amd.so (the C ABI of include/cmtj_amd.h).

The numpy dtypes below mirror the C structs field by field; their sizes and
key offsets are checked against the library's own ``cmtj_abi`` at load time.
The library must exist: nothing here falls back to another implementation.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libcmtj_amd.so"

ABI_VERSION = 6
MAX_LAYERS = 8
MAX_STACK = 64
MAX_GROUP = 64
INTERACTION_NULL, INTERACTION_KANAO, INTERACTION_NOMURA, INTERACTION_TENSOR = 0, 1, 2, 3
NSLOT = 30
ENERGY_CHANNELS = 23

# enum cmtj_update_type
CONSTANT, PULSE, SINE, STEP, POSINE, HALFSINE, TRAPEZOID, GAUSSIMPULSE, GAUSSSTEP = range(9)
TABLE = 101

# enum cmtj_slot
SLOT_HEXT, SLOT_HOE, SLOT_HDMI, SLOT_HRES = 0, 3, 6, 9
SLOT_K, SLOT_K2 = 12, 13
SLOT_J_TOP, SLOT_J_BOTTOM, SLOT_J2_TOP, SLOT_J2_BOTTOM = 14, 15, 16, 17
SLOT_CURRENT, SLOT_FL, SLOT_DL, SLOT_FL2, SLOT_DL2 = 18, 19, 20, 21, 22
SLOT_TEMPERATURE = 23
SLOT_IDMI_TOP, SLOT_IDMI_BOTTOM = 24, 27

# enum cmtj_reference / cmtj_solver / cmtj_mr_mode
REF_NONE, REF_FIXED, REF_TOP, REF_BOTTOM = range(4)
EULER_HEUN, RK4, DORMAND_PRINCE, HEUN = range(4)
MR_NONE, MR_CLASSIC, MR_STRIP = range(3)

# enum cmtj_log_bits
LOG_MX, LOG_MYZ, LOG_MR, LOG_ENERGY, LOG_TIME = 1, 2, 4, 8, 16
LOG_FULL = LOG_MX | LOG_MYZ | LOG_MR

FP64, FP32 = 0, 1
LOG_F64, LOG_F32 = 0, 1

OK, ERR_INVALID, ERR_HIP, ERR_NAN, ERR_LIMIT, ERR_DP_REJECTED = 0, 1, 2, 3, 4, 5

DRIVER_DTYPE = np.dtype(
    [
        ("type", "<i4"),
        ("table_index", "<i4"),
        ("constant", "<f8"),
        ("amplitude", "<f8"),
        ("frequency", "<f8"),
        ("phase", "<f8"),
        ("period", "<f8"),
        ("cycle", "<f8"),
        ("t_start", "<f8"),
        ("t_stop", "<f8"),
        ("edge_time", "<f8"),
        ("steady_time", "<f8"),
    ],
    align=True,
)

LAYER_DTYPE = np.dtype(
    [
        ("mag", "<f8", (3,)),
        ("anis", "<f8", (3,)),
        ("Ms", "<f8"),
        ("thickness", "<f8"),
        ("cell_surface", "<f8"),
        ("damping", "<f8"),
        ("demag", "<f8", (9,)),
        ("dipole_top", "<f8", (9,)),
        ("dipole_bottom", "<f8", (9,)),
        ("ref", "<f8", (3,)),
        ("ref2", "<f8", (3,)),
        ("kappa", "<f8"),
        ("beta", "<f8"),
        ("slonczewski", "<f8"),
        ("polarisation", "<f8"),
        ("fl_torque", "<f8"),
        ("dl_torque", "<f8"),
        ("ref_type", "<i4"),
        ("include_stt", "<i4"),
        ("include_sot", "<i4"),
        ("dynamic_sot", "<i4"),
        ("alternative_stt", "<i4"),
        ("temperature_set", "<i4"),
        ("drv", DRIVER_DTYPE, (NSLOT,)),
        ("dp_abs_tol", "<f8"),
        ("dp_rel_tol", "<f8"),
        ("dp_max_factor", "<f8"),
        ("dp_min_factor", "<f8"),
        ("dp_safety", "<f8"),
        ("dp_use_pid", "<i4"),
        ("dp_pad", "<i4"),
        ("dp_kp", "<f8"),
        ("dp_ki", "<f8"),
        ("dp_kd", "<f8"),
        ("dp_integral_error", "<f8"),
        ("dp_prev_error_ratio", "<f8"),
    ],
    align=True,
)

JUNCTION_DTYPE = np.dtype(
    [
        ("n_layers", "<i4"),
        ("mr_mode", "<i4"),
        ("solver", "<i4"),
        ("log_mask", "<i4"),
        ("n_steps", "<u4"),
        ("write_every", "<u4"),
        ("n_entries", "<u4"),
        ("layer_offset", "<u4"),
        ("seed", "<u8"),
        ("stream", "<u8"),
        ("log_offset", "<u8"),
        ("dt", "<f8"),
        ("Rp", "<f8"),
        ("Rap", "<f8"),
        ("Rx0", "<f8", (MAX_LAYERS,)),
        ("Ry0", "<f8", (MAX_LAYERS,)),
        ("AMR_X", "<f8", (MAX_LAYERS,)),
        ("AMR_Y", "<f8", (MAX_LAYERS,)),
        ("SMR_X", "<f8", (MAX_LAYERS,)),
        ("SMR_Y", "<f8", (MAX_LAYERS,)),
        ("AHE", "<f8", (MAX_LAYERS,)),
        ("total_time", "<f8"),
        ("write_freq", "<f8"),
    ],
    align=True,
)

CONSTANTS_DTYPE = np.dtype(
    [("mu0", "<f8"), ("gyro", "<f8"), ("TtoAm", "<f8"), ("hbar", "<f8"), ("echarge", "<f8"), ("kB", "<f8")],
    align=True,
)

STATUS_DTYPE = np.dtype(
    [("code", "<i4"), ("steps_done", "<u4"), ("entries_done", "<u4"), ("nan_layer", "<i4")], align=True
)

STACK_DTYPE = np.dtype(
    [
        ("series", "<i4"),
        ("kcl", "<i4"),
        ("n_junctions", "<i4"),
        ("two_layer", "<i4"),
        ("delayed", "<i4"),
        ("pad_", "<i4"),
        ("phase_offset", "<f8"),
        ("pol", "<f8", (3,)),
        ("current", DRIVER_DTYPE),
        ("coupling", "<f8", (MAX_STACK,)),
        ("top", "<i4", (MAX_STACK,)),
        ("bottom", "<i4", (MAX_STACK,)),
    ],
    align=True,
)

GROUP_DTYPE = np.dtype(
    [
        ("n_elements", "<i4"),
        ("interaction", "<i4"),
        ("top", "<i4", (MAX_GROUP,)),
        ("pos", "<f8", (MAX_GROUP, 3)),
    ],
    align=True,
)

STATE_DTYPE = np.dtype(
    [("m", "<f8", (MAX_LAYERS, 3)), ("step", "<u4"), ("entries", "<u4"), ("code", "<i4"), ("nan_layer", "<i4")],
    align=True,
)


class _AbiInfo(ctypes.Structure):
    _fields_ = [
        (n, ctypes.c_int32)
        for n in (
            "version",
            "max_layers",
            "nslot",
            "sizeof_driver",
            "sizeof_layer",
            "sizeof_junction",
            "sizeof_constants",
            "sizeof_status",
            "sizeof_state",
            "sizeof_stack",
            "sizeof_group",
            "offsetof_layer_drv",
            "offsetof_layer_ref_type",
            "offsetof_junction_dt",
            "offsetof_junction_AHE",
        )
    ]


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("n_instances", ctypes.c_uint64),
        ("log_elements", ctypes.c_uint64),
        ("total_steps", ctypes.c_uint64),
        ("max_steps", ctypes.c_uint64),
        ("precision", ctypes.c_int32),
        ("log_dtype", ctypes.c_int32),
        ("max_layers", ctypes.c_int32),
        ("n_kernels", ctypes.c_int32),
        ("lanes_per_instance", ctypes.c_int32),
        ("fp64_groups", ctypes.c_int32),
        ("kernel_name", ctypes.c_char * 128),
    ]


# every entry point declared in include/cmtj_amd.h
EXPORTS = (
    "cmtj_abi",
    "cmtj_last_error",
    "cmtj_last_kernels",
    "cmtj_constants_get",
    "cmtj_constants_set",
    "cmtj_constants_reset",
    "cmtj_driver_eval",
    "cmtj_run",
    "cmtj_run_noalloc",
    "cmtj_plan_create",
    "cmtj_plan_arena_size",
    "cmtj_group_key",
    "cmtj_host_alloc",
    "cmtj_host_free",
    "cmtj_plan_create_in",
    "cmtj_plan_execute",
    "cmtj_plan_execute_slice",
    "cmtj_plan_graph_create",
    "cmtj_plan_graph_launch",
    "cmtj_plan_query",
    "cmtj_plan_destroy",
    "cmtj_vsd_mix",
    "cmtj_run_stacks",
    "cmtj_run_groups",
)

_P = ctypes.c_void_p
_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP engine first "
            "(python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_GLOBAL)
    lib.cmtj_abi.argtypes = [ctypes.POINTER(_AbiInfo)]
    lib.cmtj_last_error.restype = ctypes.c_char_p
    lib.cmtj_last_kernels.restype = ctypes.c_char_p
    lib.cmtj_constants_get.argtypes = [_P]
    lib.cmtj_constants_set.argtypes = [_P]
    lib.cmtj_driver_eval.argtypes = [_P, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    lib.cmtj_run.argtypes = [_P, _P, _P, _P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, _P]
    lib.cmtj_run_noalloc.argtypes = [_P, _P, _P, _P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, _P, _P]
    lib.cmtj_plan_create.argtypes = [
        _P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, _P, ctypes.c_size_t,
        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
    ]
    lib.cmtj_plan_create.restype = _P
    lib.cmtj_plan_arena_size.argtypes = [
        _P, ctypes.c_size_t, _P, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
        ctypes.POINTER(ctypes.c_size_t),
    ]
    lib.cmtj_plan_create_in.argtypes = [
        _P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, _P, ctypes.c_size_t,
        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P, ctypes.c_size_t, _P, ctypes.c_int32,
    ]
    lib.cmtj_host_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(_P), ctypes.POINTER(_P)]
    lib.cmtj_host_free.argtypes = [_P]
    lib.cmtj_plan_create_in.restype = _P
    lib.cmtj_group_key.argtypes = [_P, _P, ctypes.POINTER(ctypes.c_uint64)]
    lib.cmtj_plan_execute.argtypes = [_P, _P, ctypes.c_size_t, _P, _P, _P]
    lib.cmtj_plan_execute_slice.argtypes = [_P, _P, ctypes.c_size_t, _P, _P, _P, ctypes.c_uint32,
                                            ctypes.POINTER(_P), ctypes.c_int32]
    lib.cmtj_plan_graph_create.argtypes = [_P, _P, ctypes.c_size_t, _P, _P]
    lib.cmtj_plan_graph_launch.argtypes = [_P, _P]
    lib.cmtj_plan_query.argtypes = [_P, ctypes.POINTER(PlanInfo)]
    lib.cmtj_plan_destroy.argtypes = [_P]
    lib.cmtj_vsd_mix.argtypes = [_P, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, _P, ctypes.c_double, _P, _P,
                                 ctypes.c_int32, _P, _P]
    lib.cmtj_run_stacks.argtypes = [_P, ctypes.c_size_t, _P, _P, ctypes.c_size_t, _P, _P, ctypes.c_size_t, _P,
                                    ctypes.c_size_t, _P, _P]
    lib.cmtj_run_groups.argtypes = [_P, ctypes.c_size_t, _P, _P, ctypes.c_size_t, _P, _P, ctypes.c_size_t, _P,
                                    ctypes.c_size_t, _P, _P]
    info = _AbiInfo()
    lib.cmtj_abi(ctypes.byref(info))
    checks = {
        "version": (info.version, ABI_VERSION),
        "max_layers": (info.max_layers, MAX_LAYERS),
        "nslot": (info.nslot, NSLOT),
        "driver": (info.sizeof_driver, DRIVER_DTYPE.itemsize),
        "layer": (info.sizeof_layer, LAYER_DTYPE.itemsize),
        "junction": (info.sizeof_junction, JUNCTION_DTYPE.itemsize),
        "constants": (info.sizeof_constants, CONSTANTS_DTYPE.itemsize),
        "status": (info.sizeof_status, STATUS_DTYPE.itemsize),
        "state": (info.sizeof_state, STATE_DTYPE.itemsize),
        "stack": (info.sizeof_stack, STACK_DTYPE.itemsize),
        "group": (info.sizeof_group, GROUP_DTYPE.itemsize),
        "layer.drv": (info.offsetof_layer_drv, LAYER_DTYPE.fields["drv"][1]),
        "layer.ref_type": (info.offsetof_layer_ref_type, LAYER_DTYPE.fields["ref_type"][1]),
        "junction.dt": (info.offsetof_junction_dt, JUNCTION_DTYPE.fields["dt"][1]),
        "junction.AHE": (info.offsetof_junction_AHE, JUNCTION_DTYPE.fields["AHE"][1]),
    }
    for k, (got, want) in checks.items():
        if got != want:
            raise ImportError(f"ABI mismatch for {k}: library {got}, binding {want}")
    _lib = lib
    return lib


def lib():
    return _load()


def last_error() -> str:
    return (lib().cmtj_last_error() or b"").decode()


def last_kernels() -> list[str]:
    """Kernels the calling thread's last engine run launched."""
    return [k for k in (lib().cmtj_last_kernels() or b"").decode().split(";") if k]


def ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    return a.ctypes.data


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {last_error()}")


def n_channels(n_layers: int, mr_mode: int, log_mask: int) -> int:
    n = 0
    if log_mask & LOG_MX:
        n += n_layers
    if log_mask & LOG_MYZ:
        n += 2 * n_layers
    if log_mask & LOG_MR:
        n += 1 if mr_mode == MR_CLASSIC else (2 if mr_mode == MR_STRIP else 0)
    if log_mask & LOG_ENERGY:
        n += ENERGY_CHANNELS * n_layers
    if log_mask & LOG_TIME:
        n += 1
    return n


def constants_get() -> np.ndarray:
    c = np.zeros((), dtype=CONSTANTS_DTYPE)
    check(lib().cmtj_constants_get(ptr(c)), "cmtj_constants_get")
    return c


def constants_set(c: np.ndarray) -> None:
    c = np.ascontiguousarray(c, dtype=CONSTANTS_DTYPE)
    check(lib().cmtj_constants_set(ptr(c)), "cmtj_constants_set")


def constants_reset() -> None:
    check(lib().cmtj_constants_reset(), "cmtj_constants_reset")


def driver_eval(rec: np.ndarray, t: float) -> float:
    out = ctypes.c_double(0.0)
    check(lib().cmtj_driver_eval(ptr(rec), float(t), ctypes.byref(out)), "cmtj_driver_eval")
    return out.value


def cat_records(arrs, dtype) -> np.ndarray:
    """Concatenation of structured records (0-d or 1-d arrays of `dtype`)
    through their bytes: numpy's structured-dtype promotion costs tens of
    microseconds per array."""
    parts = [np.ascontiguousarray(a, dtype=dtype).reshape(-1).view(np.uint8) for a in arrs]
    if not parts:
        return np.zeros(0, dtype=dtype)
    return np.concatenate(parts).view(dtype)


def run_junction(j: np.ndarray, layers: np.ndarray, consts: np.ndarray, tables: np.ndarray | None,
                 noalloc: bool = False, stream: int | None = None, pid_out: np.ndarray | None = None):
    """One junction through cmtj_run (or cmtj_run_noalloc on `stream`, a
    hipStream_t handle, None = default stream); returns (log [ch, entries],
    mag [L,3], status).  pid_out (float64 [L, 2]): receives the final PID
    controller state (integral, prev) of every layer of a Dormand-Prince run
    whose layers use it (left alone otherwise)."""
    j = np.ascontiguousarray(j, dtype=JUNCTION_DTYPE).reshape(())
    layers = np.ascontiguousarray(layers, dtype=LAYER_DTYPE)
    nch = n_channels(int(j["n_layers"]), int(j["mr_mode"]), int(j["log_mask"]))
    ne = int(j["n_entries"])
    nl = int(j["n_layers"])
    log = np.empty((nch, ne), dtype=np.float64)
    out = np.zeros(5 * nl, dtype=np.float64)  # m [L,3], then the PID state [L,2]
    mag = out[: 3 * nl].reshape(nl, 3)
    st = np.zeros((), dtype=STATUS_DTYPE)
    tabs = None if tables is None or len(tables) == 0 else np.ascontiguousarray(tables, dtype=np.float64)
    args = (ptr(j), ptr(layers), ptr(consts), ptr(tabs), 0 if tabs is None else tabs.size,
            ptr(log) if log.size else None, log.size, ptr(out), ptr(st))
    if noalloc:
        check(lib().cmtj_run_noalloc(*args, stream), "cmtj_run_noalloc")
    else:
        check(lib().cmtj_run(*args), "cmtj_run")
    if pid_out is not None and int(j["solver"]) == DORMAND_PRINCE and np.any(layers["dp_use_pid"][:nl]):
        pid_out[...] = out[3 * nl:].reshape(nl, 2)
    return log, mag, st


def run_stacks(stacks: np.ndarray, js: np.ndarray, layers: np.ndarray, consts: np.ndarray,
               tables: np.ndarray | None = None):
    """Stacks through cmtj_run_stacks (js[s * members + j]: member j of stack
    s); js[q]["log_offset"] must place each member's block of n_channels + 2
    channels.  Returns (log, mag [n, 8, 3], status [n])."""
    stacks = np.ascontiguousarray(stacks, dtype=STACK_DTYPE).reshape(-1)
    js = np.ascontiguousarray(js, dtype=JUNCTION_DTYPE)
    layers = np.ascontiguousarray(layers, dtype=LAYER_DTYPE)
    n = js.size
    need = 0
    for q in range(n):
        nch = n_channels(int(js[q]["n_layers"]), int(js[q]["mr_mode"]), int(js[q]["log_mask"])) + 2
        need = max(need, int(js[q]["log_offset"]) + nch * int(js[q]["n_entries"]))
    log = np.zeros(need, dtype=np.float64)
    mag = np.zeros((n, MAX_LAYERS, 3), dtype=np.float64)
    st = np.zeros(n, dtype=STATUS_DTYPE)
    tabs = None if tables is None or len(tables) == 0 else np.ascontiguousarray(tables, dtype=np.float64)
    check(lib().cmtj_run_stacks(ptr(stacks), stacks.size, ptr(js), ptr(layers), layers.size, ptr(consts),
                                ptr(tabs), 0 if tabs is None else tabs.size,
                                ptr(log) if log.size else None, log.size, ptr(mag), ptr(st)), "cmtj_run_stacks")
    return log, mag, st


def run_groups(groups: np.ndarray, js: np.ndarray, layers: np.ndarray, consts: np.ndarray,
               pair_terms: np.ndarray | None = None):
    """Dipole-coupled groups through cmtj_run_groups (js[g * elements + e]:
    element e of group g); js[q]["log_offset"] must place each element's
    block of n_channels channels; pair_terms [groups][n][12][n] for tensor
    interactions.  Returns (log, mag [n, 8, 3], status [n])."""
    groups = np.ascontiguousarray(groups, dtype=GROUP_DTYPE).reshape(-1)
    js = np.ascontiguousarray(js, dtype=JUNCTION_DTYPE)
    layers = np.ascontiguousarray(layers, dtype=LAYER_DTYPE)
    n = js.size
    need = 0
    for q in range(n):
        nch = n_channels(int(js[q]["n_layers"]), int(js[q]["mr_mode"]), int(js[q]["log_mask"]))
        need = max(need, int(js[q]["log_offset"]) + nch * int(js[q]["n_entries"]))
    log = np.zeros(need, dtype=np.float64)
    mag = np.zeros((n, MAX_LAYERS, 3), dtype=np.float64)
    st = np.zeros(n, dtype=STATUS_DTYPE)
    pt = None if pair_terms is None else np.ascontiguousarray(pair_terms, dtype=np.float64).reshape(-1)
    check(lib().cmtj_run_groups(ptr(groups), groups.size, ptr(js), ptr(layers), layers.size, ptr(consts),
                                ptr(pt), 0 if pt is None else pt.size,
                                ptr(log) if log.size else None, log.size, ptr(mag), ptr(st)), "cmtj_run_groups")
    return log, mag, st


def plan_arena_size(js, layers, n_table: int, precision: int, log_dtype: int, lanes_per_instance: int = 0) -> int:
    js = np.ascontiguousarray(js, dtype=JUNCTION_DTYPE)
    layers = np.ascontiguousarray(layers, dtype=LAYER_DTYPE)
    out = ctypes.c_size_t(0)
    check(lib().cmtj_plan_arena_size(ptr(js), js.size, ptr(layers), layers.size, n_table, precision, log_dtype,
                                     lanes_per_instance, ctypes.byref(out)), "cmtj_plan_arena_size")
    return int(out.value)


def group_key(jd, layers) -> int:
    """Structure-signature key of one junction descriptor (jd: one record,
    layers: its layer rows)."""
    jd = np.ascontiguousarray(np.asarray(jd, dtype=JUNCTION_DTYPE).reshape(1))
    layers = np.ascontiguousarray(layers, dtype=LAYER_DTYPE)
    out = ctypes.c_uint64(0)
    check(lib().cmtj_group_key(ptr(jd), ptr(layers), ctypes.byref(out)), "cmtj_group_key")
    return int(out.value)


def vsd_mix(R, freq, dt: float, b, a, stream: int = 0):
    """Mixing voltage of each row of R (cmtj_vsd_mix): R a float64 device
    tensor [n, len] with unit column stride, freq [n] frequencies; b, a the
    filter coefficients.  Returns a float64 device tensor [n]."""
    import torch

    if R.dtype != torch.float64 or R.dim() != 2 or (R.shape[1] > 1 and R.stride(1) != 1):
        raise ValueError("R must be a 2-D float64 tensor with contiguous rows")
    n, length = R.shape
    f = torch.as_tensor(np.asarray(freq, dtype=np.float64).ravel(), device=R.device)
    if f.numel() != n:
        raise ValueError("one frequency per series")
    bb = np.ascontiguousarray(b, dtype=np.float64).ravel()
    aa = np.ascontiguousarray(a, dtype=np.float64).ravel()
    taps = max(len(bb), len(aa))
    bb = np.pad(bb, (0, taps - len(bb)))
    aa = np.pad(aa, (0, taps - len(aa)))
    out = torch.empty(n, dtype=torch.float64, device=R.device)
    check(lib().cmtj_vsd_mix(R.data_ptr(), R.stride(0) if n > 1 else length, n, length, f.data_ptr(), float(dt),
                             ptr(bb), ptr(aa), taps, out.data_ptr(), stream or None), "cmtj_vsd_mix")
    return out


class HostBuffer:
    """Pinned host memory mapped for the device (cmtj_host_alloc): `array`
    is a uint8 numpy view of it, `device` the address kernels use."""

    def __init__(self, nbytes: int) -> None:
        h, d = _P(), _P()
        check(lib().cmtj_host_alloc(int(nbytes), ctypes.byref(h), ctypes.byref(d)), "cmtj_host_alloc")
        self.host, self.device, self.nbytes = int(h.value), int(d.value), int(nbytes)
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * self.nbytes).from_address(self.host))

    def close(self) -> None:
        if self.host:
            self.array = None
            lib().cmtj_host_free(self.host)
            self.host = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """Owner of a cmtj_plan (device-resident descriptors of a batch).  With
    `arena` = (device pointer, bytes, stream) the descriptors go into that
    caller-owned memory, which must outlive the plan's executions."""

    def __init__(self, js, layers, consts, tables, precision: int, log_dtype: int, lanes_per_instance: int = 0,
                 arena=None):
        self._js = np.ascontiguousarray(js, dtype=JUNCTION_DTYPE)
        self._layers = np.ascontiguousarray(layers, dtype=LAYER_DTYPE)
        tabs = None if tables is None or len(tables) == 0 else np.ascontiguousarray(tables, dtype=np.float64)
        if arena is None:
            self._h = lib().cmtj_plan_create(
                ptr(self._js), self._js.size, ptr(self._layers), self._layers.size, ptr(consts),
                ptr(tabs), 0 if tabs is None else tabs.size, precision, log_dtype, lanes_per_instance,
            )
        else:
            d_arena, nbytes, stream = arena[:3]
            on_host = len(arena) > 3 and bool(arena[3])
            self._h = lib().cmtj_plan_create_in(
                ptr(self._js), self._js.size, ptr(self._layers), self._layers.size, ptr(consts),
                ptr(tabs), 0 if tabs is None else tabs.size, precision, log_dtype, lanes_per_instance,
                d_arena, nbytes, stream, int(on_host),
            )
        if not self._h:
            raise RuntimeError(f"cmtj_plan_create: {last_error()}")
        self.info = PlanInfo()
        check(lib().cmtj_plan_query(self._h, ctypes.byref(self.info)), "cmtj_plan_query")

    def execute(self, d_log: int | None, log_capacity: int, d_mag: int | None, d_status: int | None, stream: int | None):
        check(
            lib().cmtj_plan_execute(self._h, d_log, log_capacity, d_mag, d_status, stream),
            "cmtj_plan_execute",
        )

    def execute_slice(self, d_log: int | None, log_capacity: int, d_mag: int | None, d_status: int | None,
                      d_state: int, max_steps: int, streams) -> None:
        arr = (_P * max(len(streams), 1))(*streams)
        check(
            lib().cmtj_plan_execute_slice(self._h, d_log, log_capacity, d_mag, d_status, d_state, max_steps,
                                          arr, len(streams)),
            "cmtj_plan_execute_slice",
        )

    def graph_create(self, d_log: int | None, log_capacity: int, d_mag: int | None, d_status: int | None) -> None:
        check(lib().cmtj_plan_graph_create(self._h, d_log, log_capacity, d_mag, d_status), "cmtj_plan_graph_create")

    def graph_launch(self, stream: int | None) -> None:
        check(lib().cmtj_plan_graph_launch(self._h, stream), "cmtj_plan_graph_launch")

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().cmtj_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


if os.environ.get("CMTJ_AMD_EAGER_LOAD", "1") == "1":
    _load()
