"""ctypes binding of ``libpermhip.so`` (the C ABI in ``include/permhip.h``).

The library sits next to this module and carries the ``lib`` prefix so it
never shadows a Python module (the rule of perm_binary_cuda.py:42-60).  There
is no CPU fallback: every entry point raises ``RuntimeError`` when the
library is missing, when no HIP device is present, or when a call returns the
negative sentinel (perm_cuda.py:125-129).
"""

from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

import numpy as np

LIBRARY_PATH = Path(os.environ.get("PERMHIP_LIBRARY", Path(__file__).resolve().parent / "libpermhip.so"))

_lock = threading.Lock()
_lib = None
# Process that first touched the HIP runtime through this module.  A child
# forked after that point inherits a runtime it cannot use (HIP does not
# support fork after initialisation), so every device call in such a child
# raises instead of entering the library.
_device_pid = None

_I64P = ctypes.POINTER(ctypes.c_int64)
_U64P = ctypes.POINTER(ctypes.c_uint64)
_U8P = ctypes.POINTER(ctypes.c_uint8)
_INTP = ctypes.POINTER(ctypes.c_int)
_I32P = ctypes.POINTER(ctypes.c_int32)
_F64P = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes)
_SIGNATURES = {
    "phip_available": (ctypes.c_int, []),
    "phip_last_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "phip_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "phip_permanent_residues": (
        ctypes.c_int,
        [_I64P, ctypes.c_int64, ctypes.c_int, _I64P, ctypes.c_int, _I64P],
    ),
    "phip_permanent_residues_noalloc": (
        ctypes.c_int,
        [_I64P, ctypes.c_int64, ctypes.c_int, _I64P, ctypes.c_int, _I64P],
    ),
    "phip_permanent_modp": (ctypes.c_int64, [_I64P, ctypes.c_int, ctypes.c_int64]),
    "phip_binary_frontier_residues": (
        ctypes.c_int,
        [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64), ctypes.c_int, ctypes.c_int, _I64P,
         ctypes.c_int, ctypes.c_int64, _I64P, _I64P],
    ),
    "phip_frontier_order": (
        ctypes.c_int,
        [ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)],
    ),
    "phip_binary_blocks": (
        ctypes.c_int,
        [ctypes.POINTER(ctypes.c_uint64), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64)],
    ),
    "phip_binary_blocks_csr": (
        ctypes.c_int,
        [_I32P, _I32P, ctypes.c_int, _I32P, _I32P, _I32P, ctypes.POINTER(ctypes.c_int)],
    ),
    "phip_affine_permanent_residues": (
        ctypes.c_int,
        [_I64P, _I64P, ctypes.c_int64, ctypes.c_int, _I64P, ctypes.c_int, _I64P],
    ),
    "phip_schur_permanent_modp": (ctypes.c_int64, [ctypes.c_int, ctypes.c_int64, ctypes.c_int64]),
    "phip_schur_necklace_orbit_stats": (ctypes.c_int, [ctypes.c_int, _U64P, _U64P]),
    "phip_binary_permanent_residues": (
        ctypes.c_int,
        [_U8P, ctypes.c_int, _I64P, ctypes.c_int, _I64P],
    ),
    "phip_binary_permanent_modp": (ctypes.c_int64, [_U8P, ctypes.c_int, ctypes.c_int64]),
    "phip_permanent_mod3_matrix": (ctypes.c_int, [_INTP, ctypes.c_int]),
    "phip_permanent_mod3": (ctypes.c_int, [_U64P, _U64P, ctypes.c_int]),
    "phip_batch_create": (
        ctypes.c_int,
        [_I64P, _I32P, ctypes.c_int, _I64P, _I32P, ctypes.POINTER(ctypes.c_void_p)],
    ),
    "phip_batch_run": (ctypes.c_int, [ctypes.c_void_p, _I64P]),
    "phip_batch_launch": (ctypes.c_int, [ctypes.c_void_p]),
    "phip_batch_query": (ctypes.c_int, [ctypes.c_void_p, _I64P, ctypes.c_int]),
    "phip_batch_set_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64]),
    "phip_batch_stats": (ctypes.c_int, [ctypes.c_void_p, _F64P, ctypes.c_int]),
    "phip_batch_plan_classes": (ctypes.c_int, [_I64P, _I32P, ctypes.c_int, _I64P, _I32P, _I32P, ctypes.c_int]),
    "phip_block_tables": (ctypes.c_int64, [_I64P, ctypes.c_int, ctypes.c_int, _I32P, ctypes.c_int64, _I32P]),
    "phip_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "phip_set_tuning": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64]),
    "phip_get_tuning": (ctypes.c_int64, [ctypes.c_int]),
    "phip_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "phip_profile_read": (ctypes.c_int, [_F64P, _U64P]),
    "phip_valu_microbench": (ctypes.c_int, [_F64P, ctypes.c_int]),
    "phip_mix_microbench": (ctypes.c_int, [_I32P, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "phip_copy_calibration": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int]),
}

TUNE_LANES_LOG2 = 1
TUNE_MIN_CBITS = 2
TUNE_BATCH_ITEMS_LOG2 = 3
# Moduli per phip_binary_frontier_residues call.
FRONTIER_MAX_PRIMES = 64
TUNE_BATCH_BLOCKS = 4
TUNE_NECKLACE_MIN_ORDER = 5
TUNE_NECKLACE_DEPTH = 6
TUNE_WIDE_MIN_TIER = 7
TUNE_AFFINE_MIN_TIER = 8
TUNE_SKIP_ZEROS_PERMILLE = 9
TUNE_EXACT_MIN_PRIMES = 10
TUNE_EXACT_WIDE_MIN_PRIMES = 11
TUNE_EXACT_COMPACT = 12
TUNE_SKIP_BLOCKS = 13
TUNE_MOD3_BLOCKS = 14
TUNE_SKIP_EXACT_MIN_PRIMES = 15
TUNE_EXACT_SIGNED = 16
BATCH_BLOCKS_PER_CU = 1
BATCH_USE_GRAPH = 2
BATCH_ITEMS_LOG2 = 3
MICROBENCH_NAMES = (
    "v_add_u32",
    "v_mul_i32_i24",
    "v_mul_lo_u32",
    "v_mul_hi_i32",
    "v_mad_i64_i32",
    "v_mad_u64_u32",
    "mont31_product",
    "v_pk_add_u16",
    "v_mul_i32_i24_sdwa",
    "v_lshl_add_u64",
    "v_ashrrev_i32",
)


def exported_symbols():
    """Names the header declares and the library must export."""
    return tuple(_SIGNATURES)


def load(path: Path | None = None):
    """Load the library once; raises when it is absent."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        target = Path(path) if path is not None else LIBRARY_PATH
        if not target.exists():
            raise RuntimeError(
                f"libpermhip.so not found at {target}; build it with "
                "`python -m permhip.build` (there is no CPU fallback)"
            )
        lib = ctypes.CDLL(str(target))
        for name, (restype, argtypes) in _SIGNATURES.items():
            try:
                fn = getattr(lib, name)
            except AttributeError as exc:
                raise RuntimeError(
                    f"{target} lacks {name}: the library is outdated; rebuild it with `python -m permhip.build`"
                ) from exc
            fn.restype = restype
            fn.argtypes = argtypes
        if path is None:
            _lib = lib
        return lib


def last_error() -> str:
    lib = load()
    buf = ctypes.create_string_buffer(1024)
    lib.phip_last_error(buf, len(buf))
    return buf.value.decode("utf-8", "replace")


class ForkedProcessError(RuntimeError):
    """A device call in a process forked after the parent initialised HIP."""


def _check_fork() -> None:
    if _device_pid is not None and os.getpid() != _device_pid:
        raise ForkedProcessError(
            f"libpermhip was initialised in process {_device_pid}; this process ({os.getpid()}) is a fork of it "
            "and cannot use the HIP runtime. Call the permanent functions with parallel=False, or start "
            "worker processes with the 'spawn' method"
        )


def _touch_device(lib) -> int:
    """phip_available(), recording the first process that initialised HIP."""
    global _device_pid
    _check_fork()
    status = lib.phip_available()
    if _device_pid is None:
        _device_pid = os.getpid()
    return status


def available() -> bool:
    """True when the library loads and a HIP device is usable in this
    process (False in a fork of a process that initialised HIP)."""
    try:
        return bool(_touch_device(load()))
    except (OSError, RuntimeError):
        return False


def require_device():
    lib = load()
    if not _touch_device(lib):
        raise RuntimeError("no HIP device is available to libpermhip (there is no CPU fallback)")
    return lib


def device_info() -> str:
    lib = require_device()
    buf = ctypes.create_string_buffer(256)
    if lib.phip_device_info(buf, len(buf)) != 0:
        raise RuntimeError(last_error())
    return buf.value.decode()


def _check(status: int, what: str) -> None:
    if status < 0:
        raise RuntimeError(f"{what} failed: {last_error()}")


def _i64ptr(a: np.ndarray):
    return a.ctypes.data_as(_I64P)


def permanent_residues(matrix: np.ndarray, moduli) -> list:
    """perm(matrix) mod each modulus, one device pass.

    ``matrix`` is either an (n, n) int64 array used for every modulus or a
    (k, n, n) int64 array holding one already-reduced copy per modulus.
    """
    lib = require_device()
    a = np.ascontiguousarray(matrix, dtype=np.int64)
    mods = np.ascontiguousarray(np.asarray(list(moduli), dtype=np.int64))
    k = int(mods.shape[0])
    if a.ndim == 2:
        n = int(a.shape[0])
        stride = 0
    elif a.ndim == 3:
        if a.shape[0] != k:
            raise ValueError("one matrix per modulus is required")
        n = int(a.shape[1])
        stride = n * n
    else:
        raise ValueError("matrix must be 2- or 3-dimensional")
    out = np.zeros(k, dtype=np.int64)
    status = lib.phip_permanent_residues(_i64ptr(a), stride, n, _i64ptr(mods), k, _i64ptr(out))
    _check(status, "phip_permanent_residues")
    return [int(v) for v in out]


def binary_blocks(masks):
    """(feasible, blocks as tuples of local row masks) of a 0/1 matrix given
    as row masks: host code in the library, no device needed."""
    lib = load()
    n = len(masks)
    arr = (ctypes.c_uint64 * max(1, n))(*[int(m) for m in masks])
    count = ctypes.c_int()
    sizes = (ctypes.c_int * max(1, n))()
    out = (ctypes.c_uint64 * max(1, n))()
    status = lib.phip_binary_blocks(arr, n, ctypes.byref(count), sizes, out)
    _check(status, "phip_binary_blocks")
    if status == 0:
        return False, ()
    blocks, pos = [], 0
    for b in range(count.value):
        k = sizes[b]
        blocks.append(tuple(int(out[pos + i]) for i in range(k)))
        pos += k
    return True, tuple(blocks)


def binary_blocks_csr(row_ptr: np.ndarray, col_idx: np.ndarray, n: int):
    """(row matching, row block ids, column block ids, block count) of a 0/1
    matrix of any order in CSR form, or None without a perfect matching:
    host code in the library, no device needed."""
    lib = load()
    ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
    idx = np.ascontiguousarray(col_idx, dtype=np.int32)
    if ptr.shape != (n + 1,):
        raise ValueError("row_ptr must hold n + 1 offsets")
    match = np.zeros(max(1, n), dtype=np.int32)
    rows = np.zeros(max(1, n), dtype=np.int32)
    cols = np.zeros(max(1, n), dtype=np.int32)
    count = ctypes.c_int()
    status = lib.phip_binary_blocks_csr(ptr.ctypes.data_as(_I32P), idx.ctypes.data_as(_I32P) if idx.size else None,
                                        int(n), match.ctypes.data_as(_I32P), rows.ctypes.data_as(_I32P),
                                        cols.ctypes.data_as(_I32P), ctypes.byref(count))
    _check(status, "phip_binary_blocks_csr")
    if status == 0:
        return None
    return match[:n].tolist(), rows[:n].tolist(), cols[:n].tolist(), int(count.value)


def frontier_order(columns):
    """(row order, maximum frontier width, frontier area) of the frontier
    DP's row order for a 0/1 matrix given as the column lists of its rows:
    host code in the library."""
    lib = load()
    n = len(columns)
    ptr = np.zeros(n + 1, dtype=np.int32)
    ptr[1:] = np.cumsum([len(c) for c in columns], dtype=np.int64)
    idx = np.ascontiguousarray(np.fromiter((j for c in columns for j in c), dtype=np.int32, count=int(ptr[-1])))
    order = np.zeros(max(1, n), dtype=np.int32)
    width = ctypes.c_int()
    area = ctypes.c_int64()
    i32 = ctypes.POINTER(ctypes.c_int32)
    status = lib.phip_frontier_order(ptr.ctypes.data_as(i32), idx.ctypes.data_as(i32) if idx.size else None, n,
                                     order.ctypes.data_as(i32), ctypes.byref(width), ctypes.byref(area))
    _check(status, "phip_frontier_order")
    return tuple(int(v) for v in order[:n]), int(width.value), int(area.value)


def binary_frontier_residues(row_slots, expiring, slots: int, primes, max_states: int):
    """Frontier-DP counts modulo each prime (at most 64 moduli, odd and
    below 2^31) for rows given as masks over live-column slots; returns
    (residues, None), or (None, (rows processed, observed states)) when a
    layer exceeds max_states."""
    lib = require_device()
    n = len(row_slots)
    rm = (ctypes.c_uint64 * max(1, n))(*[int(m) for m in row_slots])
    ex = (ctypes.c_uint64 * max(1, n))(*[int(m) for m in expiring])
    ps = np.ascontiguousarray(np.asarray(list(primes), dtype=np.int64))
    out = np.zeros(max(1, len(ps)), dtype=np.int64)
    info = np.zeros(2, dtype=np.int64)
    status = lib.phip_binary_frontier_residues(rm, ex, n, int(slots), _i64ptr(ps), len(ps), int(max_states),
                                               _i64ptr(out), _i64ptr(info))
    _check(status, "phip_binary_frontier_residues")
    if status == 1:
        return None, (int(info[0]), int(info[1]))
    return [int(v) for v in out[: len(ps)]], None


def affine_permanent_residues(constant: np.ndarray, linear: np.ndarray, moduli) -> list:
    """Coefficient tuples of per(C + x L) modulo each modulus, one device
    pass.  C and L are (n, n) int64 arrays used for every modulus, or
    (k, n, n) arrays holding one already-reduced pair per modulus."""
    lib = require_device()
    c = np.ascontiguousarray(constant, dtype=np.int64)
    l = np.ascontiguousarray(linear, dtype=np.int64)
    if c.shape != l.shape:
        raise ValueError("constant and linear matrices must have the same shape")
    mods = np.ascontiguousarray(np.asarray(list(moduli), dtype=np.int64))
    k = int(mods.shape[0])
    if c.ndim == 2:
        n = int(c.shape[0])
        stride = 0
    elif c.ndim == 3:
        if c.shape[0] != k:
            raise ValueError("one matrix pair per modulus is required")
        n = int(c.shape[1])
        stride = n * n
    else:
        raise ValueError("matrices must be 2- or 3-dimensional")
    if n == 0:
        c = l = np.zeros(1, dtype=np.int64)
    out = np.zeros(max(1, k * (n + 1)), dtype=np.int64)
    status = lib.phip_affine_permanent_residues(_i64ptr(c), _i64ptr(l), stride, n, _i64ptr(mods), k, _i64ptr(out))
    _check(status, "phip_affine_permanent_residues")
    return [tuple(int(v) for v in out[i * (n + 1):(i + 1) * (n + 1)]) for i in range(k)]


def schur_permanent_modp(n: int, omega: int, p: int) -> int:
    lib = require_device()
    result = lib.phip_schur_permanent_modp(int(n), int(omega), int(p))
    _check(result, "phip_schur_permanent_modp")
    return int(result)


def schur_necklace_orbit_stats(n: int):
    """(representatives, words) of the device's necklace walk for S_n."""
    lib = require_device()
    reps = ctypes.c_uint64()
    words = ctypes.c_uint64()
    _check(lib.phip_schur_necklace_orbit_stats(int(n), ctypes.byref(reps), ctypes.byref(words)),
           "phip_schur_necklace_orbit_stats")
    return int(reps.value), int(words.value)


def binary_permanent_residues(matrix: np.ndarray, primes) -> list:
    lib = require_device()
    a = np.ascontiguousarray(matrix, dtype=np.uint8)
    n = int(a.shape[0]) if a.ndim == 2 else 0
    ps = np.ascontiguousarray(np.asarray(list(primes), dtype=np.int64))
    out = np.zeros(len(ps), dtype=np.int64)
    ptr = a.ctypes.data_as(_U8P) if n else ctypes.cast(None, _U8P)
    status = lib.phip_binary_permanent_residues(ptr, n, _i64ptr(ps), len(ps), _i64ptr(out))
    _check(status, "phip_binary_permanent_residues")
    return [int(v) for v in out]


def permanent_mod3(matrix: np.ndarray) -> int:
    """Takes an (n, n) matrix already reduced modulo 3 (values 0, 1, 2)."""
    lib = require_device()
    a = np.ascontiguousarray(matrix, dtype=np.int32)
    n = int(a.shape[0])
    result = lib.phip_permanent_mod3_matrix(a.ctypes.data_as(_INTP), n)
    _check(result, "phip_permanent_mod3_matrix")
    return int(result)


def set_tuning(key: int, value: int) -> None:
    _check(load().phip_set_tuning(int(key), int(value)), "phip_set_tuning")


def get_tuning(key: int) -> int:
    value = load().phip_get_tuning(int(key))
    _check(value, "phip_get_tuning")
    return int(value)


def profile_enable(on: bool) -> None:
    _check(require_device().phip_profile_enable(1 if on else 0), "phip_profile_enable")


def profile_read():
    """(summed kernel milliseconds, launches) of the single-matrix kernel."""
    ms = ctypes.c_double()
    launches = ctypes.c_uint64()
    _check(require_device().phip_profile_read(ctypes.byref(ms), ctypes.byref(launches)), "phip_profile_read")
    return ms.value, launches.value


def valu_microbench() -> dict:
    """{instruction: lane-operations per second} over the whole device."""
    out = (ctypes.c_double * len(MICROBENCH_NAMES))()
    count = require_device().phip_valu_microbench(out, len(out))
    _check(count, "phip_valu_microbench")
    return {name: out[i] for i, name in enumerate(MICROBENCH_NAMES[:count])}


MIX_TYPES = ("v_mad_i64_i32", "v_mul_lo_u32", "v_pk_add_u16", "v_mul_i32_i24_sdwa", "v_add_u32")


def mix_microbench(rounds, iters: int = 64) -> float:
    """Lane-iterations per second of a synthetic kernel issuing rounds[t]
    rounds of 8 independent instructions of each type in MIX_TYPES."""
    arr = (ctypes.c_int32 * 5)(*[int(r) for r in rounds])
    out = ctypes.c_double()
    _check(require_device().phip_mix_microbench(arr, int(iters), ctypes.byref(out)), "phip_mix_microbench")
    return out.value


def copy_calibration(nbytes: int, reps: int = 1) -> None:
    _check(require_device().phip_copy_calibration(int(nbytes), int(reps)), "phip_copy_calibration")


def _batch_description(matrices, moduli_lists):
    mats = [np.ascontiguousarray(m, dtype=np.int64) for m in matrices]
    if len(mats) != len(moduli_lists):
        raise ValueError("one modulus list per matrix is required")
    orders = np.array([m.shape[0] for m in mats], dtype=np.int32)
    entries = (np.concatenate([m.ravel() for m in mats]) if mats else np.zeros(1, dtype=np.int64)).astype(np.int64)
    counts = np.array([len(ms) for ms in moduli_lists], dtype=np.int32)
    flat = [int(q) for ms in moduli_lists for q in ms]
    moduli = np.array(flat if flat else [0], dtype=np.int64)
    return mats, entries, orders, counts, moduli, len(flat)


def batch_plan_classes(matrices, moduli_lists) -> list:
    """(group, prime-set size, variant) of the kernel class that would
    compute each residue of this batch, in input order: host code, no
    device."""
    lib = load()
    mats, entries, orders, counts, moduli, total = _batch_description(matrices, moduli_lists)
    out = np.zeros(max(1, total), dtype=np.int32)
    count = lib.phip_batch_plan_classes(entries.ctypes.data_as(_I64P), orders.ctypes.data_as(_I32P), len(mats),
                                        moduli.ctypes.data_as(_I64P), counts.ctypes.data_as(_I32P),
                                        out.ctypes.data_as(_I32P), total)
    _check(count, "phip_batch_plan_classes")
    return [(int(c) // 256, (int(c) // 8) % 32, int(c) % 8) for c in out[:count]]


def block_tables(matrix: np.ndarray, mod3: bool = False):
    """The block zero-skipping walk's tables for one 0/1 matrix, or with
    mod3 the modulo-3 walk's for entries in {-1, 0, 1} (host code, no
    device): (uint32 words, rows tested, rows of even degree, row-sum
    group)."""
    lib = load()
    a = np.ascontiguousarray(matrix, dtype=np.int64)
    n = int(a.shape[0])
    meta = np.zeros(4, dtype=np.int32)
    words = lib.phip_block_tables(a.ctypes.data_as(_I64P), n, int(mod3), None, 0, meta.ctypes.data_as(_I32P))
    _check(words, "phip_block_tables")
    out = np.zeros(int(words), dtype=np.int32)
    got = lib.phip_block_tables(a.ctypes.data_as(_I64P), n, int(mod3), out.ctypes.data_as(_I32P), int(words),
                                meta.ctypes.data_as(_I32P))
    _check(got, "phip_block_tables")
    return out.view(np.uint32), int(meta[0]), int(meta[1]), int(meta[2])


class Batch:
    """Device-resident batch plan (phip_batch_*)."""

    def __init__(self, matrices, moduli_lists):
        lib = require_device()
        mats, entries, orders, counts, moduli, total = _batch_description(matrices, moduli_lists)
        self._keep = (entries, orders, counts, moduli)
        self.counts = [int(c) for c in counts]
        self.total = total
        handle = ctypes.c_void_p()
        status = lib.phip_batch_create(
            entries.ctypes.data_as(_I64P),
            orders.ctypes.data_as(_I32P),
            len(mats),
            moduli.ctypes.data_as(_I64P),
            counts.ctypes.data_as(_I32P),
            ctypes.byref(handle),
        )
        _check(status, "phip_batch_create")
        self._lib = lib
        self._handle = handle
        self._out = np.zeros(max(1, self.total), dtype=np.int64)

    def set_option(self, option: int, value: int) -> None:
        _check(self._lib.phip_batch_set_option(self._handle, option, int(value)), "phip_batch_set_option")

    def run(self, fetch: bool = True):
        """Run the plan; returns the flat residue list (input order) or None."""
        out = self._out.ctypes.data_as(_I64P) if fetch else ctypes.cast(None, _I64P)
        _check(self._lib.phip_batch_run(self._handle, out), "phip_batch_run")
        if not fetch:
            return None
        return [int(v) for v in self._out[: self.total]]

    def launch(self) -> None:
        """Start a run without waiting (see poll)."""
        _check(self._lib.phip_batch_launch(self._handle), "phip_batch_launch")

    def poll(self, block: bool = False):
        """Residue list once the launched run has finished, else None."""
        out = self._out.ctypes.data_as(_I64P)
        status = self._lib.phip_batch_query(self._handle, out, 1 if block else 0)
        _check(status, "phip_batch_query")
        if status == 0:
            return None
        return [int(v) for v in self._out[: self.total]]

    def stats(self) -> dict:
        keys = ("last_ms", "states", "mulmods", "jobs", "items", "launches", "waves", "group", "residues", "set",
                "chunk_bits_min", "chunk_bits_max", "items_log2", "blocks_per_cu", "variant", "resident_blocks_per_cu",
                "skip_test_words", "skip_block_rows")
        vals = (ctypes.c_double * len(keys))()
        count = self._lib.phip_batch_stats(self._handle, vals, len(keys))
        _check(count, "phip_batch_stats")
        return {k: vals[i] for i, k in enumerate(keys[:count])}

    def close(self) -> None:
        if getattr(self, "_handle", None):
            self._lib.phip_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
