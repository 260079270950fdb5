"""Python host driver for the MI355X Ouroboros allocator (libouro.so, C ABI in include/ouro_capi.h).

The allocator itself is device code: kernels call malloc/free through include/ouro/instances.hpp.
This module drives the library's batched entry points and its trace harness from Python, with the
reference's host vocabulary (Ouroboros::initialize, getDeviceMemoryManager,
include/device/Ouroboros.cuh:281,297). The library must be built (ouroboros_amd/build.py); there
is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libouro.so"

VARIANTS = {"PQ": 0, "CQ": 1, "VAPQ": 2, "VACQ": 3, "VLPQ": 4, "VLCQ": 5,
            # pool geometries of the batched form's tuner (queue kind / chunk bytes / classes / capacity)
            "VAPQ-c8k-q10": 6, "VAPQ-c8k-q8": 7, "VAPQ-c16k-q10": 8, "VAPQ-c16k-q8": 9,
            "VAPQ-cap4k": 10, "PQ-cap1m": 11}
VARIANT_NAMES = {v: k for k, v in VARIANTS.items()}
MAX_CLASSES = 16

ERR_BITS = {
    1 << 0: "OOM",
    1 << 1: "QUEUE_FULL",
    1 << 2: "VSTART",
    1 << 3: "TIMEOUT",
    1 << 4: "BAD_FREE",
    1 << 5: "TOO_LARGE",
    1 << 6: "REUSE_FULL",
    1 << 7: "LOCATOR",
    1 << 8: "CORRUPT",
}


class OuroError(RuntimeError):
    pass


class State(C.Structure):
    _fields_ = [
        ("allocation_size", C.c_uint64),
        ("max_chunks", C.c_uint64),
        ("data_offset", C.c_uint64),
        ("next_free_chunk", C.c_uint32),
        ("error", C.c_uint32),
        ("variant", C.c_uint32),
        ("num_classes", C.c_uint32),
        ("page_size", C.c_uint32 * MAX_CLASSES),
        ("pages_per_chunk", C.c_uint32 * MAX_CLASSES),
        ("sem_count", C.c_int64 * MAX_CLASSES),
        ("sem_expected", C.c_uint32 * MAX_CLASSES),
        ("sem_reserved", C.c_uint32 * MAX_CLASSES),
        ("front", C.c_uint32 * MAX_CLASSES),
        ("back", C.c_uint32 * MAX_CLASSES),
        ("queue_count", C.c_int32 * MAX_CLASSES),
        ("carved", C.c_uint32 * MAX_CLASSES),
        ("front_qc", C.c_uint32 * MAX_CLASSES),
        ("kind", C.c_uint32 * MAX_CLASSES),
        ("reuse_count", C.c_int32 * 2),
    ]


class CheckResult(C.Structure):
    _fields_ = [
        ("violations", C.c_uint64),
        ("chunks_checked", C.c_uint64),
        ("pages_free", C.c_uint64),
        ("first_class", C.c_int32),
        ("first_chunk", C.c_int32),
        ("message", C.c_char * 160),
    ]


class TraceConfig(C.Structure):
    _fields_ = [
        ("threads", C.c_uint32),
        ("rounds", C.c_uint32),
        ("agg", C.c_uint32),
        ("block", C.c_uint32),
        ("check_rounds", C.c_uint32),
        ("latency", C.c_uint32),
        ("hold_permille", C.c_uint32),
        ("pad", C.c_uint32),
        ("hold_seed", C.c_uint64),
        ("sizes", C.POINTER(C.c_uint32)),
        ("round_sizes", C.POINTER(C.c_uint32)),
        ("offsets_out", C.POINTER(C.c_uint64)),
    ]


class TraceResult(C.Structure):
    _fields_ = [
        ("alloc_ms", C.c_double),
        ("free_ms", C.c_double),
        ("write_ms", C.c_double),
        ("verify_ms", C.c_double),
        ("mallocs", C.c_uint64),
        ("frees", C.c_uint64),
        ("null_returns", C.c_uint64),
        ("verify_errors", C.c_uint64),
        ("bounds_errors", C.c_uint64),
        ("overlap_errors", C.c_uint64),
        ("align_errors", C.c_uint64),
        ("touch_bytes", C.c_uint64),
        ("peak_chunks", C.c_uint64),
        ("peak_live_chunks", C.c_uint64),
        ("ideal_chunks", C.c_uint64),
        ("lat_samples", C.c_uint64),
        ("lat_p50_ns", C.c_double),
        ("lat_p99_ns", C.c_double),
        ("lat_max_ns", C.c_double),
        ("error_word", C.c_uint32),
        ("pad", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}
        d["errors"] = decode_errors(self.error_word)
        return d


# Entry points and their C signatures (include/ouro_capi.h).
class BatchConfig(C.Structure):
    _fields_ = [
        ("requests", C.c_uint32),
        ("rounds", C.c_uint32),
        ("warmup", C.c_uint32),
        ("agg", C.c_int32),
        ("block", C.c_uint32),
        ("grid", C.c_uint32),
        ("graph", C.c_int32),
        ("check", C.c_int32),
        ("sizes", C.POINTER(C.c_uint32)),
        ("offsets_out", C.POINTER(C.c_uint64)),
        ("latency", C.c_int32),
        ("tile", C.c_uint32),
        ("kernel_timing", C.c_int32),
        ("pad", C.c_uint32),
    ]


class BatchResult(C.Structure):
    _fields_ = [
        ("alloc_ms", C.c_double),
        ("free_ms", C.c_double),
        ("touch_ms", C.c_double),
        ("round_ms", C.c_double),
        ("mallocs", C.c_uint64),
        ("null_returns", C.c_uint64),
        ("verify_errors", C.c_uint64),
        ("bounds_errors", C.c_uint64),
        ("overlap_errors", C.c_uint64),
        ("align_errors", C.c_uint64),
        ("touch_bytes", C.c_uint64),
        ("error_word", C.c_uint32),
        ("grid", C.c_uint32),
        ("peak_chunks", C.c_uint32),
        ("graph_used", C.c_uint32),
        ("lat_p50_ns", C.c_double),
        ("lat_p99_ns", C.c_double),
        ("lat_max_ns", C.c_double),
        ("lat_samples", C.c_uint64),
        ("alloc_kernel_ms", C.c_double),
        ("free_kernel_ms", C.c_double),
        ("kernel_rounds", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["errors"] = decode_errors(self.error_word)
        t = (self.alloc_ms + self.free_ms) * 1e-3
        d["mops"] = 2.0 * self.mallocs / t / 1e6 if t > 0 else 0.0
        return d


class BatchWorkload(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("workloads", C.c_uint32),
        ("lo", C.c_uint32),
        ("hi", C.c_uint32),
        ("max_allocs", C.c_uint32),
        ("classes", C.c_uint32),
        ("pad", C.c_uint32),
        ("mu", C.c_double),
        ("sigma", C.c_double),
        ("s_lo", C.c_double),
        ("s_hi", C.c_double),
    ]


class VariantInfo(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("classes", C.c_uint32),
        ("smallest_page", C.c_uint32),
        ("largest_page", C.c_uint32),
        ("chunk_bytes", C.c_uint32),
        ("queue_entries", C.c_uint32),
        ("page_capacity", C.c_uint64),
    ]


class ServeConfig(C.Structure):
    _fields_ = [
        ("rate", C.c_double),
        ("zipf", C.c_double),
        ("seed", C.c_uint64),
        ("requests", C.c_uint32),
        ("slots", C.c_uint32),
        ("agg", C.c_int32),
        ("min_log2", C.c_uint32),
        ("max_log2", C.c_uint32),
        ("classes", C.c_uint32),
        ("admit", C.c_int32),
        ("max_batch", C.c_uint32),
        ("results_out", C.POINTER(C.c_ulonglong)),
        ("latency_out", C.POINTER(C.c_double)),
        ("tile", C.c_uint32),
        ("pad", C.c_uint32),
    ]


class ServeResult(C.Structure):
    _fields_ = [
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("mean_ms", C.c_double),
        ("max_ms", C.c_double),
        ("wall_ms", C.c_double),
        ("offered_ms", C.c_double),
        ("mops", C.c_double),
        ("achieved_rate", C.c_double),
        ("mallocs", C.c_uint64),
        ("mismatches", C.c_uint64),
        ("nulls", C.c_uint64),
        ("completed", C.c_uint32),
        ("max_in_flight", C.c_uint32),
        ("deferred", C.c_uint32),
        ("error_word", C.c_uint32),
        ("launches", C.c_uint32),
        ("pad", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["errors"] = decode_errors(self.error_word)
        return d


SIGNATURES = {
    "ouro_serve": (C.c_int, [C.c_void_p, C.POINTER(ServeConfig), C.POINTER(ServeResult)]),
    "ouro_serve_expected": (C.c_int, [C.POINTER(ServeConfig), C.c_uint32, C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]),
    "ouro_variant_geometry": (C.c_int, [C.c_int, C.POINTER(VariantInfo)]),
    "ouro_gen_batch": (C.c_int, [C.POINTER(BatchWorkload), C.POINTER(C.c_uint32), C.c_uint64,
                                 C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
    "ouro_run_batch": (C.c_int, [C.c_void_p, C.POINTER(BatchConfig), C.POINTER(BatchResult)]),
    "ouro_run_batch_static": (C.c_int, [C.c_void_p, C.POINTER(BatchConfig), C.c_void_p, C.POINTER(BatchResult)]),
    "ouro_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "ouro_destroy": (C.c_int, [C.c_void_p]),
    "ouro_device_manager": (C.c_void_p, [C.c_void_p]),
    "ouro_data_base": (C.c_void_p, [C.c_void_p]),
    "ouro_layout": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "ouro_malloc_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]),
    "ouro_free_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]),
    "ouro_query": (C.c_int, [C.c_void_p, C.POINTER(State)]),
    "ouro_check_quiescent": (C.c_int, [C.c_void_p, C.POINTER(CheckResult)]),
    "ouro_snapshot": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ouro_snapshot_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "ouro_restore": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ouro_reinitialize": (C.c_int, [C.c_void_p, C.c_double]),
    "ouro_set_recycling": (C.c_int, [C.c_void_p, C.c_uint32]),
    "ouro_preload": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]),
    "ouro_set_statistics": (C.c_int, [C.c_void_p, C.c_int]),
    "ouro_statistics": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "ouro_run_trace": (C.c_int, [C.c_void_p, C.POINTER(TraceConfig), C.POINTER(TraceResult)]),
    "ouro_gen_sizes": (C.c_int, [C.c_uint64, C.c_uint32, C.c_int, C.c_double, C.c_uint32, C.POINTER(C.c_uint32)]),
    "ouro_sem_share": (C.c_int, [C.c_uint64, C.c_int64, C.c_int64, C.POINTER(C.c_uint32)]),
    "ouro_sem_make": (C.c_uint64, [C.c_int64, C.c_uint64]),
    "ouro_abi_sizes": (C.c_int, [C.POINTER(C.c_uint32), C.c_uint32]),
    "ouro_set_origin": (C.c_int, [C.c_void_p, C.c_uint32]),
    "ouro_bench_atomics": (C.c_int, [C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_double)]),
    "ouro_bench_copy": (C.c_int, [C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
    "ouro_bench_floor": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_uint32)]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libouro.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise OuroError(f"{_LIB_PATH} is missing: build it with ouroboros_amd/build.py")
        _lib = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def decode_errors(word: int) -> list[str]:
    return [name for bit, name in ERR_BITS.items() if word & bit]


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise OuroError(f"{what} failed with status {rc}")


def layout(variant: str | int, pool_bytes: int) -> tuple[int, int]:
    """(max_chunks, allocation_size) for a pool (MemoryInitialization.cuh:156-166); no device needed."""
    v = VARIANTS[variant] if isinstance(variant, str) else variant
    mc, asz = C.c_uint64(), C.c_uint64()
    _check(lib().ouro_layout(v, pool_bytes, C.byref(mc), C.byref(asz)), "ouro_layout")
    return mc.value, asz.value


def gen_sizes(seed: int, n: int, dist: str = "uniform", param: float = 0.0, classes: int = 10) -> np.ndarray:
    """Host trace generator: 'fixed' (param = bytes), 'uniform' over classes, 'zipf' (param = s)."""
    d = {"fixed": 0, "uniform": 1, "zipf": 2}[dist]
    out = np.empty(n, dtype=np.uint32)
    _check(
        lib().ouro_gen_sizes(seed, n, d, float(param), classes, out.ctypes.data_as(C.POINTER(C.c_uint32))),
        "ouro_gen_sizes",
    )
    return out


def variant_geometry(variant: str) -> dict:
    """Kind, classes, page range, chunk bytes and per-class queue capacity of a variant."""
    g = VariantInfo()
    _check(lib().ouro_variant_geometry(VARIANTS[variant], C.byref(g)), "ouro_variant_geometry")
    d = {name: getattr(g, name) for name, _ in g._fields_}
    d["kind_name"] = VARIANT_NAMES[g.kind]
    return d


def gen_batch(seed: int = 42, workloads: int = 256, mu: float = 10.0, sigma: float = 1.5, lo: int = 64,
              hi: int = 1 << 20, s_lo: float = 0.8, s_hi: float = 1.6, max_allocs: int = 4,
              classes: int = 10) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The batch workload set (default: 256 workloads, lognormal(10, 1.5) requests clamped to
    [64, 2^20], Zipf exponents in [0.8, 1.6], 1-4 allocations per request). Returns (flat allocation
    sizes, requests per workload, Zipf exponent per workload)."""
    w = BatchWorkload(seed=seed, workloads=workloads, lo=lo, hi=hi, max_allocs=max_allocs, classes=classes,
                      pad=0, mu=mu, sigma=sigma, s_lo=s_lo, s_hi=s_hi)
    n = C.c_uint64()
    req = np.empty(workloads, dtype=np.uint32)
    zs = np.empty(workloads, dtype=np.float64)
    _check(lib().ouro_gen_batch(C.byref(w), None, 0, C.byref(n), req.ctypes.data_as(C.POINTER(C.c_uint32)),
                                zs.ctypes.data_as(C.POINTER(C.c_double))), "ouro_gen_batch")
    sizes = np.empty(n.value, dtype=np.uint32)
    _check(lib().ouro_gen_batch(C.byref(w), sizes.ctypes.data_as(C.POINTER(C.c_uint32)), n.value, C.byref(n),
                                None, None), "ouro_gen_batch")
    return sizes, req, zs


def serve_config(requests: int, rate: float, slots: int = 8, agg: int = 64, min_log2: int = 10,
                 max_log2: int = 18, zipf: float = 1.1, classes: int = 10, seed: int = 1,
                 admit: bool = True, max_batch: int = 1, tile: int = 1) -> ServeConfig:
    cfg = ServeConfig()
    cfg.rate, cfg.zipf, cfg.seed, cfg.requests, cfg.slots, cfg.agg = rate, zipf, seed, requests, slots, agg
    cfg.min_log2, cfg.max_log2, cfg.classes, cfg.admit = min_log2, max_log2, classes, int(admit)
    cfg.max_batch = max_batch
    cfg.tile = tile
    return cfg


def serve_expected(cfg: ServeConfig, r: int) -> tuple[int, int]:
    """(threads, intact words) of request r when every allocation succeeds (host computation)."""
    t, w = C.c_uint64(), C.c_uint64()
    _check(lib().ouro_serve_expected(C.byref(cfg), r, C.byref(t), C.byref(w)), "ouro_serve_expected")
    return t.value, w.value


def sem_share(before: int, num: int, ppc: int) -> int:
    """Chunks the device makes a claim of `num` pages carve, given the semaphore word it displaced."""
    k = C.c_uint32()
    _check(lib().ouro_sem_share(before, num, ppc, C.byref(k)), "ouro_sem_share")
    return k.value


def sem_make(count: int, expected: int) -> int:
    """The device's packing of (count, expected) into the semaphore word."""
    return lib().ouro_sem_make(count, expected)


def bench_atomics(threads: int = 1 << 20, iters: int = 64, pattern: str = "distinct") -> float:
    """Device-scope u32 fetch-add-with-return rate (atomics/s)."""
    p = {"distinct": 0, "per_wave": 1, "same": 2}[pattern]
    r = C.c_double()
    _check(lib().ouro_bench_atomics(threads, iters, p, C.byref(r)), "ouro_bench_atomics")
    return r.value


def bench_copy(nbytes: int = 1 << 30, reps: int = 10) -> float:
    """Device copy bandwidth in GB/s (read + write bytes)."""
    r = C.c_double()
    _check(lib().ouro_bench_copy(nbytes, reps, C.byref(r)), "ouro_bench_copy")
    return r.value


@dataclass
class TraceOutput:
    result: dict
    offsets: np.ndarray | None


class MemoryManager:
    """One allocator pool on the current HIP device (MultiOuro<variant>)."""

    def __init__(self, variant: str = "PQ", pool_bytes: int = 512 << 20, origin: int = 0):
        """origin: start every class queue's 32-bit positions there (a multiple of 2048; test hook
        for the position wrap)."""
        self.variant = variant
        self.pool_bytes = pool_bytes
        h = C.c_void_p()
        _check(lib().ouro_create(VARIANTS[variant], pool_bytes, C.byref(h)), f"ouro_create({variant})")
        self._h = h
        if origin:
            _check(lib().ouro_set_origin(self._h, origin), "ouro_set_origin")

    # reference host API names
    def getDeviceMemoryManager(self) -> int:  # noqa: N802 - reference name
        return lib().ouro_device_manager(self._h)

    def data_base(self) -> int:
        return lib().ouro_data_base(self._h)

    def close(self) -> None:
        if self._h:
            lib().ouro_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def query(self) -> State:
        st = State()
        _check(lib().ouro_query(self._h, C.byref(st)), "ouro_query")
        return st

    def check_quiescent(self) -> CheckResult:
        r = CheckResult()
        _check(lib().ouro_check_quiescent(self._h, C.byref(r)), "ouro_check_quiescent")
        return r

    def set_statistics(self, on: bool = True) -> None:
        """Zero the statistics counters and start counting (on), or stop."""
        _check(lib().ouro_set_statistics(self._h, int(on)), "ouro_set_statistics")

    def bench_floor(self, sizes: np.ndarray, block: int = 1024, tile: int = 1, rounds: int = 10,
                    grid: int = 0) -> dict:
        """The floor kernel under the batched malloc at the allocator's launch shape: average kernel ms."""
        sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        ms, g = C.c_double(), C.c_uint32()
        _check(lib().ouro_bench_floor(self._h, sizes.ctypes.data_as(C.POINTER(C.c_uint32)), len(sizes), block, grid,
                                      tile, rounds, C.byref(ms), C.byref(g)), "ouro_bench_floor")
        return {"kernel_ms": ms.value, "grid": g.value}

    def set_recycling(self, min_queued: int | None = 0) -> None:
        """Chunk recycling for the chunk-queue kinds: fully free chunks go back to the reuse ring while
        their class queue holds more than min_queued chunks; None switches it off (the default)."""
        v = 0xFFFFFFFF if min_queued is None else int(min_queued)
        _check(lib().ouro_set_recycling(self._h, v), "ouro_set_recycling")

    def statistics(self) -> dict:
        """pageAllocCount, pageFreeCount, chunkAllocationCount, chunkReuseCount (Statistics.h:5-17)."""
        out = (C.c_uint64 * 4)()
        _check(lib().ouro_statistics(self._h, out), "ouro_statistics")
        return dict(zip(("page_alloc", "page_free", "chunk_alloc", "chunk_reuse"), list(out)))

    def preload(self, class_index: int, available_pages: int) -> int:
        """enqueueInitialChunk on a chunk taken from the pool: its last `available_pages` pages become
        free pages of the class (counted over the whole chain); returns the chunk index."""
        c = C.c_uint32()
        _check(lib().ouro_preload(self._h, class_index, available_pages, C.byref(c)), "ouro_preload")
        return c.value

    def reinitialize(self, overallocation_factor: float) -> None:
        """A fresh pool of overallocation_factor times the original size (blocks handed out before
        become invalid; the device manager moves)."""
        _check(lib().ouro_reinitialize(self._h, overallocation_factor), "ouro_reinitialize")

    def snapshot(self) -> np.ndarray:
        """Host copy of the complete allocator state (the pool and its chunk class bytes)."""
        n = C.c_uint64()
        _check(lib().ouro_snapshot_size(self._h, C.byref(n)), "ouro_snapshot_size")
        buf = np.empty(n.value, dtype=np.uint8)
        _check(lib().ouro_snapshot(self._h, buf.ctypes.data_as(C.c_void_p)), "ouro_snapshot")
        return buf

    def restore(self, snap: np.ndarray) -> None:
        """Continue from a snapshot of a pool of the same variant and size."""
        snap = np.ascontiguousarray(snap, dtype=np.uint8)
        _check(lib().ouro_restore(self._h, snap.ctypes.data_as(C.c_void_p)), "ouro_restore")

    def malloc_batch(self, d_sizes: int, d_out: int, n: int, agg: int = 64, stream: int = 0) -> None:
        _check(lib().ouro_malloc_batch(self._h, d_sizes, d_out, n, agg, stream), "ouro_malloc_batch")

    def free_batch(self, d_ptrs: int, n: int, agg: int = 64, stream: int = 0) -> None:
        _check(lib().ouro_free_batch(self._h, d_ptrs, n, agg, stream), "ouro_free_batch")

    def run_batch(
        self,
        sizes: np.ndarray,
        rounds: int,
        warmup: int = 1,
        agg: int = 0,
        block: int = 256,
        grid: int = 0,
        graph: bool = True,
        check: bool = False,
        offsets: np.ndarray | None = None,
        latency: bool = False,
        tile: int = 1,
        kernel_timing: bool = False,
    ) -> dict:
        """The batched form: `sizes` is a flat request list served by persistent block-aggregated
        kernels (agg 0, `tile` = 1, 2 or 4 requests per thread and call) or the per-thread kernels
        (agg 1, 16, 64); rounds replay a captured graph. `offsets`, if given, receives the validation
        round's blocks (byte offsets, UINT64_MAX = null); `latency` adds an untimed round that samples
        the wall-clock time of each allocation call; `kernel_timing` (agg 0) adds as many rounds again that
        time each persistent kernel by its own dispatch's events (alloc_kernel_ms, free_kernel_ms)."""
        sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        cfg = BatchConfig()
        cfg.requests = len(sizes)
        cfg.rounds = rounds
        cfg.warmup = warmup
        cfg.agg = agg
        cfg.block = block
        cfg.grid = grid
        cfg.graph = int(graph)
        cfg.check = int(check or offsets is not None)
        cfg.latency = int(latency)
        cfg.tile = tile
        cfg.kernel_timing = int(kernel_timing)
        cfg.sizes = sizes.ctypes.data_as(C.POINTER(C.c_uint32))
        if offsets is not None:
            assert offsets.dtype == np.uint64 and offsets.shape == sizes.shape and offsets.flags.c_contiguous
            cfg.offsets_out = offsets.ctypes.data_as(C.POINTER(C.c_uint64))
        res = BatchResult()
        _check(lib().ouro_run_batch(self._h, C.byref(cfg), C.byref(res)), "ouro_run_batch")
        return res.as_dict()

    def run_batch_static(self, sizes: np.ndarray, rounds: int, warmup: int = 0, agg: int = 0, block: int = 256,
                         grid: int = 0, check: bool = False, offsets: np.ndarray | None = None, tile: int = 1,
                         stream: int = 0) -> dict:
        """run_batch without allocating: request sizes, block pointers and counters live in the
        library's static device scratch (at most 2^20 requests), kernels run on `stream`; no timings,
        no graph, no host overlap check."""
        sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        cfg = BatchConfig()
        cfg.requests, cfg.rounds, cfg.warmup, cfg.agg, cfg.block, cfg.grid = len(sizes), rounds, warmup, agg, block, grid
        cfg.check = int(check or offsets is not None)
        cfg.tile = tile
        cfg.sizes = sizes.ctypes.data_as(C.POINTER(C.c_uint32))
        if offsets is not None:
            assert offsets.dtype == np.uint64 and offsets.shape == sizes.shape and offsets.flags.c_contiguous
            cfg.offsets_out = offsets.ctypes.data_as(C.POINTER(C.c_uint64))
        res = BatchResult()
        _check(lib().ouro_run_batch_static(self._h, C.byref(cfg), stream, C.byref(res)), "ouro_run_batch_static")
        return res.as_dict()

    def serve(
        self,
        requests: int,
        rate: float,
        slots: int = 8,
        agg: int = 64,
        min_log2: int = 10,
        max_log2: int = 18,
        zipf: float = 1.1,
        classes: int = 10,
        seed: int = 1,
        admit: bool = True,
        per_request: bool = False,
        max_batch: int = 1,
        tile: int = 1,
    ) -> dict:
        """The serving form: Poisson arrivals at `rate` per second, each request 2^U[min_log2, max_log2]
        threads allocating Zipf(zipf) class sizes; up to `slots` launches in flight, each carrying up
        to `max_batch` of the requests that have arrived.
        agg 0 serves each request block-wide (malloc_tile / free_tile, `tile` request threads per GPU
        thread); slots = 1, agg = 1 is the plain per-request launch path. With per_request the result holds
        `results` ([requests, 4]: words intact, words not, nulls, threads) and `latency_ms`."""
        cfg = serve_config(requests, rate, slots, agg, min_log2, max_log2, zipf, classes, seed, admit, max_batch,
                           tile)
        res_buf = np.zeros((requests, 4), np.uint64) if per_request else None
        lat_buf = np.zeros(requests, np.float64) if per_request else None
        if per_request:
            cfg.results_out = res_buf.ctypes.data_as(C.POINTER(C.c_ulonglong))
            cfg.latency_out = lat_buf.ctypes.data_as(C.POINTER(C.c_double))
        res = ServeResult()
        _check(lib().ouro_serve(self._h, C.byref(cfg), C.byref(res)), "ouro_serve")
        d = res.as_dict()
        if per_request:
            d["results"] = res_buf
            d["latency_ms"] = lat_buf
        return d

    def run_trace(
        self,
        sizes: np.ndarray,
        rounds: int,
        agg: int = 64,
        block: int = 256,
        check_rounds: bool = False,
        latency: bool = False,
        keep_offsets: bool = False,
        hold_permille: int = 0,
        hold_seed: int = 0,
    ) -> TraceOutput:
        """Run `rounds` alloc / write / verify / free rounds (src/main.cu:146-167).

        `sizes` is either [threads] (fixed across rounds) or [rounds, threads]."""
        sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        per_round = sizes.ndim == 2
        n = sizes.shape[-1]
        cfg = TraceConfig()
        cfg.threads = n
        cfg.rounds = rounds
        cfg.agg = agg
        cfg.block = block
        cfg.check_rounds = int(check_rounds)
        cfg.latency = int(latency)
        cfg.hold_permille = hold_permille
        cfg.hold_seed = hold_seed
        ptr = sizes.ctypes.data_as(C.POINTER(C.c_uint32))
        if per_round:
            assert sizes.shape[0] == rounds
            cfg.round_sizes = ptr
        else:
            cfg.sizes = ptr
        offsets = None
        if keep_offsets:
            offsets = np.empty((rounds, n), dtype=np.uint64)
            cfg.offsets_out = offsets.ctypes.data_as(C.POINTER(C.c_uint64))
        res = TraceResult()
        _check(lib().ouro_run_trace(self._h, C.byref(cfg), C.byref(res)), "ouro_run_trace")
        return TraceOutput(res.as_dict(), offsets)
