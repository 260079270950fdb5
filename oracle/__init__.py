"""CPU checkers for the MI355X Ouroboros allocator -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package, and only as
the checker or the CPU baseline. The product (ouroboros_amd, libouro.so) never uses it.

- Sim: ctypes binding of ouro_sim.c, the sequential restatement of the reference's queue logic.
- glibc_trace: the host-core baseline (glibc malloc/free over the identical trace).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
VARIANTS = {"PQ": 0, "CQ": 1, "VAPQ": 2, "VACQ": 3, "VLPQ": 4, "VLCQ": 5}
NONE = np.uint64(0xFFFFFFFFFFFFFFFF)


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


class SimState(C.Structure):
    _fields_ = [
        ("next_free_chunk", C.c_uint32),
        ("error", C.c_uint32),
        ("num_classes", C.c_uint32),
        ("sem_count", C.c_int64 * 16),
        ("sem_expected", C.c_uint32 * 16),
        ("sem_reserved", C.c_uint32 * 16),
        ("front", C.c_uint32 * 16),
        ("back", C.c_uint32 * 16),
        ("queue_count", C.c_int32 * 16),
        ("carved", C.c_uint32 * 16),
        ("reuse_count", C.c_int32 * 2),
        ("reuse_front", C.c_uint32 * 2),
        ("reuse_back", C.c_uint32 * 2),
    ]


_sim = None
_cpu = None


def _load_sim() -> C.CDLL:
    global _sim
    if _sim is None:
        path = HERE / "libouro_sim.so"
        if not path.exists():
            build()
        lib = C.CDLL(str(path))
        u64p = C.POINTER(C.c_uint64)
        lib.sim_create.restype = C.c_void_p
        lib.sim_create.argtypes = [C.c_int, C.c_uint64]
        lib.sim_destroy.argtypes = [C.c_void_p]
        lib.sim_layout.argtypes = [C.c_int, C.c_uint64, u64p, u64p]
        lib.sim_wave_malloc.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_uint32), u64p]
        lib.sim_wave_free.argtypes = [C.c_void_p, C.c_int, u64p]
        lib.sim_get_state.argtypes = [C.c_void_p, C.POINTER(SimState)]
        lib.sim_chunk_info.restype = C.c_int
        lib.sim_chunk_info.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_int32), u64p]
        lib.sim_class_of.restype = C.c_int
        lib.sim_chunk_queue_pos.argtypes = [C.c_void_p, C.c_uint32]
        lib.sim_chunk_queue_pos.restype = C.c_uint64
        lib.sim_set_origin.argtypes = [C.c_void_p, C.c_uint32]
        lib.sim_set_origin.restype = C.c_int
        lib.sim_set_agg.argtypes = [C.c_void_p, C.c_int]
        lib.sim_set_agg.restype = C.c_int
        lib.sim_set_recycling.argtypes = [C.c_void_p, C.c_uint32]
        lib.sim_set_recycling.restype = C.c_int
        lib.sim_preload.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        lib.sim_preload.restype = C.c_uint32
        lib.sim_queue_window.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32]
        lib.sim_queue_window.restype = C.c_uint32
        lib.sim_class_of.argtypes = [C.c_uint32, C.c_uint32]
        lib.sim_set_block_order.argtypes = [C.c_void_p, C.c_int]
        lib.sim_sem_make.restype = C.c_uint64
        lib.sim_sem_make.argtypes = [C.c_int64, C.c_uint64]
        lib.sim_sem_count.restype = C.c_int64
        lib.sim_sem_count.argtypes = [C.c_uint64]
        lib.sim_sem_expected.restype = C.c_uint64
        lib.sim_sem_expected.argtypes = [C.c_uint64]
        lib.sim_sem_share.restype = C.c_uint32
        lib.sim_sem_share.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64]
        lib.sim_locate.restype = C.c_uint32
        lib.sim_locate.argtypes = [C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32]
        _sim = lib
    return _sim


def layout(variant: str, pool_bytes: int) -> tuple[int, int]:
    mc, asz = C.c_uint64(), C.c_uint64()
    _load_sim().sim_layout(VARIANTS[variant], pool_bytes, C.byref(mc), C.byref(asz))
    return mc.value, asz.value


def class_of(size: int, smallest: int = 16) -> int:
    return _load_sim().sim_class_of(size, smallest)


def locate(bits: np.ndarray, idx: int) -> int:
    bits = np.ascontiguousarray(bits, dtype=np.uint32)
    return _load_sim().sim_locate(bits.ctypes.data_as(C.POINTER(C.c_uint32)), len(bits), idx)


def sem_make(count: int, expected: int) -> int:
    return _load_sim().sim_sem_make(count, expected)


def sem_fields(v: int) -> tuple[int, int]:
    lib = _load_sim()
    return lib.sim_sem_count(v), lib.sim_sem_expected(v)


def sem_share(count: int, expected: int, num: int, ppc: int) -> int:
    """Chunks a claim of `num` pages carves when the class word holds (count, expected)."""
    return _load_sim().sim_sem_share(count, expected, num, ppc)


class Sim:
    """One simulated pool (MultiOuro<variant>)."""

    def __init__(self, variant: str, pool_bytes: int, origin: int = 0, agg: int = 64):
        self._lib = _load_sim()
        self.kind = variant
        self._h = self._lib.sim_create(VARIANTS[variant], pool_bytes)
        if not self._h:
            raise ValueError(variant)
        if origin and self._lib.sim_set_origin(self._h, origin) != 0:
            raise ValueError(f"origin {origin:#x}")
        if self._lib.sim_set_agg(self._h, agg) != 0:
            raise ValueError(f"aggregation width {agg}")

    def set_recycling(self, min_queued: int | None = 0) -> None:
        """Chunk recycling (chunk-queue kinds): None switches it off."""
        v = 0xFFFFFFFF if min_queued is None else int(min_queued)
        if self._lib.sim_set_recycling(self._h, v) != 0:
            raise ValueError("recycling applies to chunk-queue kinds only")

    def close(self) -> None:
        if self._h:
            self._lib.sim_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def wave_malloc(self, sizes, block_tile: bool = False) -> np.ndarray:
        """One wave's malloc (<= 64 lanes), or with block_tile (block order only) one block-wide call
        of a 64-thread workgroup over up to 4 requests per thread."""
        sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        assert len(sizes) <= (256 if block_tile else 64)
        out = np.empty(len(sizes), dtype=np.uint64)
        self._lib.sim_wave_malloc(self._h, len(sizes), sizes.ctypes.data_as(C.POINTER(C.c_uint32)),
                                  out.ctypes.data_as(C.POINTER(C.c_uint64)))
        return out

    def wave_free(self, offsets) -> None:
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        assert len(offsets) <= 64
        self._lib.sim_wave_free(self._h, len(offsets), offsets.ctypes.data_as(C.POINTER(C.c_uint64)))

    def preload(self, k: int, available: int) -> int:
        """enqueueInitialChunk on a chunk taken from the pool; returns the chunk index."""
        return self._lib.sim_preload(self._h, k, available)

    def state(self) -> SimState:
        st = SimState()
        self._lib.sim_get_state(self._h, C.byref(st))
        return st

    def chunk(self, c: int):
        ps, cnt = C.c_uint32(), C.c_int32()
        mask = np.zeros(8, dtype=np.uint64)
        t = self._lib.sim_chunk_info(self._h, c, C.byref(ps), C.byref(cnt), mask.ctypes.data_as(C.POINTER(C.c_uint64)))
        return t, ps.value, cnt.value, mask

    def queue_pos(self, c: int) -> int:
        return self._lib.sim_chunk_queue_pos(self._h, c)

    def window(self, k: int) -> np.ndarray:
        """Slot values of class k's queue window [front, back)."""
        n = self._lib.sim_queue_window(self._h, k, None, 0)
        out = np.empty(max(n, 1), dtype=np.uint32)
        self._lib.sim_queue_window(self._h, k, out.ctypes.data_as(C.POINTER(C.c_uint32)), n)
        return out[:n]

    def block_trace(self, sizes, rounds: int, tile: int = 1) -> np.ndarray:
        """The batched form run by one 64-thread persistent workgroup: each round allocates the flat
        request list tile by tile (64 * tile requests, block-wide calls), then frees it the same way
        (a block-wide free of a tile releases its sub-tiles of 64 in order, which is a wave free
        each). Returns offsets [rounds, requests]."""
        sizes = np.asarray(sizes, dtype=np.uint32)
        w = 64 * tile
        self._lib.sim_set_block_order(self._h, 1)
        try:
            out = []
            for _ in range(rounds):
                offs = np.concatenate([self.wave_malloc(sizes[t:t + w], block_tile=True)
                                       for t in range(0, len(sizes), w)])
                for t in range(0, len(sizes), 64):
                    self.wave_free(offs[t:t + 64])
                out.append(offs)
        finally:
            self._lib.sim_set_block_order(self._h, 0)
        return np.stack(out)

    def trace(self, sizes, rounds: int) -> np.ndarray:
        """Single-wave trace: `rounds` x (wave malloc, wave free) over <= 64 lanes.
        sizes: [lanes] or [rounds, lanes]. Returns offsets [rounds, lanes]."""
        sizes = np.asarray(sizes, dtype=np.uint32)
        out = []
        for r in range(rounds):
            sz = sizes[r] if sizes.ndim == 2 else sizes
            offs = self.wave_malloc(sz)
            out.append(offs)
            self.wave_free(offs)
        return np.stack(out)


def glibc_trace(sizes: np.ndarray, rounds: int, threads: int) -> dict:
    """glibc malloc/free over the identical trace on `threads` host threads (the CPU baseline)."""
    global _cpu
    if _cpu is None:
        path = HERE / "libcpu_baseline.so"
        if not path.exists():
            build()
        _cpu = C.CDLL(str(path))
        _cpu.cpu_glibc_trace.restype = C.c_int
        _cpu.cpu_glibc_trace.argtypes = [C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32, C.c_int,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
    a, f, e = C.c_double(), C.c_double(), C.c_uint64()
    _cpu.cpu_glibc_trace(sizes.ctypes.data_as(C.POINTER(C.c_uint32)), len(sizes), rounds, threads,
                         C.byref(a), C.byref(f), C.byref(e))
    return {"alloc_s": a.value, "free_s": f.value, "errors": e.value}
