"""Autotuner of the batched form: for one request list it picks the pool geometry (queue kind, chunk
size, class count, queue capacity), the aggregation width and the launch shape of the persistent
kernels (workgroup size, workgroups, requests per thread) by measuring them on the device.

Search: every feasible (variant, aggregation, block, tile, grid) candidate runs a short batch on a strided
sample of the requests (same class mix, fewer of them); the best quarter runs again on the whole list
and the fastest malloc+free Mops/s wins. Infeasible geometries are dropped before any launch: a class
queue must be able to hold every page its class needs at once, and the largest page must cover the
largest request. Chunk bytes of 4096 cannot be built on the pool's 8 KiB base chunk and are not part
of the space.

The runner is injectable so the search logic is testable without a GPU.
"""
from __future__ import annotations

import hashlib
import json
import math
from dataclasses import asdict, dataclass, field
from pathlib import Path
from typing import Callable, Iterable

import numpy as np

# The geometry axis: name -> (queue kind, chunk bytes, classes, queue capacity label)
GEOMETRIES = {
    "PQ": ("PQ", 8192, 12, "2^23 slots"),
    "CQ": ("CQ", 8192, 12, "2^23 chunks"),
    "VAPQ": ("VAPQ", 8192, 12, "16384 queue chunks"),
    "VACQ": ("VACQ", 8192, 12, "16384 queue chunks"),
    "VLPQ": ("VLPQ", 8192, 12, "linked"),
    "VLCQ": ("VLCQ", 8192, 12, "linked"),
    "VAPQ-c8k-q10": ("VAPQ", 8192, 10, "16384 queue chunks"),
    "VAPQ-c8k-q8": ("VAPQ", 8192, 8, "16384 queue chunks"),
    "VAPQ-c16k-q10": ("VAPQ", 16384, 10, "16384 queue chunks"),
    "VAPQ-c16k-q8": ("VAPQ", 16384, 8, "16384 queue chunks"),
    "VAPQ-cap4k": ("VAPQ", 8192, 12, "4096 queue chunks"),
    "PQ-cap1m": ("PQ", 8192, 12, "2^20 slots"),
}
AGGREGATION = (0, 1, 16, 64)  # 0 = block-wide (persistent kernels)
BLOCKS = (256, 512, 1024)     # block-wide path; the per-wave kernels use 256
TILES = (1, 2, 4)             # block-wide path: requests per thread and block-wide call


@dataclass(frozen=True)
class Candidate:
    variant: str
    agg: int
    block: int
    grid: int = 0  # persistent workgroups; 0 = every CU at its occupancy limit
    tile: int = 1  # block-wide path: requests per thread and call

    def label(self) -> str:
        a = "block" if self.agg == 0 else f"w{self.agg}"
        g = "" if self.agg != 0 else f"/g{self.grid or 'occ'}"
        t = f"/t{self.tile}" if self.tile != 1 else ""
        return f"{self.variant}/{a}/b{self.block}{t}{g}"


@dataclass
class Trial:
    candidate: Candidate
    stage: int
    requests: int
    mops: float
    ok: bool
    note: str = ""


@dataclass
class TuneResult:
    best: Candidate
    best_mops: float
    trials: list[Trial] = field(default_factory=list)
    excluded: dict[str, str] = field(default_factory=dict)

    def to_json(self) -> dict:
        return {
            "best": asdict(self.best),
            "best_label": self.best.label(),
            "best_mops": self.best_mops,
            "excluded": self.excluded,
            "trials": [dict(asdict(t), label=t.candidate.label()) for t in self.trials],
        }


def class_pages(sizes: np.ndarray, smallest: int, classes: int) -> np.ndarray:
    """Pages each class needs when every request is live at once."""
    q = np.maximum(np.ceil(np.log2(np.maximum(sizes, 1))).astype(np.int64) - int(math.log2(smallest)), 0)
    return np.bincount(q, minlength=classes)


def feasible(geometry: dict, sizes: np.ndarray, pool_bytes: int) -> str:
    """Empty string if the geometry can serve the request list, else the reason it cannot."""
    if int(sizes.max()) > geometry["largest_page"]:
        return f"largest page {geometry['largest_page']} B < largest request {int(sizes.max())} B"
    pages = class_pages(sizes, geometry["smallest_page"], geometry["classes"])
    chunk_kind = geometry["kind"] in (1, 3, 5)
    if chunk_kind:
        ppc = np.array([max(1, geometry["chunk_bytes"] // (geometry["smallest_page"] << q))
                        for q in range(len(pages))])
        need = int(np.max(-(-pages // ppc)))
    else:
        need = int(pages.max())
    if need > geometry["page_capacity"]:
        unit = "chunks" if chunk_kind else "pages"
        return f"a class needs {need} {unit} at once, its queue holds {geometry['page_capacity']}"
    # the data region must hold the live set (pages rounded up to whole chunks per class)
    ppc = np.array([max(1, geometry["chunk_bytes"] // (geometry["smallest_page"] << q)) for q in range(len(pages))])
    chunks = int(np.sum(-(-pages // ppc)) * (geometry["chunk_bytes"] // 8192))
    if chunks * 8320 > pool_bytes * 0.95:
        return f"the live set needs {chunks} base chunks, more than the pool holds"
    return ""


def space(variants: Iterable[str], grids: Iterable[int] = (0,), tiles: Iterable[int] = TILES) -> list[Candidate]:
    out = []
    tiles = tuple(tiles)
    for v in variants:
        for agg in AGGREGATION:
            if agg == 0:
                out += [Candidate(v, 0, b, g, t) for b in BLOCKS for t in tiles for g in grids]
            else:
                out.append(Candidate(v, agg, 256))
    return out


Runner = Callable[[Candidate, np.ndarray], tuple[float, bool, str]]


def device_runner(pool_bytes: int, rounds: int = 3, warmup: int = 2) -> Runner:
    """Measure a candidate on the device: Mops/s of `rounds` graph-replayed batch rounds. A fresh
    pool first serves the list once untimed, so the candidate that happens to run first on a pool
    does not pay for carving its chunks."""
    import ouroboros_amd as ob

    pools: dict[str, object] = {}

    def run(c: Candidate, sizes: np.ndarray) -> tuple[float, bool, str]:
        mm = pools.get(c.variant)
        if mm is None:
            for old in pools.values():
                old.close()
            pools.clear()
            mm = pools[c.variant] = ob.MemoryManager(c.variant, pool_bytes)
            try:
                mm.run_batch(sizes, rounds=1, warmup=0, agg=0, block=256, graph=False)
            except RuntimeError:
                pass
        try:
            # persistent candidates are ranked by their kernels' own durations (the events of each
            # dispatch), the per-thread ones by stream events around their launches
            r = mm.run_batch(sizes, rounds=rounds, warmup=warmup, agg=c.agg, block=c.block, grid=c.grid, graph=False,
                             tile=c.tile, kernel_timing=c.agg == 0)
        except RuntimeError as e:  # e.g. a block size the persistent kernels cannot launch
            return 0.0, False, str(e)
        ok = r["error_word"] == 0 and r["null_returns"] == 0 and r["verify_errors"] == 0
        mops = r["mops"]
        if r.get("kernel_rounds"):
            t = (r["alloc_kernel_ms"] + r["free_kernel_ms"]) / r["kernel_rounds"] * 1e-3
            mops = 2.0 * len(sizes) / t / 1e6 if t > 0 else 0.0
        return mops, ok, ",".join(r["errors"])

    run.close = lambda: [p.close() for p in pools.values()]  # type: ignore[attr-defined]
    return run


def tune(sizes: np.ndarray, pool_bytes: int, variants: Iterable[str] | None = None, runner: Runner | None = None,
         sample: int = 8, keep: float = 0.25, grids: Iterable[int] = (0,), geometry=None,
         tiles: Iterable[int] = TILES) -> TuneResult:
    """Pick the fastest feasible candidate for `sizes` (see the module docstring)."""
    import ouroboros_amd as ob

    sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
    variants = list(variants or GEOMETRIES)
    geometry = geometry or ob.variant_geometry
    excluded = {}
    usable = []
    for v in variants:
        why = feasible(geometry(v), sizes, pool_bytes)
        if why:
            excluded[v] = why
        else:
            usable.append(v)
    if not usable:
        raise ValueError(f"no geometry can serve this request list: {excluded}")
    own_runner = runner is None
    runner = runner or device_runner(pool_bytes)
    trials: list[Trial] = []
    try:
        # stage 1: a strided sample keeps the class mix
        part = sizes[::sample] if sample > 1 else sizes
        cands = space(usable, grids, tiles)
        scored = []
        for c in sorted(cands, key=lambda c: c.variant):  # one pool per variant at a time
            mops, ok, note = runner(c, part)
            trials.append(Trial(c, 1, len(part), mops, ok, note))
            if ok:
                scored.append((mops, c))
        if not scored:
            raise RuntimeError("every candidate failed on the sample")
        scored.sort(key=lambda t: -t[0])
        finalists = [c for _, c in scored[: max(1, int(math.ceil(len(scored) * keep)))]]
        # stage 2: the whole request list
        best, best_mops = None, -1.0
        for c in sorted(finalists, key=lambda c: c.variant):
            mops, ok, note = runner(c, sizes)
            trials.append(Trial(c, 2, len(sizes), mops, ok, note))
            if ok and mops > best_mops:
                best, best_mops = c, mops
        if best is None:
            raise RuntimeError("every finalist failed on the whole request list")
    finally:
        if own_runner and hasattr(runner, "close"):
            runner.close()
    return TuneResult(best, best_mops, trials, excluded)


def workload_key(sizes: np.ndarray) -> str:
    """Cache key of a request list: its length and class histogram."""
    h = np.bincount(np.log2(np.maximum(sizes, 1)).astype(np.int64), minlength=16)
    return hashlib.sha1(f"{len(sizes)}:{h.tolist()}".encode()).hexdigest()[:16]


def load_cached(path: Path, sizes: np.ndarray) -> Candidate | None:
    try:
        data = json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return None
    entry = data.get(workload_key(sizes))
    return Candidate(**entry["best"]) if entry else None


def store(path: Path, sizes: np.ndarray, result: TuneResult) -> None:
    path = Path(path)
    try:
        data = json.loads(path.read_text())
    except (OSError, ValueError):
        data = {}
    data[workload_key(sizes)] = result.to_json()
    path.write_text(json.dumps(data, indent=1))
