#!/usr/bin/env python3
"""Benchmark: malloc+free Mops/s on one MI355X, 1,048,576 threads, mixed sizes.

The workload is the 1M-thread mixed-size trace: 1,048,576 requests whose sizes are uniform over the
ten power-of-two classes 16 B .. 8 KiB (SplitMix64, seed 42) on a 4 GiB pool. A "step" is one round
of it: every request is allocated, its block is written with the requester's id and read back, and
every block is freed. Mops/s = (mallocs + frees) / (alloc-kernel time + free-kernel time), both
kernels timed with HIP events on the stream they run on; the write / verify passes are timed apart
and reported as HBM GB/s.

The headline is the batched form: persistent kernels with block-wide aggregation, the round captured
once in a hipGraph and replayed, with the pool geometry, aggregation width and launch shape chosen
by the autotuner for this trace. Beside it: the plain per-thread path (the reference's own
aggregation, one kernel thread per request), the 64-lane wave-aggregated path, p50 / p99 allocation
latency, the atomic roofline, the batch workload set (256 workloads on a 32 GiB pool), the serving
form (Poisson request arrivals over a 64 GiB pool, p50 / p99 request latency and Mops/s against the
plain per-request launch path) and a glibc malloc/free baseline on the host cores.

Prints one JSON line. Only N = 1 is supported.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

import ouroboros_amd as ob  # noqa: E402
from ouroboros_amd import tune  # noqa: E402

THREADS = 1 << 20
POOL = 4 << 30
SEED = 42
HBM_PEAK_GBS = 8000.0
# bytes a malloc and a free must touch at minimum: request size in (4) + pointer out (8) +
# queue slot read and clear (4 + 4); free: pointer in (8) + chunk header word (4) + slot write (4)
ALG_BYTES_MALLOC = 20
ALG_BYTES_FREE = 16
KINDS = ["PQ", "CQ", "VAPQ", "VACQ", "VLPQ", "VLCQ"]


def log(msg: str) -> None:
    print(f"# {msg}", file=sys.stderr, flush=True)


def mops(res: dict, n: int, steps: int) -> float:
    t = (res["alloc_ms"] + res["free_ms"]) * 1e-3
    return 2.0 * n * steps / t / 1e6 if t > 0 else 0.0


def kernel_mops(res: dict, n: int) -> float:
    """Batched form: (mallocs + frees) / (alloc + free kernel time) from the kernel-only rounds (each
    persistent kernel timed by its own dispatch's events); stream-event timing when there are none."""
    kr = res.get("kernel_rounds") or 0
    if not kr:
        return res["mops"]
    t = (res["alloc_kernel_ms"] + res["free_kernel_ms"]) / kr * 1e-3
    return 2.0 * n / t / 1e6 if t > 0 else 0.0


def run_trace(variant: str, agg: int, sizes: np.ndarray, steps: int, warmup: int) -> dict:
    """The per-thread kernels (one request per kernel thread), launched one by one."""
    with ob.MemoryManager(variant, POOL) as mm:
        if warmup:
            mm.run_trace(sizes, warmup, agg=agg)
        res = dict(mm.run_trace(sizes, steps, agg=agg).result)
    res["mops"] = mops(res, len(sizes), steps)
    return res


def latency(variant: str, agg: int, sizes: np.ndarray) -> dict:
    with ob.MemoryManager(variant, POOL) as mm:
        mm.run_trace(sizes, 1, agg=agg)
        r = mm.run_trace(sizes, 3, agg=agg, latency=True).result
    return {"p50_ns": r["lat_p50_ns"], "p99_ns": r["lat_p99_ns"], "max_ns": r["lat_max_ns"], "samples": r["lat_samples"]}


# host threads one GPU's share of the box may use (its CPU share is 16 threads per GPU, whatever
# nproc reports for the whole machine)
CPU_SHARE = 16


def cpu_baseline(sizes: np.ndarray, rounds: int) -> dict:
    import oracle

    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    cores = max(1, min(CPU_SHARE, affinity))
    r = oracle.glibc_trace(sizes, rounds, cores)
    t = r["alloc_s"] + r["free_s"]
    return {
        "value": round(2.0 * len(sizes) * rounds / t / 1e6, 3) if t > 0 else None,
        "unit": "Mops/s",
        "cores": cores,
        "nproc": nproc,
        "affinity_cpus": affinity,
        "kind": "port",
        "sample": f"glibc malloc/free over the identical trace: {len(sizes):,} requests x {rounds} rounds (the "
        f"GPU's timed round count), one thread per core on the {cores} host threads of one GPU's CPU share "
        f"(nproc {nproc}), malloc and free phases timed separately (errors={r['errors']})",
    }


def batch_config4(steps: int) -> dict:
    """The batch workload set: 256 workloads, lognormal request counts, per-workload Zipf classes,
    1-4 allocations per request, 32 GiB pool; tuned, then measured on the whole set."""
    sizes, req, zs = ob.gen_batch(seed=SEED)
    pool = 32 << 30
    t0 = time.perf_counter()
    res = tune.tune(sizes, pool, sample=16, keep=0.15)
    tune_s = time.perf_counter() - t0
    c = res.best
    with ob.MemoryManager(c.variant, pool) as mm:
        r = mm.run_batch(sizes, rounds=max(2, steps // 4), warmup=1, agg=c.agg, block=c.block, grid=c.grid,
                         graph=True, check=True, tile=c.tile, kernel_timing=c.agg == 0)
    return {
        "workloads": len(req),
        "requests": int(req.sum()),
        "allocations": int(len(sizes)),
        "mean_bytes": round(float(sizes.mean()), 1),
        "pool_bytes": pool,
        "tuned": c.label(),
        "tune_seconds": round(tune_s, 1),
        "excluded": res.excluded,
        "mops": round(kernel_mops(r, len(sizes)), 2),
        "graph_replay_mops": round(r["mops"], 2),
        "alloc_ms_per_round": round(r["alloc_ms"] / max(1, r["mallocs"] // len(sizes)), 3),
        "free_ms_per_round": round(r["free_ms"] / max(1, r["mallocs"] // len(sizes)), 3),
        "touch_gbs": round(r["touch_bytes"] * (r["mallocs"] // len(sizes)) / (r["touch_ms"] * 1e-3) / 1e9, 1)
        if r["touch_ms"] else None,
        "invariants": {k: r[k] for k in ("verify_errors", "bounds_errors", "overlap_errors", "align_errors",
                                          "null_returns")} | {"device_error_bits": r["errors"]},
    }


def stress_config3(rounds: int = 4) -> dict:
    """Mixed-size skewed stress on the chunk-based virtualized-array queue: 16 GiB pool, 4,194,304
    threads, Zipf(1.2) sizes over 16 B .. 8 KiB, half of the blocks held across each round boundary;
    the live set is checked every round. Fragmentation = chunks holding a live block at the worst
    round end over the ideal packing (per class, ceil(live pages / pages per chunk))."""
    n, pool = 1 << 22, 16 << 30
    sizes = ob.gen_sizes(SEED, n, "zipf", 1.2)
    out = {}
    for agg in (1, 64):
        with ob.MemoryManager("VACQ", pool) as mm:
            r = mm.run_trace(sizes, rounds, agg=agg, check_rounds=True, hold_permille=500, hold_seed=SEED).result
            chk = mm.check_quiescent()
        out[f"agg{agg}"] = {
            "mops": round(mops(r, n, rounds), 2),
            "peak_live_chunks": r["peak_live_chunks"],
            "ideal_chunks": r["ideal_chunks"],
            "fragmentation": round(r["peak_live_chunks"] / r["ideal_chunks"], 4) if r["ideal_chunks"] else None,
            "peak_bump_chunks": r["peak_chunks"],
            "invariants": {k: r[k] for k in ("verify_errors", "bounds_errors", "overlap_errors", "align_errors",
                                              "null_returns")} | {"device_error_bits": r["errors"],
                                                                  "quiescent_violations": int(chk.violations)},
        }
    return {"kind": "VACQ", "threads": n, "pool_bytes": pool, "rounds": rounds, "zipf": 1.2, "held": 0.5,
            "runs": out}


SERVE_RATES = (2000.0, 5000.0, 10000.0, 20000.0, 50000.0)
TABLE_TILES = (1, 2, 4)


def serve_config5(kind: str, requests: int) -> dict:
    """The serving form on a 64 GiB device-resident pool: Poisson arrivals at each offered load, each
    request one kernel of 2^U[10, 18] threads allocating Zipf(1.1) sizes over 16 B .. 8 KiB, writing,
    verifying and freeing. Continuous batching (4 launch slots, each launch carrying up to 64 of the
    requests that have arrived, admission by the device free-page count) with block-wide and with
    64-lane aggregation, against the plain per-request launch path (one request per launch, one at a
    time, per-thread aggregation), on the same arrival schedule."""
    pool = 64 << 30
    rows = {}
    with ob.MemoryManager(kind, pool) as mm:
        for agg in (0, 1, 64):  # warm the code objects and the pool's classes
            mm.serve(16, rate=1e5, slots=4, agg=agg, max_batch=4)
        for rate in SERVE_RATES:
            row = {}
            for name, kw in (("batching_block", dict(slots=4, agg=0, admit=True, max_batch=64)),
                             ("batching", dict(slots=4, agg=64, admit=True, max_batch=64)),
                             ("plain", dict(slots=1, agg=1, admit=False, max_batch=1))):
                log(f"serve {kind} rate={rate:.0f}/s {name} ...")
                r = mm.serve(requests, rate=rate, zipf=1.1, min_log2=10, max_log2=18, seed=SEED, **kw)
                row[name] = {
                    "p50_ms": round(r["p50_ms"], 4), "p99_ms": round(r["p99_ms"], 4),
                    "mops": round(r["mops"], 2), "achieved_req_per_s": round(r["achieved_rate"], 1),
                    "max_in_flight": r["max_in_flight"], "deferred": r["deferred"], "launches": r["launches"],
                    "mismatches": r["mismatches"], "nulls": r["nulls"], "device_error_bits": r["errors"],
                }
            p = row["plain"]
            for name in ("batching_block", "batching"):
                b = row[name]
                row[f"p99_speedup_{name}"] = round(p["p99_ms"] / b["p99_ms"], 3) if b["p99_ms"] else None
                row[f"mops_speedup_{name}"] = round(b["mops"] / p["mops"], 3) if p["mops"] else None
            rows[f"{rate:.0f}"] = row
        chk = mm.check_quiescent()
    return {"kind": kind, "pool_bytes": pool, "requests_per_load": requests,
            "quiescent_violations": int(chk.violations), "loads": rows}


def counted_atomics(variant: str, block: int, tile: int) -> dict | None:
    """Atomic requests per round counted by the L2 (rocprofv3 TCC_ATOMIC_sum, median per dispatch of
    the allocation and the free kernel) in the committed counter pass of this configuration."""
    try:
        t = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    a = t.get("configs", {}).get(f"{variant}/b{block}/t{tile}", {}).get("atomics")
    if not a or "k_malloc_persist" not in a or "k_free_persist" not in a:
        return None
    per_kernel = {k: int(a[k].get("TCC_ATOMIC_sum", 0)) for k in ("k_malloc_persist", "k_free_persist")}
    return {"per_round": sum(per_kernel.values()), "per_kernel": per_kernel,
            "source": str(TRAFFIC_FILE.relative_to(ROOT)), "counter": "TCC_ATOMIC_sum"}


def atomic_roofline(best, kind: str, sizes: np.ndarray, head: dict, steps: int, distinct: float,
                    same_wave: float) -> dict:
    """Atomic RMWs the headline form issues per round against the measured agent-scope u32
    fetch-add rates: the L2's own count from the committed counter pass where it covers the headline
    configuration, else the count the protocol implies. The block-wide form on standard page rings
    issues four per class present in a tile (semaphore claim and front_ add on malloc, back_ add and
    signal on free) and hands ring slots over with plain loads and stores; the per-thread forms
    issue at least two per malloc+free pair (a slot or mask-bit RMW on each side)."""
    round_s = (head["alloc_ms"] + head["free_ms"]) / steps * 1e-3
    if best.agg == 0 and kind == "PQ":
        per_tile = best.block * best.tile
        q = np.maximum(np.ceil(np.log2(np.maximum(sizes, 1))).astype(np.int64) - 4, 0)
        tiles = np.arange(len(sizes)) // per_tile
        present = len(np.unique(tiles * 64 + q))  # (tile, class) pairs with requests
        issued = 4 * present
        how = "4 per (tile, class) with requests: semaphore claim + front_ add, back_ add + signal"
    else:
        issued = 2 * len(sizes)
        how = "at least 2 per malloc+free pair (slot or mask-bit RMW on each side)"
    counted = counted_atomics(best.variant, best.block, best.tile) if best.agg == 0 else None
    implied = issued
    if counted:
        issued = counted["per_round"]
        how = f"counted by the L2 ({counted['counter']}, {counted['source']}); the protocol implies {implied}: " + how
    rate = issued / round_s if round_s else None
    return {
        "atomics_per_round": int(issued),
        "atomics_per_round_implied": int(implied),
        "counted": counted,
        "atomics_per_malloc_free_pair": round(issued / len(sizes), 5),
        "achieved_atomics_per_s": rate,
        "peak_atomics_per_s": distinct,
        "frac": round(rate / distinct, 5) if rate and distinct else None,
        "per_wave_same_address_atomics_per_s": same_wave,
        "note": f"{how}; peak = measured distinct-address agent-scope u32 fetch-add-with-return rate",
    }


COPY_BYTES = 1 << 30
TRAFFIC_FILE = ROOT / "profiles" / "traffic.json"


def probe(args, sizes: np.ndarray) -> int:
    """The profiled command: one device copy of COPY_BYTES (read + write, the counters' calibration),
    then `steps` rounds of the batched form on --variant / --block / --tile (persistent kernels, grid =
    every CU at its occupancy limit), followed by the runner's kernel-only timing rounds (up to 20
    plain launches, each kernel timed by its own dispatch's events), nothing else."""
    variant = args.variant or "PQ"
    gbs = ob.bench_copy(COPY_BYTES, 1)
    if args.stress:
        # the held-block stress (4,194,304 threads, Zipf 1.2, half held across rounds, 16 GiB) on the
        # per-thread kernels at --agg (default 1), --variant (default VACQ)
        n, pool, agg = 1 << 22, 16 << 30, args.agg or 1
        variant = args.variant or "VACQ"
        zs = ob.gen_sizes(SEED, n, "zipf", 1.2)
        with ob.MemoryManager(variant, pool) as mm:
            r = mm.run_trace(zs, args.steps, agg=agg, hold_permille=500, hold_seed=SEED).result
        print(json.dumps({"probe": "stress", "variant": variant, "agg": agg, "copy_gbs": round(gbs, 1),
                          "mops": round(mops(r, n, args.steps), 1), "alloc_ms": r["alloc_ms"] / args.steps,
                          "free_ms": r["free_ms"] / args.steps}), flush=True)
        return 0
    if args.agg:
        # the per-thread kernels (k_malloc / k_free, one request per kernel thread) at this width
        with ob.MemoryManager(variant, POOL) as mm:
            mm.run_trace(sizes, max(1, args.warmup), agg=args.agg)
            r = mm.run_trace(sizes, args.steps, agg=args.agg).result
        print(json.dumps({"probe": variant, "agg": args.agg, "copy_gbs": round(gbs, 1), "mops": round(mops(r, len(sizes), args.steps), 1),
                          "alloc_ms": r["alloc_ms"] / args.steps, "free_ms": r["free_ms"] / args.steps}), flush=True)
        return 0
    with ob.MemoryManager(variant, POOL) as mm:
        r = mm.run_batch(sizes, rounds=args.steps, warmup=args.warmup, agg=0, block=args.block,
                         graph=bool(args.graph), tile=args.tile, kernel_timing=True)
        floor = mm.bench_floor(sizes, block=args.block, tile=args.tile, rounds=args.steps)
    kr = max(1, r["kernel_rounds"])
    print(json.dumps({"probe": variant, "block": args.block, "tile": args.tile, "copy_gbs": round(gbs, 1),
                      "mops": round(r["mops"], 1), "alloc_kernel_ms": r["alloc_kernel_ms"] / kr,
                      "free_kernel_ms": r["free_kernel_ms"] / kr, "kernel_rounds": r["kernel_rounds"],
                      "floor_kernel_ms": floor["kernel_ms"], "floor_grid": floor["grid"],
                      "alloc_ms": r["alloc_ms"] / args.steps, "free_ms": r["free_ms"] / args.steps}), flush=True)
    return 0


def measured_traffic(kernel: str, variant: str, block: int, tile: int) -> dict | None:
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE / WRITE_SIZE passes (each scaled
    by what the same counter reads for the calibration copy), when they cover this configuration."""
    try:
        t = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    t = t.get("configs", {}).get(f"{variant}/b{block}/t{tile}", {})
    k = t.get("kernels", {}).get(kernel)
    if not k:
        return None
    return {"bytes": k["fetch_bytes"] + k["write_bytes"], "fetch_bytes": k["fetch_bytes"],
            "write_bytes": k["write_bytes"], "source": str(TRAFFIC_FILE.relative_to(ROOT)),
            "profile": t.get("profile")}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--variant", default=None, help="skip the tuner: batched form on this variant, block-wide")
    ap.add_argument("--block", type=int, default=1024, help="workgroup size with --variant")
    ap.add_argument("--tile", type=int, default=1, help="requests per thread and call with --variant (1, 2, 4)")
    ap.add_argument("--probe", action="store_true",
                    help="profiling run: a 1 GiB device copy (counter calibration), then the headline rounds only")
    ap.add_argument("--agg", type=int, default=0, help="with --probe: 1, 16 or 64 = the per-thread kernels at that "
                    "aggregation width instead of the batched form")
    ap.add_argument("--graph", type=int, default=1, help="with --probe: 1 = graph-replayed rounds as the headline, "
                    "0 = plain launches (counter passes)")
    ap.add_argument("--stress", action="store_true", help="with --probe: the held-block stress (4M threads, Zipf "
                    "1.2, half held, 16 GiB) on the per-thread kernels at --agg (default 1), --variant (default VACQ)")
    ap.add_argument("--no-table", action="store_true", help="skip the per-kind plain / wave / batched table")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config4", action="store_true", help="skip the 256-workload batch set")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving form")
    ap.add_argument("--no-config3", action="store_true", help="skip the held-block stress run")
    ap.add_argument("--serve-requests", type=int, default=2000, help="requests per offered load and mode")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X only; --gpus {args.gpus} is not supported (the allocator "
              "shards as one independent pool per GPU)", file=sys.stderr)
        return 2
    ob.lib()  # fails loudly if libouro.so is missing
    sizes = ob.gen_sizes(SEED, THREADS, "uniform")
    steps, warmup = args.steps, args.warmup

    # ---- the batched form: tuned, then measured
    t0 = time.perf_counter()
    if args.probe:
        return probe(args, sizes)
    if args.variant:
        best = tune.Candidate(args.variant, 0, args.block, 0, args.tile)
        tuned = None
    else:
        log("tuning the batched form on the 1M-thread trace ...")
        tuned = tune.tune(sizes, POOL, sample=8)
        best = tuned.best
    tune_s = time.perf_counter() - t0
    log(f"headline {best.label()} ...")
    with ob.MemoryManager(best.variant, POOL) as mm:
        head = mm.run_batch(sizes, rounds=steps, warmup=warmup, agg=best.agg, block=best.block, grid=best.grid,
                            graph=True, check=True, tile=best.tile, kernel_timing=True)
        chk = mm.check_quiescent()
        # allocation-call latency of the same form: one more round, sampled on lane 0 of every wave
        lat_head = mm.run_batch(sizes, rounds=1, warmup=0, agg=best.agg, block=best.block, grid=best.grid,
                                graph=False, latency=True, tile=best.tile)
        # the floor under the allocation kernel: the same launch, loads, LDS rank and barriers, no allocator
        floor = mm.bench_floor(sizes, block=best.block, tile=best.tile, rounds=steps) if best.agg == 0 else None
        # the same geometry with one request per thread and call: shorter calls, more of them
        lat_t1 = None
        if best.agg == 0 and best.tile != 1:
            lat_t1 = mm.run_batch(sizes, rounds=1, warmup=1, agg=0, block=best.block, grid=best.grid,
                                  graph=False, latency=True, tile=1)
    kind = ob.variant_geometry(best.variant)["kind_name"]

    # ---- the plain per-thread path (reference aggregation) on the same queue kind, and the best of all kinds
    log(f"plain {kind} agg=1 ...")
    plain = run_trace(kind, 1, sizes, steps, warmup)
    log(f"wave-aggregated {kind} agg=64 ...")
    wave = run_trace(kind, 64, sizes, steps, warmup)

    table = {}
    table_tile = best.tile if best.agg == 0 else 1
    if not args.no_table:
        for v in KINDS:
            row = {}
            for agg in (1, 64):
                if v == kind:
                    r = plain if agg == 1 else wave
                else:
                    log(f"table {v} agg={agg} ...")
                    r = run_trace(v, agg, sizes, max(4, steps // 4), 1)
                row[f"agg{agg}"] = {"mops": round(r["mops"], 2), "errors": r["errors"], "verify_errors": r["verify_errors"]}
            # the batched form at each request tile (a tuned parameter of the form; the headline's tile
            # is tuned for PQ), reported at the fastest
            by_tile, best_b, best_t = {}, None, table_tile
            for t in TABLE_TILES:
                log(f"table {v} batched tile={t} ...")
                with ob.MemoryManager(v, POOL) as mm:
                    # like the headline: a validation round and two warm-up rounds before the timed ones
                    b = mm.run_batch(sizes, rounds=max(4, steps // 2), warmup=2, agg=0, block=1024, graph=True,
                                     check=True, tile=t, kernel_timing=True)
                by_tile[t] = round(kernel_mops(b, len(sizes)), 2)
                if best_b is None or by_tile[t] > by_tile[best_t]:
                    best_b, best_t = b, t
            row["batched_block1024"] = {"mops": by_tile[best_t], "tile": best_t, "mops_by_tile": by_tile,
                                        "graph_replay_mops": round(best_b["mops"], 2), "errors": best_b["errors"]}
            table[v] = row
    best_plain = max([plain["mops"]] + [table[v]["agg1"]["mops"] for v in table])

    lat = latency(kind, 64, sizes)
    atom_distinct = ob.bench_atomics(1 << 20, 64, "distinct")
    atom_same_wave = ob.bench_atomics(1 << 20, 16, "per_wave")
    cfg4 = None if args.no_config4 else batch_config4(steps)
    cfg5 = None if args.no_serve else serve_config5(kind, args.serve_requests)
    cfg3 = None if args.no_config3 else stress_config3()

    # ---- roofline of the dominant kernel of the headline
    # kernel-only durations (events recorded by each kernel's own dispatch, plain-launch rounds after
    # the timed ones) where the form has them; otherwise the timed rounds' per-kernel events
    if head.get("kernel_rounds"):
        alloc_avg_s = head["alloc_kernel_ms"] / head["kernel_rounds"] * 1e-3
        free_avg_s = head["free_kernel_ms"] / head["kernel_rounds"] * 1e-3
        launch_timing = (f"kernel-only: start / end events of each kernel's own dispatch (hipExtLaunchKernel), "
                         f"{head['kernel_rounds']} plain-launch rounds after the timed ones")
    else:
        alloc_avg_s = head["alloc_ms"] / steps * 1e-3
        free_avg_s = head["free_ms"] / steps * 1e-3
        launch_timing = "hipEvents around each kernel of the timed rounds"
    dominant = "alloc" if alloc_avg_s >= free_avg_s else "free"
    dom_s = max(alloc_avg_s, free_avg_s)
    alg_bytes = THREADS * (ALG_BYTES_MALLOC if dominant == "alloc" else ALG_BYTES_FREE)
    achieved_gbs = alg_bytes / dom_s / 1e9
    touch_gbs = head["touch_bytes"] / (head["touch_ms"] / steps * 1e-3) / 1e9 if head["touch_ms"] else None
    cpu = None if args.no_cpu else cpu_baseline(sizes, steps)
    kern = "k_malloc_persist" if dominant == "alloc" else "k_free_persist"
    if best.agg != 0:
        kern = "k_malloc" if dominant == "alloc" else "k_free"
    traffic = measured_traffic(kern, best.variant, best.block, best.tile) if best.agg == 0 else None

    # the headline: (mallocs + frees) / (alloc-kernel time + free-kernel time), the kernel times taken by
    # the start / end events of each kernel's own dispatch over `steps` rounds; the per-thread table
    # and the plain path time their ms-long kernels with stream events around each launch
    if head.get("kernel_rounds"):
        step_ms = (head["alloc_kernel_ms"] + head["free_kernel_ms"]) / head["kernel_rounds"]
        timing_note = (f"kernel-only: start / end events of each allocation and free kernel's own dispatch "
                       f"(hipExtLaunchKernel), {head['kernel_rounds']} rounds after the graph-replayed ones")
    else:
        step_ms = (head["alloc_ms"] + head["free_ms"]) / steps
        timing_note = "stream events around each allocation and free kernel launch"
    head_mops = 2.0 * THREADS / (step_ms * 1e-3) / 1e6 if step_ms > 0 else 0.0
    line = {
        "metric": "malloc+free Mops/s on 1 MI355X (1M threads, mixed sizes); p99 alloc latency",
        "value": round(head_mops, 3),
        "unit": "Mops/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "1,048,576 requests, sizes uniform over the 16 B..8 KiB power-of-two classes (SplitMix64 "
            "seed 42), 4 GiB pool; per step alloc / write id / verify / free; batched form: persistent "
            "block-aggregated kernels, hipGraph-replayed rounds, autotuned geometry and launch",
            "tuned": best.label(),
            "variant": best.variant,
            "aggregation": "block" if best.agg == 0 else best.agg,
            "block": best.block,
            "tile": best.tile,
            "grid": head["grid"],
            "requests": THREADS,
            "gpu_threads": head["grid"] * best.block,
            "requests_per_thread_per_call": best.tile if best.agg == 0 else 1,
            "pool_bytes": POOL,
        },
        "plain_mops": round(plain["mops"], 3),
        "plain_variant": kind,
        "speedup_vs_plain": round(head_mops / plain["mops"], 3) if plain["mops"] else None,
        "speedup_vs_best_plain": round(head_mops / best_plain, 3) if best_plain else None,
        "timing": timing_note,
        "graph_replay": {"alloc_ms_per_step": round(head["alloc_ms"] / steps, 4),
                         "free_ms_per_step": round(head["free_ms"] / steps, 4),
                         "mops": round(head["mops"], 3),
                         "note": "the same rounds replayed as three captured graphs (alloc, touch, free), "
                         "events recorded on the stream between the graph launches: each phase includes "
                         "its graph's launch"} if head.get("graph_used") else None,
        "wave64_mops": round(wave["mops"], 3),
        "p99_alloc_latency_ns": round(lat_head["lat_p99_ns"], 1),
        "p50_alloc_latency_ns": round(lat_head["lat_p50_ns"], 1),
        "latency_unit": "per allocation call of the headline form, which is also the latency of every request "
        "it serves: a call returns all of its requests' blocks at once",
        "latency_note": f"headline form ({best.label()}): wall_clock64 around each malloc_tile call on lane 0 "
        f"of every wave (one call serves {best.tile} request(s) per thread, {best.block * best.tile} per "
        f"workgroup), {lat_head['lat_samples']} samples, the whole 1M-request round in flight; the per-thread "
        "device malloc of the wave-aggregated path is p99_alloc_latency_ns_wave64",
        "p99_alloc_latency_ns_tile1": round(lat_t1["lat_p99_ns"], 1) if lat_t1 else None,
        "p50_alloc_latency_ns_tile1": round(lat_t1["lat_p50_ns"], 1) if lat_t1 else None,
        "p99_alloc_latency_ns_wave64": round(lat["p99_ns"], 1),
        "p50_alloc_latency_ns_wave64": round(lat["p50_ns"], 1),
        "latency_note_wave64": f"{kind} device malloc<64>, per-thread kernel, lane 0 of every wave, "
        f"{lat['samples']} samples",
        "touch_pass_gbs": round(touch_gbs, 1) if touch_gbs else None,
        "tune_seconds": round(tune_s, 1),
        "invariants": {
            "verify_errors": head["verify_errors"],
            "bounds_errors": head["bounds_errors"],
            "overlap_errors": head["overlap_errors"],
            "align_errors": head["align_errors"],
            "null_returns": head["null_returns"],
            "device_error_bits": head["errors"],
            "quiescent_violations": int(chk.violations),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"{kern}<{best.variant}>",
            "achieved": round(achieved_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved_gbs / HBM_PEAK_GBS, 5),
            "traffic": traffic["bytes"] if traffic else None,
            "traffic_detail": traffic,
            "avg_launch_ms": round(dom_s * 1e3, 4),
            "launch_timing": launch_timing,
            "avg_launch_ms_graph_phase": round((head["alloc_ms"] if dominant == "alloc" else head["free_ms"]) / steps, 4),
            "algorithmic_bytes_per_launch": alg_bytes,
            "floor_kernel_ms": round(floor["kernel_ms"], 4) if floor else None,
            "alloc_kernel_ms": round(alloc_avg_s * 1e3, 4),
            "free_kernel_ms": round(free_avg_s * 1e3, 4),
            "floor_note": "k_null_persist at the allocation kernel's launch shape: loads the sizes, ranks them "
            "per class in LDS with the same barriers and writes a made-up pointer per request, touching no "
            "allocator state; alloc_kernel_ms / floor_kernel_ms is how far the allocator sits above it",
        },
        "atomic_roofline": atomic_roofline(best, kind, sizes, head, steps, atom_distinct, atom_same_wave),
        "cpu_baseline": cpu,
        "stress_config3": cfg3,
        "batch_config4": cfg4,
        "serve_config5": cfg5,
        "tuner": None if tuned is None else {
            "best": tuned.best.label(),
            "excluded": tuned.excluded,
            "stage2": {t.candidate.label(): round(t.mops, 1) for t in tuned.trials if t.stage == 2},
        },
        "table": table,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
