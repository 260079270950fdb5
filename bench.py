#!/usr/bin/env python3
"""Benchmark: malloc+free Mops/s on one MI355X, 1,048,576 threads, mixed sizes (config 2).

A "step" is one round of the config-2 trace: 1,048,576 device threads each malloc a block of their
size class (uniform over 16 B .. 8 KiB, seed 42), write their id into every word, read it back and
free it. Mops/s = (mallocs + frees) / (alloc-kernel time + free-kernel time), the kernels timed
with HIP events on the stream they run on (inside libouro's trace harness). The touch passes
(write + verify) are timed separately and reported as HBM GB/s.

Prints one JSON line (the driver contract). Only N = 1 is supported.
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

THREADS = 1 << 20
POOL = 4 << 30
SEED = 42
HBM_PEAK_GBS = 8000.0
# bytes a malloc and a free must touch at minimum: request size in (4) + pointer out (8) +
# queue slot read and clear (4 + 4); free: pointer in (8) + chunk header word (4) + slot write (4)
ALG_BYTES_MALLOC = 20
ALG_BYTES_FREE = 16


def mops(res: dict, n: int, steps: int) -> float:
    t = (res["alloc_ms"] + res["free_ms"]) * 1e-3
    return 2.0 * n * steps / t / 1e6 if t > 0 else 0.0


def run(variant: str, agg: int, sizes: np.ndarray, steps: int, warmup: int, check: bool = False) -> dict:
    with ob.MemoryManager(variant, POOL) as mm:
        if warmup:
            mm.run_trace(sizes, warmup, agg=agg)
        t0 = time.perf_counter()
        out = mm.run_trace(sizes, steps, agg=agg)
        wall = time.perf_counter() - t0
        res = dict(out.result)
        res["wall_s"] = wall
        if check:
            chk = mm.check_quiescent()
            res["quiescent_violations"] = int(chk.violations)
            res["quiescent_message"] = chk.message.decode()
    res["mops"] = mops(res, len(sizes), steps)
    return res


def latency(variant: str, agg: int, sizes: np.ndarray) -> dict:
    with ob.MemoryManager(variant, POOL) as mm:
        mm.run_trace(sizes, 1, agg=agg)
        out = mm.run_trace(sizes, 3, agg=agg, latency=True)
    r = out.result
    return {"p50_ns": r["lat_p50_ns"], "p99_ns": r["lat_p99_ns"], "max_ns": r["lat_max_ns"], "samples": r["lat_samples"]}


def cpu_baseline(sizes: np.ndarray, rounds: int) -> dict:
    import oracle

    cores = min(16, os.cpu_count() or 1)
    r = oracle.glibc_trace(sizes, rounds, cores)
    t = r["alloc_s"] + r["free_s"]
    return {
        "value": round(2.0 * len(sizes) * rounds / t / 1e6, 3) if t > 0 else None,
        "unit": "Mops/s",
        "cores": cores,
        "kind": "port",
        "sample": f"glibc malloc/free over the identical trace: {len(sizes):,} requests x {rounds} rounds, "
        f"one thread per core, malloc and free phases timed separately (errors={r['errors']})",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--variant", default="VAPQ", choices=list(ob.VARIANTS))
    ap.add_argument("--no-table", action="store_true", help="skip the six-variant plain/aggregated table")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X only; --gpus {args.gpus} is not supported (the allocator "
              "shards as one independent pool per GPU, see DESIGN.md)", file=sys.stderr)
        return 2
    ob.lib()  # fails loudly if libouro.so is missing
    sizes = ob.gen_sizes(SEED, THREADS, "uniform")
    steps, warmup = args.steps, args.warmup

    print(f"# headline {args.variant} agg=64 ...", file=sys.stderr, flush=True)
    head = run(args.variant, 64, sizes, steps, warmup, check=True)
    print(f"# plain {args.variant} agg=1 ...", file=sys.stderr, flush=True)
    plain = run(args.variant, 1, sizes, steps, warmup)

    table = {}
    if not args.no_table:
        for v in ob.VARIANTS:
            for agg in (1, 64):
                if v == args.variant:
                    r = head if agg == 64 else plain
                else:
                    print(f"# table {v} agg={agg} ...", file=sys.stderr, flush=True)
                    r = run(v, agg, sizes, max(4, steps // 4), 1)
                table[f"{v}/agg{agg}"] = {
                    "mops": round(r["mops"], 2),
                    "alloc_ms_per_round": round(r["alloc_ms"] / max(1, r["mallocs"] // THREADS), 4),
                    "free_ms_per_round": round(r["free_ms"] / max(1, r["mallocs"] // THREADS), 4),
                    "errors": r["errors"],
                    "verify_errors": r["verify_errors"],
                }

    lat = latency(args.variant, 64, sizes)
    atom_distinct = ob.bench_atomics(1 << 20, 64, "distinct")
    atom_same_wave = ob.bench_atomics(1 << 20, 16, "per_wave")

    alloc_avg_s = head["alloc_ms"] / steps * 1e-3
    free_avg_s = head["free_ms"] / steps * 1e-3
    dominant = "alloc" if alloc_avg_s >= free_avg_s else "free"
    dom_s = max(alloc_avg_s, free_avg_s)
    alg_bytes = THREADS * (ALG_BYTES_MALLOC if dominant == "alloc" else ALG_BYTES_FREE)
    achieved_gbs = alg_bytes / dom_s / 1e9
    pairs_per_s = head["mops"] * 1e6 / 2.0
    touch_gbs = head["touch_bytes"] / ((head["write_ms"] + head["verify_ms"]) * 1e-3) / 1e9

    cpu = None if args.no_cpu else cpu_baseline(sizes, 3)

    line = {
        "metric": "malloc+free Mops/s on 1 MI355X (1M threads, mixed sizes); p99 alloc latency",
        "value": round(head["mops"], 3),
        "unit": "Mops/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round((head["alloc_ms"] + head["free_ms"]) / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"config 2: MultiOuro{args.variant}, 4 GiB pool, 1,048,576 threads, sizes uniform over "
            f"16 B..8 KiB power-of-two classes (SplitMix64 seed 42), alloc/write/verify/free per step, "
            f"64-lane wave-aggregated path",
            "variant": args.variant,
            "aggregation_width": 64,
            "threads": THREADS,
            "pool_bytes": POOL,
        },
        "plain_mops": round(plain["mops"], 3),
        "speedup_vs_plain": round(head["mops"] / plain["mops"], 3) if plain["mops"] else None,
        "p99_alloc_latency_ns": round(lat["p99_ns"], 1),
        "p50_alloc_latency_ns": round(lat["p50_ns"], 1),
        "latency_samples": lat["samples"],
        "touch_pass_gbs": round(touch_gbs, 1),
        "wall_ms_per_step": round(head["wall_s"] / steps * 1e3, 3),
        "invariants": {
            "verify_errors": head["verify_errors"],
            "bounds_errors": head["bounds_errors"],
            "null_returns": head["null_returns"],
            "device_error_bits": head["errors"],
            "quiescent_violations": head.get("quiescent_violations"),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_malloc<MultiOuro{args.variant}, 64>" if dominant == "alloc" else f"k_free<MultiOuro{args.variant}, 64>",
            "achieved": round(achieved_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved_gbs / HBM_PEAK_GBS, 5),
            "traffic": None,
            "avg_launch_ms": round(dom_s * 1e3, 4),
            "algorithmic_bytes_per_launch": alg_bytes,
        },
        "atomic_roofline": {
            "achieved_atomics_per_s": pairs_per_s * 2.0,
            "peak_atomics_per_s": atom_distinct,
            "frac": round(pairs_per_s * 2.0 / atom_distinct, 4) if atom_distinct else None,
            "per_wave_same_address_atomics_per_s": atom_same_wave,
            "note": "2 algorithmic atomic RMWs per malloc+free pair vs the measured distinct-address "
            "agent-scope u32 fetch-add-with-return rate",
        },
        "cpu_baseline": cpu,
        "table": table,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
