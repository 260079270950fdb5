"""Tuner search logic and feasibility rules, on the CPU (a stand-in runner replaces the device)."""
import numpy as np

import ouroboros_amd as ob
from ouroboros_amd import tune


def test_geometries_match_the_library():
    for name, (kind, chunk, classes, _) in tune.GEOMETRIES.items():
        g = ob.variant_geometry(name)
        assert g["kind_name"] == kind and g["chunk_bytes"] == chunk and g["classes"] == classes, name
        assert g["largest_page"] >= 8192


def test_feasibility_rules():
    # 16M requests of 16 B need 16M pages in class 0 at once: the standard ring (2^23) and the
    # 4096-queue-chunk VA ring (2^23 - 4096) cannot hold them, the 16384-queue-chunk ring and the
    # linked queues can, and chunk queues count chunks (16M / 512)
    sizes = np.full(1 << 24, 16, np.uint32)
    pool = 32 << 30
    why = {v: tune.feasible(ob.variant_geometry(v), sizes, pool) for v in tune.GEOMETRIES}
    assert why["PQ"] and why["PQ-cap1m"] and why["VAPQ-cap4k"]
    assert not why["VAPQ"] and not why["VLPQ"] and not why["CQ"]
    # a 32 KiB request needs instance 2 of the chains: single-instance geometries drop out
    big = np.array([32768, 16], np.uint32)
    assert tune.feasible(ob.variant_geometry("VAPQ-c8k-q10"), big, pool)
    assert not tune.feasible(ob.variant_geometry("VAPQ"), big, pool)
    # a live set larger than the pool
    assert tune.feasible(ob.variant_geometry("VAPQ"), np.full(1 << 20, 8192, np.uint32), 4 << 30)


def test_search_keeps_the_best_quarter_and_picks_the_fastest():
    sizes = ob.gen_sizes(42, 4096, "uniform")
    seen = []

    def runner(c, part):
        seen.append((c, len(part)))
        # a made-up landscape: block-wide beats per-wave, 512-thread blocks are best, VLPQ fails
        if c.variant == "VLPQ":
            return 0.0, False, "TIMEOUT"
        base = {0: 10.0, 64: 5.0, 16: 3.0, 1: 1.0}[c.agg] + (0.5 if c.block == 512 else 0.0)
        return base + (0.25 if c.variant == "VAPQ" else 0.0) + (0.01 if len(part) == len(sizes) else 0.0), True, ""

    variants = ["PQ", "VAPQ", "VLPQ", "CQ"]
    res = tune.tune(sizes, 4 << 30, variants=variants, runner=runner, sample=8)
    assert res.best == tune.Candidate("VAPQ", 0, 512, 0)
    stage1 = [t for t in res.trials if t.stage == 1]
    stage2 = [t for t in res.trials if t.stage == 2]
    assert len(stage1) == len(tune.space(variants)) and all(t.requests == 512 for t in stage1)
    ok1 = sum(t.ok for t in stage1)
    assert len(stage2) == int(np.ceil(ok1 * 0.25)) and all(t.requests == 4096 for t in stage2)
    # runs are grouped by variant, so each pool is built once
    order = [c.variant for c, _ in seen[: len(stage1)]]
    assert order == sorted(order)


def test_cache_round_trip(tmp_path):
    sizes = ob.gen_sizes(1, 1000, "uniform")
    res = tune.TuneResult(tune.Candidate("VAPQ", 0, 256, 0), 123.0)
    p = tmp_path / "tuned.json"
    tune.store(p, sizes, res)
    assert tune.load_cached(p, sizes) == res.best
    assert tune.load_cached(p, ob.gen_sizes(2, 1000, "zipf", 1.2)) is None


def test_batch_generator_restatement():
    # the batch workload generator against a numpy restatement of its definition, small sizes
    M = (1 << 64) - 1

    def sm(x):
        x = (x + 0x9E3779B97F4A7C15) & M
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
        return x ^ (x >> 31)

    def keyed(seed, i):
        return sm(sm(seed) ^ ((i * 0xD1B54A32D192ED03) & M))

    def unit(r):
        return (r >> 11) * (1.0 / 9007199254740992.0)

    seed, W = 5, 6
    sizes, req, zs = ob.gen_batch(seed=seed, workloads=W, mu=4.0, sigma=1.0, lo=3, hi=200)
    flat = []
    for k in range(W):
        u1 = 1.0 - unit(keyed(seed ^ 0x5851F42D4C957F2D, k))
        u2 = unit(keyed(seed ^ 0x14057B7EF767814F, k))
        z = np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)
        r = int(min(200, max(3, np.floor(np.exp(4.0 + z) + 0.5))))
        s = 0.8 + 0.8 * unit(keyed(seed ^ 0x2545F4914F6CDD1D, k))
        assert req[k] == r and abs(zs[k] - s) < 1e-12
        cdf = np.cumsum([(q + 1) ** -s for q in range(10)])
        cdf /= cdf[-1]
        ws = seed ^ ((0x9E3779B97F4A7C15 * (k + 1)) & M)
        for i in range(r):
            for j in range(1 + keyed(ws, i) % 4):
                u = unit(keyed(ws ^ 0xD6E8FEB86659FD93, i * 4 + j))
                q = 0
                while q + 1 < 10 and u >= cdf[q]:
                    q += 1
                flat.append(16 << q)
    assert np.array_equal(sizes, np.array(flat, np.uint32))
