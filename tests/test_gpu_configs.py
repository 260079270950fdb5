"""Every workload the benchmark reports, at its full size on the device (MI355X), through the C ABI,
with the size-independent checks the domain offers: every block on the page grid of its class, no two
live blocks overlapping, every written word read back, no null return where the pool has room, no
device error bit, and a quiescent pool (queues, semaphores and bitmasks holding exactly the carved
pages) once everything is freed.

- the 1M-request mixed trace (seed 42, 4 GiB): the headline batched form and the per-thread paths of
  all six queue kinds at widths 1 and 64 (src/main.cu:146-172 runs this harness loop);
- the skewed held-block stress (VACQ, 16 GiB, 4,194,304 threads, Zipf 1.2, half the blocks held
  across each round boundary, live set checked every round);
- the 256-workload batch set (32 GiB, one pinned tuner candidate, graph-replayed rounds);
- the serving form on a 64 GiB pool (2^10..2^18 threads per request, arrivals at 50k req/s): every
  request returns exactly what the host computes for it, alone and in the mix."""
import numpy as np
import pytest

import ouroboros_amd as ob

pytestmark = pytest.mark.gpu

ALL = ["PQ", "CQ", "VAPQ", "VACQ", "VLPQ", "VLCQ"]
N = 1 << 20
POOL = 4 << 30


def assert_clean(r, nulls=0):
    assert r["verify_errors"] == 0 and r["bounds_errors"] == 0, r
    assert r["overlap_errors"] == 0 and r["align_errors"] == 0, r
    assert r["null_returns"] == nulls, r
    assert r["error_word"] == 0, r["errors"]


@pytest.fixture(scope="module")
def trace_1m():
    return ob.gen_sizes(42, N, "uniform")


@pytest.mark.parametrize("variant", ["PQ", "PQ-cap1m"])
def test_headline_batched_form_at_full_size(trace_1m, variant):
    # the benchmark's headline form: persistent 1024-thread workgroups, 4 requests per thread and
    # call, graph-replayed rounds; the validation round's 1M blocks are checked on the host
    print(f"headline {variant} block 1024 tile 4, 1M requests, 4 GiB", flush=True)
    with ob.MemoryManager(variant, POOL) as mm:
        offs = np.empty(N, np.uint64)
        r = mm.run_batch(trace_1m, rounds=4, warmup=1, agg=0, block=1024, graph=True, offsets=offs, tile=4)
        assert_clean(r)
        assert r["graph_used"] == 1 and r["mallocs"] == 4 * N
        assert (offs != np.uint64(0xFFFFFFFFFFFFFFFF)).all()
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("kind", ALL)
@pytest.mark.parametrize("agg", [1, 64])
def test_per_thread_paths_at_full_size(trace_1m, kind, agg):
    # one kernel thread per request (the reference's own path at width 1), live set checked each round
    print(f"per-thread {kind} agg={agg}, 1M requests, 4 GiB", flush=True)
    with ob.MemoryManager(kind, POOL) as mm:
        r = mm.run_trace(trace_1m, 3, agg=agg, check_rounds=True).result
        assert_clean(r)
        assert r["mallocs"] == 3 * N and r["frees"] == 3 * N
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("agg", [1, 64])
def test_held_block_stress_at_full_size(agg):
    # VACQ, 16 GiB, 4,194,304 threads, Zipf(1.2) over the class ranks, half of the blocks held across
    # each round boundary; the live set (this round's blocks plus the held ones) is checked every
    # round, and the chunks holding a live block stay within 1% of the ideal packing
    n = 1 << 22
    print(f"held-block stress VACQ agg={agg}, {n} threads, 16 GiB", flush=True)
    sizes = ob.gen_sizes(42, n, "zipf", 1.2)
    with ob.MemoryManager("VACQ", 16 << 30) as mm:
        r = mm.run_trace(sizes, 4, agg=agg, check_rounds=True, hold_permille=500, hold_seed=42).result
        assert_clean(r)
        assert r["frees"] == r["mallocs"] == 4 * n
        assert 0 < r["ideal_chunks"] <= r["peak_live_chunks"] <= 1.01 * r["ideal_chunks"]
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()


def test_batch_workload_set_at_full_size():
    # 256 workloads (lognormal request counts clamped to [64, 2^20], per-workload Zipf exponents in
    # [0.8, 1.6], 1-4 allocations per request) on a 32 GiB pool, served by one pinned configuration
    # of the tuner's space (VA page queues with 16 KiB chunks and 8 classes, block-wide, 1024 threads,
    # 4 requests per thread), rounds replayed from a captured graph
    sizes, req, _ = ob.gen_batch(seed=42)
    assert len(req) == 256 and 64 <= req.min() and req.max() <= 1 << 20
    print(f"batch set: {len(sizes)} allocations, 32 GiB", flush=True)
    with ob.MemoryManager("VAPQ-c16k-q8", 32 << 30) as mm:
        r = mm.run_batch(sizes, rounds=2, warmup=1, agg=0, block=1024, graph=True, check=True, tile=4)
        assert_clean(r)
        assert r["graph_used"] == 1 and r["mallocs"] == 2 * len(sizes)
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("kind", ["PQ", "VACQ"])
def test_serving_at_full_size(kind):
    # 64 GiB device-resident pool, requests of 2^U[10, 18] threads with Zipf(1.1) sizes: at 50k req/s
    # offered (continuous batching, admission by the device free-page count, block-wide and 64-lane)
    # every request returns exactly the words the host computes for it, as it does alone
    n = 100
    kw = dict(min_log2=10, max_log2=18, zipf=1.1, classes=10, seed=42)
    cfg = ob.serve_config(n, rate=50000.0, **kw)
    want = np.array([ob.serve_expected(cfg, r) for r in range(n)], np.uint64)
    print(f"serving {kind}, 64 GiB, {n} requests", flush=True)
    with ob.MemoryManager(kind, 64 << 30) as mm:
        alone = mm.serve(n, rate=2000.0, slots=1, agg=1, admit=False, per_request=True, **kw)
        block = mm.serve(n, rate=50000.0, slots=4, agg=0, admit=True, per_request=True, max_batch=64, **kw)
        wave = mm.serve(n, rate=50000.0, slots=4, agg=64, admit=True, per_request=True, max_batch=64, **kw)
        for name, r in (("alone", alone), ("block", block), ("wave64", wave)):
            assert r["error_word"] == 0, (name, r["errors"])
            assert r["completed"] == n and r["mismatches"] == 0 and r["nulls"] == 0, (name, r)
            assert np.array_equal(r["results"][:, 3], want[:, 0]), name
            assert np.array_equal(r["results"][:, 0], want[:, 1]), name
        assert np.array_equal(alone["results"], block["results"])
        assert np.array_equal(alone["results"], wave["results"])
        assert block["max_in_flight"] > 1 and block["launches"] < n
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()
