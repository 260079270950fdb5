"""The batched form on the device (MI355X): persistent kernels with block-wide aggregation and
hipGraph-replayed rounds, through the C ABI (ouro_run_batch).

A single 64-thread persistent workgroup replays deterministically, so its blocks must equal the
oracle's block-order replay bit for bit; concurrent runs must pass the invariant checkers (no overlap,
page alignment, every written id read back, a quiescent pool afterwards)."""
import numpy as np
import pytest

import oracle
import ouroboros_amd as ob

pytestmark = pytest.mark.gpu

ALL = ["PQ", "CQ", "VAPQ", "VACQ", "VLPQ", "VLCQ"]


def assert_clean(r):
    assert r["verify_errors"] == 0 and r["bounds_errors"] == 0, r
    assert r["overlap_errors"] == 0 and r["align_errors"] == 0, r
    assert r["null_returns"] == 0, r
    assert r["error_word"] == 0, r["errors"]


@pytest.mark.parametrize("tile", [1, 2, 4])
@pytest.mark.parametrize("kind", ALL)
def test_single_workgroup_matches_oracle(kind, tile):
    # 640 requests, mixed classes (both instances), served tile by tile by one 64-thread workgroup,
    # `tile` requests per thread and block-wide call (640 = 2.5 tiles of 256 at tile 4: the last one
    # is partly empty)
    sizes = (16 << (np.arange(640) * 7 % 12)).astype(np.uint32)
    print(f"single-workgroup {kind} tile={tile}", flush=True)
    sim = oracle.Sim(kind, 512 << 20)
    want = sim.block_trace(sizes, 6, tile=tile)
    assert sim.state().error == 0
    with ob.MemoryManager(kind, 512 << 20) as mm:
        for call in range(3):
            # each call runs its validation round (observed) and then one timed round
            got = np.empty(len(sizes), dtype=np.uint64)
            res = mm.run_batch(sizes, rounds=1, warmup=0, agg=0, block=64, grid=1, graph=False, offsets=got,
                               tile=tile)
            assert_clean(res)
            w = want[2 * call]
            mism = np.flatnonzero(got != w)
            assert len(mism) == 0, f"round {2 * call} request {mism[0]}: device {got[mism[0]]} oracle {w[mism[0]]}"
        assert mm.check_quiescent().violations == 0


@pytest.mark.parametrize("kind", ALL)
@pytest.mark.parametrize("block,tile", [(256, 1), (1024, 1), (256, 4), (1024, 2)])
def test_persistent_block_aggregated_invariants(kind, block, tile):
    print(f"persistent {kind} block={block} tile={tile}", flush=True)
    sizes = ob.gen_sizes(42, 65536, "uniform")
    with ob.MemoryManager(kind, 1 << 30) as mm:
        res = mm.run_batch(sizes, rounds=3, warmup=1, agg=0, block=block, graph=True, check=True, tile=tile)
        assert res["graph_used"] == 1
        assert_clean(res)
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("kind", ["VAPQ", "VLPQ"])
@pytest.mark.parametrize("block,tile", [(256, 2), (256, 4), (1024, 4)])
def test_virtual_page_queue_tiles_at_1m(kind, block, tile):
    # 1M requests on one pool after rounds at one request per thread: with several requests per
    # thread, frees into VA / VL page queues must not wait on a queue chunk that another workgroup
    # only pre-allocates later (every sub-tile's pre-allocations run before any put of the tile), so
    # no bounded wait may expire
    print(f"virtual page queue tiles {kind} block={block} tile={tile}", flush=True)
    sizes = ob.gen_sizes(42, 1 << 20, "uniform")
    with ob.MemoryManager(kind, 4 << 30) as mm:
        warm = mm.run_batch(sizes, rounds=2, warmup=1, agg=0, block=block, graph=True)
        assert_clean(warm)
        res = mm.run_batch(sizes, rounds=4, warmup=1, agg=0, block=block, graph=True, check=True, tile=tile)
        assert_clean(res)
        assert mm.check_quiescent().violations == 0


@pytest.mark.parametrize("kind", ["VAPQ", "VACQ", "VLPQ", "VLCQ", "VAPQ-c16k-q8"])
def test_virtualized_tiles_over_many_rounds(kind):
    # twelve 1M-request rounds of the tiled batched form: the tile-wide puts and takes recycle queue
    # chunks round after round (their counts and pointer-ring entries updated sub-tile by sub-tile),
    # with every round's live set checked and the queues quiescent at the end
    print(f"virtualized tiles over rounds {kind}", flush=True)
    sizes = ob.gen_sizes(7, 1 << 20, "uniform")
    with ob.MemoryManager(kind, 4 << 30) as mm:
        res = mm.run_batch(sizes, rounds=12, warmup=1, agg=0, block=1024, graph=True, check=True, tile=4)
        assert_clean(res)
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()
        st = mm.query()
        assert st.error == 0, ob.decode_errors(st.error)


def test_kernel_only_timing_rounds():
    # the runner's extra plain-launch rounds, each persistent kernel timed by its own dispatch events
    print("kernel-only timing", flush=True)
    sizes = ob.gen_sizes(11, 65536, "uniform")
    with ob.MemoryManager("PQ", 1 << 30) as mm:
        r = mm.run_batch(sizes, rounds=3, warmup=1, agg=0, block=1024, graph=True, tile=4, kernel_timing=True)
        assert_clean(r)
        assert r["kernel_rounds"] == 3 and r["alloc_kernel_ms"] > 0 and r["free_kernel_ms"] > 0
        assert mm.check_quiescent().violations == 0


@pytest.mark.parametrize("agg", [1, 16, 64])
def test_batch_runner_per_wave_widths(agg):
    print(f"batch per-wave agg={agg}", flush=True)
    sizes = ob.gen_sizes(7, 65536, "uniform")
    with ob.MemoryManager("VAPQ", 1 << 30) as mm:
        res = mm.run_batch(sizes, rounds=2, warmup=1, agg=agg, block=256, graph=True, check=True)
        assert_clean(res)
        assert mm.check_quiescent().violations == 0


@pytest.mark.parametrize("agg,tile", [(0, 1), (0, 2), (64, 1)])
def test_batch_latency_round(agg, tile):
    print(f"batch latency agg={agg} tile={tile}", flush=True)
    sizes = ob.gen_sizes(5, 65536, "uniform")
    with ob.MemoryManager("PQ", 1 << 30) as mm:
        r = mm.run_batch(sizes, rounds=1, warmup=1, agg=agg, block=1024 if agg == 0 else 256, graph=False,
                         latency=True, tile=tile)
        assert_clean(r)
        assert r["lat_samples"] == len(sizes) // 64
        assert 0 < r["lat_p50_ns"] <= r["lat_p99_ns"] <= r["lat_max_ns"] < 1e9
        assert mm.check_quiescent().violations == 0


def test_graph_replay_without_warmup():
    # a captured round with no untimed round before it (its events are first recorded inside the
    # capture unless the runner records them once beforehand)
    print("graph without warm-up", flush=True)
    sizes = ob.gen_sizes(9, 65536, "uniform")
    with ob.MemoryManager("PQ", 1 << 30) as mm:
        r = mm.run_batch(sizes, rounds=2, warmup=0, agg=0, block=1024, graph=True, tile=4)
        assert r["graph_used"] == 1 and r["mallocs"] == 2 * len(sizes)
        assert_clean(r)
        assert mm.check_quiescent().violations == 0
