"""Chunk recycling on the device (the flash of ChunkQueue::freePage, ChunkQueue_impl.cuh:163-199):
single-wave traces equal the oracle's replay with recycling on, a class that only fits in recycled
chunks is served, and concurrent runs of every form stay exact while chunks change class."""
import numpy as np
import pytest

import oracle
import ouroboros_amd as ob

pytestmark = pytest.mark.gpu

CHUNK_KINDS = ["CQ", "VACQ", "VLCQ"]


def _assert_clean(r):
    assert r["verify_errors"] == 0 and r["bounds_errors"] == 0, r
    assert r["overlap_errors"] == 0 and r["align_errors"] == 0, r
    assert r["error_word"] == 0, r["errors"]


def _class_switch_rounds(seed, rounds, lanes=64):
    # every round one or two classes across the wave, changing from round to round, so chunks freed
    # whole are recycled and carved for the next round's classes
    rng = np.random.default_rng(seed)
    out = np.zeros((rounds, lanes), np.uint32)
    for r in range(rounds):
        a, b = rng.integers(0, 12, size=2)
        out[r] = np.where(np.arange(lanes) < lanes // 2, 16 << a, 16 << b)
    return out


@pytest.mark.parametrize("kind", CHUNK_KINDS)
@pytest.mark.parametrize("keep", [0, 3])
def test_single_wave_with_recycling_bit_exact(kind, keep):
    sizes = _class_switch_rounds(9, 40)
    print(f"recycling single-wave {kind} keep={keep}", flush=True)
    with ob.MemoryManager(kind, 512 << 20) as mm:
        mm.set_recycling(keep)
        mm.set_statistics(True)
        out = mm.run_trace(sizes, 40, agg=64, keep_offsets=True, check_rounds=True)
        st = mm.query()
        stats = mm.statistics()
        chk = mm.check_quiescent()
    _assert_clean(out.result)
    sim = oracle.Sim(kind, 512 << 20)
    sim.set_recycling(keep)
    want = sim.trace(sizes, 40)
    ss = sim.state()
    assert ss.error == 0
    mism = np.argwhere(out.offsets != want)
    assert len(mism) == 0, f"first mismatch at (round, lane) {mism[0]}"
    assert st.next_free_chunk == ss.next_free_chunk
    assert list(st.reuse_count) == list(ss.reuse_count)
    assert stats["chunk_reuse"] > 0
    assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("kind", CHUNK_KINDS)
def test_per_lane_claims_with_recycling_bit_exact(kind):
    # W = 1 (claims reserve on {count, gen} with a CAS while recycling is on); lanes request pairwise
    # different classes, so the trace has one answer
    rng = np.random.default_rng(4)
    sizes = np.stack([(16 << rng.permutation(12)[:12]).astype(np.uint32) for _ in range(30)])
    print(f"recycling per-lane {kind}", flush=True)
    with ob.MemoryManager(kind, 512 << 20) as mm:
        mm.set_recycling(0)
        out = mm.run_trace(sizes, 30, agg=1, keep_offsets=True, check_rounds=True)
        chk = mm.check_quiescent()
    _assert_clean(out.result)
    sim = oracle.Sim(kind, 512 << 20, agg=1)
    sim.set_recycling(0)
    want = sim.trace(sizes, 30)
    mism = np.argwhere(out.offsets != want)
    assert len(mism) == 0, f"first mismatch at (round, lane) {mism[0]}"
    assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("kind", CHUNK_KINDS)
@pytest.mark.parametrize("form", ["agg1", "agg64", "block_t4"])
def test_another_class_fits_only_with_recycling(kind, form):
    # 8 KiB blocks over 70 % of the pool, all freed, then 4 KiB blocks needing 70 % of the pool again:
    # served in full with recycling, out of memory without
    pool = (400 << 20) if kind == "CQ" else (24 << 20)
    mc = oracle.layout(kind, pool)[0]
    n_a = int(mc * 0.7) // 64 * 64
    a = np.full(n_a, 8192, np.uint32)
    b = np.full(2 * n_a, 4096, np.uint32)
    res = {}
    for recycle in (False, True):
        print(f"class change {kind} {form} recycle={recycle}", flush=True)
        with ob.MemoryManager(kind, pool) as mm:
            if recycle:
                mm.set_recycling(0)
            if form == "block_t4":
                ra = mm.run_batch(a, rounds=1, warmup=0, agg=0, block=1024, graph=False, check=True, tile=4)
                rb = mm.run_batch(b, rounds=1, warmup=0, agg=0, block=1024, graph=False, check=True, tile=4)
            else:
                w = int(form[3:])
                ra = mm.run_trace(a, 1, agg=w, check_rounds=True).result
                rb = mm.run_trace(b, 1, agg=w, check_rounds=True).result
            assert ra["null_returns"] == 0
            res[recycle] = rb
            if recycle:
                _assert_clean(rb)
                chk = mm.check_quiescent()
                assert chk.violations == 0, chk.message.decode()
    assert res[False]["null_returns"] > 0 and "OOM" in res[False]["errors"]
    assert res[True]["null_returns"] == 0


@pytest.mark.parametrize("kind", CHUNK_KINDS)
@pytest.mark.parametrize("form", ["agg1", "agg64", "block_t4"])
def test_concurrent_recycling_stays_exact(kind, form):
    # a 262,144-request uniform trace with half the blocks held into the next round (per-thread forms)
    # or whole rounds (block form), recycling at a low fill level so chunks leave and come back all the
    # time: no overlap, every block reads back, and the pool ends quiescent
    sizes = ob.gen_sizes(17, 1 << 18, "uniform")
    print(f"recycling concurrent {kind} {form}", flush=True)
    with ob.MemoryManager(kind, 2 << 30) as mm:
        mm.set_recycling(4)
        mm.set_statistics(True)
        if form == "block_t4":
            r = mm.run_batch(sizes, rounds=3, warmup=1, agg=0, block=1024, graph=False, check=True, tile=4)
        else:
            r = mm.run_trace(sizes, 4, agg=int(form[3:]), check_rounds=True, hold_permille=500, hold_seed=8).result
        _assert_clean(r)
        assert r["null_returns"] == 0
        assert mm.statistics()["chunk_reuse"] > 0
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()
