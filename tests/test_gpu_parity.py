"""Device parity tests (MI355X): the HIP allocator through the C ABI (libouro.so) against the oracle.

Single-wavefront traces must equal the oracle's sequential replay bit for bit; concurrent runs must
pass the invariant checkers (no overlap, page-grid alignment, every written id read back, and a
quiescent state that holds exactly the carved pages)."""
import numpy as np
import pytest

import oracle
import ouroboros_amd as ob

pytestmark = pytest.mark.gpu

PAGE_KINDS = ["PQ", "VAPQ", "VLPQ"]
CHUNK_KINDS = ["CQ", "VACQ", "VLCQ"]
ALL = PAGE_KINDS + CHUNK_KINDS


def run_single_wave(kind, sizes, rounds, agg, pool=512 << 20):
    print(f"single-wave {kind} agg={agg} lanes={sizes.shape[-1]} rounds={rounds}", flush=True)
    with ob.MemoryManager(kind, pool) as mm:
        out = mm.run_trace(sizes, rounds, agg=agg, keep_offsets=True, check_rounds=True)
        st = mm.query()
    return out, st


def assert_clean(r):
    assert r["verify_errors"] == 0
    assert r["bounds_errors"] == 0
    assert r["overlap_errors"] == 0
    assert r["align_errors"] == 0
    assert r["error_word"] == 0, r["errors"]


CASES = [
    ("config1-16B", lambda: np.full(64, 16, np.uint32), 10),
    ("uniform-seed42", lambda: ob.gen_sizes(42, 64, "uniform"), 10),
    ("class0-40rounds", lambda: np.full(64, 16, np.uint32), 40),
    ("ragged-37-lanes", lambda: ob.gen_sizes(7, 37, "uniform"), 6),
    ("with-instance2", lambda: (16 << (np.arange(64) % 12)).astype(np.uint32), 8),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("kind,agg", [(k, 64) for k in ALL] + [(k, 1) for k in PAGE_KINDS])
def test_single_wave_bit_exact(kind, agg, case):
    name, mk, rounds = case
    sizes = mk()
    out, st = run_single_wave(kind, sizes, rounds, agg)
    assert_clean(out.result)
    sim = oracle.Sim(kind, 512 << 20)
    want = sim.trace(sizes, rounds)
    assert sim.state().error == 0
    got = out.offsets
    mism = np.argwhere(got != want)
    assert len(mism) == 0, f"first mismatch at (round, lane) {mism[0]}: device {got[tuple(mism[0])]} oracle {want[tuple(mism[0])]}"
    assert st.next_free_chunk == sim.state().next_free_chunk


def test_config1_closed_form_on_device():
    # 8,192 threads x 16 B would interleave 128 waves; the single-wave sub-trace is threads 0..63
    sizes = np.full(64, 16, np.uint32)
    out, _ = run_single_wave("PQ", sizes, 10, 1)
    want = np.array([[128 + 16 * ((64 * r + l) % 512) for l in range(64)] for r in range(10)], dtype=np.uint64)
    assert (out.offsets == want).all()


@pytest.mark.parametrize("kind", ALL)
@pytest.mark.parametrize("agg", [1, 16, 64])
def test_concurrent_mixed_invariants(kind, agg):
    print(f"concurrent {kind} agg={agg}", flush=True)
    sizes = ob.gen_sizes(42, 65536, "uniform")
    with ob.MemoryManager(kind, 1 << 30) as mm:
        out = mm.run_trace(sizes, 4, agg=agg, check_rounds=True)
        r = out.result
        assert_clean(r)
        assert r["null_returns"] == 0
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("kind", ["CQ", "VACQ", "VLCQ"])
@pytest.mark.parametrize("form", ["agg1", "agg64", "block_t4"])
def test_chunk_queue_front_keeps_up(kind, form):
    # chunk queues remove emptied chunks wherever they sit and compact front_ over the tombstones;
    # across rounds every class window back_ - front_ must stay within twice its live entries plus the
    # 64-position relocation slack (and a margin for compaction in flight). A compaction that falls
    # behind lets it grow by a fixed amount every round; a partly used chunk left at front_ did so in
    # the classes of 256 and 512 pages per chunk, by 200-400 positions a round, until the per-thread
    # claims scanned thousands of dead positions (3 -> 80 ms per 1M-request round)
    print(f"front keeps up {kind} {form}", flush=True)
    sizes = ob.gen_sizes(42, 262144, "uniform")
    excess = []
    with ob.MemoryManager(kind, 2 << 30) as mm:
        for _ in range(16):
            if form == "block_t4":
                r = mm.run_batch(sizes, rounds=1, warmup=0, agg=0, block=1024, graph=False, tile=4)
            else:
                r = mm.run_trace(sizes, 1, agg=int(form[3:])).result
            assert r["error_word"] == 0, r["errors"]
            st = mm.query()
            excess.append(max((st.back[k] - st.front[k]) % (1 << 32) - 2 * st.queue_count[k] for k in range(10)))
        assert mm.check_quiescent().violations == 0
    assert max(excess) <= 64 + 512, excess


@pytest.mark.parametrize("kind", ["CQ", "VACQ", "VLCQ"])
@pytest.mark.parametrize("agg", [1, 64])
def test_chunk_queue_window_bounded_with_held_blocks(kind, agg):
    # blocks held across rounds leave chunks with few free pages; claims that start at spread
    # positions pass some of them over, and front relocation must keep each class window within
    # twice its queued chunks plus the 64-position slack (before it, the window of the 8 KiB class
    # grew by some 50,000 positions a round)
    print(f"held-block window {kind} agg={agg}", flush=True)
    sizes = ob.gen_sizes(42, 1 << 20, "zipf", 1.2)
    with ob.MemoryManager(kind, 4 << 30) as mm:
        r = mm.run_trace(sizes, 12, agg=agg, hold_permille=500, hold_seed=42).result
        assert r["error_word"] == 0 and r["null_returns"] == 0 and r["verify_errors"] == 0, r
        st = mm.query()
        for k in range(10):
            length = (st.back[k] - st.front[k]) % (1 << 32)
            assert length <= 2 * st.queue_count[k] + 64 + 8192, (k, length, st.queue_count[k])
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("kind", ["VACQ", "CQ", "VAPQ"])
def test_held_blocks_across_rounds(kind):
    # skewed sizes (Zipf 1.2 over the class ranks) and half of the blocks held across the round
    # boundary, so live blocks interleave with frees; the live set is checked every round and the
    # chunks holding a live block never drop below the ideal packing
    print(f"held {kind}", flush=True)
    sizes = ob.gen_sizes(7, 65536, "zipf", 1.2)
    with ob.MemoryManager(kind, 1 << 30) as mm:
        out = mm.run_trace(sizes, 5, agg=64, check_rounds=True, hold_permille=500, hold_seed=3)
        r = out.result
        assert_clean(r)
        assert r["null_returns"] == 0
        assert r["frees"] == r["mallocs"]
        assert 0 < r["ideal_chunks"] <= r["peak_live_chunks"] <= r["peak_chunks"]
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()


def test_config1_full_concurrent():
    # config 1: MultiOuroPQ, 512 MiB, 8,192 threads x 16 B, 10 rounds, plain path
    with ob.MemoryManager("PQ", 512 << 20) as mm:
        out = mm.run_trace(np.full(8192, 16, np.uint32), 10, agg=1, check_rounds=True)
        assert_clean(out.result)
        assert out.result["peak_chunks"] == 16  # 8,192 x 16 B = 16 chunks of 512 pages
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("kind", ALL)
def test_out_of_memory_concurrent(kind):
    # 16,384 threads x 8 KiB against a pool of a few thousand chunks: the pool runs out mid-round.
    # Requests beyond it get nullptr (no fault, no time-out), the OOM bit is set, the blocks handed
    # out never overlap and read back intact, and once everything is freed the pool is quiescent.
    print(f"oom {kind}", flush=True)
    pool = (400 << 20) if kind in ("PQ", "CQ") else (24 << 20)
    mc, _ = oracle.layout(kind, pool)
    n = 16384
    with ob.MemoryManager(kind, pool) as mm:
        out = mm.run_trace(np.full(n, 8192, np.uint32), 2, agg=64, check_rounds=True)
        r = out.result
        assert r["verify_errors"] == 0 and r["bounds_errors"] == 0
        assert r["overlap_errors"] == 0 and r["align_errors"] == 0
        assert "OOM" in r["errors"] and "TIMEOUT" not in r["errors"], r["errors"]
        served = 2 * n - r["null_returns"]
        assert 0 < served <= 2 * (mc - 1)
        chk = mm.check_quiescent()
        # the only violation allowed is the sticky error word itself
        assert chk.violations == 1 and "error word" in chk.message.decode(), chk.message.decode()


@pytest.mark.parametrize("kind", ALL)
@pytest.mark.parametrize("tile", [1, 4])
def test_out_of_memory_block_wide(kind, tile):
    # the batched form running out of pool mid-round: block-wide claims the pool cannot supply end in
    # nullptr (page queues: null page entries keep every promised position filled; chunk queues: the
    # claim gives its unserved pages back), never in a time-out, and what is handed out never overlaps
    print(f"oom block {kind} tile={tile}", flush=True)
    pool = (400 << 20) if kind in ("PQ", "CQ") else (24 << 20)
    mc, _ = oracle.layout(kind, pool)
    n = 16384
    with ob.MemoryManager(kind, pool) as mm:
        r = mm.run_batch(np.full(n, 8192, np.uint32), rounds=2, warmup=0, agg=0, block=1024, graph=False,
                         check=True, tile=tile)
        assert r["verify_errors"] == 0 and r["bounds_errors"] == 0
        assert r["overlap_errors"] == 0 and r["align_errors"] == 0
        assert "OOM" in r["errors"] and "TIMEOUT" not in r["errors"], r["errors"]
        served = 2 * n - r["null_returns"]
        assert 0 < served <= 2 * (mc - 1)
        chk = mm.check_quiescent()
        assert chk.violations == 1 and "error word" in chk.message.decode(), chk.message.decode()


WRAP = [
    # (name, sizes, rounds, origin): the class queues' 32-bit positions cross 2^32 during the trace
    ("class9", lambda: np.full(64, 8192, np.uint32), 40, (1 << 32) - 2048),
    ("uniform", lambda: (16 << np.random.default_rng(42).integers(0, 10, size=(400, 64))).astype(np.uint32), 400,
     (1 << 32) - 2048),
    ("with-instance2", lambda: (16 << (np.arange(64) % 12)).astype(np.uint32), 500, (1 << 32) - 2048),
]


@pytest.mark.parametrize("case", WRAP, ids=[c[0] for c in WRAP])
@pytest.mark.parametrize("kind,agg", [(k, 64) for k in ALL] + [("VLPQ", 1)])
def test_position_wrap_single_wave(kind, agg, case):
    # started just below 2^32 (ouro_set_origin), the device must hand out exactly the blocks the
    # oracle hands out from position 0: linked queues locate queue chunks across the wrap, chunk
    # queues take positions 0xFFFFFFFE / 0xFFFFFFFF without confusing them with queue_pos sentinels
    name, mk, rounds, origin = case
    sizes = mk()
    print(f"wrap {kind} agg={agg} {name} origin={origin:#x}", flush=True)
    with ob.MemoryManager(kind, 512 << 20, origin=origin) as mm:
        out = mm.run_trace(sizes, rounds, agg=agg, keep_offsets=True, check_rounds=True)
        st = mm.query()
        chk = mm.check_quiescent()
    assert_clean(out.result)
    want = oracle.Sim(kind, 512 << 20).trace(sizes, rounds)
    mism = np.argwhere(out.offsets != want)
    assert len(mism) == 0, f"first mismatch at (round, lane) {mism[0]}"
    assert any(st.back[k] < origin for k in range(st.num_classes)), "no class crossed the wrap"
    assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("kind", ALL)
@pytest.mark.parametrize("form", ["agg1", "agg64", "block_t4"])
def test_position_wrap_concurrent(kind, form):
    # 65,536 mixed requests per round from positions just below 2^32: every class crosses the wrap in
    # the first rounds, under full concurrency, with clean invariants and a quiescent pool afterwards
    print(f"wrap concurrent {kind} {form}", flush=True)
    sizes = ob.gen_sizes(42, 65536, "uniform")
    origin = (1 << 32) - 4096
    with ob.MemoryManager(kind, 1 << 30, origin=origin) as mm:
        if form == "block_t4":
            r = mm.run_batch(sizes, rounds=3, warmup=1, agg=0, block=1024, graph=False, check=True, tile=4)
        else:
            r = mm.run_trace(sizes, 4, agg=int(form[3:]), check_rounds=True).result
        assert_clean(r)
        assert r["null_returns"] == 0
        st = mm.query()
        # page queues move every position through the queue; chunk queues only move chunks that empty,
        # so there the classes of at most four pages per chunk are the ones that cross
        crossing = range(10) if kind in PAGE_KINDS else range(7, 10)
        assert all(st.back[k] < origin for k in crossing), "a class did not cross the wrap"
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()


def _distinct_class_rounds(seed, rounds, lanes=12):
    # a wave of `lanes` threads; every round a new assignment of pairwise different classes of both
    # instances to the lanes (the harness has no idle lane: a size of 0 is a request for 16 bytes)
    rng = np.random.default_rng(seed)
    return np.stack([(16 << rng.permutation(12)[:lanes]).astype(np.uint32) for _ in range(rounds)])


def _segment_class_rounds(seed, rounds):
    # every round: all 64 lanes, one class per 16-lane segment, the four classes pairwise different
    rng = np.random.default_rng(seed)
    sizes = np.zeros((rounds, 64), np.uint32)
    for r in range(rounds):
        cls = rng.permutation(12)[:4]
        sizes[r] = np.repeat(16 << cls, 16).astype(np.uint32)
    return sizes


@pytest.mark.parametrize("kind", CHUNK_KINDS)
@pytest.mark.parametrize("seed,lanes", [(1, 12), (2, 10)])
def test_per_lane_chunk_claims_bit_exact(kind, seed, lanes):
    # the plain path of the chunk kinds (W = 1: every lane walks its class queue and claims its page
    # bit on its own, the reference's per-thread ChunkAccess::allocPage): with pairwise different
    # classes in a wave the claims touch disjoint queues, so the trace has one answer, the oracle's
    sizes = _distinct_class_rounds(seed, 30, lanes)
    print(f"per-lane chunk claims {kind} seed={seed} lanes={lanes}", flush=True)
    with ob.MemoryManager(kind, 512 << 20) as mm:
        out = mm.run_trace(sizes, 30, agg=1, keep_offsets=True, check_rounds=True)
        chk = mm.check_quiescent()
    assert_clean(out.result)
    want = oracle.Sim(kind, 512 << 20, agg=1).trace(sizes, 30)
    mism = np.argwhere(out.offsets != want)
    assert len(mism) == 0, f"first mismatch at (round, lane) {mism[0]}"
    assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("kind", ALL)
def test_width16_segments_bit_exact(kind):
    # W = 16: groups inside 16-lane segments. One class per segment, pairwise different, gives groups
    # that touch disjoint class queues, so the trace is the oracle's group replay
    sizes = _segment_class_rounds(3, 24)
    print(f"width-16 segments {kind}", flush=True)
    with ob.MemoryManager(kind, 512 << 20) as mm:
        out = mm.run_trace(sizes, 24, agg=16, keep_offsets=True, check_rounds=True)
        chk = mm.check_quiescent()
    assert_clean(out.result)
    want = oracle.Sim(kind, 512 << 20).trace(sizes, 24)
    mism = np.argwhere(out.offsets != want)
    assert len(mism) == 0, f"first mismatch at (round, lane) {mism[0]}"
    assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("kind", ALL)
def test_snapshot_restore_continues_bit_exact(kind):
    # a snapshot taken after round k of a single-wave trace, restored into a fresh pool, continues with
    # exactly the blocks the uninterrupted run hands out (the oracle's rounds k..R); a snapshot of
    # another variant or pool size is refused
    sizes = (16 << (np.arange(64) % 12)).astype(np.uint32)
    rounds, k = 12, 5
    print(f"snapshot/restore {kind}", flush=True)
    want = oracle.Sim(kind, 512 << 20).trace(sizes, rounds)
    with ob.MemoryManager(kind, 512 << 20) as a:
        first = a.run_trace(sizes, k, agg=64, keep_offsets=True).offsets
        snap = a.snapshot()
    assert np.array_equal(first, want[:k])
    with ob.MemoryManager(kind, 512 << 20) as b:
        b.restore(snap)
        rest = b.run_trace(sizes, rounds - k, agg=64, keep_offsets=True, check_rounds=True)
        assert_clean(rest.result)
        mism = np.argwhere(rest.offsets != want[k:])
        assert len(mism) == 0, f"first mismatch at (round, lane) {mism[0]}"
        assert b.check_quiescent().violations == 0
    other = "VACQ" if kind != "VACQ" else "PQ"
    with ob.MemoryManager(other, 512 << 20) as c, pytest.raises(ob.OuroError):
        c.restore(snap)
    with ob.MemoryManager(kind, 640 << 20) as d, pytest.raises(ob.OuroError):
        d.restore(snap)


def _free_host_pointers(mm, ptrs):
    # hand blocks back through ouro_free_batch from a device copy of their addresses (HIP runtime)
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    d = C.c_void_p()
    arr = np.ascontiguousarray(np.asarray(ptrs, np.uint64))
    assert hip.hipMalloc(C.byref(d), C.c_size_t(max(8, arr.nbytes))) == 0
    try:
        assert hip.hipMemcpy(d, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0  # host -> device
        mm.free_batch(d.value, len(arr), 64)
        assert hip.hipDeviceSynchronize() == 0
    finally:
        hip.hipFree(d)


@pytest.mark.parametrize("kind", ALL)
def test_preload_matches_oracle(kind):
    # enqueueInitialChunk through ouro_preload: chunks taken from the pool for four classes (one of
    # the second instance), only their last pages free; the single-wave trace that follows must hand
    # out exactly the oracle's blocks, never one of the caller's pages, and once the caller frees its
    # pages the pool is quiescent
    pre = [(0, 100), (3, 40), (9, 0), (10, 1)]  # (class over the chain, available pages)
    sizes = (16 << (np.arange(64) % 12)).astype(np.uint32)
    sim = oracle.Sim(kind, 512 << 20)
    chunks = [sim.preload(k, a) for k, a in pre]
    want = sim.trace(sizes, 6)
    assert sim.state().error == 0
    print(f"preload {kind}", flush=True)
    with ob.MemoryManager(kind, 512 << 20) as mm:
        assert [mm.preload(k, a) for k, a in pre] == chunks
        out = mm.run_trace(sizes, 6, agg=64, keep_offsets=True, check_rounds=True)
        assert_clean(out.result)
        mism = np.argwhere(out.offsets != want)
        assert len(mism) == 0, f"first mismatch at (round, lane) {mism[0]}"
        taken = []
        for (k, a), c in zip(pre, chunks):
            ps = 16 << k
            ppc = (8192 if k < 10 else 32768) // ps
            taken += [c * 8320 + 128 + p * ps for p in range(ppc - a)]
        assert not set(taken) & set(int(o) for o in out.offsets.ravel())
        _free_host_pointers(mm, [mm.data_base() + t for t in taken])
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()
