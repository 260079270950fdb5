"""CPU tests that pin the oracle (oracle/ouro_sim.c) against facts it does not compute itself:
the survey's pool-partition table, closed forms, brute force and cross-variant invariances."""
import numpy as np
import pytest

import oracle

KINDS = ["PQ", "CQ", "VAPQ", "VACQ", "VLPQ", "VLCQ"]
NONE = oracle.NONE
STRIDE = 8320


def test_layout_matches_survey_table():
    # chunk counts per pool size, survey table restated from MemoryInitialization.cuh:156-166
    table = {
        "PQ": [16068, 467763, 2016429, 4081317, 8211093],
        "VAPQ": [64369, 516064, 2064730, 4129618, 8259394],
        "VLPQ": [64464, 516159, 2064825, 4129713, 8259489],
    }
    pools = [512 << 20, 4 << 30, 16 << 30, 32 << 30, 64 << 30]
    for v, want in table.items():
        assert [oracle.layout(v, p)[0] for p in pools] == want
    # cudaMalloc asks for slightly more than 4 GiB: the locator is added on top
    assert oracle.layout("PQ", 4 << 30)[1] == 4_295_041_920
    # chunk kinds share the page kinds' geometry
    for a, b in [("CQ", "PQ"), ("VACQ", "VAPQ"), ("VLCQ", "VLPQ")]:
        assert oracle.layout(a, 4 << 30) == oracle.layout(b, 4 << 30)


def test_size_class_closed_form_exhaustive():
    # QueueIndex<16,8192>::getQueueIndex = max(ceil(log2 size) - 4, 0), independent restatement
    for size in range(1, 32769):
        want = max((size - 1).bit_length() - 4, 0)
        assert oracle.class_of(size, 16) == want
    # instance 2 (16 KiB smallest page)
    assert oracle.class_of(16384, 16384) == 0 and oracle.class_of(16385, 16384) == 1
    assert oracle.class_of(32768, 16384) == 1


def test_semaphore_word_round_trip_and_identities():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        c = int(rng.integers(-(1 << 38), 1 << 38))
        e = int(rng.integers(0, 1 << 24))
        v = oracle.sem_make(c, e)
        assert oracle.sem_fields(v) == (c, e)
        # signal(N) and a claim of N: a plain 64-bit add or subtract moves only the count
        n = int(rng.integers(0, 1 << 20))
        assert oracle.sem_fields((v + n) % (1 << 64)) == (c + n, e)
        assert oracle.sem_fields((v - n) % (1 << 64)) == (c - n, e)
        # signalExpected(N) (BulkSemaphore_impl.cuh:305-308): count += N, expected -= N
        if e >= n:
            v2 = (v + n - (n << 40)) % (1 << 64)
            assert oracle.sem_fields(v2) == (c + n, e - n)


def _ceil_div(a, b):
    return -(-a // b)


def test_semaphore_debt_line_carves_exactly_what_claims_need():
    # claims only, from a cold class: the shares telescope to ceil(total / ppc) chunks, and each
    # share is the number of chunk boundaries its own claim crosses
    rng = np.random.default_rng(5)
    for _ in range(300):
        ppc = int(2 ** rng.integers(0, 10))
        c, e, total, chunks = 0, 0, 0, 0
        for _ in range(int(rng.integers(1, 200))):
            num = int(rng.integers(1, 65))
            k = oracle.sem_share(c, e, num, ppc)
            assert k == _ceil_div(total + num, ppc) - _ceil_div(total, ppc)
            c -= num
            e += k * ppc
            total += num
            chunks += k
        assert chunks == _ceil_div(total, ppc)


def test_semaphore_debt_line_never_leaves_a_claim_uncovered():
    # random interleavings of claims, promise publications and frees: promised-but-unpublished
    # chunks always cover the debt, so once every promise is published no claim is short of a page
    rng = np.random.default_rng(9)
    for _ in range(300):
        ppc = int(2 ** rng.integers(0, 10))
        c, e = int(rng.integers(0, 2 * ppc)), 0
        pending = []
        for _ in range(300):
            r = rng.random()
            if r < 0.5:
                num = int(rng.integers(1, 65))
                k = oracle.sem_share(c, e, num, ppc)
                assert 0 <= k <= _ceil_div(num, ppc) + 1
                c -= num
                e += k * ppc
                if k:
                    pending.append(k)
            elif r < 0.8 and pending:
                k = pending.pop(int(rng.integers(0, len(pending))))
                c += k * ppc
                e -= k * ppc
            elif r < 0.9:
                c += 1  # a free
            debt = -(c + e)
            assert sum(pending) >= _ceil_div(max(debt, 0), ppc)
        while pending:
            k = pending.pop()
            c += k * ppc
            e -= k * ppc
        assert e == 0 and c >= 0


def test_locator_brute_force():
    rng = np.random.default_rng(7)
    for trial in range(200):
        nwords = int(rng.integers(1, 6))
        bits = rng.integers(0, 2**32, size=nwords, dtype=np.uint64).astype(np.uint32)
        bits &= rng.integers(0, 2**32, size=nwords, dtype=np.uint64).astype(np.uint32)  # sparser
        if trial % 3 == 0:
            bits[:] = 0
            bits[0] = 1
        for idx in range(nwords * 32):
            want = max([i for i in range(idx + 1) if (int(bits[i >> 5]) >> (i & 31)) & 1], default=0xFFFFFFFF)
            assert oracle.locate(bits, idx) == want


def closed_form_pq(rounds, lanes, chunk):
    # PageQueue FIFO: round r, lane l gets page (64 r + l) mod 512 of the single carved chunk
    return np.array([[chunk * STRIDE + 128 + 16 * ((lanes * r + l) % 512) for l in range(lanes)]
                     for r in range(rounds)], dtype=np.uint64)


@pytest.mark.parametrize("kind,first_chunk", [("PQ", 0), ("VAPQ", 12), ("VLPQ", 12)])
def test_config1_single_wave_closed_form(kind, first_chunk):
    # 40 rounds push the virtualized queues across queue-chunk boundaries and through reclamation;
    # the FIFO page order must not depend on how the queue is stored
    s = oracle.Sim(kind, 512 << 20)
    t = s.trace(np.full(64, 16), 40)
    assert (t == closed_form_pq(40, 64, first_chunk)).all()
    st = s.state()
    assert st.error == 0
    assert st.sem_count[0] == 512 and st.carved[0] == 512


@pytest.mark.parametrize("kind", ["CQ", "VACQ", "VLCQ"])
def test_chunk_kinds_single_wave_closed_form(kind):
    # group claim takes the lowest free bits from word (size - count) / 64: with every page back
    # before the next round, lane l always receives page l of the first data chunk
    first = 0 if kind == "CQ" else 12
    s = oracle.Sim(kind, 512 << 20)
    t = s.trace(np.full(64, 16), 12)
    want = np.array([[first * STRIDE + 128 + 16 * l for l in range(64)]] * 12, dtype=np.uint64)
    assert (t == want).all()
    t_, ps, cnt, mask = s.chunk(first)
    assert (t_, ps, cnt) == (2, 16, 512) and (mask == np.uint64(0xFFFFFFFFFFFFFFFF)).all()


def test_va_first_carve_preallocates_queue_chunk_13():
    # PageQueueVA init takes queue chunks 0..11; the first carve (data chunk 12) is enqueued at
    # virtual position 0 and pre-allocates queue chunk 13 (PageQueueVA_impl.cuh:43-53)
    s = oracle.Sim("VAPQ", 512 << 20)
    off = s.wave_malloc(np.full(1, 16))
    assert off[0] == 12 * STRIDE + 128
    assert s.state().next_free_chunk == 14
    t, ps, _, _ = s.chunk(13)
    assert t == 3 and ps == 0


# MultiOuro chains: instance 1 pages 16 B .. 8 KiB on 8 KiB chunks, instance 2 pages 16 and 32 KiB on
# 32 KiB chunks spanning four base chunks (InstanceDefinitions.cuh:17-34)
CHAIN = [(16, 8192, 8192), (16384, 32768, 32768)]  # (smallest page, largest page, chunk bytes)


def run_start(off, ps, chunk):
    """Base chunk c0 whose run puts a page of size ps at byte offset `off` (offset = c0 * 8320 + 128 +
    p * ps, p < chunk / ps), or None: a block never covers its run's header or reaches past its data."""
    idx = off // STRIDE
    for j in range(chunk // 8192):
        c0 = idx - j
        if c0 < 0:
            break
        inside = off - c0 * STRIDE
        if inside >= 128 and (inside - 128) % ps == 0 and inside - 128 + ps <= chunk:
            return c0
    return None


def test_page_grid_check_rejects_misplaced_blocks():
    # the checker itself: blocks on a run's page grid pass; a block over a chunk header, off the grid,
    # or reaching past a run's data does not
    assert run_start(5 * STRIDE + 128 + 3 * 64, 64, 8192) == 5
    assert run_start(5 * STRIDE + 64, 64, 8192) is None            # inside the header
    assert run_start(5 * STRIDE + 128 + 8, 64, 8192) is None        # off the grid
    assert run_start(5 * STRIDE + 128 + 8192, 8192, 8192) is None   # the next chunk's header
    # 32 KiB runs of four base chunks: page 1 of a 16 KiB class lies in the run's third base chunk
    assert run_start(7 * STRIDE + 128 + 16384, 16384, 32768) == 7
    assert run_start(7 * STRIDE + 128 + 32768, 16384, 32768) is None  # past the run's data
    assert run_start(7 * STRIDE + 128 + 8192, 16384, 32768) is None   # not on the 16 KiB grid


def _offset_checks(offs, sizes):
    """every block on the page grid of a run of its instance, and no overlap among live blocks"""
    iv = []
    for o, sz in zip(offs, sizes):
        if o == NONE:
            continue
        ps = 16
        while ps < sz:
            ps <<= 1
        o = int(o)
        inst = [c for c in CHAIN if c[0] <= ps <= c[1]]
        assert inst and run_start(o, ps, inst[0][2]) is not None, (o, ps)
        iv.append((o, o + ps))
    iv.sort()
    for a, b in zip(iv, iv[1:]):
        assert b[0] >= a[1], "overlapping live blocks"


@pytest.mark.parametrize("kind", KINDS)
def test_sequential_waves_invariants(kind):
    # many waves replayed one after another, mixed sizes (incl. instance 2), live blocks held across
    # waves: no overlap, page-grid alignment, and a quiescent state that holds exactly the carved pages
    rng = np.random.default_rng(3)
    s = oracle.Sim(kind, 512 << 20)
    live = []
    for step in range(300):
        n = int(rng.integers(1, 65))
        sizes = (16 << rng.integers(0, 12, size=n)).astype(np.uint32)
        offs = s.wave_malloc(sizes)
        live.append((offs, sizes))
        all_offs = np.concatenate([o for o, _ in live])
        all_sizes = np.concatenate([z for _, z in live])
        _offset_checks(all_offs, all_sizes)
        if len(live) > 4 or rng.random() < 0.3:
            o, _ = live.pop(int(rng.integers(0, len(live))))
            s.wave_free(o)
    for o, _ in live:
        s.wave_free(o)
    st = s.state()
    assert st.error == 0
    for k in range(st.num_classes):
        assert st.sem_expected[k] == 0
        if kind in ("PQ", "VAPQ", "VLPQ"):
            assert st.back[k] - st.front[k] == st.sem_count[k] == st.carved[k]
        else:
            ppc = (512 >> k) if k < 10 else (2 >> (k - 10))
            assert st.sem_count[k] == st.queue_count[k] * ppc == st.carved[k] * ppc


TOMB = 0xFFFFFFFD


def _check_chunk_queues(s, st):
    """Every class chunk with free pages is queued exactly once, at the position its queue_pos names;
    exhausted chunks are not queued; removed entries are tombstones; front_ is compacted."""
    class_ps = [16 << k for k in range(10)] + [16384, 32768]
    for k in range(st.num_classes):
        w = s.window(k)
        live = {}
        for i, val in enumerate(w):
            assert val != 0xFFFFFFFF, (k, i, "unwritten slot inside the window")
            if val == TOMB:
                continue
            t, ps, _, _ = s.chunk(int(val))
            assert t == 2 and ps == class_ps[k], (k, i, val)
            assert s.queue_pos(int(val)) == (st.front[k] + i) & 0xFFFFFFFF, (k, i, val)
            assert int(val) not in live, (k, val, "queued twice")
            live[int(val)] = i
        assert len(w) == 0 or w[0] != TOMB, (k, "front_ not compacted")
        assert st.queue_count[k] == len(live)
        for c in range(st.next_free_chunk):
            t, ps, cnt, _ = s.chunk(c)
            if t != 2 or ps != class_ps[k]:
                continue
            assert (cnt > 0) == (c in live), (k, c, cnt)


@pytest.mark.parametrize("kind", ["CQ", "VACQ", "VLCQ"])
def test_chunk_queue_removal_keeps_window_exact(kind):
    # blocks held across waves exhaust chunks in the middle of the queue window; each is removed
    # where it sits and re-enqueued at back_ by the free that gives it pages again
    rng = np.random.default_rng(11)
    s = oracle.Sim(kind, 512 << 20)
    live = []
    for step in range(400):
        n = int(rng.integers(1, 65))
        sizes = (16 << rng.integers(0, 12, size=n)).astype(np.uint32)
        live.append(s.wave_malloc(sizes))
        if len(live) > 6 or rng.random() < 0.4:
            s.wave_free(live.pop(int(rng.integers(0, len(live)))))
        if step % 25 == 0:
            st = s.state()
            assert st.error == 0
            _check_chunk_queues(s, st)
    for o in live:
        s.wave_free(o)
    st = s.state()
    assert st.error == 0
    _check_chunk_queues(s, st)


def test_out_of_memory_returns_nullptr_and_sets_error():
    # a small pool, 8 KiB pages (one per chunk): the group that outruns the pool gets the pages that
    # exist (a prefix in lane order) and nullptr for the rest, later requests get nullptr. In total
    # exactly maxChunks - 1 pages are handed out, the reference's bound chunk + 1 < maxChunks
    # (Ouroboros.cuh:133); the OOM bit is set instead of using an out-of-range chunk, and after
    # everything is freed the class holds exactly its carved pages (no null entry is left behind).
    mc, _ = oracle.layout("PQ", 400 << 20)  # standard queues take 384.5 MiB of the pool
    s = oracle.Sim("PQ", 400 << 20)
    got, partial, live = 0, 0, []
    for _ in range(mc // 64 + 2):
        offs = s.wave_malloc(np.full(64, 8192))
        ok = offs != NONE
        assert (ok == (np.arange(64) < ok.sum())).all()
        partial += int(ok.any() and not ok.all())
        got += int(ok.sum())
        live.append(offs)
    for _ in range(128):
        offs = s.wave_malloc(np.full(1, 8192))
        got += int((offs != NONE).sum())
        live.append(offs)
    assert got == mc - 1 and partial <= 1
    st = s.state()
    assert st.error & 1
    for o in live:
        s.wave_free(o)
    st = s.state()
    k = 9  # the 8 KiB class
    assert st.back[k] - st.front[k] == st.sem_count[k] == st.carved[k] == mc - 1
    assert st.sem_expected[k] == 0


@pytest.mark.parametrize("kind", KINDS)
def test_out_of_memory_every_kind_recovers_after_frees(kind):
    # every queue kind: the pool runs out, requests past it get nullptr without a time-out (queue
    # chunks of virtualized queues come from their reserve), and after all frees the class holds
    # exactly its carved pages, so a second round serves exactly as many requests again
    pool = (400 << 20) if kind in ("PQ", "CQ") else (24 << 20)
    s = oracle.Sim(kind, pool)
    served = []
    for _ in range(2):
        live = [s.wave_malloc(np.full(64, 8192)) for _ in range(64)]
        served.append(int(sum((o != NONE).sum() for o in live)))
        for o in live:
            s.wave_free(o)
        st = s.state()
        assert st.error & 1 and not st.error & 8, st.error  # OOM, no TIMEOUT
        k = 9
        assert st.sem_expected[k] == 0
        if kind in ("PQ", "VAPQ", "VLPQ"):
            assert st.back[k] - st.front[k] == st.sem_count[k] == st.carved[k]
        else:
            assert st.sem_count[k] == st.queue_count[k] == st.carved[k]
    assert served[0] == served[1] > 0


GOLDEN = np.load(__import__("pathlib").Path(__file__).with_name("golden") / "single_wave.npz")


@pytest.mark.parametrize("kind", KINDS)
def test_oracle_reproduces_committed_fixtures(kind):
    cases = sorted({k.split("__")[0] for k in GOLDEN.files})
    for case in cases:
        sizes = GOLDEN[f"{case}__sizes"]
        want = GOLDEN[f"{case}__{kind}"]
        got = oracle.Sim(kind, 512 << 20).trace(sizes, want.shape[0])
        assert np.array_equal(got, want), (case, kind)


def test_trace_generator_matches_fixture_restatement():
    import ouroboros_amd as ob

    assert np.array_equal(GOLDEN["uniform_seed42__sizes"], ob.gen_sizes(42, 64, "uniform"))
    assert np.array_equal(GOLDEN["ragged_37__sizes"], ob.gen_sizes(7, 37, "uniform"))


@pytest.mark.parametrize("tile", [1, 2, 4])
@pytest.mark.parametrize("kind", KINDS)
def test_block_order_replay_invariants(kind, tile):
    # the block-wide replay (class leaders in class order; chunk queues claimed 64 positions at a
    # time by a whole wave; `tile` requests per thread and call): every round's blocks are disjoint
    # and on the page grid, the oracle raises no error, and after the last free every chunk queue
    # holds exactly its chunks with free pages
    sizes = (16 << (np.arange(1280) * 7 % 12)).astype(np.uint32)
    s = oracle.Sim(kind, 512 << 20)
    rounds = s.block_trace(sizes, 5, tile=tile)
    assert rounds.shape == (5, len(sizes))
    for r in range(5):
        _offset_checks(rounds[r], sizes)
    st = s.state()
    assert st.error == 0
    for k in range(st.num_classes):
        assert st.sem_expected[k] == 0
    if kind in ("CQ", "VACQ", "VLCQ"):
        _check_chunk_queues(s, st)


WRAP_CASES = [
    # (name, sizes, rounds, origin): every case runs its classes' 32-bit positions across 2^32
    ("class0", np.full(64, 16, np.uint32), 140, (1 << 32) - 8192),
    ("class9", np.full(64, 8192, np.uint32), 40, (1 << 32) - 2048),
    ("uniform", None, 400, (1 << 32) - 2048),
    ("with-instance2", (16 << (np.arange(64) % 12)).astype(np.uint32), 500, (1 << 32) - 2048),
]


def _wrap_sizes(case):
    name, sizes, rounds, origin = case
    if sizes is None:
        rng = np.random.default_rng(42)
        sizes = (16 << rng.integers(0, 10, size=(rounds, 64))).astype(np.uint32)
    return sizes


@pytest.mark.parametrize("case", WRAP_CASES, ids=[c[0] for c in WRAP_CASES])
@pytest.mark.parametrize("kind", KINDS)
def test_position_wrap_is_invisible(kind, case):
    # queue positions are 32-bit and wrap: started just below 2^32, every class queue crosses the
    # wrap within a few rounds (chunk queues through the positions 0xFFFFFFFE and 0xFFFFFFFF, which a
    # 32-bit queue_pos would confuse with its sentinels; linked queues through a queue chunk whose
    # virtual start is 0), and the trace must hand out exactly the blocks it hands out from 0
    name, _, rounds, origin = case
    sizes = _wrap_sizes(case)
    a = oracle.Sim(kind, 512 << 20)
    b = oracle.Sim(kind, 512 << 20, origin=origin)
    ta, tb = a.trace(sizes, rounds), b.trace(sizes, rounds)
    sa, sb = a.state(), b.state()
    assert sa.error == 0 and sb.error == 0, (sa.error, sb.error)
    assert np.array_equal(ta, tb)
    if not (name == "class0" and kind in ("CQ", "VACQ", "VLCQ")):  # one chunk stays queued there
        assert any(sb.back[k] < origin for k in range(sb.num_classes)), "no class crossed the wrap"
    for k in range(sa.num_classes):
        assert (sb.back[k] - sb.front[k]) % (1 << 32) == (sa.back[k] - sa.front[k]) % (1 << 32)
        assert sa.sem_count[k] == sb.sem_count[k] and sa.carved[k] == sb.carved[k]
    if kind in ("CQ", "VACQ", "VLCQ"):
        _check_chunk_queues(b, sb)


@pytest.mark.parametrize("kind", ["CQ", "VLCQ"])
def test_chunk_queue_window_stays_bounded_under_held_blocks(kind):
    # skewed sizes and blocks held for a random number of waves: claims start at spread positions,
    # so live chunks near front_ get passed over while the chunks behind them cycle through back_;
    # front relocation keeps the window within twice the queued chunks plus the 64-position slack
    # (without it the window grows without bound: about 8,400 dead positions after 20,000 waves)
    rng = np.random.default_rng(5)
    s = oracle.Sim(kind, 512 << 20)
    live = []
    worst = 0
    for step in range(14000):
        sizes = (16 << np.minimum(rng.zipf(1.3, size=64) - 1, 9)).astype(np.uint32)
        live.append(s.wave_malloc(sizes))
        while len(live) > 40:
            s.wave_free(live.pop(int(rng.integers(0, len(live)))))
        if step % 100 == 0:
            st = s.state()
            assert st.error == 0
            for k in range(10):
                length = (st.back[k] - st.front[k]) % (1 << 32)
                worst = max(worst, length - 2 * st.queue_count[k])
    assert 0 < worst <= 64 + 64, worst  # reaches the slack (relocation is exercised), never far past it
    for o in live:
        s.wave_free(o)
    _check_chunk_queues(s, s.state())


def _w1_first_page(lane, q):
    # the per-lane claim's first page in a fresh chunk of class q (ChunkAccess_impl.cuh:85-125 as
    # restated on 64-bit words): word (lane + 1) mod words, the first free bit at or after lane mod 64,
    # cyclically; a fresh chunk's words hold min(ppc, 64) free bits from bit 0
    ppc = 512 >> q
    words = max(1, ppc // 64)
    bits = min(ppc, 64)
    off = lane % 64
    return ((lane + 1) % words) * 64 + (off if off < bits else 0)


@pytest.mark.parametrize("kind,first", [("CQ", 0), ("VACQ", 12), ("VLCQ", 12)])
def test_per_lane_chunk_claim_closed_form(kind, first):
    # one lane alone in its wave takes the first page of a freshly carved chunk by the per-lane rule
    for lane in (0, 5, 37, 63):
        for q in (0, 2, 3, 5, 9):
            s = oracle.Sim(kind, 512 << 20, agg=1)
            sizes = np.zeros(64, np.uint32)
            sizes[lane] = 16 << q
            off = s.wave_malloc(sizes)
            want = first * STRIDE + 128 + _w1_first_page(lane, q) * (16 << q)
            assert off[lane] == want, (lane, q, off[lane], want)
            assert (off[np.arange(64) != lane] == NONE).all()


@pytest.mark.parametrize("kind", ["CQ", "VACQ", "VLCQ"])
def test_per_lane_chunk_claims_distinct_classes(kind):
    # waves whose lanes request pairwise different classes (both instances), blocks held across
    # waves: the per-lane claims keep the live set disjoint and on the page grid, and every chunk
    # queue ends holding exactly its chunks with free pages
    rng = np.random.default_rng(17)
    s = oracle.Sim(kind, 512 << 20, agg=1)
    live = []
    for step in range(400):
        sizes = np.zeros(64, np.uint32)
        lanes = rng.choice(64, size=12, replace=False)
        sizes[lanes] = (16 << rng.permutation(12)).astype(np.uint32)
        offs = s.wave_malloc(sizes)  # size 0: an idle lane of the replay
        assert (offs[sizes > 0] != NONE).all()
        live.append((offs, sizes))
        _offset_checks(np.concatenate([o for o, _ in live]), np.concatenate([z for _, z in live]))
        if len(live) > 5 or rng.random() < 0.4:
            s.wave_free(live.pop(int(rng.integers(0, len(live))))[0])
    for o, _ in live:
        s.wave_free(o)
    st = s.state()
    assert st.error == 0
    _check_chunk_queues(s, st)


@pytest.mark.parametrize("kind,first", [("PQ", 0), ("VAPQ", 12), ("VLPQ", 12)])
def test_preload_closed_form_page_kinds(kind, first):
    # enqueueInitialChunk (PageQueue_impl.cuh:40-53): the last `available` pages of the preloaded chunk
    # are the class's free pages in page order, so a wave takes them first, FIFO; the caller's pages
    # come back through free() and are handed out after them
    s = oracle.Sim(kind, 512 << 20)
    avail = 100
    assert s.preload(0, avail) == first
    st = s.state()
    assert st.sem_count[0] == avail and st.back[0] - st.front[0] == avail and st.carved[0] == 512
    offs = s.wave_malloc(np.full(64, 16))
    assert (offs == np.array([first * STRIDE + 128 + 16 * (512 - avail + l) for l in range(64)], np.uint64)).all()
    s.wave_free(offs)
    taken = np.array([first * STRIDE + 128 + 16 * p for p in range(512 - avail)], np.uint64)
    for i in range(0, len(taken), 64):
        s.wave_free(taken[i:i + 64])
    st = s.state()
    assert st.error == 0 and st.sem_count[0] == 512 and st.back[0] - st.front[0] == 512
    # FIFO: the 36 untouched preloaded pages, then the wave's own 64, then the caller's in order
    offs = s.wave_malloc(np.full(64, 16))
    want = [512 - avail + 64 + l for l in range(36)] + [512 - avail + l for l in range(28)]
    assert (offs == np.array([first * STRIDE + 128 + 16 * p for p in want], np.uint64)).all()


@pytest.mark.parametrize("kind,first", [("CQ", 0), ("VACQ", 12), ("VLCQ", 12)])
def test_preload_chunk_kinds(kind, first):
    # ChunkQueue_impl.cuh:72-78: the chunk joins the queue with `available` free pages; its mask holds
    # exactly pages [ppc - available, ppc), which the group claim takes lowest bit first
    s = oracle.Sim(kind, 512 << 20)
    assert s.preload(3, 40) == first  # class 3: 64 pages per chunk
    t, ps, cnt, mask = s.chunk(first)
    assert (t, ps, cnt) == (2, 128, 40) and int(mask[0]) == ((1 << 64) - 1) ^ ((1 << 24) - 1)
    offs = s.wave_malloc(np.full(40, 128))
    assert (offs == np.array([first * STRIDE + 128 + 128 * (24 + l) for l in range(40)], np.uint64)).all()
    st = s.state()
    assert st.error == 0 and st.sem_count[3] == 0 and st.queue_count[3] == 0


def test_oracle_constants_match_the_device():
    # the oracle restates the device's scan, compaction and relocation bounds: a bound changed on one
    # side only changes which chunk a claim finds first, which only a GPU trace that reaches the bound
    # would show. Read both sources and compare the pairs.
    import re
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    dev = (root / "include/ouro/detail/instance.hpp").read_text() + (root / "include/ouro/detail/layout.hpp").read_text()
    sim = (root / "oracle/ouro_sim.c").read_text()
    pairs = {"kPeek": "PEEK", "kCompactSteps": "COMPACT_STEPS", "kWaveCompactSteps": "WAVE_COMPACT_STEPS",
             "kRedrawAfter": "REDRAW_AFTER", "kMaxRedraws": "MAX_REDRAWS", "kRelocSlack": "RELOC_SLACK",
             "kMaxRelocate": "MAX_RELOCATE", "kOldHold": "OLD_HOLD", "kQCReserve": "QC_RESERVE",
             "kVAEntries": "VA_ENTRIES", "kReuseEntries": "REUSE_ENTRIES"}
    for d, o in pairs.items():
        dv = re.search(rf"constexpr uint32_t {d} = (\d+);", dev)
        ov = re.search(rf"#define {o} (\d+)u", sim)
        assert dv and ov, (d, o)
        assert int(dv.group(1)) == int(ov.group(1)), (d, dv.group(1), o, ov.group(1))
