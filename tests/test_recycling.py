"""Chunk recycling of the chunk-queue kinds in the oracle (ChunkQueue::freePage's flash,
ChunkQueue_impl.cuh:163-199): closed forms for a single wave, the class-change property that only
recycling can satisfy, the fill-level threshold, and the queue / semaphore invariants under a random
workload with blocks held across waves."""
import numpy as np
import pytest

import oracle

NONE = oracle.NONE
TOMB = 0xFFFFFFFD
CHUNK_KINDS = ["CQ", "VACQ", "VLCQ"]
CLASS_PS = [16 << k for k in range(10)] + [16384, 32768]
STRIDE, META = 8320, 128


def _class_chunks(s, st, k):
    out = []
    for c in range(st.next_free_chunk):
        t, ps, cnt, _ = s.chunk(c)
        if t == 2 and ps == CLASS_PS[k]:
            out.append((c, cnt))
    return out


def _check(s):
    """Every chunk of a class with free pages is queued once at its queue_pos; the semaphore holds
    exactly the class's free pages; carved == the class's chunks; recycled chunks (type 0) are not in
    any queue and are counted by the reuse ring."""
    st = s.state()
    assert st.error == 0, st.error
    recycled = 0
    for c in range(st.next_free_chunk):
        t, ps, _, _ = s.chunk(c)
        recycled += t == 0 and s.queue_pos(c) == 0xFFFFFFFFFFFFFFFD
    for k in range(st.num_classes):
        w = s.window(k)
        live = {}
        for i, val in enumerate(w):
            if val == TOMB:
                continue
            t, ps, _, _ = s.chunk(int(val))
            assert t == 2 and ps == CLASS_PS[k], (k, i, val, t, ps)
            assert s.queue_pos(int(val)) == (st.front[k] + i) & 0xFFFFFFFF
            assert int(val) not in live
            live[int(val)] = i
        chunks = _class_chunks(s, st, k)
        assert st.carved[k] == len(chunks), (k, st.carved[k], len(chunks))
        assert st.queue_count[k] == len(live)
        assert st.sem_count[k] == sum(cnt for _, cnt in chunks), k
        for c, cnt in chunks:
            assert (cnt > 0) == (c in live), (k, c, cnt)
    # CQ has no queue chunks: its rings hold exactly the recycled data chunks
    if s.kind == "CQ":
        assert sum(st.reuse_count) == recycled
    else:
        assert sum(st.reuse_count) >= recycled
    return st


def test_single_wave_closed_form_cq():
    # round 1: 64 lanes x 8 KiB carve chunks 0..63 (one page each, queue positions 0..63); the free
    # makes each full, and with recycling on every one goes to the reuse ring in lane order. Round 2:
    # 64 lanes x 4 KiB (two pages per chunk) carve 32 chunks, which are the first 32 recycled ones,
    # lane i taking page i % 2 of chunk i // 2 (a group claim takes each chunk's lowest bits first)
    s = oracle.Sim("CQ", 512 << 20)
    s.set_recycling(0)
    o1 = s.wave_malloc(np.full(64, 8192))
    assert (o1 == np.arange(64) * STRIDE + META).all()
    s.wave_free(o1)
    st = s.state()
    assert st.reuse_count[0] == 64 and st.carved[9] == 0 and st.queue_count[9] == 0
    assert st.front[9] == st.back[9] == 64  # recycled before being queued again: no tombstone left
    o2 = s.wave_malloc(np.full(64, 4096))
    i = np.arange(64)
    assert (o2 == (i // 2) * STRIDE + META + (i % 2) * 4096).all()
    assert s.state().next_free_chunk == 64
    s.wave_free(o2)
    _check(s)
    # without recycling the same two rounds need 96 chunks
    t = oracle.Sim("CQ", 512 << 20)
    t.wave_free(t.wave_malloc(np.full(64, 8192)))
    o2b = t.wave_malloc(np.full(64, 4096))
    assert (o2b == (64 + i // 2) * STRIDE + META + (i % 2) * 4096).all()


@pytest.mark.parametrize("kind", CHUNK_KINDS)
def test_another_class_fits_only_with_recycling(kind):
    # class A (8 KiB) takes 70 % of the pool and frees everything; class B (4 KiB, two pages per chunk)
    # then needs 70 % of the pool again: without recycling A's chunks stay in A's queue and B runs out,
    # with recycling they come back through the reuse ring and B is served in full
    pool = (400 << 20) if kind == "CQ" else (24 << 20)
    mc, _ = oracle.layout(kind, pool)
    waves = int(mc * 0.7) // 64
    results = {}
    for recycle in (False, True):
        s = oracle.Sim(kind, pool)
        if recycle:
            s.set_recycling(0)
        live = [s.wave_malloc(np.full(64, 8192)) for _ in range(waves)]
        assert all((o != NONE).all() for o in live)
        for o in live:
            s.wave_free(o)
        live = [s.wave_malloc(np.full(64, 4096)) for _ in range(2 * waves)]
        served = int(sum((o != NONE).sum() for o in live))
        st = s.state()
        results[recycle] = (served, st.error)
        for o in live:
            s.wave_free(o)
        if recycle:
            _check(s)
    assert results[False][0] < 2 * waves * 64 and results[False][1] & 1  # OOM
    assert results[True] == (2 * waves * 64, 0)


@pytest.mark.parametrize("kind", CHUNK_KINDS)
@pytest.mark.parametrize("keep", [0, 5, 16, 17, 100])
def test_threshold_keeps_min_queued_chunks(kind, keep):
    # a free recycles only while the class queue holds more than `keep` chunks (the reference's
    # count_ > lower_fill_level), counting the exhausted chunk the free brings back. Ten waves of 64
    # x 1 KiB take 8 whole chunks each; each wave's free brings its 8 chunks back at once, and they
    # are all recycled when the queue plus one exceeds `keep`, else all queued. So the queue ends with
    # 8 * ceil(keep / 8) chunks (at most 80), every page of them in the semaphore
    s = oracle.Sim(kind, 512 << 20)
    s.set_recycling(keep)
    live = [s.wave_malloc(np.full(64, 1024)) for _ in range(10)]  # 8 pages per chunk: 80 chunks
    for o in live:
        s.wave_free(o)
    st = _check(s)
    k, left = 6, min(80, 8 * -(-keep // 8))
    assert st.queue_count[k] == left and st.carved[k] == left
    assert st.sem_count[k] == 8 * left


@pytest.mark.parametrize("kind", CHUNK_KINDS)
def test_random_workload_with_recycling_keeps_invariants(kind):
    # mixed classes of both instances, blocks held across waves: chunks are recycled and carved for
    # other classes (or queue chunks) again and again; the queues, semaphores and carve counts stay
    # exact, and every block handed out lies on its chunk's page grid without overlapping another
    rng = np.random.default_rng(5)
    s = oracle.Sim(kind, 512 << 20)
    s.set_recycling(2)
    live = []
    recycled_total = 0
    classes_seen = {}
    for step in range(500):
        n = int(rng.integers(1, 65))
        cls = rng.integers(0, 12, size=n) if step % 100 < 50 else rng.integers(6, 10, size=n)
        sizes = (16 << cls).astype(np.uint32)
        offs = s.wave_malloc(sizes)
        assert (offs != NONE).all()
        live.append((offs, sizes))
        if len(live) > 8 or rng.random() < 0.45:
            o, _ = live.pop(int(rng.integers(0, len(live))))
            before = sum(s.state().reuse_count)
            s.wave_free(o)
            recycled_total += max(0, sum(s.state().reuse_count) - before)
        if step % 50 == 0:
            st = _check(s)
            for c in range(st.next_free_chunk):
                t, ps, _, _ = s.chunk(c)
                if t == 2:
                    classes_seen.setdefault(c, set()).add(ps)
            iv = sorted((int(a), int(a) + int(z)) for o, sz in live for a, z in zip(o, sz))  # powers of two
            assert all(iv[i][0] >= iv[i - 1][1] for i in range(1, len(iv)))
    for o, _ in live:
        s.wave_free(o)
    _check(s)
    assert recycled_total > 0
    assert any(len(v) > 1 for v in classes_seen.values())  # some chunk served two classes


def test_page_kinds_refuse_recycling():
    for kind in ("PQ", "VAPQ", "VLPQ"):
        with pytest.raises(ValueError):
            oracle.Sim(kind, 512 << 20).set_recycling(0)
