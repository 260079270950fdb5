"""The allocation-free entry point (ouro_run_batch_static) against ouro_run_batch on the device.

A single 64-thread persistent workgroup replays deterministically, so the static variant must hand out
exactly the blocks ouro_run_batch hands out, call after call (the second call smaller than the first,
on the state the first left), with the same counters; a concurrent run through the static scratch
must pass the invariant checks and leave the pool quiescent."""
import numpy as np
import pytest

import ouroboros_amd as ob

pytestmark = pytest.mark.gpu

KINDS = ["PQ", "CQ", "VAPQ", "VACQ", "VLPQ", "VLCQ"]
COUNTS = ("verify_errors", "bounds_errors", "null_returns", "mallocs", "touch_bytes", "error_word", "peak_chunks")


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("tile", [1, 4])
def test_static_round_equals_run_batch(kind, tile):
    print(f"static vs run_batch {kind} tile={tile}", flush=True)
    first = (16 << (np.arange(640) * 7 % 12)).astype(np.uint32)
    second = (16 << (np.arange(200) * 5 % 12)).astype(np.uint32)
    got, want = [], []
    with ob.MemoryManager(kind, 512 << 20) as a, ob.MemoryManager(kind, 512 << 20) as b:
        for sizes in (first, second):
            oa = np.empty(len(sizes), np.uint64)
            ob_ = np.empty(len(sizes), np.uint64)
            ra = a.run_batch(sizes, rounds=1, warmup=0, agg=0, block=64, grid=1, graph=False, offsets=oa, tile=tile)
            rb = b.run_batch_static(sizes, rounds=1, warmup=0, agg=0, block=64, grid=1, offsets=ob_, tile=tile)
            assert ra["error_word"] == 0 and ra["verify_errors"] == 0 and ra["null_returns"] == 0, ra
            assert {k: rb[k] for k in COUNTS} == {k: ra[k] for k in COUNTS}
            want.append(oa)
            got.append(ob_)
        assert a.check_quiescent().violations == 0
        assert b.check_quiescent().violations == 0
    for g, w in zip(got, want):
        mism = np.flatnonzero(g != w)
        assert len(mism) == 0, f"request {mism[0]}: static {g[mism[0]]} run_batch {w[mism[0]]}"


@pytest.mark.parametrize("agg,tile", [(0, 4), (64, 1)])
def test_static_round_concurrent(agg, tile):
    print(f"static concurrent agg={agg} tile={tile}", flush=True)
    with ob.MemoryManager("VAPQ", 4 << 30) as mm:  # 1M requests hold about 1.7 GB at once
        for n in (1 << 20, 1 << 16):
            sizes = ob.gen_sizes(42, n, "uniform")
            r = mm.run_batch_static(sizes, rounds=2, warmup=1, agg=agg, block=1024 if agg == 0 else 256, check=True,
                                    tile=tile)
            assert r["verify_errors"] == 0 and r["bounds_errors"] == 0 and r["null_returns"] == 0, r
            assert r["error_word"] == 0 and r["mallocs"] == 2 * n
        assert mm.check_quiescent().violations == 0


def test_static_round_rejects_what_it_cannot_serve():
    with ob.MemoryManager("PQ", 512 << 20) as mm:
        with pytest.raises(ob.OuroError):
            mm.run_batch_static(np.full((1 << 20) + 1, 16, np.uint32), rounds=1)
