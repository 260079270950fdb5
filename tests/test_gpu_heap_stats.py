"""Requests above the chain's largest page (the device heap, Ouroboros.cuh:398-408 and
Ouroboros_impl.cuh:98-104) and the statistics counters (Statistics.h:5-17) on the device."""
import numpy as np
import pytest

import ouroboros_amd as ob

pytestmark = pytest.mark.gpu


def _sizes_with_large(n, seed=3):
    sizes = ob.gen_sizes(seed, n, "uniform").copy()
    big = np.array([40000, 65536, 100000, 32769], np.uint32)
    sizes[::16] = big[np.arange(len(sizes[::16])) % len(big)]
    return sizes


@pytest.mark.parametrize("kind", ["PQ", "VACQ", "VLPQ"])
@pytest.mark.parametrize("agg", [1, 64])
def test_sizes_above_the_chain_use_the_device_heap(kind, agg):
    # every 16th request is above 32 KiB: it gets a device-heap block of the next power of two, which
    # its thread fills and reads back like any other, and free() hands it back to the heap; the pool
    # itself ends quiescent and no error bit is set
    sizes = _sizes_with_large(4096)
    print(f"device heap {kind} agg={agg}", flush=True)
    with ob.MemoryManager(kind, 1 << 30) as mm:
        r = mm.run_trace(sizes, 3, agg=agg, check_rounds=True).result
        assert r["verify_errors"] == 0 and r["bounds_errors"] == 0 and r["null_returns"] == 0, r
        assert r["overlap_errors"] == 0 and r["align_errors"] == 0, r
        assert r["error_word"] == 0, r["errors"]
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()


@pytest.mark.parametrize("tile", [1, 4])
def test_sizes_above_the_chain_block_wide(tile):
    sizes = _sizes_with_large(16384)
    print(f"device heap block-wide tile={tile}", flush=True)
    with ob.MemoryManager("PQ", 1 << 30) as mm:
        r = mm.run_batch(sizes, rounds=2, warmup=1, agg=0, block=256, graph=False, check=True, tile=tile)
        assert r["verify_errors"] == 0 and r["bounds_errors"] == 0 and r["null_returns"] == 0, r
        assert r["overlap_errors"] == 0 and r["align_errors"] == 0 and r["error_word"] == 0, r
        assert mm.check_quiescent().violations == 0


@pytest.mark.parametrize("kind", ["PQ", "CQ", "VAPQ", "VACQ", "VLPQ", "VLCQ"])
@pytest.mark.parametrize("form", ["agg1", "agg64", "block_t4"])
def test_statistics_count_exactly(kind, form):
    # one atomic per wave or workgroup, yet exact: every request served and every block returned is
    # counted, and the chunk counters reconcile with the bump pointer and the reuse ring (instance-1
    # sizes only, so every chunk handed out is one base chunk; VA / VL pools start with 12 queue chunks)
    n = 65536
    sizes = ob.gen_sizes(42, n, "uniform")
    print(f"statistics {kind} {form}", flush=True)
    with ob.MemoryManager(kind, 1 << 30) as mm:
        mm.set_statistics(True)
        if form == "block_t4":
            r = mm.run_batch(sizes, rounds=2, warmup=0, agg=0, block=1024, graph=True, check=True, tile=4)
            rounds = 3  # the validation round and two timed ones
        else:
            r = mm.run_trace(sizes, 3, agg=int(form[3:])).result
            rounds = 3
        assert r["error_word"] == 0 and r["null_returns"] == 0, r
        st = mm.statistics()
        q = mm.query()
        assert st["page_alloc"] == rounds * n and st["page_free"] == rounds * n, st
        first_data = 0 if kind in ("PQ", "CQ") else 12
        assert st["chunk_alloc"] == q.next_free_chunk - first_data + st["chunk_reuse"] - q.reuse_count[0], (st, q.next_free_chunk)
        if kind in ("PQ", "CQ"):
            assert st["chunk_reuse"] == 0
        mm.set_statistics(False)
        mm.run_trace(sizes, 1, agg=64)
        assert mm.statistics() == st  # switched off: nothing counted


@pytest.mark.parametrize("kind", ["PQ", "VAPQ", "VLCQ"])
def test_reinitialize_grows_the_pool(kind):
    # a pool that runs out, reinitialized at eight times its size, serves the same trace in full
    pool = (400 << 20) if kind == "PQ" else (24 << 20)
    sizes = np.full(16384, 8192, np.uint32)
    print(f"reinitialize {kind}", flush=True)
    with ob.MemoryManager(kind, pool) as mm:
        small = mm.query().max_chunks
        r = mm.run_trace(sizes, 1, agg=64).result
        assert r["null_returns"] > 0 and "OOM" in r["errors"]
        mm.reinitialize(8.0)
        st = mm.query()
        assert st.max_chunks > 4 * small and st.error == 0 and st.next_free_chunk == (0 if kind == "PQ" else 12)
        r = mm.run_trace(sizes, 2, agg=64, check_rounds=True).result
        assert r["null_returns"] == 0 and r["error_word"] == 0 and r["verify_errors"] == 0, r
        assert mm.check_quiescent().violations == 0
