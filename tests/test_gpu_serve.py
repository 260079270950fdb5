"""The serving form on the device: Poisson-arriving request kernels over one device-resident pool.

A request's result (words read back intact, words not, null returns) depends only on its seed and
thread count, never on what ran beside it, so every request must return exactly the words the host
computes for it (ouro_serve_expected), whether it ran alone (one slot, arrivals far apart) or in a
mix of overlapping requests (many slots, arrivals faster than service)."""
import numpy as np
import pytest

import ouroboros_amd as ob

pytestmark = pytest.mark.gpu


def expected(n, **kw):
    cfg = ob.serve_config(n, **kw)
    return np.array([ob.serve_expected(cfg, r) for r in range(n)], np.uint64)


@pytest.mark.parametrize("kind", ["PQ", "VAPQ", "VLCQ"])
def test_results_same_alone_and_in_the_mix(kind):
    n = 48
    kw = dict(min_log2=10, max_log2=14, zipf=1.1, classes=10, seed=11)
    want = expected(n, rate=1000.0, **kw)
    print(f"serve {kind}", flush=True)
    with ob.MemoryManager(kind, 2 << 30) as mm:
        alone = mm.serve(n, rate=2000.0, slots=1, agg=1, admit=False, per_request=True, **kw)
        mix = mm.serve(n, rate=1e6, slots=16, agg=64, admit=True, per_request=True, **kw)
        block_mix = mm.serve(n, rate=1e6, slots=4, agg=0, admit=True, per_request=True, max_batch=16, **kw)
        tile_mix = mm.serve(n, rate=1e6, slots=4, agg=0, admit=True, per_request=True, max_batch=16, tile=4, **kw)
        for name, r in (("alone", alone), ("mix", mix), ("block mix", block_mix), ("tile mix", tile_mix)):
            assert r["error_word"] == 0, (name, r["errors"])
            assert r["completed"] == n and r["mismatches"] == 0 and r["nulls"] == 0, (name, r)
            res = r["results"]
            assert np.array_equal(res[:, 3], want[:, 0]), name
            assert np.array_equal(res[:, 0], want[:, 1]), name
        assert np.array_equal(alone["results"], mix["results"])
        assert np.array_equal(alone["results"], block_mix["results"])
        assert np.array_equal(alone["results"], tile_mix["results"])
        assert mix["max_in_flight"] > 1 and block_mix["max_in_flight"] > 1
        assert alone["launches"] == n and block_mix["launches"] < n
        assert mm.check_quiescent().violations == 0


def test_admission_holds_back_requests_beyond_the_pool():
    # a 512 MiB pool (about 500 MiB of chunks) and requests of up to 2^18 threads, some 150 MB each:
    # the admission test must defer requests while others run, and every request must still complete
    # intact
    n = 32
    kw = dict(min_log2=16, max_log2=18, zipf=1.1, classes=10, seed=3)
    want = expected(n, rate=1000.0, **kw)
    print("serve admission", flush=True)
    with ob.MemoryManager("VAPQ", 512 << 20) as mm:
        r = mm.serve(n, rate=1e6, slots=16, agg=64, admit=True, per_request=True, max_batch=8, **kw)
        assert r["error_word"] == 0, r["errors"]
        assert r["completed"] == n and r["mismatches"] == 0 and r["nulls"] == 0, r
        assert np.array_equal(r["results"][:, 0], want[:, 1])
        assert r["deferred"] > 0
        assert mm.check_quiescent().violations == 0
