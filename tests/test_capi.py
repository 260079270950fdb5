"""CPU tests of the C ABI library: it loads, exports every entry point include/*.h declares, and its
pure-host functions (pool arithmetic, trace generator) agree with the oracle and with statistics."""
import re
from pathlib import Path

import numpy as np

import oracle
import ouroboros_amd as ob

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        names |= set(re.findall(r"^\s*(?:int|void\*|void)\s+(ouro_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = ob.lib()
    names = declared_symbols()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert names <= set(ob.SIGNATURES), names - set(ob.SIGNATURES)


def test_struct_mirrors_match_the_c_layout():
    # every ctypes mirror has the size the library was compiled with (a field added on one side only
    # would make the library write past the Python buffer)
    mirrors = [ob.VariantInfo, ob.State, ob.CheckResult, ob.TraceConfig, ob.TraceResult, ob.BatchConfig,
               ob.BatchResult, ob.BatchWorkload, ob.ServeConfig, ob.ServeResult]
    got = (ob.C.c_uint32 * len(mirrors))()
    assert ob.lib().ouro_abi_sizes(got, len(mirrors)) == 0
    assert [ob.C.sizeof(m) for m in mirrors] == list(got)


def test_layout_agrees_with_oracle():
    for v in ["PQ", "CQ", "VAPQ", "VACQ", "VLPQ", "VLCQ"]:
        for pool in [400 << 20, 512 << 20, 4 << 30, 64 << 30]:
            assert ob.layout(v, pool) == oracle.layout(v, pool)
    # the tuner's geometries: smaller queue arrays leave more of the pool to chunks
    assert ob.layout("PQ-cap1m", 4 << 30)[0] > ob.layout("PQ", 4 << 30)[0]
    assert ob.layout("VAPQ-cap4k", 4 << 30)[0] > ob.layout("VAPQ", 4 << 30)[0]


def test_trace_generator_is_deterministic_and_uniform():
    a = ob.gen_sizes(42, 1 << 20, "uniform")
    b = ob.gen_sizes(42, 1 << 20, "uniform")
    assert np.array_equal(a, b)
    assert set(np.unique(a)) == {16 << q for q in range(10)}
    counts = np.bincount(np.log2(a).astype(int) - 4, minlength=10)
    expected = len(a) / 10
    chi2 = ((counts - expected) ** 2 / expected).sum()
    assert chi2 < 40  # 9 dof; p ~ 1e-5
    # mean 1636.8 B (uniform over the ten power-of-two classes)
    assert abs(a.mean() - 1636.8) / 1636.8 < 0.01


def test_zipf_generator_matches_rank_probabilities():
    s = 1.2
    a = ob.gen_sizes(9, 1 << 20, "zipf", s)
    p = np.array([(q + 1) ** -s for q in range(10)])
    p /= p.sum()
    got = np.bincount(np.log2(a).astype(int) - 4, minlength=10) / len(a)
    assert np.allclose(got, p, atol=3e-3)
    assert abs(p[0] - 0.405) < 1e-3  # rank-1 probability quoted for config 3


def test_semaphore_share_matches_oracle():
    # the device's carve decision and word packing, compiled for the host, against the oracle's
    # restatement. Regression: a negative debt once became a huge chunk count (a uint32 wrap) and
    # moved the bump pointer backwards; every share stays within [0, ceil(num / ppc) + 1].
    rng = np.random.default_rng(11)
    for _ in range(20000):
        c = int(rng.integers(-(1 << 30), 1 << 30)) if rng.random() < 0.3 else int(rng.integers(-2000, 2000))
        e = int(rng.integers(0, 1 << 24)) if rng.random() < 0.2 else int(rng.integers(0, 4096))
        num = int(rng.integers(0, 1025))
        ppc = int(2 ** rng.integers(0, 10))
        word = ob.sem_make(c, e)
        assert word == oracle.sem_make(c, e)
        k = ob.sem_share(word, num, ppc)
        assert k == oracle.sem_share(c, e, num, ppc)
        assert 0 <= k <= -(-num // ppc) + 1


def test_serve_expected_restatement():
    # ouro_serve_expected against a restatement of the request kernel's draws: thread i of request r
    # draws class q from the top 32 bits of keyed(seed_r, i) against Zipf thresholds and reads back
    # (16 << q) / 4 words
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

    seed, zipf, classes = 9, 1.1, 10
    cfg = ob.serve_config(8, 1000.0, min_log2=4, max_log2=9, zipf=zipf, classes=classes, seed=seed)
    p = np.array([(q + 1) ** -zipf for q in range(classes)])
    p /= p.sum()
    thr = [min(4294967295.0, np.floor(c * 4294967296.0)) for c in np.cumsum(p)]
    thr[-1] = 0xFFFFFFFF
    for r in range(8):
        e = 4 + 5 * unit(keyed(seed ^ 0xE7037ED1A0B428DB, r))
        n = int(np.floor(2.0 ** e))
        sr = keyed(seed ^ 0x8EBC6AF09C88C6E3, r)
        words = 0
        for i in range(n):
            x = keyed(sr, i) >> 32
            q = 0
            while q + 1 < classes and x >= thr[q]:
                q += 1
            words += (16 << q) // 4
        assert ob.serve_expected(cfg, r) == (n, words)


def test_pools_beyond_the_chunk_index_range_are_rejected():
    # MemoryIndex keeps 23 chunk bits: a pool of more than 2^23 - 1 chunks would alias page handles
    # (and chunk 2^23 - 1 would encode page handles equal to the queue sentinels), so ouro_create
    # refuses it before touching the device; 64 GiB (8,259,489 VA chunks) stays inside the range
    assert ob.layout("VAPQ", 64 << 30)[0] < (1 << 23) - 1
    assert ob.layout("PQ", 72 << 30)[0] > (1 << 23) - 1
    h = ob.C.c_void_p()
    assert ob.lib().ouro_create(ob.VARIANTS["PQ"], 72 << 30, ob.C.byref(h)) == 1  # OURO_E_INVALID
    assert not h.value
