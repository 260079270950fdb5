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


def test_layout_agrees_with_oracle():
    for v in ob.VARIANTS:
        for pool in [400 << 20, 512 << 20, 4 << 30, 64 << 30]:
            assert ob.layout(v, pool) == oracle.layout(v, pool)


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


def test_semaphore_decision_bounds():
    # Regression: a count driven below zero by optimistic takes once produced a negative chunk
    # count that wrapped to ~4e9 as uint32 and moved the bump pointer backwards. The device's
    # decision function (run here on the host) must keep 1 <= k <= max_k and every field in range.
    rng = np.random.default_rng(11)
    F = (1 << 19) - 1
    for _ in range(20000):
        c = int(rng.integers(-(1 << 24), 1 << 24)) if rng.random() < 0.5 else int(rng.integers(-200, 200))
        e = int(rng.integers(0, F + 1)) if rng.random() < 0.3 else int(rng.integers(0, 2048))
        r = int(rng.integers(0, F + 1)) if rng.random() < 0.3 else int(rng.integers(0, 2048))
        num = int(rng.integers(1, 65))
        ppc = int(2 ** rng.integers(0, 10))
        max_k = int(rng.integers(1, 65))
        word = oracle.sem_make(c, e, r)
        mode, k, nxt = ob.sem_decide(word, num, ppc, max_k)
        c2, e2, r2 = oracle.sem_fields(nxt)
        assert 0 <= e2 <= F and 0 <= r2 <= F
        if mode == 0:  # carve
            assert 1 <= k <= max_k
            assert (c2, e2, r2) == (c, e + k * ppc, r)
            assert max(c, 0) + e - r < num
        elif mode == 1:  # take
            assert c >= num and (c2, e2, r2) == (c - num, e, r)
        elif mode == 2:  # reserve
            assert (c2, e2, r2) == (c, e, r + num)
        else:  # hold: nothing changes
            assert mode == 3 and nxt == word
    # the case behind the fault: negative count, expected already covering the group
    mode, k, _ = ob.sem_decide(oracle.sem_make(-50000, 600, 0), 6, 512, 64)
    assert mode == 2  # reserve, not a zero- or negative-sized carve
