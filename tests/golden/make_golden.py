"""Writes tests/golden/single_wave.npz: single-wavefront traces of the oracle for every queue kind.

The reference ships no golden vectors (it has no tests), so these fixtures are the oracle's own
replays, committed so that a later change to the oracle or to the device code is caught against a
fixed answer. Page-kind traces of 16 B requests are also checked against the closed form in
tests/test_oracle.py.

    python tests/golden/make_golden.py
"""
from pathlib import Path
import sys

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402

KINDS = ["PQ", "CQ", "VAPQ", "VACQ", "VLPQ", "VLCQ"]


def splitmix(seed: int, n: int) -> np.ndarray:
    # numpy restatement of ouro_gen_sizes(dist=uniform): 16 << (keyed(seed, i) % 10)
    M = (1 << 64) - 1

    def sm(x):
        x = (x + 0x9E3779B97F4A7C15) & M
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
        return x ^ (x >> 31)

    s = sm(seed)
    return np.array([16 << (sm(s ^ ((i * 0xD1B54A32D192ED03) & M)) % 10) for i in range(n)], dtype=np.uint32)


CASES = {
    "config1_16B": (np.full(64, 16, np.uint32), 10),
    "uniform_seed42": (splitmix(42, 64), 10),
    "class0_40rounds": (np.full(64, 16, np.uint32), 40),
    "ragged_37": (splitmix(7, 37), 6),
    "instance2": ((16 << (np.arange(64) % 12)).astype(np.uint32), 8),
}


def main() -> None:
    out = {}
    for name, (sizes, rounds) in CASES.items():
        out[f"{name}__sizes"] = sizes
        for k in KINDS:
            s = oracle.Sim(k, 512 << 20)
            out[f"{name}__{k}"] = s.trace(sizes, rounds)
            assert s.state().error == 0, (name, k)
    np.savez_compressed(Path(__file__).with_name("single_wave.npz"), **out)
    print(f"wrote {len(out)} arrays")


if __name__ == "__main__":
    main()
