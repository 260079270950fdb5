"""Build libouro.so in-tree: one translation unit per allocator instantiation (the six MultiOuro*
chains and the tuner's pool geometries) plus the C ABI, compiled for gfx950 with hipcc in parallel and
linked into ouroboros_amd/libouro.so."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_obj"
LIB = PKG / "libouro.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
NUM_VARIANTS = 12  # the six MultiOuro chains and six pool geometries (csrc/variant.hip)
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        # hipcc's host pass warns about the gfx-only target feature; only real failures surface
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stderr[-4000:]}")


def build(jobs: int | None = None, verbose: bool = False, extra_flags: list[str] | None = None,
          out: Path | None = None) -> Path:
    """extra_flags / out: alternative builds (e.g. -DOURO_DIAG into a separate library file)."""
    extra_flags = extra_flags or []
    lib = out or LIB
    obj_dir = OBJ if not extra_flags else OBJ.with_name("_obj_" + lib.stem)
    obj_dir.mkdir(exist_ok=True)
    units = [(CSRC / "variant.hip", obj_dir / f"variant{v}.o", [f"-DOURO_VARIANT={v}"]) for v in range(NUM_VARIANTS)]
    units.append((CSRC / "capi.hip", obj_dir / "capi.o", []))
    units.append((CSRC / "serve.hip", obj_dir / "serve.o", []))
    cmds = [[HIPCC, *FLAGS, *extra_flags, *extra, "-c", str(src), "-o", str(obj)] for src, obj, extra in units]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = lib.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *[str(o) for _, o, _ in units]])
    tmp.replace(lib)
    if verbose:
        print(f"built {lib}", file=sys.stderr)
    if not extra_flags and out is None:
        build_dropin()
    return lib


DROPIN = ROOT / "tests" / "dropin"


def build_dropin() -> Path:
    """The test kernels written only against include/ouro/instances.hpp (tests/dropin/dropin.hip)."""
    so = DROPIN / "libdropin.so"
    tmp = so.with_suffix(".so.tmp")
    _run([HIPCC, *FLAGS, "-shared", "-o", str(tmp), str(DROPIN / "dropin.hip")])
    tmp.replace(so)
    return so


if __name__ == "__main__":
    build(verbose=True)
