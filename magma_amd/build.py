"""Build libmagma_mi355x.so (HIP kernels + C ABI) for gfx950 with hipcc.

Usage: python -m magma_amd.build [--jobs N] [--verbose]

Every ``csrc/*.hip`` file is compiled to an object under ``magma_amd/_build``
and linked into ``magma_amd/libmagma_mi355x.so``. The build is complete each
time (no incremental skipping) so that a call always reflects the sources.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libmagma_mi355x.so"
ARCH = os.environ.get("MAGMA_MI355X_ARCH", "gfx950")

# -packed-fp32-ops: never emit packed FP32 VALU instructions (v_pk_fma_f32 and
# friends); scalar and float2 code is compiled to unpacked instructions.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libmagma_mi355x.so")


def _compile(src: Path, verbose: bool, incremental: bool = False) -> Path:
    obj = BUILD / (src.stem + ".o")
    if incremental and obj.exists():
        newest = max([src.stat().st_mtime] + [h.stat().st_mtime for h in CSRC.glob("*.h")] +
                     [(PKG.parent / "include" / "magma_mi355x.h").stat().st_mtime])
        if obj.stat().st_mtime > newest:
            return obj
    cmd = [hipcc(), *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)
    return obj


def build(jobs: int | None = None, verbose: bool = False, incremental: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    jobs = jobs or min(len(sources), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, incremental), sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--incremental", action="store_true",
                    help="skip sources whose object is newer than the source and every header")
    args = ap.parse_args()
    print(build(args.jobs, args.verbose, args.incremental))


if __name__ == "__main__":
    main()
