"""Build the library's HIP sources as host C++ under AddressSanitizer, for the
CPU emulation in tests/emu.

The sources are rewritten mechanically (nothing else changes):
  * `kernel<<<grid, block, shm, stream>>>(args);` becomes
    `EMU_LAUNCH(grid, block, shm, stream) << [=]() { kernel(args); };`
  * `extern __shared__ ... unsigned char name[];` becomes a pointer to the
    emulated workgroup's dynamic LDS buffer;
  * the public header is included from the repository's include/.
The output goes to tests/emu/_build/ (libemu.so).

Usage: python tests/emu/build_emu.py [--src DIR]
"""
from __future__ import annotations

import argparse
import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
BUILD = HERE / "_build"


def clang() -> str | None:
    for c in (os.environ.get("EMU_CXX"), "/opt/rocm/llvm/bin/clang++", shutil.which("clang++")):
        if c and Path(c).exists():
            return c
    return None


def asan_runtime(cxx: str) -> str | None:
    out = subprocess.run([cxx, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    p = out.stdout.strip()
    return p if p and Path(p).exists() else None


def _match_back_angle(s: str, i: int) -> int:
    depth = 0
    while i >= 0:
        if s[i] == '>':
            depth += 1
        elif s[i] == '<':
            depth -= 1
            if depth == 0:
                return i
        i -= 1
    raise ValueError("unbalanced <>")


def _match_paren(s: str, i: int) -> int:
    depth = 0
    while i < len(s):
        if s[i] == '(':
            depth += 1
        elif s[i] == ')':
            depth -= 1
            if depth == 0:
                return i
        i += 1
    raise ValueError("unbalanced ()")


def rewrite_launches(s: str) -> str:
    out, pos = [], 0
    while True:
        k = s.find("<<<", pos)
        if k < 0:
            out.append(s[pos:])
            return "".join(out)
        e = s.find(">>>", k)
        cfg = s[k + 3:e]
        j = k - 1
        while s[j].isspace():
            j -= 1
        if s[j] == '>':
            j = _match_back_angle(s, j) - 1
        while j >= 0 and (s[j].isalnum() or s[j] in "_:"):
            j -= 1
        kstart = j + 1
        kernel = s[kstart:k].strip()
        p = e + 3
        while s[p].isspace():
            p += 1
        if s[p] != '(':
            raise ValueError("launch without argument list near: " + s[k - 40:p + 10])
        q = _match_paren(s, p)
        out.append(s[pos:kstart])
        out.append(f"EMU_LAUNCH({cfg}) << [=]() {{ {kernel}({s[p + 1:q]}); }}")
        pos = q + 1


DYN = re.compile(r"extern\s+__shared__\s+(?:__align__\(\d+\)\s+)?unsigned\s+char\s+(\w+)\[\];")
# static LDS declarations (per workgroup storage in the emulation)
SHV = re.compile(r"__shared__\s+(?:__align__\(\d+\)\s+)?([\w:<>, ]+?)\s+(\w+)((?:\[[^\];]+\])*)\s*;")
_SHV_ID = [0]


def _shvar(m: re.Match) -> str:
    _SHV_ID[0] += 1
    return f"auto& {m.group(2)} = emu::shvar<{m.group(1)}{m.group(3)}>({_SHV_ID[0]});"


def transform(src: Path, dst: Path) -> list[Path]:
    dst.mkdir(parents=True, exist_ok=True)
    cpps = []
    for f in sorted(src.glob("*.hip")) + sorted(src.glob("*.h")):
        s = f.read_text()
        s = DYN.sub(r"unsigned char* \1 = emu::dyn_shared();", s)
        s = SHV.sub(_shvar, s)
        s = rewrite_launches(s)
        s = s.replace('"../../include/magma_mi355x.h"', "<magma_mi355x.h>")
        out = dst / f.name.replace(".hip", ".cpp")
        if not out.exists() or out.read_text() != s:
            out.write_text(s)
        if out.suffix == ".cpp":
            cpps.append(out)
    return cpps


def build(src: Path | None = None) -> Path:
    cxx = clang()
    if cxx is None:
        raise RuntimeError("no clang++ for the emulation build")
    src = src or (ROOT / "magma_amd" / "csrc")
    gen = BUILD / "src"
    objd = BUILD / "obj"
    objd.mkdir(parents=True, exist_ok=True)
    cpps = transform(src, gen) + [HERE / "emu.cpp"]
    flags = ["-w", "-std=c++20", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-fPIC",
             f"-I{HERE / 'include'}", f"-I{ROOT / 'include'}"]

    headers = [h for h in gen.glob("*.h")] + [HERE / "include" / "hip" / "hip_runtime.h"]
    newest_h = max(h.stat().st_mtime for h in headers)

    def compile_one(c: Path) -> Path:
        o = objd / (c.stem + ".o")
        if o.exists() and o.stat().st_mtime > max(c.stat().st_mtime, newest_h):
            return o
        r = subprocess.run([cxx, *flags, "-c", str(c), "-o", str(o)], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile of {c.name} failed:\n{r.stderr[-4000:]}")
        return o

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, cpps))
    lib = BUILD / "libemu.so"
    r = subprocess.run([cxx, "-shared", "-fsanitize=address", "-pthread", *map(str, objs), "-o", str(lib)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    return lib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--src", type=Path, default=None)
    args = ap.parse_args()
    print(build(args.src))


if __name__ == "__main__":
    sys.exit(main())
