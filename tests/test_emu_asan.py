"""The library's kernels, built as host C++ and run lane by lane in the CPU
emulation of tests/emu under AddressSanitizer: every global-memory access of
the LU, solve and refinement paths is checked against the exact buffer sizes
on the same inputs as the GPU tests, and the results are checked numerically.
Needs clang++ with the ASan runtime (ROCm's LLVM); skipped otherwise."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / "emu"))

import build_emu  # noqa: E402

SCENARIOS = ["lu", "overflow", "singular", "nonconv", "acc:fp32:ir:100", "acc:bf16:gmres:100",
             "acc:fp16:ir:10000", "acc:bf16x3:ir:100:256", "acc:fp16:gmres:100:177", "dhgesv:200:fp32:ir",
             "dhgesv:257:bf16:gmres", "vbplan:5", "vbplan:2"]


@pytest.fixture(scope="module")
def emu_lib():
    cxx = build_emu.clang()
    if cxx is None or build_emu.asan_runtime(cxx) is None:
        pytest.skip("no clang++ with the ASan runtime")
    build_emu.build()
    return build_emu.asan_runtime(cxx)


@pytest.mark.timeout(900)
def test_kernels_under_asan(emu_lib):
    env = dict(os.environ, LD_PRELOAD=emu_lib, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, str(HERE / "emu" / "scenarios.py"), *SCENARIOS], env=env,
                       capture_output=True, text=True, timeout=880)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out, out[-6000:]
    assert r.returncode == 0, out[-6000:]
    assert "all scenarios passed" in out


# The single-matrix LU's panel is a cooperative kernel whose workgroups wait
# on one another, so the emulation runs several workgroups interleaved.
DENSE_SCENARIOS = ["dgetrf:300:280", "dgetrf:600:700", "sgetrf:520:530", "sgetrfp:200:200:bf16x3",
                   "sgetrfp:300:260:fp16", "sgetrfp:260:300:bf16", "dgetrs:300:3:0", "dgetrs:300:3:1",
                   "dgetrs:400:70:1", "dhgesv:600:fp32:irdtrs", "dhgesv:300:fp32:gmgmstrs",
                   "sgetrfp:200:200:fp8"]


@pytest.mark.timeout(900)
def test_dense_kernels_under_asan(emu_lib):
    env = dict(os.environ, LD_PRELOAD=emu_lib, ASAN_OPTIONS="detect_leaks=0", EMU_WG_WINDOW="5")
    r = subprocess.run([sys.executable, str(HERE / "emu" / "scenarios.py"), *DENSE_SCENARIOS], env=env,
                       capture_output=True, text=True, timeout=880)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out, out[-6000:]
    assert r.returncode == 0, out[-6000:]
    assert "all scenarios passed" in out
