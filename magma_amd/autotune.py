"""Autotuner for variable-size batches: picks, per n-bucket, the kernel
variant of the vbatched plan (register-tiled LU for n <= 128, or the blocked
LU with panel width 32, 16 or 8, or two 32-wide panels per 64-wide block) by
timing each valid variant on the bucket's own systems, replayed as a hipGraph.
Results are cached in a JSON file keyed by device (architecture and CU count)
and bucket id; `load_variants` reads the cache and tunes only buckets that are
missing from it.

The bucket boundaries follow the plan's (magma_amd/csrc/vbatched.hip):
n <= 128 in steps of 16, then 64 up to 512, 128 up to 1024, 256 up to 2048.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np
import torch

from .batched import IpivV, VBatch, VbatchedPlan

VARIANT_NAMES = {1: "tile", 2: "blocked nb=32", 3: "blocked nb=16", 4: "blocked nb=8", 5: "blocked 2x32"}
CACHE = Path(os.environ.get("MAGMA_AMD_TUNING", Path(__file__).resolve().parent / "tuning" / "vbatched.json"))


def bucket_of(n: np.ndarray) -> np.ndarray:
    """Bucket id of each square order (vectorised form of vb_bucket_of)."""
    n = np.asarray(n, dtype=np.int64)
    b = np.full(n.shape, -1, dtype=np.int64)
    b = np.where((n >= 1) & (n <= 128), (n + 15) // 16 - 1, b)
    b = np.where((n > 128) & (n <= 512), 8 + (n - 128 + 63) // 64 - 1, b)
    b = np.where((n > 512) & (n <= 1024), 14 + (n - 512 + 127) // 128 - 1, b)
    b = np.where((n > 1024) & (n <= 2048), 18 + (n - 1024 + 255) // 256 - 1, b)
    return b


def device_key() -> str:
    """Architecture and CU count (e.g. "gfx950-256cu"): stable across driver
    and profiler environments, unlike the marketing name."""
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    arch = getattr(p, "gcnArchName", "") or p.name
    return f"{arch.split(':')[0]}-{p.multi_processor_count}cu"


def _make_inputs(n, nrhs, seed, dev):
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    A = VBatch(n, n, device=dev)
    A.data.uniform_(-1.0, 1.0, generator=gen)
    B = VBatch(n, nrhs, device=dev)
    B.data.uniform_(-1.0, 1.0, generator=gen)
    return A, B, IpivV(n, dev), torch.zeros(len(n), dtype=torch.int32, device=dev)


def time_plan(plan: VbatchedPlan, A: VBatch, B: VBatch, ipiv: IpivV, info: torch.Tensor, reps: int = 3) -> float:
    """Average seconds per graph replay, the inputs restored before each."""
    A0, B0 = A.data.clone(), B.data.clone()
    plan.capture(A, B, ipiv, info)
    times = []
    for _ in range(reps + 1):
        A.data.copy_(A0)
        B.data.copy_(B0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        plan.replay()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / 1e3)
    return float(np.mean(times[1:]))


def tune_buckets(n, nrhs, buckets=None, max_systems: int = 4096, reps: int = 3, seed: int = 1234) -> dict:
    """Time every valid variant of each bucket (on up to max_systems of the
    bucket's own systems) and return {bucket: (variant, {variant: seconds})}."""
    n = np.asarray(n, dtype=np.int32)
    nrhs = np.asarray(nrhs, dtype=np.int32)
    ids = bucket_of(n)
    dev = torch.device("cuda", torch.cuda.current_device())
    out = {}
    for b in sorted(set(int(x) for x in ids if x >= 0)):
        if buckets is not None and b not in buckets:
            continue
        idx = np.nonzero(ids == b)[0][:max_systems]
        sn, sr = n[idx], nrhs[idx]
        A, B, ipiv, info = _make_inputs(sn, sr, seed + b, dev)
        timings = {}
        for v in (1, 2, 3, 4, 5):
            plan = VbatchedPlan(sn, sr, variants={b: v})
            got = [x for x in plan.buckets() if x["id"] == b]
            if not got or got[0]["variant"] != v:
                plan.close()
                continue                        # not valid for this bucket
            timings[v] = time_plan(plan, A, B, ipiv, info, reps)
            plan.close()
        best = min(timings, key=timings.get)
        out[b] = (best, timings)
        del A, B, ipiv, info
        torch.cuda.empty_cache()
    return out


def load_variants(n=None, nrhs=None, path: Path = CACHE, tune_missing: bool = True) -> dict:
    """{bucket id: variant} for this device from the cache; buckets of the
    given sizes that are missing are tuned now and written back."""
    table = {}
    if path.exists():
        table = json.loads(path.read_text())
    key = device_key()
    mine = {int(k): int(v["variant"]) for k, v in table.get(key, {}).items()}
    if n is not None and tune_missing:
        need = sorted(set(int(b) for b in bucket_of(np.asarray(n)) if b >= 0) - set(mine))
        if need:
            res = tune_buckets(n, nrhs if nrhs is not None else np.ones(len(n), dtype=np.int32), buckets=need)
            entry = table.setdefault(key, {})
            for b, (v, t) in res.items():
                entry[str(b)] = {"variant": v, "name": VARIANT_NAMES[v],
                                 "seconds": {VARIANT_NAMES[k]: round(s, 6) for k, s in t.items()}}
                mine[b] = v
            path.parent.mkdir(parents=True, exist_ok=True)
            path.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    return mine
