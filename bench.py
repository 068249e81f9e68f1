"""Benchmark of the batched FP64 solve (MAGMA magma_dgesv_batched drop-in) on
one MI355X, with the host-LAPACK CPU path timed beside it.

Workload (one "step"): 20,000 systems A_s X_s = B_s, A_s 128x128 FP64 with
i.i.d. U[-1,1] entries (column-major, ldda = 128), B_s 128x4 U[-1,1];
magma_dgesv_batched factors (getrf) and solves (getrs) all of them. Every
timed step consumes its own pre-generated input set, so the timed region holds
only the solver; inputs are resident in HBM before timing starts.

Prints one JSON line (see the keys below). Run: python bench.py [--steps K]
[--warmup W] [--gpus 1].
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("solves/s + GFLOP/s: BF16-LU+IR vs plain FP64 batched dgesv at ≤1e-14 backward err")
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def flops_getrf(n: float) -> float:
    # testing/flops.h:80-91 (square)
    fmuls = 0.5 * n * (n * (n - n / 3.0 - 1.0) + n) + (2.0 / 3.0) * n
    fadds = 0.5 * n * (n * (n - n / 3.0) - n) + n / 6.0
    return fmuls + fadds


def flops_getrs(n: float, nrhs: float) -> float:
    return nrhs * n * n + nrhs * n * (n - 1.0)


def algorithmic_bytes(n: int, nrhs: int) -> int:
    # A read + LU written (FP64), B read + X written, int32 ipiv + info.
    return 16 * n * n + 16 * n * nrhs + 4 * n + 4


def measured_traffic(workload: str):
    """HBM bytes per launch of the dominant kernel from the committed PMC runs
    (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each scaled by what the
    counter reads for a device copy of known size), or None when this
    workload has not been profiled."""
    path = ROOT / "profiles" / "traffic.json"
    try:
        table = json.loads(path.read_text())
    except (OSError, ValueError):
        return None
    rec = table.get(workload)
    if rec is None:
        return None
    kib = rec["fetch_kib"] * rec["fetch_scale"] + rec["write_kib"] * rec["write_scale"]
    return int(kib * 1024)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=20000)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--nrhs", type=int, default=4)
    ap.add_argument("--cpu-sample", type=int, default=20000,
                    help="systems timed with host LAPACK for the CPU baseline (0 = skip)")
    ap.add_argument("--seed", type=int, default=20250101)
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="before the workload, copy a buffer of this many MiB once on the device (a kernel of "
                         "known traffic that scales the FETCH_SIZE / WRITE_SIZE counters of a --pmc run)")
    return ap.parse_args()


def cpu_baseline(A_host, B_host, n, nrhs, min_secs=10.0):
    """Host LAPACK dgesv over a sample of the same workload on the host cores
    (testing/testing_zgesv_batched.cpp:158-180 CPU path)."""
    import numpy as np
    import oracle

    lib = oracle.host_lapack()
    if lib is None:
        return None
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    # Repeat passes over the sample until about min_secs of wall time (a few
    # seconds of every core's work), so the rate is not a sub-second blip.
    secs, passes, X = 0.0, 0, None
    while secs < min_secs or passes == 0:
        LU, ipiv, info, X, s = oracle.host_dgesv_batch(A_host, B_host, cores)
        secs += s
        passes += 1
    batch = A_host.shape[0]
    be = oracle.gesv_backward_error(A_host, X, B_host, n)
    solved = batch * passes
    return {
        "value": solved / secs,
        "unit": "solves/s",
        "cores": cores,
        "kind": "reference",
        "sample": (f"{passes} passes over {batch} systems of the same workload ({n}x{n} FP64, {nrhs} RHS, "
                   f"U[-1,1]); host LAPACK dgesv per system, OpenMP schedule(dynamic) over the batch on {cores} "
                   f"threads, LAPACK sequential ({lib.cb_path().decode()}); {secs:.2f} s wall in total, "
                   f"max tester backward error {float(np.nanmax(be)):.2e}"),
        "gflops": solved * (flops_getrf(n) + flops_getrs(n, nrhs)) / secs / 1e9,
    }


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent systems shard by batch index across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    import torch
    from magma_amd import batched as mb

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    batch, n, nrhs = args.batch, args.n, args.nrhs
    K, W = args.steps, args.warmup
    queue = mb.default_queue(dev)
    lib = mb._ffi.load()

    if args.copy_calib_mib > 0:
        words = args.copy_calib_mib * (1 << 20) // 8
        src = torch.ones(words, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy of {args.copy_calib_mib} MiB done", file=sys.stderr, flush=True)
        del src, dst

    # Input sets: one per timed step (plus warm-up reuse), generated on device.
    gen = torch.Generator(device=dev)
    sets = []
    nsets = max(K, 1)
    free, total = torch.cuda.mem_get_info()
    per_set = batch * (n * n + n * nrhs) * 8 + batch * (n + 1) * 4 + 3 * batch * 8
    if nsets * per_set > 0.6 * free:
        nsets = max(1, int(0.6 * free // per_set))

    def make_set(i):
        gen.manual_seed(args.seed + i)
        A = torch.rand((batch, n, n), dtype=torch.float64, device=dev, generator=gen) * 2 - 1
        B = torch.rand((batch, nrhs, n), dtype=torch.float64, device=dev, generator=gen) * 2 - 1
        ipiv = torch.zeros((batch, n), dtype=torch.int32, device=dev)
        info = torch.zeros(batch, dtype=torch.int32, device=dev)
        pA = mb.pointer_array(A, queue)
        pB = mb.pointer_array(B, queue)
        pP = mb.pointer_array(ipiv.view(batch, 1, n), queue)
        return dict(A=A, B=B, ipiv=ipiv, info=info, pA=pA, pB=pB, pP=pP)

    for i in range(nsets):
        sets.append(make_set(i))
    torch.cuda.synchronize()

    def step(s):
        rc = lib.magma_dgesv_batched(n, nrhs, s["pA"].data_ptr(), n, s["pP"].data_ptr(), s["pB"].data_ptr(), n,
                                     s["info"].data_ptr(), batch, queue.handle)
        if rc != 0:
            raise RuntimeError(f"magma_dgesv_batched returned {rc}")

    def refresh(i):
        gen.manual_seed(args.seed + i)
        sets[i]["A"].copy_(torch.rand((batch, n, n), dtype=torch.float64, device=dev, generator=gen) * 2 - 1)
        sets[i]["B"].copy_(torch.rand((batch, nrhs, n), dtype=torch.float64, device=dev, generator=gen) * 2 - 1)

    print(f"[bench] warmup {W} steps", file=sys.stderr, flush=True)
    for w in range(W):
        step(sets[w % nsets])
    torch.cuda.synchronize()
    for i in range(min(W, nsets)):
        refresh(i)
    torch.cuda.synchronize()

    # Timed region: exactly K steps, inputs resident, one event pair per step
    # on the queue's stream (= torch's current stream).
    print(f"[bench] timing {K} steps", file=sys.stderr, flush=True)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        if k >= nsets:
            raise RuntimeError("not enough device memory for one input set per timed step")
        ev[k][0].record()
        step(sets[k])
        ev[k][1].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    ev_ms = [a.elapsed_time(b) for a, b in ev]
    ms_per_step = wall / K * 1e3
    solves_per_s = batch * K / wall
    flop = batch * (flops_getrf(n) + flops_getrs(n, nrhs))
    gflops = flop * K / wall / 1e9

    # Accuracy of the measured outputs (checker arithmetic in FP64 torch).
    ref_gen = torch.Generator(device=dev)
    ref_gen.manual_seed(args.seed + 0)
    A0 = torch.rand((batch, n, n), dtype=torch.float64, device=dev, generator=ref_gen) * 2 - 1
    B0 = torch.rand((batch, nrhs, n), dtype=torch.float64, device=dev, generator=ref_gen) * 2 - 1
    X = sets[0]["B"]
    R = B0 - torch.bmm(X, A0)     # (X^T A^T)^T: rows are RHS, so R^T = B - A X
    Ainf = A0.abs().sum(1).amax(1)            # ||A||_inf (row sums of A = column sums of A^T)
    Xinf = X.abs().sum(1).amax(1)             # ||X||_inf
    rinf = R.abs().sum(1).amax(1)
    bwd = (rinf / (Ainf * Xinf)).max().item()
    tester = (rinf / (n * Ainf * Xinf)).max().item()
    n_info = int((sets[0]["info"] != 0).sum().item())
    del A0, B0, R

    # Roofline of the dominant launch group (getrf + getrs of one step).
    bytes_step = batch * algorithmic_bytes(n, nrhs)
    ev_avg_s = sum(ev_ms) / len(ev_ms) / 1e3
    achieved = bytes_step / ev_avg_s / 1e9
    workload = f"dgesv_batched {batch} x {n}x{n} FP64, nrhs={nrhs}, ldda={n}"
    fused = n <= 128 and nrhs <= 16
    roof = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(workload),
        "kernel": (f"getrf_tile_kernel<{(n + 15) // 16}, true> (one fused LU + solve launch per call)" if fused
                   else "magma_dgesv_batched launch group (getrf + getrs)"),
        "event_ms_avg": round(ev_avg_s * 1e3, 4),
        "algorithmic_bytes_per_launch": bytes_step,
        "fp64_tflops": round(flop / ev_avg_s / 1e12, 2),
        "fp64_frac": round(flop / ev_avg_s / 1e12 / FP64_PEAK_TFLOPS, 4),
    }

    cpu = None
    if args.cpu_sample > 0:
        print(f"[bench] cpu baseline on {args.cpu_sample} systems", file=sys.stderr, flush=True)
        m = min(args.cpu_sample, batch)
        refresh(0)
        torch.cuda.synchronize()
        A_host = sets[0]["A"][:m].cpu().numpy()
        B_host = sets[0]["B"][:m].cpu().numpy()
        cpu = cpu_baseline(A_host, B_host, n, nrhs)

    out = {
        "metric": METRIC,
        "value": round(solves_per_s, 1),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (U[-1,1] from torch Philox, seeded)",
        "config": {"workload": workload,
                   "batch": batch, "n": n, "nrhs": nrhs, "solver": "fp64 getrf+getrs"},
        "gflops": round(gflops, 1),
        "backward_error_max": bwd,
        "tester_error_max": tester,
        "nonzero_info": n_info,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
