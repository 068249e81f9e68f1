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
    ap.add_argument("--mixed", type=int, default=1,
                    help="also time BF16-LU + FP64 refinement against FP64 dgesv on 2,000 x 256^2 (0 = skip)")
    ap.add_argument("--adjoint", type=int, default=1,
                    help="also time the differentiable solve (forward + backward) on 8,192 x 64^2 (0 = skip)")
    ap.add_argument("--dense", type=int, default=1,
                    help="also run the dense mixed-precision leg (one system of order 24,576 at three condition "
                         "numbers); 0 skips")
    ap.add_argument("--dense-n", type=int, default=24576)
    ap.add_argument("--dense-conds", default="1e2,1e4,1e6",
                    help="condition numbers of the dense leg, comma separated")
    ap.add_argument("--dense-variants", default="bf16:ir,bf16:gmres,fp16:ir,fp16:gmres,bf16x3:ir,bf16x3:gmres",
                    help="facto:solver pairs of the dense leg, comma separated")
    ap.add_argument("--c1", type=int, default=1,
                    help="also run the small CPU-path leg (1,000 x 32^2, host LAPACK beside the GPU); 0 skips")
    ap.add_argument("--vbatched", type=int, default=1,
                    help="also run the variable-size leg (50,000 systems, n ~ U[8,512], nrhs ~ U[1,16]); 0 skips")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="before the workload, copy a buffer of this many MiB once on the device (a kernel of "
                         "known traffic that scales the FETCH_SIZE / WRITE_SIZE counters of a --pmc run)")
    return ap.parse_args()


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads() -> int:
    """Threads for the host LAPACK baseline: the cores this process may use,
    capped at 16 (the host share of one GPU on the measurement boxes)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    return max(1, min(cores, 16))


def cpu_baseline(A_host, B_host, n, nrhs, min_secs=10.0):
    """Host LAPACK dgesv over a sample of the same workload on the host cores
    (testing/testing_zgesv_batched.cpp:158-180 CPU path)."""
    import numpy as np
    import oracle

    lib = oracle.host_lapack()
    if lib is None:
        return None
    cores = cpu_threads()
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
        "cpu_model": cpu_model(),
        "host_cores_visible": os.cpu_count(),
    }


def c1_leg(torch, mb, lib, queue, dev, seed, batch=1000, n=32, nrhs=1, steps=20):
    """The reference's CPU path at small size (lapackf77_dgesv looped over the
    batch, testing_zgesv_batched.cpp:158-180) on 1,000 systems 32x32 with one
    RHS, timed on the host cores, with the GPU's fused tile LU + solve of the
    same batch beside it (inputs resident, HIP events on the queue stream)."""
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    A = torch.rand((batch, n, n), dtype=torch.float64, device=dev, generator=gen) * 2 - 1
    B = torch.rand((batch, nrhs, n), dtype=torch.float64, device=dev, generator=gen) * 2 - 1
    A_host, B_host = A.cpu().numpy(), B.cpu().numpy()
    sets = [(A.clone(), B.clone()) for _ in range(steps + 1)]
    mb.dgesv_batched(*sets[0])
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for k in range(1, steps + 1):
        mb.dgesv_batched(*sets[k])
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / steps
    X = sets[1][1]
    be = float(backward_errors(torch, A, X, B).max())
    nbytes = batch * algorithmic_bytes(n, nrhs)
    cpu = cpu_baseline(A_host, B_host, n, nrhs, min_secs=5.0)
    return {"workload": f"dgesv {batch} x {n}x{n} FP64, nrhs={nrhs} (the reference's CPU path)",
            "gpu": {"solves_per_s": round(batch / t, 1), "us_per_call": round(t * 1e6, 2),
                    "backward_error_max": be,
                    "roofline": {"bound": "hbm", "achieved": round(nbytes / t / 1e9, 1), "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": round(nbytes / t / 1e9 / HBM_PEAK_GBS, 4),
                                 "algorithmic_bytes_per_launch": nbytes}},
            "cpu": cpu}


def svd_geo_batch(torch, batch, n, cond, gen, dev):
    """A = U diag(sigma) V^T with Haar U, V and sigma_i = cond^(-(i-1)/(n-1))
    (testing/magma_generate.cpp, matrix svd_geo), generated on the device."""
    def haar():
        Z = torch.randn(batch, n, n, dtype=torch.float64, device=dev, generator=gen)
        out = torch.empty_like(Z)
        for c in range(0, batch, 125):      # the batched QR's workspace is sized per call
            Q, R = torch.linalg.qr(Z[c:c + 125])
            out[c:c + 125] = Q * torch.sign(torch.diagonal(R, dim1=1, dim2=2))[:, None, :]
        return out
    sig = cond ** (-torch.arange(n, dtype=torch.float64, device=dev) / max(n - 1, 1))
    U = haar() * sig[None, None, :]
    Vt = haar().transpose(1, 2).contiguous()
    return torch.bmm(U, Vt)


def backward_errors(torch, A_cm, X_cm, B_cm):
    """||B - A X||_inf / (||A||_inf ||X||_inf) per system; all three in the
    column-major (batch, ncols, ld) layout of the C interface."""
    A = A_cm.transpose(1, 2)
    X = X_cm.transpose(1, 2)
    R = B_cm.transpose(1, 2) - torch.bmm(A, X.contiguous())
    rn = R.abs().sum(2).amax(1)
    an = A.abs().sum(2).amax(1)
    xn = X.abs().sum(2).amax(1)
    return rn / (an * xn)


def mixed_precision_leg(torch, mb, lib, queue, dev, seed, steps=3, batch=2000, n=256, nrhs=1,
                        conds=(1e2, 1e4, 1e6)):
    """The batch half of the mixed-precision workload at each condition
    number (one run per cond)."""
    return {f"cond_{c:.0e}": mixed_precision_batch(torch, mb, lib, queue, dev, seed, steps, batch, n, nrhs, c)
            for c in conds}


def mixed_precision_batch(torch, mb, lib, queue, dev, seed, steps, batch, n, nrhs, cond):
    """BF16- and FP16-update LU + FP64 refinement (classical and FGMRES)
    against FP64 dgesv_batched on the same batch of svd_geo systems: solves/s,
    FP64-credited GFLOP/s, refinement steps and the worst backward / forward
    errors."""
    from magma_amd import iteref
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    A = svd_geo_batch(torch, batch, n, cond, gen, dev)          # read as column-major
    B = torch.rand((batch, nrhs, n), dtype=torch.float64, device=dev, generator=gen) * 2 - 1
    flop = batch * (flops_getrf(n) + flops_getrs(n, nrhs))
    res = {"workload": f"{batch} x {n}x{n} svd_geo cond {cond:g}, nrhs={nrhs}", "steps": steps}

    # FP64 reference path: every step factors its own copy
    copies = [(A.clone(), B.clone()) for _ in range(steps + 1)]
    mb.dgesv_batched(*copies[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(1, steps + 1):
        mb.dgesv_batched(*copies[k])
    torch.cuda.synchronize()
    t64 = (time.perf_counter() - t0) / steps
    X64 = copies[1][1]
    be64 = float(backward_errors(torch, A, X64, B).max())
    del copies

    out = {}
    for facto in ("bf16", "fp16"):
        for solver in ("ir", "gmres"):
            X, it, info = iteref.dxgesv_iteref_batched(A, B, facto=facto, solver=solver, itermax=100)   # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                X, it, info = iteref.dxgesv_iteref_batched(A, B, facto=facto, solver=solver, itermax=100)
            torch.cuda.synchronize()
            tir = (time.perf_counter() - t0) / steps
            be = float(backward_errors(torch, A, X, B).max())
            fe = float(((X - X64).abs().amax(dim=(1, 2)) / X64.abs().amax(dim=(1, 2))).max())
            itc = it.cpu()
            out[f"{facto}_{solver}"] = {
                "solves_per_s": round(batch / tir, 1), "ms": round(tir * 1e3, 3),
                "gflops_fp64_credited": round(flop / tir / 1e9, 1),
                "iters_mean": round(float(itc.clamp(min=0).double().mean()), 2),
                "iters_max": int(itc.max()), "fallbacks": int((itc < 0).sum()), "backward_error_max": be,
                "forward_error_vs_fp64_max": fe, "speedup_vs_fp64": round(t64 / tir, 3)}
    res.update({"fp64_dgesv": {"solves_per_s": round(batch / t64, 1), "ms": round(t64 * 1e3, 3),
                               "gflops": round(flop / t64 / 1e9, 1), "backward_error_max": be64},
                "lu_refinement": out})
    return res


def haar_cols(torch, n, gen, dev):
    """Haar-distributed orthogonal n x n matrix: QR of a Gaussian matrix with
    the sign of R's diagonal moved into Q."""
    Z = torch.randn(n, n, dtype=torch.float64, device=dev, generator=gen)
    Q, R = torch.linalg.qr(Z)
    Q *= torch.sign(torch.diagonal(R))[None, :]
    del Z, R
    return Q


def dense_leg(torch, dev, seed, n=24576, conds=(1e2, 1e4, 1e6),
              variants=(("bf16", "ir"), ("bf16", "gmres"), ("fp16", "ir"), ("fp16", "gmres"), ("bf16x3", "ir"),
                        ("bf16x3", "gmres"))):
    """The dense half of the mixed-precision workload: one system of order n,
    A = U diag(sigma) V^T (Haar U and V, generated once; sigma geometric to
    cond), b ~ U[-1,1]. FP64 magma_dgetrf_gpu + magma_dgetrs_gpu against the
    refinement solver magma_dxgesv_gmres_gpu with BF16 / FP16 updates and
    classical or GMRES refinement; each call is timed whole (host clock,
    device synchronised), as a caller sees it. The low-precision LU alone is
    timed through magma_sgetrf_prec_gpu for its MFMA fraction."""
    from magma_amd import dense, iteref
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    t_gen = time.perf_counter()
    U = haar_cols(torch, n, gen, dev)
    V = haar_cols(torch, n, gen, dev)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t_gen
    flop = flops_getrf(n) + flops_getrs(n, 1)
    out = {"workload": f"one FP64 system n={n}, svd_geo (Haar U, V), nrhs=1", "generation_s": round(t_gen, 2),
           "fp64_credited_gflop": round(flop / 1e9, 1)}

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    # low-precision LU alone (MFMA fraction); the matrix does not matter for its speed
    Af = (U * 0.5).T.contiguous().float()
    lu_only = {}
    # FP8 (e4m3) is an experiment: its MFMA peak is 5.0 PFLOP/s
    for facto, peak in (("bf16", 2500.0), ("fp16", 2500.0), ("bf16x3", 2500.0 / 3), ("fp8", 5000.0)):
        W = Af.clone()
        (ipiv, info), t = timed(lambda: dense.sgetrf_prec_gpu(W, facto))
        lu_only[facto] = {"ms": round(t * 1e3, 2), "tflops": round(flops_getrf(n) / t / 1e12, 2),
                          "frac_of_mfma_peak": round(flops_getrf(n) / t / 1e12 / peak, 4)}
        del W
    del Af
    out["low_precision_lu"] = lu_only
    for cond in conds:
        sig = cond ** (-torch.arange(n, dtype=torch.float64, device=dev) / max(n - 1, 1))
        A = ((U * sig[None, :]) @ V.T.contiguous()).T.contiguous()       # column storage of U S V^T
        b = torch.rand((1, n), dtype=torch.float64, device=dev, generator=gen) * 2 - 1
        LU = A.clone()
        (ipiv, info), t_lu = timed(lambda: dense.getrf_gpu(LU))
        X64 = b.clone()
        _, t_rs = timed(lambda: dense.getrs_gpu(LU, ipiv, X64))
        del LU
        t64 = t_lu + t_rs
        rec = {"fp64": {"ms": round(t64 * 1e3, 2), "getrf_ms": round(t_lu * 1e3, 2),
                        "getrs_ms": round(t_rs * 1e3, 2), "solves_per_s": round(1 / t64, 3),
                        "tflops": round(flop / t64 / 1e12, 2),
                        "getrf_fp64_frac": round(flops_getrf(n) / t_lu / 1e12 / FP64_PEAK_TFLOPS, 4),
                        "backward_error": dense_be(torch, A, X64, b)}}
        for facto, solver in variants:
            Ac = A.clone()
            (X, _, _, it, info), t = timed(lambda: iteref.dhgesv_iteref_gpu(Ac, b.clone(), facto=facto,
                                                                            solver=solver))
            del Ac
            fe = float((X - X64).abs().max() / X64.abs().max())
            rec[f"{facto}_{solver}"] = {"ms": round(t * 1e3, 2), "solves_per_s": round(1 / t, 3),
                                        "tflops_fp64_credited": round(flop / t / 1e12, 2), "iter": it,
                                        "info": info, "backward_error": dense_be(torch, A, X, b),
                                        "forward_error_vs_fp64": fe, "speedup_vs_fp64": round(t64 / t, 3)}
        out[f"cond_{cond:.0e}"] = rec
        del A
        torch.cuda.empty_cache()
    del U, V
    torch.cuda.empty_cache()
    return out


def dense_be(torch, A_cols, X, B):
    """max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf), FP64 on the device."""
    A = A_cols.T
    R = B.T - A @ X.T
    return float((R.abs().amax(0) / (A.abs().sum(1).max() * X.T.abs().amax(0))).max())


def adjoint_leg(torch, dev, seed, steps=5, batch=8192, n=64, nrhs=8):
    """Forward + backward of the differentiable solve, FP64 and FP32."""
    from magma_amd.autograd import BatchedSolve
    res = {"workload": f"{batch} x {n}x{n} (U[-1,1] + 8I), nrhs={nrhs}, loss = sum(W * X)", "steps": steps}
    gen = torch.Generator(device=dev)
    for dt, name in ((torch.float64, "fp64"), (torch.float32, "fp32")):
        gen.manual_seed(seed)
        A = (torch.rand((batch, n, n), dtype=dt, device=dev, generator=gen) * 2 - 1
             + 8 * torch.eye(n, dtype=dt, device=dev))
        B = torch.rand((batch, n, nrhs), dtype=dt, device=dev, generator=gen) * 2 - 1
        W = torch.rand((batch, n, nrhs), dtype=dt, device=dev, generator=gen) * 2 - 1

        def step():
            a = A.detach().requires_grad_(True)
            b = B.detach().requires_grad_(True)
            X, _ = BatchedSolve.apply(a, b)
            ga, gb = torch.autograd.grad(X, (a, b), W)
            return X, ga, gb

        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            X, ga, gb = step()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / steps
        e = 8 if dt == torch.float64 else 4
        # forward: A read + LU written, B read + X written, ipiv; backward:
        # LU read, G read + grad_B written + X read, grad_A written
        nbytes = batch * (2 * e * n * n + 2 * e * n * nrhs + 4 * n + e * n * n + 3 * e * n * nrhs + e * n * n)
        res[name] = {"systems_per_s": round(batch / t, 1), "ms_fwd_bwd": round(t * 1e3, 3),
                     "roofline": {"bound": "hbm", "achieved": round(nbytes / t / 1e9, 1), "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": round(nbytes / t / 1e9 / HBM_PEAK_GBS, 4),
                                  "algorithmic_bytes_per_step": nbytes}}
    return res


def vbatched_leg(torch, mb, dev, seed, steps=3, batch=50000, nmin=8, nmax=512, rmax=16, spot=48):
    """Variable-size solve: n_i ~ U[nmin, nmax], nrhs_i ~ U[1, rmax], bucketed
    by the plan with the autotuned variant per bucket, captured once as a
    hipGraph and replayed. Inputs are restored from pristine copies before each
    replay, outside the timed region (events bracket only the replay)."""
    import numpy as np
    from magma_amd import autotune
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, batch).astype(np.int32)
    rs = rng.integers(1, rmax + 1, batch).astype(np.int32)
    t_tune = time.perf_counter()
    variants = autotune.load_variants(ns, rs)
    t_tune = time.perf_counter() - t_tune
    A = mb.VBatch(ns, ns, device=dev)
    B = mb.VBatch(ns, rs, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    A.data.uniform_(-1.0, 1.0, generator=gen)
    B.data.uniform_(-1.0, 1.0, generator=gen)
    ipiv = mb.IpivV(ns, dev)
    info = torch.zeros(batch, dtype=torch.int32, device=dev)
    A0, B0 = A.data.clone(), B.data.clone()
    plan = mb.VbatchedPlan(ns, rs, variants=variants)
    plan.capture(A, B, ipiv, info)
    times = []
    for k in range(steps + 1):                        # replay 0 is warm-up
        A.data.copy_(A0)
        B.data.copy_(B0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        plan.replay()
        e1.record()
        torch.cuda.synchronize()
        if k:
            times.append(e0.elapsed_time(e1) / 1e3)
    t = sum(times) / len(times)
    n64, r64 = ns.astype(np.float64), rs.astype(np.float64)
    flop = float(sum(flops_getrf(x) + flops_getrs(x, r) for x, r in zip(n64, r64)))
    # bytes: A read + LU written, B read + X written, ipiv and info written
    nbytes = float((16 * (n64 * n64 + n64 * r64) + 4 * n64 + 4).sum())
    # backward error on a spot sample, FP64 numpy
    worst = 0.0
    for i in rng.choice(batch, size=min(spot, batch), replace=False):
        m, r = int(ns[i]), int(rs[i])
        a0, b0 = int(A.offs[i]), int(B.offs[i])
        Ai = A0[a0:a0 + m * m].view(m, m).T.cpu().numpy()
        Bi = B0[b0:b0 + m * r].view(r, m).T.cpu().numpy()
        Xi = B.get(int(i))
        rinf = np.abs(Bi - Ai @ Xi).sum(1).max()
        worst = max(worst, rinf / (np.abs(Ai).sum(1).max() * np.abs(Xi).sum(1).max()))
    buckets = [{"id": b["id"], "variant": autotune.VARIANT_NAMES.get(b["variant"], str(b["variant"])),
                "systems": b["count"], "max_n": b["max_n"]} for b in plan.buckets()]
    n_info = int((info != 0).sum().item())
    plan.close()
    del A, B, A0, B0, ipiv, info
    torch.cuda.empty_cache()
    return {"workload": f"dgesv_vbatched {batch} systems, n ~ U[{nmin},{nmax}], nrhs ~ U[1,{rmax}], FP64, "
                        "one hipGraph replay per step",
            "steps": steps, "ms_per_replay": round(t * 1e3, 3), "solves_per_s": round(batch / t, 1),
            "fp64_tflops": round(flop / t / 1e12, 3), "fp64_frac": round(flop / t / 1e12 / FP64_PEAK_TFLOPS, 4),
            "hbm_gbs": round(nbytes / t / 1e9, 1), "hbm_frac": round(nbytes / t / 1e9 / HBM_PEAK_GBS, 4),
            "floor_ms": {"fp64_mfma": round(flop / (FP64_PEAK_TFLOPS * 1e12) * 1e3, 3),
                         "hbm": round(nbytes / (HBM_PEAK_GBS * 1e9) * 1e3, 3)},
            "backward_error_max_spot": float(worst), "spot_systems": int(min(spot, batch)),
            "nonzero_info": n_info, "tuning_s": round(t_tune, 2), "buckets": buckets}


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

    A_host = B_host = None
    if args.cpu_sample > 0:
        m = min(args.cpu_sample, batch)
        refresh(0)
        torch.cuda.synchronize()
        A_host = sets[0]["A"][:m].cpu().numpy()
        B_host = sets[0]["B"][:m].cpu().numpy()
    del sets
    torch.cuda.empty_cache()

    mixed = None
    if args.mixed:
        print("[bench] mixed-precision leg", file=sys.stderr, flush=True)
        mixed = mixed_precision_leg(torch, mb, lib, queue, dev, args.seed)
    adjoint = None
    if args.adjoint:
        print("[bench] adjoint leg", file=sys.stderr, flush=True)
        adjoint = adjoint_leg(torch, dev, args.seed)

    dense_res = None
    if args.dense:
        print("[bench] dense mixed-precision leg", file=sys.stderr, flush=True)
        dense_res = dense_leg(torch, dev, args.seed, n=args.dense_n,
                              conds=tuple(float(c) for c in args.dense_conds.split(",") if c),
                              variants=tuple(tuple(v.split(":")) for v in args.dense_variants.split(",") if v))
    c1 = None
    if args.c1:
        print("[bench] small CPU-path leg", file=sys.stderr, flush=True)
        c1 = c1_leg(torch, mb, lib, queue, dev, args.seed)

    vb = None
    if args.vbatched:
        print("[bench] variable-size leg", file=sys.stderr, flush=True)
        vb = vbatched_leg(torch, mb, dev, args.seed)

    cpu = None
    if A_host is not None:
        print(f"[bench] cpu baseline on {A_host.shape[0]} systems", file=sys.stderr, flush=True)
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
        "mixed_precision": mixed,
        "dense_mixed_precision": dense_res,
        "small_cpu_path": c1,
        "adjoint": adjoint,
        "vbatched": vb,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
