"""Scenarios run against the CPU-emulated library (tests/emu/_build/libemu.so,
built by build_emu.py) under AddressSanitizer, through the same C entry points
and with the same inputs as the GPU tests. Every buffer comes from malloc with
its exact size, so an out-of-bounds access by a kernel is reported by ASan.
Each scenario also checks its numerical result and the process exits non-zero
on the first failure.

Run (ASan runtime preloaded, as the test does):
  LD_PRELOAD=<libclang_rt.asan-x86_64.so> ASAN_OPTIONS=detect_leaks=0 \
      python tests/emu/scenarios.py lu suite
"""
import ctypes
import sys
import time
from pathlib import Path

import numpy as np

LIB = ctypes.CDLL(str(Path(__file__).resolve().parent / "_build" / "libemu.so"))
FAILURES = []


def check(cond, what):
    if not cond:
        FAILURES.append(what)
        print("FAIL:", what, flush=True)
libc = ctypes.CDLL("libc.so.6")
libc.malloc.restype = ctypes.c_void_p
libc.malloc.argtypes = [ctypes.c_size_t]
libc.free.argtypes = [ctypes.c_void_p]
vp, i32 = ctypes.c_void_p, ctypes.c_int

PREC = {"fp32": 900, "fp16": 905, "bf16": 940, "bf16x3": 941, "fp8": 942}
SOLV = {"ir": 920, "gmres": 924, "irstrs": 920, "irdtrs": 921, "irgmstrs": 922, "irgmdtrs": 923, "gmstrs": 924,
        "gmdtrs": 925, "gmgmstrs": 926, "gmgmdtrs": 927}


class Buf:
    def __init__(self, arr):
        arr = np.ascontiguousarray(arr)
        self.dtype, self.shape = arr.dtype, arr.shape
        self.nbytes = arr.nbytes
        self.ptr = libc.malloc(max(self.nbytes, 1))
        ctypes.memmove(self.ptr, arr.ctypes.data, self.nbytes)

    def get(self):
        out = np.empty(self.shape, self.dtype)
        ctypes.memmove(out.ctypes.data, self.ptr, self.nbytes)
        return out

    def at(self, byte_off):
        return self.ptr + byte_off


def ptr_array(buf, count, stride_bytes):
    return Buf(np.array([buf.ptr + i * stride_bytes for i in range(count)], dtype=np.uint64))


def queue():
    q = vp()
    LIB.magma_queue_create_internal.argtypes = [i32, ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_char_p, i32]
    LIB.magma_queue_create_internal(0, ctypes.byref(q), b"emu", b"emu", 0)
    return q


def svd_geo(rng, n, cond):
    def haar(k):
        Q, R = np.linalg.qr(rng.standard_normal((k, k)))
        return Q * np.sign(np.diag(R))[None, :]
    sig = cond ** (-np.arange(n) / max(n - 1, 1))
    return (haar(n) * sig[None, :]) @ haar(n).T


def iteref_batched(q, A, B, facto, solver, itermax=50):
    """A: (batch, n, n) column-major storage (A[s] is the transpose of the
    row-major matrix), B: (batch, nrhs, n)."""
    batch, n, _ = A.shape
    nrhs = B.shape[1]
    bA, bB = Buf(A), Buf(B)
    bX = Buf(np.zeros_like(B))
    pA, pB, pX = ptr_array(bA, batch, n * n * 8), ptr_array(bB, batch, n * nrhs * 8), ptr_array(bX, batch, n * nrhs * 8)
    it, info = Buf(np.zeros(batch, np.int32)), Buf(np.zeros(batch, np.int32))
    fn = LIB.magma_dxgesv_iteref_batched
    fn.argtypes = [i32, i32, vp, i32, vp, i32, vp, i32, i32, i32, i32, vp, vp, i32, vp]
    rc = fn(n, nrhs, pA.ptr, n, pB.ptr, n, pX.ptr, n, PREC[facto], SOLV[solver], itermax, it.ptr, info.ptr, batch, q)
    return rc, bX.get(), it.get(), info.get(), bA.get()


def backward(A, X, B):
    n = A.shape[1]
    out = []
    for s in range(A.shape[0]):
        Ar = A[s].T
        Xr = X[s].T
        Br = B[s].T
        R = Br - Ar @ Xr
        out.append(np.abs(R).sum(0).max() / (np.abs(Ar).sum(1).max() * np.abs(Xr).sum(0).max() + 1e-300))
    return np.array(out)


def scenario_overflow(q):
    rng = np.random.default_rng(2)
    n = 40
    A = np.stack([svd_geo(rng, n, 10.0).T.copy() for _ in range(3)])
    B = rng.uniform(-1, 1, (3, 1, n))
    A[1] *= 1e300
    rc, X, it, info, _ = iteref_batched(q, A, B, "fp32", "ir")
    be = backward(A, X, B)
    print("overflow rc", rc, "iter", it, "info", info, "be", be, flush=True)
    check(rc == 0 and it[1] == -2 and it[0] >= 0 and it[2] >= 0 and be.max() < 1e-14, "overflow")


def scenario_accuracy(q, facto, solver, cond, batch=6, n=96, nrhs=2):
    rng = np.random.default_rng(int(cond) % 97 + len(facto))
    A = np.stack([svd_geo(rng, n, cond).T.copy() for _ in range(batch)])
    B = rng.uniform(-1, 1, (batch, nrhs, n))
    t0 = time.time()
    rc, X, it, info, A_after = iteref_batched(q, A, B, facto, solver, itermax=60)
    be = backward(A, X, B).max()
    print(f"acc {facto} {solver} {cond:g} rc {rc} iter {it} info {info} be {be:.2e} "
          f"A unchanged {np.array_equal(A, A_after)} ({time.time() - t0:.1f}s)", flush=True)
    check(rc == 0 and (info == 0).all() and be <= 1e-14, f"accuracy {facto} {solver} {cond:g}")
    if solver == "gmres" or facto in ("fp32", "bf16x3"):
        check((it >= 0).all(), f"converged {facto} {solver} {cond:g}")


def scenario_unchanged(q):
    rng = np.random.default_rng(1)
    A = np.stack([svd_geo(rng, 64, 1e2).T.copy() for _ in range(4)])
    B = rng.uniform(-1, 1, (4, 1, 64))
    rc, X, it, info, A2 = iteref_batched(q, A, B, "bf16x3", "ir")
    print("unchanged rc", rc, "iter", it, "A same", np.array_equal(A, A2), flush=True)
    check(rc == 0 and (it >= 0).all() and np.array_equal(A, A2), "inputs unchanged")


def scenario_singular(q):
    rng = np.random.default_rng(3)
    n = 32
    A = np.stack([svd_geo(rng, n, 10.0).T.copy() for _ in range(2)])
    B = rng.uniform(-1, 1, (2, 1, n))
    A[0][5] = 0.0
    rc, X, it, info, _ = iteref_batched(q, A, B, "fp32", "ir")
    print("singular rc", rc, "iter", it, "info", info, flush=True)
    check(rc == 0 and it[0] == -3 and info[0] == 6 and info[1] == 0, "singular")


def scenario_nonconv(q):
    rng = np.random.default_rng(4)
    n = 128
    A = np.stack([svd_geo(rng, n, 1e6).T.copy() for _ in range(2)])
    B = rng.uniform(-1, 1, (2, 1, n))
    rc, X, it, info, _ = iteref_batched(q, A, B, "bf16", "ir", itermax=2)
    be = backward(A, X, B)
    print("nonconv rc", rc, "iter", it, "be", be, flush=True)
    check(rc == 0 and (it < 0).all() and be.max() < 1e-14, "non-convergence fallback")


def scenario_dhgesv(n, facto, solver):
    rng = np.random.default_rng(n)
    A = svd_geo(rng, n, 1e4)
    b = rng.uniform(-1, 1, (n, 1))
    dA, dB = Buf(np.ascontiguousarray(A.T)), Buf(np.ascontiguousarray(b.T))
    dX = Buf(np.zeros((1, n)))
    dipiv = Buf(np.zeros(n, np.int32))
    dworkd = Buf(np.zeros(n))
    dworks = Buf(np.zeros(n * (n + 1), np.float32))
    ipiv = (ctypes.c_int * n)()
    it, info = ctypes.c_int(0), ctypes.c_int(0)
    ft = ctypes.c_double(0)
    if facto == "fp16" and solver == "gmres":
        fn = LIB.magma_dhgesv_iteref_gpu
        fn.argtypes = [i32, i32, i32, vp, i32, vp, vp, vp, i32, vp, i32, vp, vp, vp, vp]
        rc = fn(111, n, 1, dA.ptr, n, ipiv, dipiv.ptr, dB.ptr, n, dX.ptr, n, dworkd.ptr, dworks.ptr,
                ctypes.byref(it), ctypes.byref(info))
    else:
        fn = LIB.magma_dxgesv_gmres_gpu
        fn.argtypes = [i32, i32, i32, vp, i32, vp, vp, vp, i32, vp, i32, vp, vp, i32, i32, vp, vp, vp]
        rc = fn(111, n, 1, dA.ptr, n, ipiv, dipiv.ptr, dB.ptr, n, dX.ptr, n, dworkd.ptr, dworks.ptr,
                PREC[facto], SOLV[solver], ctypes.byref(it), ctypes.byref(info), ctypes.byref(ft))
    x = dX.get().T
    r = b - A @ x
    be = np.abs(r).max() / (np.abs(A).sum(1).max() * np.abs(x).max())
    print(f"dhgesv n={n} {facto} {solver} rc {rc} iter {it.value} info {info.value} be {be:.2e}", flush=True)
    check(rc == 0 and info.value == 0 and it.value >= 0 and be <= 1e-14, f"dhgesv {n} {facto} {solver}")


def suite(q):
    for cond in (1e2, 1e4):
        for solver in ("ir", "gmres"):
            for facto in ("fp32", "bf16x3", "fp16", "bf16"):
                scenario_accuracy(q, facto, solver, cond)
    scenario_unchanged(q)
    scenario_overflow(q)
    scenario_singular(q)
    scenario_nonconv(q)
    for n, facto, solver in [(300, "fp16", "gmres"), (257, "bf16", "gmres"), (200, "fp32", "ir")]:
        scenario_dhgesv(n, facto, solver)


def main():
    q = queue()
    which = sys.argv[1:] or ["overflow"]
    for w in which:
        if w == "suite":
            suite(q)
        elif w.startswith("getrf:"):
            _, m, n = w.split(":")
            scenario_getrf(q, int(m), int(n), batch=1)
        elif w.startswith("sgesv:"):
            _, n, r = w.split(":")
            scenario_gesv(q, int(n), int(r), prec="s")
        elif w == "lu":
            lu_suite(q)
        elif w == "dense":
            dense_suite()
        elif w.startswith("dgetrf:"):
            _, m, n = w.split(":")
            scenario_dense_getrf(int(m), int(n))
        elif w.startswith("sgetrfp:"):
            _, m, n, facto = w.split(":")
            scenario_dense_sgetrf_prec(int(m), int(n), facto)
        elif w.startswith("sgetrf:"):
            _, m, n = w.split(":")
            scenario_dense_getrf(int(m), int(n), prec="s")
        elif w.startswith("dgetrs:"):
            _, n, r, t = w.split(":")
            scenario_dense_getrs(int(n), int(r), int(t))
        elif w == "overflow":
            scenario_overflow(q)
        elif w == "singular":
            scenario_singular(q)
        elif w == "nonconv":
            scenario_nonconv(q)
        elif w.startswith("dhgesv:"):
            _, n, facto, solver = w.split(":")
            scenario_dhgesv(int(n), facto, solver)
        elif w.startswith("vbplan:"):
            scenario_vbplan(q, int(w.split(":")[1]))
        elif w.startswith("acc:"):
            f = w.split(":")
            scenario_accuracy(q, f[1], f[2], float(f[3]), **({"n": int(f[4]), "batch": 2} if len(f) > 4 else {}))
        else:
            raise SystemExit("unknown scenario " + w)



def lu_residual(Arow, LUcm, ipiv, m, n):
    """||P A - L U|| / ||A|| for a column-major LU of an m x n matrix."""
    LU = LUcm.T[:m, :n]
    k = min(m, n)
    L = np.tril(LU[:, :k], -1) + np.eye(m, k)
    U = np.triu(LU[:k, :])
    PA = Arow.copy()
    for j in range(k):
        p = ipiv[j] - 1
        PA[[j, p]] = PA[[p, j]]
    return np.abs(PA - L @ U).max() / max(np.abs(Arow).max(), 1e-300)


def scenario_getrf(q, m, n, batch=2, prec="d"):
    rng = np.random.default_rng(m * 1000 + n)
    dt = np.float64 if prec == "d" else np.float32
    Arow = rng.uniform(-1, 1, (batch, m, n))
    cm = np.ascontiguousarray(np.transpose(Arow, (0, 2, 1))).astype(dt)   # (batch, n, ld=m)
    bA = Buf(cm)
    k = min(m, n)
    bP = Buf(np.zeros((batch, k), np.int32))
    info = Buf(np.zeros(batch, np.int32))
    pA = ptr_array(bA, batch, m * n * cm.itemsize)
    pP = ptr_array(bP, batch, k * 4)
    fn = getattr(LIB, f"magma_{prec}getrf_batched")
    fn.argtypes = [i32, i32, vp, i32, vp, vp, i32, vp]
    rc = fn(m, n, pA.ptr, m, pP.ptr, info.ptr, batch, q)
    LU, P = bA.get(), bP.get()
    res = max(lu_residual(Arow[s], LU[s].astype(np.float64), P[s], m, n) for s in range(batch))
    print(f"{prec}getrf {m}x{n} rc {rc} info {info.get()} res {res:.2e}", flush=True)
    check(rc == 0 and res < (1e-13 if prec == "d" else 1e-4), f"{prec}getrf {m}x{n}")


def scenario_gesv(q, n, nrhs, batch=2, prec="d"):
    rng = np.random.default_rng(n * 7 + nrhs)
    dt = np.float64 if prec == "d" else np.float32
    Arow = rng.uniform(-1, 1, (batch, n, n))
    Brow = rng.uniform(-1, 1, (batch, n, nrhs))
    bA = Buf(np.ascontiguousarray(np.transpose(Arow, (0, 2, 1))).astype(dt))
    bB = Buf(np.ascontiguousarray(np.transpose(Brow, (0, 2, 1))).astype(dt))
    bP = Buf(np.zeros((batch, n), np.int32))
    info = Buf(np.zeros(batch, np.int32))
    isz = np.dtype(dt).itemsize
    pA, pB, pP = ptr_array(bA, batch, n * n * isz), ptr_array(bB, batch, n * nrhs * isz), ptr_array(bP, batch, n * 4)
    fn = getattr(LIB, f"magma_{prec}gesv_batched")
    fn.argtypes = [i32, i32, vp, i32, vp, vp, i32, vp, i32, vp]
    rc = fn(n, nrhs, pA.ptr, n, pP.ptr, pB.ptr, n, info.ptr, batch, q)
    X = np.transpose(bB.get().astype(np.float64), (0, 2, 1))
    be = max(np.abs(Brow[s] - Arow[s] @ X[s]).max() / (np.abs(Arow[s]).sum(1).max() * np.abs(X[s]).max())
             for s in range(batch))
    print(f"{prec}gesv n={n} nrhs={nrhs} rc {rc} info {info.get()} be {be:.2e}", flush=True)
    check(rc == 0 and be < (1e-14 if prec == "d" else 1e-5), f"{prec}gesv {n} {nrhs}")


def scenario_getrs(q, n, nrhs, trans, batch=2):
    rng = np.random.default_rng(n * 11 + nrhs + trans)
    Arow = rng.uniform(-1, 1, (batch, n, n))
    Brow = rng.uniform(-1, 1, (batch, n, nrhs))
    bA = Buf(np.ascontiguousarray(np.transpose(Arow, (0, 2, 1))))
    bP = Buf(np.zeros((batch, n), np.int32))
    info = Buf(np.zeros(batch, np.int32))
    pA, pP = ptr_array(bA, batch, n * n * 8), ptr_array(bP, batch, n * 4)
    LIB.magma_dgetrf_batched.argtypes = [i32, i32, vp, i32, vp, vp, i32, vp]
    LIB.magma_dgetrf_batched(n, n, pA.ptr, n, pP.ptr, info.ptr, batch, q)
    bB = Buf(np.ascontiguousarray(np.transpose(Brow, (0, 2, 1))))
    pB = ptr_array(bB, batch, n * nrhs * 8)
    fn = LIB.magma_dgetrs_batched
    fn.argtypes = [i32, i32, i32, vp, i32, vp, vp, i32, i32, vp]
    rc = fn(112 if trans else 111, n, nrhs, pA.ptr, n, pP.ptr, pB.ptr, n, batch, q)
    X = np.transpose(bB.get(), (0, 2, 1))
    Aop = np.transpose(Arow, (0, 2, 1)) if trans else Arow
    be = max(np.abs(Brow[s] - Aop[s] @ X[s]).max() / (np.abs(Aop[s]).sum(1).max() * np.abs(X[s]).max())
             for s in range(batch))
    print(f"dgetrs n={n} nrhs={nrhs} trans={trans} rc {rc} be {be:.2e}", flush=True)
    check(rc == 0 and be < 1e-14, f"dgetrs {n} {nrhs} {trans}")


def scenario_vbplan(q, variant, ns=(129, 140, 200, 256, 177, 64, 30, 300, 290), rs=(1, 3, 16, 2, 5, 4, 1, 12, 3)):
    """Variable-size plan executed with one kernel choice for every bucket
    (bucket-invalid choices fall back to the built-in rule)."""
    rng = np.random.default_rng(variant * 13 + len(ns))
    batch = len(ns)
    mats, rhs, bA, bB = [], [], [], []
    for n, r in zip(ns, rs):
        M = rng.uniform(-1, 1, (n, n)); R = rng.uniform(-1, 1, (n, r))
        mats.append(M); rhs.append(R)
        bA.append(Buf(np.ascontiguousarray(M.T))); bB.append(Buf(np.ascontiguousarray(R.T)))
    bP = [Buf(np.zeros(n, np.int32)) for n in ns]
    pA = Buf(np.array([x.ptr for x in bA], dtype=np.uint64))
    pB = Buf(np.array([x.ptr for x in bB], dtype=np.uint64))
    pP = Buf(np.array([x.ptr for x in bP], dtype=np.uint64))
    dn = Buf(np.array(ns, np.int32)); dr = Buf(np.array(rs, np.int32)); info = Buf(np.zeros(batch, np.int32))
    hn = np.array(ns, np.int32); hr = np.array(rs, np.int32); hv = np.full(22, variant, np.int32)
    plan = vp()
    LIB.magma_dgesv_vbatched_plan_create.argtypes = [vp, vp, i32, vp, vp, ctypes.POINTER(vp)]
    rc = LIB.magma_dgesv_vbatched_plan_create(hn.ctypes.data, hr.ctypes.data, batch, hv.ctypes.data, q,
                                              ctypes.byref(plan))
    check(rc == 0, f"vbplan create {variant}")
    LIB.magma_dgesv_vbatched_plan_exec.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    rc = LIB.magma_dgesv_vbatched_plan_exec(plan, dn.ptr, dr.ptr, pA.ptr, dn.ptr, pP.ptr, pB.ptr, dn.ptr,
                                            info.ptr, q)
    LIB.magma_queue_sync_internal.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, i32]
    LIB.magma_queue_sync_internal(q, b"", b"", 0)
    LIB.magma_vbatched_plan_destroy.argtypes = [vp]
    LIB.magma_vbatched_plan_destroy(plan)
    worst = 0.0
    for i, (n, r) in enumerate(zip(ns, rs)):
        X = bB[i].get().T
        be = np.abs(rhs[i] - mats[i] @ X).sum(1).max() / (np.abs(mats[i]).sum(1).max() * np.abs(X).sum(1).max())
        worst = max(worst, be)
    print(f"vbplan variant {variant} rc {rc} info {info.get()} be {worst:.2e}", flush=True)
    check(rc == 0 and (info.get() == 0).all() and worst < 1e-14, f"vbplan {variant}")


def lu_suite(q):
    for n in (1, 5, 16, 17, 33, 40, 48, 64, 96, 100, 127, 128):
        scenario_getrf(q, n, n)
    for m, n in ((129, 129), (150, 150), (256, 256), (300, 300), (40, 24), (24, 40), (96, 64), (33, 7), (7, 33)):
        scenario_getrf(q, m, n)
    for n, r in ((32, 1), (64, 2), (100, 5), (128, 4), (128, 16), (129, 3), (200, 2)):
        scenario_gesv(q, n, r)
    for n, r in ((8, 3), (33, 4), (128, 9), (200, 2)):
        for t in (0, 1):
            scenario_getrs(q, n, r, t)
    for n, r in ((16, 1), (64, 3), (128, 4), (129, 2), (200, 2), (256, 3), (300, 5)):
        scenario_gesv(q, n, r, prec="s")
    for m, n in ((60, 60), (200, 200), (300, 300)):
        scenario_getrf(q, m, n, prec="s")


def np_getf2(Arow):
    """Unblocked LU with partial pivoting (LAPACK dgetf2 order, ties to the
    first index): returns ipiv (1-based)."""
    A = Arow.astype(np.float64).copy()
    m, n = A.shape
    ipiv = np.zeros(min(m, n), np.int32)
    for j in range(min(m, n)):
        p = j + int(np.argmax(np.abs(A[j:, j])))
        ipiv[j] = p + 1
        if A[p, j] != 0:
            A[[j, p]] = A[[p, j]]
            A[j + 1:, j] *= 1.0 / A[j, j]
        A[j + 1:, j + 1:] -= np.outer(A[j + 1:, j], A[j, j + 1:])
    return ipiv


def scenario_dense_getrf(m, n, prec="d"):
    """magma_[ds]getrf_gpu on one m x n matrix: LU residual, pivots against
    the unblocked LU on a well separated input."""
    rng = np.random.default_rng(m * 31 + n)
    dt = np.float64 if prec == "d" else np.float32
    Arow = rng.uniform(-1, 1, (m, n))
    bA = Buf(np.ascontiguousarray(Arow.T).astype(dt))
    k = min(m, n)
    ipiv = (ctypes.c_int * max(k, 1))()
    info = ctypes.c_int(-7)
    fn = getattr(LIB, f"magma_{prec}getrf_gpu")
    fn.argtypes = [i32, i32, vp, i32, vp, ctypes.POINTER(ctypes.c_int)]
    rc = fn(m, n, bA.ptr, m, ipiv, ctypes.byref(info))
    P = np.array(ipiv[:k], np.int32)
    res = lu_residual(Arow, bA.get().astype(np.float64), P, m, n)
    same = np.array_equal(P, np_getf2(Arow)) if prec == "d" else True
    print(f"dense {prec}getrf {m}x{n} rc {rc} info {info.value} res {res:.2e} pivots-equal {same}", flush=True)
    check(rc == 0 and info.value == 0 and res < (3e-16 * max(m, n, 300) if prec == "d" else 2e-4) and same,
          f"dense {prec}getrf {m}x{n}")


def scenario_dense_sgetrf_prec(m, n, facto):
    """magma_sgetrf_prec_gpu with 16-bit trailing updates: LU residual at
    the level of the update format."""
    rng = np.random.default_rng(m * 13 + n + len(facto))
    Arow = rng.uniform(-1, 1, (m, n))
    bA = Buf(np.ascontiguousarray(Arow.T).astype(np.float32))
    k = min(m, n)
    ipiv = (ctypes.c_int * max(k, 1))()
    info = ctypes.c_int(-7)
    LIB.magma_sgetrf_prec_gpu.argtypes = [i32, i32, vp, i32, vp, ctypes.POINTER(ctypes.c_int), i32]
    rc = LIB.magma_sgetrf_prec_gpu(m, n, bA.ptr, m, ipiv, ctypes.byref(info), PREC[facto])
    P = np.array(ipiv[:k], np.int32)
    res = lu_residual(Arow, bA.get().astype(np.float64), P, m, n)
    # the residual grows like the update format's unit roundoff times n
    # (measured ratio res / (u n) is 0.02-0.15 for all three formats); an
    # indexing mistake in the update gives O(1) or worse
    u = {"fp32": 2.0**-24, "bf16x3": 2.0**-16, "fp16": 2.0**-11, "bf16": 2.0**-8, "fp8": 2.0**-4}[facto]
    tol = max(2e-4, 0.5 * u * max(m, n))
    print(f"dense sgetrf {facto} {m}x{n} rc {rc} info {info.value} res {res:.2e}", flush=True)
    check(rc == 0 and info.value == 0 and res < tol, f"dense sgetrf {facto} {m}x{n}")


def scenario_dense_getrs(n, nrhs, trans):
    rng = np.random.default_rng(n * 17 + nrhs + trans)
    Arow = rng.uniform(-1, 1, (n, n))
    Brow = rng.uniform(-1, 1, (n, nrhs))
    bA = Buf(np.ascontiguousarray(Arow.T))
    ipiv = (ctypes.c_int * n)()
    info = ctypes.c_int(0)
    LIB.magma_dgetrf_gpu.argtypes = [i32, i32, vp, i32, vp, ctypes.POINTER(ctypes.c_int)]
    LIB.magma_dgetrf_gpu(n, n, bA.ptr, n, ipiv, ctypes.byref(info))
    bB = Buf(np.ascontiguousarray(Brow.T))
    LIB.magma_dgetrs_gpu.argtypes = [i32, i32, i32, vp, i32, vp, vp, i32, ctypes.POINTER(ctypes.c_int)]
    rc = LIB.magma_dgetrs_gpu(112 if trans else 111, n, nrhs, bA.ptr, n, ipiv, bB.ptr, n, ctypes.byref(info))
    X = bB.get().T
    Aop = Arow.T if trans else Arow
    be = np.abs(Brow - Aop @ X).max() / (np.abs(Aop).sum(1).max() * np.abs(X).max())
    print(f"dense dgetrs n={n} nrhs={nrhs} trans={trans} rc {rc} be {be:.2e}", flush=True)
    check(rc == 0 and be < 1e-14, f"dense dgetrs {n} {nrhs} {trans}")


def dense_suite():
    for m, n in ((300, 300), (600, 600), (520, 300), (300, 520), (777, 777)):
        scenario_dense_getrf(m, n)
    for m, n in ((600, 600), (700, 450)):
        scenario_dense_getrf(m, n, prec="s")
    for n, r, t in ((600, 1, 0), (600, 3, 1), (333, 9, 0), (333, 2, 1)):
        scenario_dense_getrs(n, r, t)
    for n, facto, solver in ((600, "bf16", "gmres"), (600, "fp16", "ir"), (700, "bf16x3", "ir")):
        scenario_dhgesv(n, facto, solver)


if __name__ == "__main__":
    main()
    if FAILURES:
        print("failed:", FAILURES, flush=True)
        sys.exit(1)
    print("all scenarios passed", flush=True)
