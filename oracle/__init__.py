"""ORACLE -- test infrastructure only.

Python access to the C restatement of LAPACK's LU semantics
(``lapack_oracle.c``) and to the host-LAPACK CPU baseline (``cpu_baseline.c``).
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, always as the checker; the product path
(``magma_amd``) never does.

Batched storage convention (same as the GPU side): a batch of m-by-n
column-major matrices with leading dimension ld is a C-contiguous array of
shape ``(batch, n, ld)``; element (i, j) of matrix s is ``A[s, j, i]``.

The error metrics restate MAGMA's testers:
  * ``gesv_backward_error``: ||B - A X||_inf / (n ||A||_inf ||X||_inf)
    (testing/testing_zgesv_batched.cpp:133-151);
  * ``lu_residual``: ||P A - L U||_F / (||A||_F n)
    (testing/testing_zgetrf_batched.cpp:41-81);
  * ``lu_forward_error``: ||LU_ref - LU||_F / (||LU_ref||_F n)
    (testing/testing_zgetrf_batched.cpp:86-113).
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = HERE / "liboracle.so"
_CB = HERE / "libcpubase.so"

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)


def build() -> None:
    """Compile the oracle libraries with the committed Makefile."""
    subprocess.run(["make", "-B", "-C", str(HERE)], check=True, capture_output=True)


def _load(path: Path) -> ctypes.CDLL:
    if not path.exists():
        build()
    return ctypes.CDLL(str(path))


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        L = _load(_LIB)
        L.orc_dgetf2.restype = ctypes.c_int
        L.orc_dgetf2.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int, _i32p, ctypes.c_int]
        L.orc_dgetf2_batch.restype = None
        L.orc_dgetf2_batch.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_long,
                                       _i32p, _i32p, ctypes.c_int]
        L.orc_dgetrs_batch.restype = None
        L.orc_dgetrs_batch.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int,
                                       ctypes.c_long, _i32p, _f64p, ctypes.c_int, ctypes.c_long, ctypes.c_int]
        L.orc_swp2pswp.restype = None
        L.orc_swp2pswp.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p]
        L.orc_flops_dgetrf.restype = ctypes.c_double
        L.orc_flops_dgetrf.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_flops_dgetrs.restype = ctypes.c_double
        L.orc_flops_dgetrs.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_idamax.restype = ctypes.c_int
        L.orc_idamax.argtypes = [ctypes.c_int, _f64p, ctypes.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def dgetrf(A: np.ndarray, m: int | None = None):
    """LU of a batch. A: (batch, n, ld) float64 column-major per matrix, or a
    single (n, ld) matrix. Returns (LU, ipiv 1-based int32 (batch, min(m,n)),
    info (batch,))."""
    single = A.ndim == 2
    A3 = np.ascontiguousarray(A[None] if single else A, dtype=np.float64).copy()
    batch, n, ld = A3.shape
    m = ld if m is None else m
    minmn = min(m, n)
    ipiv = np.zeros((batch, max(minmn, 1)), dtype=np.int32)
    info = np.zeros(batch, dtype=np.int32)
    if batch and minmn:
        lib().orc_dgetf2_batch(m, n, _p(A3, _f64p), ld, n * ld, _p(ipiv, _i32p), _p(info, _i32p), batch)
    ipiv = ipiv[:, :minmn]
    if single:
        return A3[0], ipiv[0], int(info[0])
    return A3, ipiv, info


def dgetrs(trans: int, LU: np.ndarray, ipiv: np.ndarray, B: np.ndarray):
    """Solve with LU factors. LU: (batch, n, lda), ipiv: (batch, n),
    B: (batch, nrhs, ldb). trans 0 = 'N', 1 = 'T'. Returns X (new array)."""
    LU3 = np.ascontiguousarray(LU, dtype=np.float64)
    piv = np.ascontiguousarray(ipiv, dtype=np.int32)
    X = np.ascontiguousarray(B, dtype=np.float64).copy()
    batch, n, lda = LU3.shape
    _, nrhs, ldb = X.shape
    if batch and n and nrhs:
        lib().orc_dgetrs_batch(int(trans), n, nrhs, _p(LU3, _f64p), lda, n * lda, _p(piv, _i32p),
                               _p(X, _f64p), ldb, nrhs * ldb, batch)
    return X


def dgesv(A: np.ndarray, B: np.ndarray):
    """Returns (LU, ipiv, info, X) for a batch."""
    LU, ipiv, info = dgetrf(A)
    X = dgetrs(0, LU, ipiv, B)
    return LU, ipiv, info, X


def swp2pswp(trans: int, ipiv: np.ndarray) -> np.ndarray:
    piv = np.ascontiguousarray(ipiv, dtype=np.int32)
    out = np.zeros_like(piv)
    lib().orc_swp2pswp(int(trans), len(piv), _p(piv, _i32p), _p(out, _i32p))
    return out


def idamax(x: np.ndarray) -> int:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return int(lib().orc_idamax(len(x), _p(x, _f64p), 1))


def flops_dgetrf(m, n) -> float:
    return float(lib().orc_flops_dgetrf(float(m), float(n)))


def flops_dgetrs(n, nrhs) -> float:
    return float(lib().orc_flops_dgetrs(float(n), float(nrhs)))


# --------------------------------------------------------------------------
# Metrics (numpy; matrices as (batch, n, ld) column-major)
# --------------------------------------------------------------------------
def to_rowmajor(A: np.ndarray, m: int | None = None) -> np.ndarray:
    """(batch, n, ld) column-major -> (batch, m, n) ordinary arrays."""
    m = A.shape[-1] if m is None else m
    return np.swapaxes(A[..., :m], -1, -2)


def from_rowmajor(M: np.ndarray, ld: int | None = None) -> np.ndarray:
    b, m, n = M.shape
    ld = m if ld is None else ld
    out = np.zeros((b, n, ld), dtype=M.dtype)
    out[..., :m] = np.swapaxes(M, -1, -2)
    return out


def perm_from_ipiv(ipiv: np.ndarray, m: int) -> np.ndarray:
    """Row map with (P^T A)[i] = A[perm[i]] for each batch entry."""
    ipiv = np.atleast_2d(ipiv)
    batch, k = ipiv.shape
    perm = np.tile(np.arange(m), (batch, 1))
    rows = np.arange(batch)
    for j in range(k):
        p = ipiv[:, j] - 1
        a = perm[rows, j].copy()
        perm[rows, j] = perm[rows, p]
        perm[rows, p] = a
    return perm


def gesv_backward_error(A: np.ndarray, X: np.ndarray, B: np.ndarray, n: int) -> np.ndarray:
    """Per-system ||B - A X||_inf / (n ||A||_inf ||X||_inf); A (batch, n, lda),
    X and B (batch, nrhs, ldb)."""
    Ar = to_rowmajor(A, n)
    Xr = to_rowmajor(X, n)
    Br = to_rowmajor(B, n)
    R = Br - Ar @ Xr
    Rn = np.abs(R).sum(axis=2).max(axis=1)
    An = np.abs(Ar).sum(axis=2).max(axis=1)
    Xn = np.abs(Xr).sum(axis=2).max(axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        return Rn / (n * An * Xn)


def lu_residual(A: np.ndarray, LU: np.ndarray, ipiv: np.ndarray, m: int, n: int) -> np.ndarray:
    """Per-matrix ||P^T A - L U||_F / (||A||_F n)."""
    Ar = to_rowmajor(A, m)
    Fr = to_rowmajor(LU, m)
    k = min(m, n)
    L = np.tril(Fr[:, :, :k], -1) + np.eye(m, k)[None]
    U = np.triu(Fr[:, :k, :])
    perm = perm_from_ipiv(ipiv, m)
    PA = np.take_along_axis(Ar, perm[:, :, None], axis=1)
    num = np.linalg.norm(PA - L @ U, axis=(1, 2))
    den = np.linalg.norm(Ar, axis=(1, 2)) * n
    with np.errstate(divide="ignore", invalid="ignore"):
        return num / den


def lu_forward_error(LU_ref: np.ndarray, LU: np.ndarray, m: int) -> np.ndarray:
    R = to_rowmajor(LU_ref, m)
    G = to_rowmajor(LU, m)
    n = R.shape[2]
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.linalg.norm(G - R, axis=(1, 2)) / (np.linalg.norm(R, axis=(1, 2)) * n)


def pivot_margin(A: np.ndarray, m: int | None = None) -> np.ndarray:
    """Smallest relative gap, over all elimination steps, between the largest
    and second-largest |candidate| of the pivot column (same recurrence as
    orc_dgetf2, in numpy). Inputs whose margin is far above rounding level are
    'tie-free': any backward-stable LU picks the same pivots on them."""
    Ar = to_rowmajor(np.asarray(A, dtype=np.float64), m).copy()
    batch, mm, n = Ar.shape
    rows = np.arange(batch)
    margin = np.full(batch, np.inf)
    for k in range(min(mm, n)):
        col = np.abs(Ar[:, k:, k])
        if col.shape[1] >= 2:
            part = -np.partition(-col, 1, axis=1)
            top, second = part[:, 0], part[:, 1]
            with np.errstate(divide="ignore", invalid="ignore"):
                g = np.where(top > 0, (top - second) / top, 0.0)
            margin = np.minimum(margin, g)
        p = k + np.argmax(col, axis=1)
        tmp = Ar[rows, k, :].copy()
        Ar[rows, k, :] = Ar[rows, p, :]
        Ar[rows, p, :] = tmp
        piv = Ar[:, k, k]
        safe = np.where(piv != 0, piv, 1.0)
        l = Ar[:, k + 1:, k] / safe[:, None]
        l = np.where((piv != 0)[:, None], l, 0.0)
        Ar[:, k + 1:, k] = l
        Ar[:, k + 1:, k + 1:] -= l[:, :, None] * Ar[:, k, None, k + 1:]
    return margin


# --------------------------------------------------------------------------
# Host LAPACK (the reference's CPU path and oracle of its testers)
# --------------------------------------------------------------------------
def lapack_candidates() -> list[str]:
    cands = []
    env = os.environ.get("MAGMA_MI355X_LAPACK")
    if env:
        cands.append(env)
    cands += ["/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so"]
    try:
        import scipy  # noqa: F401
        d = Path(scipy.__file__).resolve().parent.parent / "scipy.libs"
        cands += sorted(glob.glob(str(d / "libscipy_openblas*.so")))
    except Exception:
        pass
    cands += ["liblapack.so.3", "libopenblas.so.0"]
    return cands


_cb = None


def host_lapack() -> ctypes.CDLL | None:
    """Open a host LAPACK; returns the baseline library or None."""
    global _cb
    if _cb is not None:
        return _cb
    L = _load(_CB)
    L.cb_open.restype = ctypes.c_int
    L.cb_open.argtypes = [ctypes.c_char_p]
    L.cb_path.restype = ctypes.c_char_p
    L.cb_max_threads.restype = ctypes.c_int
    L.cb_dgesv_batch.restype = ctypes.c_double
    L.cb_dgesv_batch.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int, _f64p, ctypes.c_int,
                                 _i32p, _i32p, ctypes.c_int, ctypes.c_int]
    L.cb_dgetrf.restype = ctypes.c_int
    L.cb_dgetrf.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int, _i32p]
    L.cb_dgetrs.restype = ctypes.c_int
    L.cb_dgetrs.argtypes = [ctypes.c_char, ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int, _i32p,
                            _f64p, ctypes.c_int]
    for c in lapack_candidates():
        if L.cb_open(c.encode()) == 0:
            _cb = L
            return L
    return None


def host_dgesv_batch(A: np.ndarray, B: np.ndarray, nthreads: int):
    """Host LAPACK dgesv over a batch (copies). Returns (LU, ipiv, info, X, seconds)."""
    L = host_lapack()
    if L is None:
        raise RuntimeError("no host LAPACK found")
    A3 = np.ascontiguousarray(A, dtype=np.float64).copy()
    X = np.ascontiguousarray(B, dtype=np.float64).copy()
    batch, n, lda = A3.shape
    _, nrhs, ldb = X.shape
    ipiv = np.zeros((batch, n), dtype=np.int32)
    info = np.zeros(batch, dtype=np.int32)
    secs = L.cb_dgesv_batch(n, nrhs, _p(A3, _f64p), lda, _p(X, _f64p), ldb, _p(ipiv, _i32p),
                            _p(info, _i32p), batch, int(nthreads))
    return A3, ipiv, info, X, secs
