"""One large matrix on the whole device: magma_[ds]getrf_gpu,
magma_dgetrf_native and magma_[ds]getrs_gpu on torch device buffers.

A matrix of m rows and n columns is a (n, ld) tensor holding its columns
(column-major, as the C ABI expects), ld >= m. Pivots are host int32 arrays
(LAPACK interchanges, 1-based), as in MAGMA's non-batched interface."""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _ffi
from .batched import MagmaError


def _i32(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _check_dense(A: torch.Tensor, dtype: torch.dtype, name: str):
    if A.dim() != 2 or not A.is_cuda or A.dtype != dtype or not A.is_contiguous():
        raise ValueError(f"{name} must be a contiguous 2-D {dtype} CUDA tensor (columns, ld)")


def getrf_gpu(A: torch.Tensor, m: int | None = None, native: bool = False):
    """LU of the m x n matrix held in A (n, ld) in place. Returns (ipiv, info)."""
    dt = A.dtype
    _check_dense(A, dt, "A")
    n, ld = A.shape
    m = ld if m is None else m
    ipiv = np.zeros(max(min(m, n), 1), np.int32)
    info = ctypes.c_int(0)
    lib = _ffi.load()
    if dt == torch.float64:
        fn = lib.magma_dgetrf_native if native else lib.magma_dgetrf_gpu
    elif dt == torch.float32:
        fn = lib.magma_sgetrf_gpu
    else:
        raise ValueError("float64 or float32 only")
    torch.cuda.synchronize(A.device)
    rc = fn(m, n, A.data_ptr(), ld, _i32(ipiv), ctypes.byref(info))
    if rc < 0:
        raise MagmaError(fn.__name__, rc)
    return ipiv[:min(m, n)], info.value


def sgetrf_prec_gpu(A: torch.Tensor, facto: str = "bf16", m: int | None = None):
    """FP32 LU of A (n, ld) in place with the trailing-update format facto
    (fp32, fp16, bf16, bf16x3): magma_sgetrf_prec_gpu; "fp16" is MAGMA's
    magma_xshgetrf_gpu with tensor-core math. Returns (ipiv, info)."""
    from .iteref import FACTO
    _check_dense(A, torch.float32, "A")
    n, ld = A.shape
    m = ld if m is None else m
    ipiv = np.zeros(max(min(m, n), 1), np.int32)
    info = ctypes.c_int(0)
    lib = _ffi.load()
    torch.cuda.synchronize(A.device)
    if facto == "fp16":
        rc = lib.magma_xshgetrf_gpu(m, n, A.data_ptr(), ld, _i32(ipiv), ctypes.byref(info),
                                    _ffi.Magma_MP_ENABLE_TC_MATH, _ffi.Magma_MP_GEMEX_I16_O32_C32)
    else:
        rc = lib.magma_sgetrf_prec_gpu(m, n, A.data_ptr(), ld, _i32(ipiv), ctypes.byref(info), FACTO[facto])
    if rc < 0:
        raise MagmaError("magma_sgetrf_prec_gpu", rc)
    return ipiv[:min(m, n)], info.value


def getrs_gpu(A: torch.Tensor, ipiv: np.ndarray, B: torch.Tensor, trans: int = _ffi.MagmaNoTrans):
    """Solve op(A) X = B with the factors of getrf_gpu; B (nrhs, ldb) is
    overwritten by X. Returns info."""
    dt = A.dtype
    _check_dense(A, dt, "A")
    _check_dense(B, dt, "B")
    n, ld = A.shape
    nrhs, ldb = B.shape
    piv = np.ascontiguousarray(ipiv, dtype=np.int32)
    info = ctypes.c_int(0)
    lib = _ffi.load()
    fn = lib.magma_dgetrs_gpu if dt == torch.float64 else lib.magma_sgetrs_gpu
    torch.cuda.synchronize(A.device)
    rc = fn(trans, n, nrhs, A.data_ptr(), ld, _i32(piv), B.data_ptr(), ldb, ctypes.byref(info))
    if rc < 0:
        raise MagmaError(fn.__name__, rc)
    return info.value
