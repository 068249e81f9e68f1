"""Mixed-precision solves with FP64 refinement (magma_dxgesv_iteref_batched,
magma_dhgesv_iteref_gpu, magma_dxgesv_gmres_gpu) on torch device buffers."""
from __future__ import annotations

import ctypes

import torch

from . import _ffi
from .batched import MagmaError, Queue, _check_batch, default_queue, pointer_array

FACTO = {
    "fp32": _ffi.Magma_PREC_SS,
    "fp16": _ffi.Magma_PREC_SHT,
    "bf16": _ffi.Magma_PREC_SBT,
    "bf16x3": _ffi.Magma_PREC_SB3T,
    "fp8": _ffi.Magma_PREC_SF8T,        # experiment: one-system drivers only
}
# short names of the refinement solvers (include/magma_types.h Magma_REFINE_*):
# "ir" IRSTRS, "gmres" GMSTRS, and the reference's own names in lower case
SOLVER = {"ir": _ffi.Magma_REFINE_IRSTRS, "gmres": _ffi.Magma_REFINE_GMSTRS,
          "irstrs": _ffi.Magma_REFINE_IRSTRS, "irdtrs": _ffi.Magma_REFINE_IRDTRS,
          "irgmstrs": _ffi.Magma_REFINE_IRGMSTRS, "irgmdtrs": _ffi.Magma_REFINE_IRGMDTRS,
          "gmstrs": _ffi.Magma_REFINE_GMSTRS, "gmdtrs": _ffi.Magma_REFINE_GMDTRS,
          "gmgmstrs": _ffi.Magma_REFINE_GMGMSTRS, "gmgmdtrs": _ffi.Magma_REFINE_GMGMDTRS}


def dxgesv_iteref_batched(A: torch.Tensor, B: torch.Tensor, X: torch.Tensor | None = None, facto: str = "bf16",
                          solver: str = "gmres", itermax: int = 50, queue: Queue | None = None):
    """Solve A_s X_s = B_s (batch, n, ld) with a low-precision LU and FP64
    refinement. A and B are read only unless a system falls back to FP64 (its
    A then holds the FP64 LU). Returns (X, iter (batch,), info (batch,))."""
    _check_batch(A, "A")
    _check_batch(B, "B")
    batch, n, ldda = A.shape
    _, nrhs, lddb = B.shape
    queue = queue or default_queue(A.device)
    if X is None:
        X = torch.empty_like(B)
    it = torch.zeros(batch, dtype=torch.int32, device=A.device)
    info = torch.zeros(batch, dtype=torch.int32, device=A.device)
    pA, pB, pX = pointer_array(A, queue), pointer_array(B, queue), pointer_array(X, queue)
    rc = _ffi.load().magma_dxgesv_iteref_batched(n, nrhs, pA.data_ptr(), ldda, pB.data_ptr(), lddb, pX.data_ptr(),
                                                 X.shape[2], FACTO[facto], SOLVER[solver], itermax,
                                                 it.data_ptr(), info.data_ptr(), batch, queue.handle)
    if rc < 0:
        raise MagmaError("magma_dxgesv_iteref_batched", rc)
    return X, it, info


def dhgesv_iteref_gpu(A: torch.Tensor, B: torch.Tensor, facto: str = "fp16", solver: str = "gmres",
                      trans: int = _ffi.MagmaNoTrans):
    """One system: A (n, ldda) column-major as a (n_cols=n, ldda) tensor, B
    (nrhs, lddb). Mirrors magma_dhgesv_iteref_gpu / magma_dxgesv_gmres_gpu.
    Returns (X, ipiv (host list), dipiv (device), iter, info)."""
    n, ldda = A.shape
    nrhs, lddb = B.shape
    X = torch.zeros_like(B)
    dipiv = torch.zeros(max(n, 1), dtype=torch.int32, device=A.device)
    dworkd = torch.zeros(max(n * nrhs, 1), dtype=torch.float64, device=A.device)
    dworks = torch.zeros(max(n * (n + nrhs), 1), dtype=torch.float32, device=A.device)
    ipiv = (ctypes.c_int * max(n, 1))()
    it, info = ctypes.c_int(0), ctypes.c_int(0)
    lib = _ffi.load()
    torch.cuda.synchronize(A.device)
    if facto == "fp32" and solver == "gmres":
        rc = lib.magma_dsgesv_iteref_gpu(trans, n, nrhs, A.data_ptr(), ldda, ipiv, dipiv.data_ptr(), B.data_ptr(),
                                         lddb, X.data_ptr(), lddb, dworkd.data_ptr(), dworks.data_ptr(),
                                         ctypes.byref(it), ctypes.byref(info))
    elif facto == "fp16" and solver == "gmres":
        rc = lib.magma_dhgesv_iteref_gpu(trans, n, nrhs, A.data_ptr(), ldda, ipiv, dipiv.data_ptr(), B.data_ptr(),
                                         lddb, X.data_ptr(), lddb, dworkd.data_ptr(), dworks.data_ptr(),
                                         ctypes.byref(it), ctypes.byref(info))
    else:
        ft = ctypes.c_double(0)
        rc = lib.magma_dxgesv_gmres_gpu(trans, n, nrhs, A.data_ptr(), ldda, ipiv, dipiv.data_ptr(), B.data_ptr(),
                                        lddb, X.data_ptr(), lddb, dworkd.data_ptr(), dworks.data_ptr(),
                                        FACTO[facto], SOLVER[solver], ctypes.byref(it), ctypes.byref(info),
                                        ctypes.byref(ft))
    if rc < 0:
        raise MagmaError("magma_dxgesv_gmres_gpu", rc)
    return X, list(ipiv)[:n], dipiv, it.value, info.value
