"""Differentiable batched solve X = A^{-1} B on the library's LU (FP64 or FP32).

Forward: one batched LU of A (partial pivoting, LAPACK semantics) and one
batched triangular solve. Backward, for a loss with dL/dX = G:
    grad_B = A^{-T} G            (a transposed solve with the stored LU + ipiv)
    grad_A = -grad_B X^T         (a rank-nrhs outer product per system,
                                  magma_adjoint_outer_batched)
No second factorisation is done: the factors saved by the forward pass are
reused (the adjoint of src/zgesv_batched.cpp's getrf + getrs).

Tensors use torch's row-major convention: A is (batch, n, n), B and X are
(batch, n, nrhs). A row-major matrix read column-major is its transpose, so
the forward pass factors the column-major A^T held by a copy of A's memory
(A^T = P L U) and solves with op = transpose; the backward's A^{-T} G is then
the non-transposed solve on the same factors.
"""
from __future__ import annotations

import torch

from . import _ffi
from .batched import MagmaError, default_queue, dgetrf_batched, dgetrs_batched, sgetrf_batched, sgetrs_batched


def _getrf(LU: torch.Tensor):
    if LU.dtype == torch.float64:
        return dgetrf_batched(LU)
    if LU.dtype == torch.float32:
        return sgetrf_batched(LU)
    raise TypeError(f"unsupported dtype {LU.dtype} (float64 or float32)")


def _getrs(trans: int, LU: torch.Tensor, ipiv: torch.Tensor, B_cm: torch.Tensor) -> None:
    (dgetrs_batched if LU.dtype == torch.float64 else sgetrs_batched)(trans, LU, ipiv, B_cm)


def _outer(gB_cm: torch.Tensor, X_cm: torch.Tensor) -> torch.Tensor:
    """-grad_B X^T (row-major (batch, n, n)) from the column-major (batch,
    nrhs, n) grad_B and X, on the library's kernel."""
    batch, nrhs, n = gB_cm.shape
    gA = torch.empty((batch, n, n), dtype=gB_cm.dtype, device=gB_cm.device)
    q = default_queue(gB_cm.device)
    rc = _ffi.load().magma_adjoint_outer_batched(gB_cm.element_size(), n, nrhs, gB_cm.data_ptr(), n, nrhs * n,
                                                 X_cm.data_ptr(), n, nrhs * n, gA.data_ptr(), n, n * n, batch,
                                                 q.handle)
    if rc < 0:
        raise MagmaError("magma_adjoint_outer_batched", rc)
    return gA


def _to_cm(M: torch.Tensor) -> torch.Tensor:
    """(batch, n, k) row-major -> (batch, k, n) column-major copy (ld = n)."""
    return M.transpose(1, 2).contiguous()


class BatchedSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A: torch.Tensor, B: torch.Tensor):
        if A.dim() != 3 or A.shape[1] != A.shape[2]:
            raise ValueError("A must be (batch, n, n)")
        if B.dim() != 3 or B.shape[:2] != A.shape[:2]:
            raise ValueError("B must be (batch, n, nrhs)")
        if A.dtype != B.dtype:
            raise TypeError("A and B must share a dtype")
        LU = A.detach().contiguous().clone()          # column-major A^T
        ipiv, info = _getrf(LU)
        X_cm = _to_cm(B.detach())                      # (batch, nrhs, n)
        _getrs(_ffi.MagmaTrans, LU, ipiv, X_cm)        # (A^T)^T X = B
        X = X_cm.transpose(1, 2)
        ctx.save_for_backward(LU, ipiv, X_cm)
        ctx.mark_non_differentiable(info)
        return X, info

    @staticmethod
    def backward(ctx, G: torch.Tensor, _ginfo):
        LU, ipiv, X_cm = ctx.saved_tensors
        gB_cm = _to_cm(G)
        _getrs(_ffi.MagmaNoTrans, LU, ipiv, gB_cm)     # A^T gB = G
        gB = gB_cm.transpose(1, 2)
        gA = _outer(gB_cm, X_cm) if ctx.needs_input_grad[0] else None
        return gA, (gB if ctx.needs_input_grad[1] else None)


def solve(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    """X = A^{-1} B for a batch of square systems, differentiable in A and B.
    A singular system (exactly zero pivot) raises."""
    X, info = BatchedSolve.apply(A, B)
    if bool((info != 0).any()):
        bad = int((info != 0).nonzero()[0, 0])
        raise RuntimeError(f"system {bad} is singular (info = {int(info[bad])})")
    return X
