"""Python mirror of MAGMA's batched LU interface over libmagma_mi355x.so.

Function names, argument order and return codes follow the C routines
(``magma_dgetrf_batched`` -> ``dgetrf_batched`` ...). Device memory is held
in torch tensors; torch is used only for allocation and streams, every
computation runs in the HIP kernels of the library.

Storage: a batch of column-major m-by-n matrices with leading dimension ld
is a contiguous float64 CUDA tensor of shape (batch, n, ld); element (i, j)
of matrix s is ``A[s, j, i]`` (MAGMA's strided layout, set up with
``magma_dset_pointer(..., batch_offset = ld*n)``).
"""
from __future__ import annotations

import ctypes

import torch

from . import _ffi

_F = b"magma_amd"


class MagmaError(RuntimeError):
    def __init__(self, routine: str, code: int):
        lib = _ffi.load()
        msg = lib.magma_strerror(code).decode()
        super().__init__(f"{routine} returned {code}: {msg}")
        self.code = code


class Queue:
    """A MAGMA queue bound to a torch stream on one device."""

    def __init__(self, device: int | torch.device | None = None, stream: torch.cuda.Stream | None = None):
        lib = _ffi.load()
        if device is None:
            device = torch.cuda.current_device()
        self.device = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
        idx = self.device.index if self.device.index is not None else torch.cuda.current_device()
        self.stream = stream or torch.cuda.current_stream(idx)
        q = ctypes.c_void_p()
        lib.magma_queue_create_from_hip_internal(idx, ctypes.c_void_p(self.stream.cuda_stream), None, None,
                                                 ctypes.byref(q), _F, _F, 0)
        if not q.value:
            raise MagmaError("magma_queue_create", -113)
        self.handle = q
        self._lib = lib

    def sync(self) -> None:
        self._lib.magma_queue_sync_internal(self.handle, _F, _F, 0)

    def close(self) -> None:
        if self.handle is not None and self.handle.value:
            self._lib.magma_queue_destroy_internal(self.handle, _F, _F, 0)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_queues: dict[tuple[int, int], Queue] = {}


def default_queue(device: torch.device | None = None) -> Queue:
    idx = (device.index if device is not None and device.index is not None else torch.cuda.current_device())
    s = torch.cuda.current_stream(idx)
    key = (idx, s.cuda_stream)
    q = _default_queues.get(key)
    if q is None:
        q = Queue(idx, s)
        _default_queues[key] = q
    return q


def _check_batch(t: torch.Tensor, name: str, dtype=torch.float64) -> None:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA (HIP) tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}")
    if t.dim() != 3 or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous (batch, ncols, ld) tensor")


def pointer_array(t: torch.Tensor, queue: Queue) -> torch.Tensor:
    """Device array of per-matrix pointers for a (batch, ncols, ld) tensor,
    built by magma_dset_pointer / magma_iset_pointer."""
    lib = _ffi.load()
    batch = t.shape[0]
    ptrs = torch.empty(max(batch, 1), dtype=torch.int64, device=t.device)
    if batch == 0:
        return ptrs
    stride = t.shape[1] * t.shape[2]
    ld = t.shape[2]
    if t.dtype == torch.float64:
        lib.magma_dset_pointer(ptrs.data_ptr(), t.data_ptr(), ld, 0, 0, stride, batch, queue.handle)
    elif t.dtype == torch.int32:
        lib.magma_iset_pointer(ptrs.data_ptr(), t.data_ptr(), ld, 0, 0, stride, batch, queue.handle)
    else:
        raise ValueError(f"no pointer-array builder for {t.dtype}")
    return ptrs


def _ipiv_view(ipiv: torch.Tensor) -> torch.Tensor:
    # (batch, k) int32 -> (batch, 1, k) so that pointer_array sees ld = k.
    return ipiv.view(ipiv.shape[0], 1, ipiv.shape[1])


def dgetrf_batched(A: torch.Tensor, m: int | None = None, ipiv: torch.Tensor | None = None,
                   info: torch.Tensor | None = None, queue: Queue | None = None):
    """In-place LU of every matrix of A (batch, n, ldda); m defaults to ldda.
    Returns (ipiv (batch, min(m,n)) int32 1-based, info (batch,) int32)."""
    _check_batch(A, "A")
    batch, n, ldda = A.shape
    m = ldda if m is None else m
    queue = queue or default_queue(A.device)
    k = max(min(m, n), 1)
    if ipiv is None:
        ipiv = torch.zeros((batch, k), dtype=torch.int32, device=A.device)
    if info is None:
        info = torch.zeros(batch, dtype=torch.int32, device=A.device)
    lib = _ffi.load()
    pA = pointer_array(A, queue)
    pP = pointer_array(_ipiv_view(ipiv), queue)
    rc = lib.magma_dgetrf_batched(m, n, pA.data_ptr(), ldda, pP.data_ptr(), info.data_ptr(), batch,
                                  queue.handle)
    if rc < 0:
        raise MagmaError("magma_dgetrf_batched", rc)
    return ipiv, info


def dgetrs_batched(trans: int, LU: torch.Tensor, ipiv: torch.Tensor, B: torch.Tensor,
                   queue: Queue | None = None) -> None:
    """Overwrite B (batch, nrhs, lddb) with op(A)^{-1} B using LU factors."""
    _check_batch(LU, "LU")
    _check_batch(B, "B")
    batch, n, ldda = LU.shape
    _, nrhs, lddb = B.shape
    queue = queue or default_queue(LU.device)
    lib = _ffi.load()
    pA = pointer_array(LU, queue)
    pP = pointer_array(_ipiv_view(ipiv), queue)
    pB = pointer_array(B, queue)
    rc = lib.magma_dgetrs_batched(trans, n, nrhs, pA.data_ptr(), ldda, pP.data_ptr(), pB.data_ptr(), lddb,
                                  batch, queue.handle)
    if rc < 0:
        raise MagmaError("magma_dgetrs_batched", rc)


def dgesv_batched(A: torch.Tensor, B: torch.Tensor, ipiv: torch.Tensor | None = None,
                  info: torch.Tensor | None = None, queue: Queue | None = None):
    """Solve A_s X_s = B_s in place (A -> LU, B -> X). Returns (ipiv, info)."""
    _check_batch(A, "A")
    _check_batch(B, "B")
    batch, n, ldda = A.shape
    _, nrhs, lddb = B.shape
    queue = queue or default_queue(A.device)
    if ipiv is None:
        ipiv = torch.zeros((batch, max(n, 1)), dtype=torch.int32, device=A.device)
    if info is None:
        info = torch.zeros(batch, dtype=torch.int32, device=A.device)
    lib = _ffi.load()
    pA = pointer_array(A, queue)
    pP = pointer_array(_ipiv_view(ipiv), queue)
    pB = pointer_array(B, queue)
    rc = lib.magma_dgesv_batched(n, nrhs, pA.data_ptr(), ldda, pP.data_ptr(), pB.data_ptr(), lddb,
                                 info.data_ptr(), batch, queue.handle)
    if rc < 0:
        raise MagmaError("magma_dgesv_batched", rc)
    return ipiv, info
