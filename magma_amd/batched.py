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


# ---------------------------------------------------------------------------
# Single precision
# ---------------------------------------------------------------------------
def sgetrf_batched(A: torch.Tensor, m: int | None = None, ipiv: torch.Tensor | None = None,
                   info: torch.Tensor | None = None, queue: Queue | None = None):
    """FP32 form of dgetrf_batched."""
    _check_batch(A, "A", torch.float32)
    batch, n, ldda = A.shape
    m = ldda if m is None else m
    queue = queue or default_queue(A.device)
    k = max(min(m, n), 1)
    ipiv = ipiv if ipiv is not None else torch.zeros((batch, k), dtype=torch.int32, device=A.device)
    info = info if info is not None else torch.zeros(batch, dtype=torch.int32, device=A.device)
    lib = _ffi.load()
    pA = _sptr(A, queue)
    pP = pointer_array(_ipiv_view(ipiv), queue)
    rc = lib.magma_sgetrf_batched(m, n, pA.data_ptr(), ldda, pP.data_ptr(), info.data_ptr(), batch, queue.handle)
    if rc < 0:
        raise MagmaError("magma_sgetrf_batched", rc)
    return ipiv, info


def sgetrs_batched(trans: int, LU: torch.Tensor, ipiv: torch.Tensor, B: torch.Tensor,
                   queue: Queue | None = None) -> None:
    _check_batch(LU, "LU", torch.float32)
    _check_batch(B, "B", torch.float32)
    batch, n, ldda = LU.shape
    _, nrhs, lddb = B.shape
    queue = queue or default_queue(LU.device)
    lib = _ffi.load()
    # keep the pointer arrays alive until the launches are queued
    pA, pP, pB = _sptr(LU, queue), pointer_array(_ipiv_view(ipiv), queue), _sptr(B, queue)
    rc = lib.magma_sgetrs_batched(trans, n, nrhs, pA.data_ptr(), ldda, pP.data_ptr(), pB.data_ptr(), lddb,
                                  batch, queue.handle)
    if rc < 0:
        raise MagmaError("magma_sgetrs_batched", rc)


def _sptr(t: torch.Tensor, queue: Queue) -> torch.Tensor:
    lib = _ffi.load()
    batch = t.shape[0]
    ptrs = torch.empty(max(batch, 1), dtype=torch.int64, device=t.device)
    if batch:
        lib.magma_sset_pointer(ptrs.data_ptr(), t.data_ptr(), t.shape[2], 0, 0, t.shape[1] * t.shape[2], batch,
                               queue.handle)
    return ptrs


# ---------------------------------------------------------------------------
# Variable size
# ---------------------------------------------------------------------------
class VBatch:
    """A batch of column-major matrices of different sizes packed into one
    flat device buffer (matrix i at element offset offs[i], leading dimension
    ld[i]), with the device pointer and size arrays MAGMA's vbatched calls
    take."""

    def __init__(self, m, n, ld=None, dtype=torch.float64, device=None, cols=None):
        import numpy as np
        self.m = np.asarray(m, dtype=np.int32)
        self.n = np.asarray(n, dtype=np.int32)
        self.ld = np.maximum(self.m, 1).astype(np.int32) if ld is None else np.asarray(ld, dtype=np.int32)
        ncols = self.n if cols is None else np.asarray(cols, dtype=np.int32)
        self.cols = ncols
        sizes = self.ld.astype(np.int64) * ncols.astype(np.int64)
        self.offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        device = device or torch.device("cuda", torch.cuda.current_device())
        self.dtype = dtype
        self.data = torch.zeros(int(sizes.sum()) + 1, dtype=dtype, device=device)
        elem = self.data.element_size()
        base = self.data.data_ptr()
        self.ptrs = torch.from_numpy(base + self.offs * elem).to(device)
        self.d_m = torch.from_numpy(self.m).to(device)
        self.d_n = torch.from_numpy(self.n).to(device)
        self.d_ld = torch.from_numpy(self.ld).to(device)

    def __len__(self):
        return len(self.m)

    def set(self, i, M):
        """Store row-major numpy matrix M (m_i x cols_i) into slot i."""
        import numpy as np
        m, c, ld, off = int(self.m[i]), int(self.cols[i]), int(self.ld[i]), int(self.offs[i])
        buf = np.zeros((c, ld))
        buf[:, :m] = np.asarray(M).T
        self.data[off:off + c * ld] = torch.from_numpy(buf.reshape(-1)).to(self.data.device, self.dtype)

    def get(self, i):
        m, c, ld, off = int(self.m[i]), int(self.cols[i]), int(self.ld[i]), int(self.offs[i])
        return self.data[off:off + c * ld].view(c, ld)[:, :m].T.cpu().numpy()


class IpivV:
    """Per-matrix ipiv buffers for a variable-size batch."""

    def __init__(self, k, device=None):
        import numpy as np
        k = np.maximum(np.asarray(k, dtype=np.int64), 1)
        self.k = k
        self.offs = np.concatenate([[0], np.cumsum(k)[:-1]]).astype(np.int64)
        device = device or torch.device("cuda", torch.cuda.current_device())
        self.data = torch.zeros(int(k.sum()), dtype=torch.int32, device=device)
        self.ptrs = torch.from_numpy(self.data.data_ptr() + self.offs * 4).to(device)

    def get(self, i, count):
        off = int(self.offs[i])
        return self.data[off:off + count].cpu().numpy()


def dgetrf_vbatched(A: VBatch, ipiv: IpivV | None = None, info: torch.Tensor | None = None,
                    queue: Queue | None = None):
    import numpy as np
    queue = queue or default_queue(A.data.device)
    ipiv = ipiv or IpivV(np.minimum(A.m, A.n), A.data.device)
    info = info if info is not None else torch.zeros(len(A), dtype=torch.int32, device=A.data.device)
    rc = _ffi.load().magma_dgetrf_vbatched(A.d_m.data_ptr(), A.d_n.data_ptr(), A.ptrs.data_ptr(),
                                           A.d_ld.data_ptr(), ipiv.ptrs.data_ptr(), info.data_ptr(), len(A),
                                           queue.handle)
    if rc < 0:
        raise MagmaError("magma_dgetrf_vbatched", rc)
    return ipiv, info


def dgetrf_vbatched_max_nocheck(A: VBatch, ipiv: IpivV | None = None, info: torch.Tensor | None = None,
                                queue: Queue | None = None):
    import numpy as np
    lib = _ffi.load()
    queue = queue or default_queue(A.data.device)
    ipiv = ipiv or IpivV(np.minimum(A.m, A.n), A.data.device)
    info = info if info is not None else torch.zeros(len(A), dtype=torch.int32, device=A.data.device)
    lwork = ctypes.c_int(-1)
    max_m, max_n = int(A.m.max()), int(A.n.max())
    mn = int(np.minimum(A.m, A.n).max())
    mxn = int((A.m.astype(np.int64) * A.n).max())
    lib.magma_dgetrf_vbatched_max_nocheck_work(None, None, max_m, max_n, mn, mxn, None, None, None, None, None,
                                               ctypes.byref(lwork), len(A), queue.handle)
    work = torch.empty(max(lwork.value, 1), dtype=torch.uint8, device=A.data.device)
    rc = lib.magma_dgetrf_vbatched_max_nocheck_work(A.d_m.data_ptr(), A.d_n.data_ptr(), max_m, max_n, mn, mxn,
                                                    A.ptrs.data_ptr(), A.d_ld.data_ptr(), ipiv.ptrs.data_ptr(),
                                                    info.data_ptr(), work.data_ptr(), ctypes.byref(lwork),
                                                    len(A), queue.handle)
    if rc < 0:
        raise MagmaError("magma_dgetrf_vbatched_max_nocheck_work", rc)
    return ipiv, info


def dgetrs_vbatched(trans: int, A: VBatch, ipiv: IpivV, B: VBatch, queue: Queue | None = None) -> None:
    queue = queue or default_queue(A.data.device)
    rc = _ffi.load().magma_dgetrs_vbatched(trans, A.d_n.data_ptr(), B.d_n.data_ptr(), int(A.n.max()),
                                           int(B.n.max()), A.ptrs.data_ptr(), A.d_ld.data_ptr(),
                                           ipiv.ptrs.data_ptr(), B.ptrs.data_ptr(), B.d_ld.data_ptr(), len(A),
                                           queue.handle)
    if rc < 0:
        raise MagmaError("magma_dgetrs_vbatched", rc)


def dgesv_vbatched(A: VBatch, B: VBatch, ipiv: IpivV | None = None, info: torch.Tensor | None = None,
                   queue: Queue | None = None):
    """B is a VBatch with m = n_i rows and n = nrhs_i columns."""
    queue = queue or default_queue(A.data.device)
    ipiv = ipiv or IpivV(A.n, A.data.device)
    info = info if info is not None else torch.zeros(len(A), dtype=torch.int32, device=A.data.device)
    rc = _ffi.load().magma_dgesv_vbatched(A.d_n.data_ptr(), B.d_n.data_ptr(), int(A.n.max()), int(B.n.max()),
                                          A.ptrs.data_ptr(), A.d_ld.data_ptr(), ipiv.ptrs.data_ptr(),
                                          B.ptrs.data_ptr(), B.d_ld.data_ptr(), info.data_ptr(), len(A),
                                          queue.handle)
    if rc < 0:
        raise MagmaError("magma_dgesv_vbatched", rc)
    return ipiv, info


class VbatchedPlan:
    """Host-side bucketing of a variable-size solve, executed or replayed as a
    hipGraph without allocation or synchronisation (magma_dgesv_vbatched_plan_*)."""

    def __init__(self, n, nrhs, variants=None, queue: Queue | None = None):
        import numpy as np
        self.queue = queue or default_queue()
        self._lib = _ffi.load()
        self._n = np.ascontiguousarray(n, dtype=np.int32)
        self._r = None if nrhs is None else np.ascontiguousarray(nrhs, dtype=np.int32)
        v = None
        if variants is not None:
            v = np.zeros(22, dtype=np.int32)
            for k, val in dict(variants).items():
                v[int(k)] = int(val)
        self._v = v
        h = ctypes.c_void_p()
        rc = self._lib.magma_dgesv_vbatched_plan_create(
            self._n.ctypes.data, None if self._r is None else self._r.ctypes.data, len(self._n),
            None if v is None else v.ctypes.data, self.queue.handle, ctypes.byref(h))
        if rc < 0:
            raise MagmaError("magma_dgesv_vbatched_plan_create", rc)
        self.handle = h

    def buckets(self):
        out = []
        info = (ctypes.c_int * 5)()
        for i in range(self._lib.magma_vbatched_plan_num_buckets(self.handle)):
            self._lib.magma_vbatched_plan_bucket(self.handle, i, info)
            out.append(dict(id=info[0], variant=info[1], count=info[2], max_n=info[3], max_nrhs=info[4]))
        return out

    def _args(self, A: VBatch, B: VBatch, ipiv: IpivV, info: torch.Tensor):
        return (A.d_n.data_ptr(), B.d_n.data_ptr(), A.ptrs.data_ptr(), A.d_ld.data_ptr(), ipiv.ptrs.data_ptr(),
                B.ptrs.data_ptr(), B.d_ld.data_ptr(), info.data_ptr())

    def execute(self, A, B, ipiv, info):
        rc = self._lib.magma_dgesv_vbatched_plan_exec(self.handle, *self._args(A, B, ipiv, info), self.queue.handle)
        if rc < 0:
            raise MagmaError("magma_dgesv_vbatched_plan_exec", rc)

    def capture(self, A, B, ipiv, info):
        rc = self._lib.magma_dgesv_vbatched_plan_capture(self.handle, *self._args(A, B, ipiv, info),
                                                         self.queue.handle)
        if rc < 0:
            raise MagmaError("magma_dgesv_vbatched_plan_capture", rc)

    def replay(self):
        rc = self._lib.magma_vbatched_plan_replay(self.handle, self.queue.handle)
        if rc < 0:
            raise MagmaError("magma_vbatched_plan_replay", rc)

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.magma_vbatched_plan_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dgetrf_nopiv_batched(A: torch.Tensor, m: int | None = None, queue: Queue | None = None):
    """In-place LU without pivoting of every matrix of A (batch, n, ldda).
    Returns info (batch,) int32 (first exactly-zero pivot, 1-based)."""
    _check_batch(A, "A")
    batch, n, ldda = A.shape
    m = ldda if m is None else m
    queue = queue or default_queue(A.device)
    info = torch.zeros(batch, dtype=torch.int32, device=A.device)
    pA = pointer_array(A, queue)
    rc = _ffi.load().magma_dgetrf_nopiv_batched(m, n, pA.data_ptr(), ldda, info.data_ptr(), batch, queue.handle)
    if rc < 0:
        raise MagmaError("magma_dgetrf_nopiv_batched", rc)
    return info


def dgesv_nopiv_batched(A: torch.Tensor, B: torch.Tensor, queue: Queue | None = None):
    """A_s X_s = B_s with the LU without pivoting (A -> LU, B -> X). Returns info."""
    _check_batch(A, "A")
    _check_batch(B, "B")
    batch, n, ldda = A.shape
    _, nrhs, lddb = B.shape
    queue = queue or default_queue(A.device)
    info = torch.zeros(batch, dtype=torch.int32, device=A.device)
    pA, pB = pointer_array(A, queue), pointer_array(B, queue)
    rc = _ffi.load().magma_dgesv_nopiv_batched(n, nrhs, pA.data_ptr(), ldda, pB.data_ptr(), lddb, info.data_ptr(),
                                               batch, queue.handle)
    if rc < 0:
        raise MagmaError("magma_dgesv_nopiv_batched", rc)
    return info


def dgesv_rbt_batched(A: torch.Tensor, B: torch.Tensor, queue: Queue | None = None):
    """A_s X_s = B_s through random butterflies and the LU without pivoting
    (magma_dgesv_rbt_batched). B -> X. Returns info."""
    _check_batch(A, "A")
    _check_batch(B, "B")
    batch, n, ldda = A.shape
    _, nrhs, lddb = B.shape
    queue = queue or default_queue(A.device)
    info = torch.zeros(batch, dtype=torch.int32, device=A.device)
    pA, pB = pointer_array(A, queue), pointer_array(B, queue)
    rc = _ffi.load().magma_dgesv_rbt_batched(n, nrhs, pA.data_ptr(), ldda, pB.data_ptr(), lddb, info.data_ptr(),
                                             batch, queue.handle)
    if rc < 0:
        raise MagmaError("magma_dgesv_rbt_batched", rc)
    return info


def dgetri_outofplace_batched(LU: torch.Tensor, ipiv: torch.Tensor, queue: Queue | None = None) -> torch.Tensor:
    """Inverse of every matrix from its LU factors and pivots
    (magma_dgetri_outofplace_batched). Returns a new (batch, n, n) tensor."""
    _check_batch(LU, "LU")
    batch, n, ldda = LU.shape
    queue = queue or default_queue(LU.device)
    inv = torch.empty((batch, n, n), dtype=torch.float64, device=LU.device)
    info = torch.zeros(batch, dtype=torch.int32, device=LU.device)
    pA, pP, pI = pointer_array(LU, queue), pointer_array(_ipiv_view(ipiv), queue), pointer_array(inv, queue)
    rc = _ffi.load().magma_dgetri_outofplace_batched(n, pA.data_ptr(), ldda, pP.data_ptr(), pI.data_ptr(), n,
                                                     info.data_ptr(), batch, queue.handle)
    if rc < 0:
        raise MagmaError("magma_dgetri_outofplace_batched", rc)
    return inv
