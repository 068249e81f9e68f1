"""magma_dgetrf_native_noalloc (scratch in a fixed device array, no
allocation) against magma_dgetrf_gpu on the same inputs: identical LU
factors, pivots and info, on two calls in a row with the second smaller,
and the size limit reported as an error without a launch."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from magma_amd import _ffi, batched, dense
    return _ffi.load(), batched.default_queue(torch.device("cuda", 0)), dense


def test_noalloc_matches_dgetrf_gpu(env):
    lib, queue, dense = env
    rng = np.random.default_rng(2024)
    for m, n in ((1700, 1500), (600, 800)):          # the second call is smaller
        Arow = rng.uniform(-1, 1, (m, n))
        Arow[:, 77] = 0.0                             # a zero pivot column: info must agree too
        ref = torch.from_numpy(np.ascontiguousarray(Arow.T)).cuda()
        ipiv_ref, info_ref = dense.getrf_gpu(ref, m=m)
        dA = torch.from_numpy(np.ascontiguousarray(Arow.T)).cuda()
        d_ipiv = torch.full((min(m, n),), -1, dtype=torch.int32, device="cuda")
        d_info = torch.full((1,), -99, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        rc = lib.magma_dgetrf_native_noalloc(m, n, dA.data_ptr(), m, d_ipiv.data_ptr(), d_info.data_ptr(),
                                             queue.handle)
        assert rc == 0
        torch.cuda.synchronize()
        assert int(d_info.item()) == info_ref == 78
        assert np.array_equal(d_ipiv.cpu().numpy(), ipiv_ref)
        assert torch.equal(dA, ref)


def test_noalloc_limits(env):
    lib, queue, dense = env
    m = 256 * 256 + 1
    dA = torch.zeros((1, m), dtype=torch.float64, device="cuda")
    d_ipiv = torch.zeros(1, dtype=torch.int32, device="cuda")
    d_info = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = lib.magma_dgetrf_native_noalloc(m, 1, dA.data_ptr(), m, d_ipiv.data_ptr(), d_info.data_ptr(),
                                         queue.handle)
    assert rc == -103                                 # MAGMA_ERR_NOT_SUPPORTED
    assert lib.magma_dgetrf_native_noalloc(4, 4, None, 3, None, None, queue.handle) == -4
    assert lib.magma_dgetrf_native_noalloc(4, 4, None, 4, None, None, None) == -7
