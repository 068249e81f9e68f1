"""GPU checks of the batched LU without pivoting, the random-butterfly solver
and the batched inverse (src/zgetrf_nopiv_batched.cpp, src/zgesv_rbt_batched.cpp,
src/zgetri_outofplace_batched.cpp) against the oracle: factors of diagonally
dominant matrices equal the oracle's (no interchange happens there), the
first zero pivot in info, backward errors of the solves, and the inverse
against the oracle's solve with the identity."""
import numpy as np
import pytest

from conftest import EPS, TESTER_TOL, uniform_batch

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mb():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from magma_amd import batched
    return batched


def dominant_batch(rng, batch, n):
    """Column-major (batch, n, n) with a dominant diagonal: partial pivoting
    picks the diagonal at every step, so the oracle's pivoted LU is the LU
    without pivoting."""
    A = uniform_batch(rng, batch, n)
    for s in range(batch):
        A[s][np.arange(n), np.arange(n)] += 2.0 * n
    return A


@pytest.mark.parametrize("n", [5, 32, 100, 200, 300, 600])
def test_getrf_nopiv_matches_oracle(mb, oracle, n):
    rng = np.random.default_rng(40 + n)
    A = dominant_batch(rng, 6, n)
    dA = torch.from_numpy(A.copy()).cuda()
    info = mb.dgetrf_nopiv_batched(dA)
    torch.cuda.synchronize()
    LUo, ipo, info_o = oracle.dgetrf(A)
    assert (ipo == np.arange(1, n + 1)).all()          # the oracle kept the diagonal
    assert (info.cpu().numpy() == 0).all()
    fe = oracle.lu_forward_error(LUo, dA.cpu().numpy(), n)
    assert fe.max() < 1e3 * EPS, fe.max()


def test_getrf_nopiv_zero_pivot(mb):
    rng = np.random.default_rng(9)
    n = 150
    A = dominant_batch(rng, 3, n)
    A[1][:, :] = 0.0
    A[1][np.arange(n), np.arange(n)] = 1.0
    A[1][37, 37] = 0.0                                 # diagonal zero at 38 (1-based), nothing below it
    dA = torch.from_numpy(A.copy()).cuda()
    info = mb.dgetrf_nopiv_batched(dA).cpu().numpy()
    assert info[0] == 0 and info[1] == 38 and info[2] == 0


@pytest.mark.parametrize("n,nrhs", [(17, 1), (128, 4), (257, 3)])
def test_gesv_nopiv(mb, oracle, n, nrhs):
    rng = np.random.default_rng(n + nrhs)
    A = dominant_batch(rng, 8, n)
    B = uniform_batch(rng, 8, n, cols=nrhs)
    dA, dB = torch.from_numpy(A.copy()).cuda(), torch.from_numpy(B.copy()).cuda()
    info = mb.dgesv_nopiv_batched(dA, dB)
    torch.cuda.synchronize()
    assert (info.cpu().numpy() == 0).all()
    be = oracle.gesv_backward_error(A, dB.cpu().numpy(), B, n)
    assert be.max() < TESTER_TOL, be.max()


@pytest.mark.parametrize("n,nrhs", [(31, 1), (100, 2), (256, 5), (301, 3)])
def test_gesv_rbt(mb, oracle, n, nrhs):
    """General (not dominant) matrices: the butterflies make the LU without
    pivoting usable; the solution must meet a backward-error bound and agree
    with the oracle's pivoted solve."""
    rng = np.random.default_rng(7 * n + nrhs)
    A = uniform_batch(rng, 10, n)
    B = uniform_batch(rng, 10, n, cols=nrhs)
    dA, dB = torch.from_numpy(A.copy()).cuda(), torch.from_numpy(B.copy()).cuda()
    info = mb.dgesv_rbt_batched(dA, dB)
    torch.cuda.synchronize()
    assert (info.cpu().numpy() == 0).all()
    X = dB.cpu().numpy()
    be = oracle.gesv_backward_error(A, X, B, n)
    # RBT without pivoting is less stable than partial pivoting; the bound is
    # a few orders above the pivoted tester threshold
    assert be.max() < 1e4 * TESTER_TOL, be.max()
    Xo = oracle.dgesv(A, B)[3]
    rel = np.abs(X - Xo).max(axis=(1, 2)) / np.abs(Xo).max(axis=(1, 2))
    assert rel.max() < 1e-6, rel.max()


@pytest.mark.parametrize("n", [1, 16, 64, 129, 300])
def test_getri_outofplace(mb, oracle, n):
    rng = np.random.default_rng(300 + n)
    A = uniform_batch(rng, 12, n)
    dA = torch.from_numpy(A.copy()).cuda()
    ipiv, info = mb.dgetrf_batched(dA)
    inv = mb.dgetri_outofplace_batched(dA, ipiv)
    torch.cuda.synchronize()
    Inv = inv.cpu().numpy()                           # column-major storage like A
    LUo, ipo, _ = oracle.dgetrf(A)
    I = np.tile(np.eye(n)[None], (12, 1, 1))
    Xo = oracle.dgetrs(0, LUo, ipo, I)
    rel = np.abs(Inv - Xo).max(axis=(1, 2)) / np.abs(Xo).max(axis=(1, 2))
    assert rel.max() < 1e-9, rel.max()
    Ar, Vr = oracle.to_rowmajor(A), oracle.to_rowmajor(Inv)
    resid = np.abs(Ar @ Vr - np.eye(n)[None]).max(axis=(1, 2))
    cond = np.abs(Ar).sum(2).max(1) * np.abs(Vr).sum(2).max(1)
    assert (resid / (cond * n) < TESTER_TOL).all()


def test_nopiv_vbatched(mb, oracle):
    rng = np.random.default_rng(5)
    ns = [3, 40, 129, 200, 77]
    mats = [dominant_batch(rng, 1, n)[0] for n in ns]
    bufs = [torch.from_numpy(M.copy()).cuda() for M in mats]
    from magma_amd import _ffi
    q = mb.default_queue(torch.device("cuda", 0))
    ptrs = torch.tensor([b.data_ptr() for b in bufs], dtype=torch.int64, device="cuda")
    dn = torch.tensor(ns, dtype=torch.int32, device="cuda")
    info = torch.zeros(len(ns), dtype=torch.int32, device="cuda")
    rc = _ffi.load().magma_dgetrf_nopiv_vbatched(dn.data_ptr(), dn.data_ptr(), ptrs.data_ptr(), dn.data_ptr(),
                                                info.data_ptr(), len(ns), q.handle)
    assert rc == 0
    assert (info.cpu().numpy() == 0).all()
    for M, b, n in zip(mats, bufs, ns):
        LUo, ipo, _ = oracle.dgetrf(M[None])
        assert oracle.lu_forward_error(LUo, b.cpu().numpy()[None], n)[0] < 1e3 * EPS
