"""GPU checks of the mixed-precision solvers: FP64-level backward error
(<= 1e-14, and the tester's 30 eps bound), forward error against the FP64
oracle solution, fallbacks (overflow, non-convergence) and the dipiv row map.
Iteration counts are not compared with MAGMA's (its FGMRES runs on the host
with a CPU panel, so counts are parity unpinned); they are reported."""
import numpy as np
import pytest

from conftest import EPS, TESTER_TOL

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ir():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from magma_amd import iteref
    return iteref


def svd_geo(rng, n, cond):
    """A = U diag(sigma) V^T, U and V Haar orthogonal, sigma_i =
    cond^(-(i-1)/(n-1)) (testing/magma_generate.cpp:135, matrix svd_geo)."""
    def haar(k):
        Q, R = np.linalg.qr(rng.standard_normal((k, k)))
        return Q * np.sign(np.diag(R))[None, :]
    sig = cond ** (-np.arange(n) / max(n - 1, 1))
    return (haar(n) * sig[None, :]) @ haar(n).T


def batch_of(rng, batch, n, cond, nrhs):
    A = np.stack([svd_geo(rng, n, cond).T.copy() for _ in range(batch)])   # column-major storage
    B = rng.uniform(-1, 1, (batch, nrhs, n))
    return A, B


def backward_errors(oracle, A, X, B, n):
    Ar = oracle.to_rowmajor(A, n)
    Xr = oracle.to_rowmajor(X, n)
    Br = oracle.to_rowmajor(B, n)
    R = Br - Ar @ Xr
    rn = np.abs(R).sum(2).max(1)
    an = np.abs(Ar).sum(2).max(1)
    xn = np.abs(Xr).sum(2).max(1)
    return rn / (an * xn)


@pytest.mark.parametrize("facto", ["fp32", "bf16x3", "fp16", "bf16"])
@pytest.mark.parametrize("solver", ["ir", "gmres"])
@pytest.mark.parametrize("cond", [1e2, 1e4])
def test_iteref_batched_accuracy(ir, oracle, facto, solver, cond):
    rng = np.random.default_rng(int(cond) % 97 + len(facto))
    batch, n, nrhs = 6, 96, 2
    A, B = batch_of(rng, batch, n, cond, nrhs)
    dA = torch.from_numpy(A.copy()).cuda()
    dB = torch.from_numpy(B.copy()).cuda()
    X, it, info = ir.dxgesv_iteref_batched(dA, dB, facto=facto, solver=solver, itermax=60)
    torch.cuda.synchronize()
    X = X.cpu().numpy()
    be = backward_errors(oracle, A, X, B, n)
    assert (info.cpu().numpy() == 0).all()
    assert be.max() <= 1e-14, (facto, solver, cond, be.max(), it.cpu().numpy())
    assert oracle.gesv_backward_error(A, X, B, n).max() < TESTER_TOL
    # forward error against the FP64 LAPACK-semantics solution
    LU, ipiv, _, Xo = oracle.dgesv(A, B)
    fe = np.abs(X - Xo).max(axis=(1, 2)) / np.abs(Xo).max(axis=(1, 2))
    assert fe.max() < 100 * cond * n * EPS, fe.max()
    iters = it.cpu().numpy()
    if facto in ("fp32", "bf16x3") and solver == "ir" and cond <= 1e4:
        assert (iters >= 0).all() and (iters <= 10).all(), iters


def test_iteref_batched_inputs_unchanged(ir):
    rng = np.random.default_rng(1)
    A, B = batch_of(rng, 4, 64, 1e2, 1)
    dA = torch.from_numpy(A.copy()).cuda()
    dB = torch.from_numpy(B.copy()).cuda()
    X, it, info = ir.dxgesv_iteref_batched(dA, dB, facto="bf16x3", solver="ir")
    torch.cuda.synchronize()
    assert (it.cpu().numpy() >= 0).all()
    assert np.array_equal(dA.cpu().numpy(), A) and np.array_equal(dB.cpu().numpy(), B)


def test_iteref_overflow_falls_back(ir, oracle):
    """Down-cast overflow (src/dxgesv_gmres_gpu.cpp:403-410): the system whose
    FP32 copy overflows skips the low-precision LU and is solved in FP64
    (iter = -2); the others refine as usual. (An earlier build ran the FP32 LU
    on the infinite copy without the first-NaN pivot rule: an all-NaN column
    left the pivot index unset and the row interchanges read a wild row.)"""
    rng = np.random.default_rng(2)
    n = 40
    A, B = batch_of(rng, 3, n, 10.0, 1)
    A[1] *= 1e300                     # beyond FP32 range -> down-cast overflow
    dA = torch.from_numpy(A.copy()).cuda()
    dB = torch.from_numpy(B.copy()).cuda()
    X, it, info = ir.dxgesv_iteref_batched(dA, dB, facto="fp32", solver="ir")
    torch.cuda.synchronize()
    it = it.cpu().numpy()
    assert it[1] == -2 and it[0] >= 0 and it[2] >= 0, it
    be = oracle.gesv_backward_error(A, X.cpu().numpy(), B, n)
    assert be.max() < TESTER_TOL


@pytest.mark.parametrize("facto", ["bf16", "fp16"])
@pytest.mark.parametrize("solver", ["ir", "gmres"])
@pytest.mark.parametrize("cond", [1e2, 1e4, 1e6])
def test_iteref_batched_n256(ir, oracle, facto, solver, cond):
    """The batch half of the mixed-precision workload at n = 256 (a few hundred
    svd_geo systems per condition number): every system either reaches FP64
    backward error (<= 1e-14) with the forward error bounded against the
    oracle's dgesv, or reports a negative iter and meets the tester's bound
    through the FP64 fallback."""
    rng = np.random.default_rng(int(np.log10(cond)) * 10 + len(facto) + len(solver))
    batch, n, nrhs = 200, 256, 1
    A, B = batch_of(rng, batch, n, cond, nrhs)
    dA = torch.from_numpy(A.copy()).cuda()
    dB = torch.from_numpy(B.copy()).cuda()
    X, it, info = ir.dxgesv_iteref_batched(dA, dB, facto=facto, solver=solver, itermax=100)
    torch.cuda.synchronize()
    it = it.cpu().numpy()
    X = X.cpu().numpy()
    assert (info.cpu().numpy() == 0).all()
    be = backward_errors(oracle, A, X, B, n)
    ok = it >= 0
    assert (be[ok] <= 1e-14).all(), be[ok].max()
    assert (be[~ok] <= TESTER_TOL * n).all()
    Xo = oracle.dgesv(A, B)[3]
    fe = np.abs(X - Xo).max(axis=(1, 2)) / np.abs(Xo).max(axis=(1, 2))
    assert fe.max() <= max(cond, 10.0) * 1e-13, fe.max()
    print(f"{facto}-{solver} cond {cond:.0e}: converged {ok.sum()}/{batch}, iter max {it.max()}, "
          f"be max {be.max():.2e}, fe max {fe.max():.2e}")


def test_iteref_singular_reports_info(ir):
    rng = np.random.default_rng(3)
    n = 32
    A, B = batch_of(rng, 2, n, 10.0, 1)
    A[0][5] = 0.0                     # column 5 zero: exactly singular
    dA = torch.from_numpy(A.copy()).cuda()
    dB = torch.from_numpy(B.copy()).cuda()
    X, it, info = ir.dxgesv_iteref_batched(dA, dB, facto="fp32", solver="ir")
    torch.cuda.synchronize()
    assert it.cpu().numpy()[0] == -3
    assert info.cpu().numpy()[0] == 6 and info.cpu().numpy()[1] == 0


def test_iteref_nonconvergence_falls_back(ir, oracle):
    """BF16 factors of a 1e6-conditioned matrix with classical refinement and
    a tiny iteration budget: reported as not converged, solved in FP64."""
    rng = np.random.default_rng(4)
    n = 128
    A, B = batch_of(rng, 2, n, 1e6, 1)
    dA = torch.from_numpy(A.copy()).cuda()
    dB = torch.from_numpy(B.copy()).cuda()
    X, it, info = ir.dxgesv_iteref_batched(dA, dB, facto="bf16", solver="ir", itermax=2)
    torch.cuda.synchronize()
    assert (it.cpu().numpy() < 0).all()
    be = oracle.gesv_backward_error(A, X.cpu().numpy(), B, n)
    assert be.max() < TESTER_TOL


@pytest.mark.parametrize("n,facto,solver", [(300, "fp16", "gmres"), (257, "bf16", "gmres"), (200, "fp32", "ir")])
def test_dhgesv_iteref_gpu(ir, oracle, n, facto, solver):
    rng = np.random.default_rng(n)
    A = svd_geo(rng, n, 1e4)
    b = rng.uniform(-1, 1, (n, 1))
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()        # (cols, ld) column-major
    dB = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    X, ipiv, dipiv, it, info = ir.dhgesv_iteref_gpu(dA, dB, facto=facto, solver=solver)
    torch.cuda.synchronize()
    x = X.cpu().numpy().T
    assert info == 0 and it >= 0, (it, info)
    r = b - A @ x
    assert np.abs(r).max() / (np.abs(A).sum(1).max() * np.abs(x).max()) <= 1e-14
    assert set(ipiv) <= set(range(1, n + 1)) and all(ipiv[k] >= k + 1 for k in range(n))
    perm = oracle.perm_from_ipiv(np.array([ipiv], dtype=np.int32), n)[0]
    assert np.array_equal(dipiv.cpu().numpy(), perm)
    # A unchanged on success (src/dxgesv_gmres_gpu.cpp doc, dA on exit)
    assert np.array_equal(dA.cpu().numpy(), A.T)
