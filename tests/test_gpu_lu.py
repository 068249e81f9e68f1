"""GPU parity of the FP64 batched LU path against the oracle (and host-LAPACK
fixtures). Mirrors testing_zgetrf_batched -c2 (pivots match, LU forward error)
and testing_zgesv_batched (backward error < 30 eps)."""
import numpy as np
import pytest

from conftest import EPS, TESTER_TOL, golden_cases, uniform_batch

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

# Factor agreement with the unblocked oracle: the register-tiled kernels
# accumulate rank-16 updates in a different order, so the LU factors agree to
# a small multiple of n * eps relative to ||LU||_F.
LU_FWD_TOL = 1e3 * EPS


@pytest.fixture(scope="module")
def mb():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from magma_amd import batched
    return batched


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run_getrf(mb, A, m=None):
    dA = to_dev(A)
    ipiv, info = mb.dgetrf_batched(dA, m=m)
    torch.cuda.synchronize()
    return dA.cpu().numpy(), ipiv.cpu().numpy(), info.cpu().numpy()


def check_getrf(oracle, A, LU, ipiv, info, m, n, tag):
    LUo, ipo, info_o = oracle.dgetrf(A, m=m)
    margin = oracle.pivot_margin(A, m)
    tie_free = margin > 1e-10
    assert np.array_equal(info, info_o), tag
    assert np.array_equal(ipiv[tie_free], ipo[tie_free]), (tag, np.nonzero(~np.all(ipiv == ipo, axis=1)))
    fe = oracle.lu_forward_error(LUo[tie_free], LU[tie_free], m)
    assert (fe < LU_FWD_TOL).all(), (tag, fe.max())
    res = oracle.lu_residual(A, LU, ipiv, m, n)
    assert (res < TESTER_TOL).all(), (tag, res.max())
    return tie_free.mean()


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 15, 16, 17, 24, 31, 32, 33, 47, 48, 63, 64, 65,
                               80, 96, 100, 112, 127, 128])
def test_getrf_square_random(mb, oracle, n):
    rng = np.random.default_rng(1000 + n)
    batch = 64
    A = uniform_batch(rng, batch, n)
    LU, ipiv, info = run_getrf(mb, A)
    frac = check_getrf(oracle, A, LU, ipiv, info, n, n, f"n={n}")
    assert frac > 0.9


@pytest.mark.parametrize("m,n", [(129, 129), (150, 150), (200, 200), (256, 256), (300, 300), (512, 512),
                                 (300, 200), (200, 300), (700, 700), (1100, 1100),
                                 (300, 400), (290, 600), (480, 300), (449, 449)])
def test_getrf_blocked_sizes(mb, oracle, m, n):
    """Sizes beyond one workgroup's LDS take the blocked panel/laswp/trsm/gemm
    path."""
    rng = np.random.default_rng(m * 7 + n)
    batch = 4 if m > 600 else 8
    A = uniform_batch(rng, batch, m, cols=n)
    LU, ipiv, info = run_getrf(mb, A, m=m)
    check_getrf(oracle, A, LU, ipiv, info, m, n, f"{m}x{n}")


@pytest.mark.parametrize("n,ld", [(20, 32), (100, 128), (128, 160)])
def test_getrf_padded_ld(mb, oracle, n, ld):
    rng = np.random.default_rng(7 + n)
    A = uniform_batch(rng, 16, n, ld=ld)
    LU, ipiv, info = run_getrf(mb, A, m=n)
    check_getrf(oracle, A, LU, ipiv, info, n, n, f"n={n},ld={ld}")
    # rows below n of each column are untouched padding
    assert np.array_equal(LU[:, :, n:], A[:, :, n:])


@pytest.mark.parametrize("m,n", [(40, 24), (24, 40), (96, 64), (64, 96), (33, 7), (7, 33)])
def test_getrf_rectangular(mb, oracle, m, n):
    rng = np.random.default_rng(m * 131 + n)
    A = uniform_batch(rng, 24, m, cols=n)
    LU, ipiv, info = run_getrf(mb, A, m=m)
    check_getrf(oracle, A, LU, ipiv, info, m, n, f"{m}x{n}")


def permuted_upper(rng, n):
    """A = P U: U upper triangular with small-integer entries (diagonal
    +-1..9, above it -4..4), row i of U placed at row perm[i]. Partial
    pivoting finds exactly one nonzero candidate per column, every multiplier
    is 0, and all arithmetic is exact: the factorisation must return U
    bit for bit, with ipiv sorting the rows back (an exact check of the row
    interchanges on integer data)."""
    U = np.triu(rng.integers(-4, 5, (n, n))).astype(np.float64)
    U[np.arange(n), np.arange(n)] = rng.integers(1, 10, n) * rng.choice([-1.0, 1.0], n)
    perm = rng.permutation(n)
    A = np.empty_like(U)
    A[perm] = U
    return A, U


def apply_ipiv(A, ipiv):
    PA = A.copy()
    for j, p in enumerate(ipiv):
        PA[[j, p - 1]] = PA[[p - 1, j]]
    return PA


@pytest.mark.parametrize("n", [16, 64, 128, 200, 300, 512, 777])
def test_getrf_exact_permuted_upper(mb, n):
    rng = np.random.default_rng(n)
    rows = [permuted_upper(rng, n) for _ in range(3)]
    A = np.stack([a.T.copy() for a, _ in rows])          # column storage
    LU, ipiv, info = run_getrf(mb, A)
    assert (info == 0).all()
    for s, (Arow, U) in enumerate(rows):
        assert np.array_equal(LU[s].T, U), f"system {s}: factors differ from U"
        assert np.array_equal(apply_ipiv(Arow, ipiv[s]), U), f"system {s}: ipiv does not sort the rows back"


def test_getrf_golden_fixtures(mb, oracle, golden):
    for name in golden_cases(golden):
        Ar = golden[f"{name}/A"]
        m, n, _ = golden[f"{name}/shape"]
        A = np.ascontiguousarray(Ar.T)[None]
        try:
            LU, ipiv, info = run_getrf(mb, A, m=m)
        except Exception as e:  # pragma: no cover - reported with the case name
            raise AssertionError(f"{name}: {e}")
        assert int(info[0]) == int(golden[f"{name}/info"]), name
        if oracle.pivot_margin(A, m)[0] > 1e-10:
            assert np.array_equal(ipiv[0], golden[f"{name}/ipiv"]), name
        ref = np.ascontiguousarray(golden[f"{name}/LU"].T)[None]
        assert oracle.lu_forward_error(ref, LU, m)[0] < LU_FWD_TOL, name


def test_getrf_singular_and_ties(mb, oracle):
    rng = np.random.default_rng(11)
    for n in (12, 40, 100, 128):
        A = uniform_batch(rng, 6, n)
        A[1, 4, :] = 0.0                 # column 4 zero -> info 5
        A[2, 2, :] = 0.0
        A[2, 9, :] = 0.0                 # first zero column 2 -> info 3
        A[3] = 0.0                       # all zero -> info 1
        A[4, :, :] = np.round(A[4] * 2)  # small-integer entries with ties
        LU, ipiv, info = run_getrf(mb, A)
        assert info[1] == 5 and info[2] == 3 and info[3] == 1, (n, info)
        assert np.isfinite(LU[:4]).all()
        LUo, ipo, info_o = oracle.dgetrf(A)
        assert np.array_equal(info, info_o)
        # exactly tied integer columns: LAPACK's first-index rule
        assert np.array_equal(ipiv[4], ipo[4]), n


@pytest.mark.parametrize("n,nrhs", [(1, 1), (8, 3), (32, 1), (33, 4), (64, 16), (100, 5), (128, 4), (128, 9)])
@pytest.mark.parametrize("trans", [0, 1])
def test_getrs(mb, oracle, n, nrhs, trans):
    rng = np.random.default_rng(n * 7 + nrhs + 100 * trans)
    batch = 32
    A = uniform_batch(rng, batch, n)
    B = uniform_batch(rng, batch, n, cols=nrhs)
    LUo, ipo, info_o = oracle.dgetrf(A)
    dLU, dpiv, dB = to_dev(LUo), to_dev(ipo.astype(np.int32)), to_dev(B)
    mb.dgetrs_batched(111 if trans == 0 else 112, dLU, dpiv, dB)
    torch.cuda.synchronize()
    X = dB.cpu().numpy()
    Xo = oracle.dgetrs(trans, LUo, ipo, B)
    scale = np.abs(Xo).max(axis=(1, 2))
    diff = np.abs(X - Xo).max(axis=(1, 2)) / scale
    assert (diff < 1e4 * n * EPS).all(), diff.max()
    Aop = A if trans == 0 else np.ascontiguousarray(np.swapaxes(A, 1, 2))
    be = oracle.gesv_backward_error(Aop, X, B, n)
    assert (be < TESTER_TOL).all(), be.max()


@pytest.mark.parametrize("n,nrhs", [(32, 1), (64, 2), (128, 4)])
def test_gesv_batched(mb, oracle, n, nrhs):
    rng = np.random.default_rng(500 + n)
    batch = 200
    A = uniform_batch(rng, batch, n)
    B = uniform_batch(rng, batch, n, cols=nrhs)
    dA, dB = to_dev(A), to_dev(B)
    ipiv, info = mb.dgesv_batched(dA, dB)
    torch.cuda.synchronize()
    X = dB.cpu().numpy()
    be = oracle.gesv_backward_error(A, X, B, n)
    assert (info.cpu().numpy() == 0).all()
    assert (be < TESTER_TOL).all(), be.max()
    LUo, ipo, _ = oracle.dgetrf(A)
    tf = oracle.pivot_margin(A) > 1e-10
    assert np.array_equal(ipiv.cpu().numpy()[tf], ipo[tf])


@pytest.mark.parametrize("n,nrhs", [(129, 3), (200, 1), (256, 4), (300, 2), (449, 17)])
def test_gesv_batched_blocked(mb, oracle, n, nrhs):
    """n > 128: magma_dgesv_batched as the blocked getrf + getrs composite,
    against the oracle's dgesv (pivots on tie-free systems, solution and
    backward error)."""
    rng = np.random.default_rng(900 + n)
    batch = 24
    A = uniform_batch(rng, batch, n)
    B = uniform_batch(rng, batch, n, cols=nrhs)
    dA, dB = to_dev(A), to_dev(B)
    ipiv, info = mb.dgesv_batched(dA, dB)
    torch.cuda.synchronize()
    X = dB.cpu().numpy()
    assert (info.cpu().numpy() == 0).all()
    be = oracle.gesv_backward_error(A, X, B, n)
    assert (be < TESTER_TOL).all(), be.max()
    LUo, ipo, info_o, Xo = oracle.dgesv(A, B)
    tf = oracle.pivot_margin(A) > 1e-10
    assert np.array_equal(ipiv.cpu().numpy()[tf], ipo[tf])
    rel = np.abs(X - Xo).max(axis=(1, 2)) / np.abs(Xo).max(axis=(1, 2))
    assert rel.max() < 1e-9, rel.max()


def test_batch_beyond_grid_limit(mb, oracle):
    """70,000 systems of order 150 (more than the 65,535 matrices one grid y/z
    extent holds): magma_dgetrf_batched + magma_dgetrs_batched split the
    launches; every system's backward error is checked on the GPU and a
    sample against the oracle, including systems past the first split."""
    batch, n, nrhs = 70000, 150, 1
    g = torch.Generator(device="cuda").manual_seed(77)
    dA = torch.rand((batch, n, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    dB = torch.rand((batch, nrhs, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    A0, B0 = dA.clone(), dB.clone()
    ipiv, info = mb.dgetrf_batched(dA)
    mb.dgetrs_batched(111, dA, ipiv, dB)
    torch.cuda.synchronize()
    assert int((info != 0).sum()) == 0
    be = torch.empty(batch, dtype=torch.float64, device="cuda")
    for c in range(0, batch, 10000):
        Ar = A0[c:c + 10000].transpose(1, 2).contiguous()
        Xr = dB[c:c + 10000].transpose(1, 2).contiguous()
        Br = B0[c:c + 10000].transpose(1, 2).contiguous()
        R = Br - torch.bmm(Ar, Xr)
        be[c:c + 10000] = R.abs().sum(2).amax(1) / (n * Ar.abs().sum(2).amax(1) * Xr.abs().sum(2).amax(1))
    assert float(be.max()) < TESTER_TOL
    sample = np.array([0, 1, 65534, 65535, 65536, 69999])
    A_s = A0[sample].cpu().numpy()
    LUo, ipo, _ = oracle.dgetrf(A_s)
    tf = oracle.pivot_margin(A_s) > 1e-10
    assert np.array_equal(ipiv.cpu().numpy()[sample][tf], ipo[tf])
    del dA, dB, A0, B0
    torch.cuda.empty_cache()


def test_gesv_config2_full_size(mb, oracle):
    """20,000 systems of order 128 with 4 right-hand sides (the fixed-size
    batched workload): every backward error below the tester threshold and
    pivots equal to the oracle on a sample."""
    rng = np.random.default_rng(2)
    batch, n, nrhs = 20000, 128, 4
    g = torch.Generator(device="cuda").manual_seed(1234)
    dA = (torch.rand((batch, n, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1)
    dB = (torch.rand((batch, nrhs, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1)
    A0 = dA.clone()
    B0 = dB.clone()
    ipiv, info = mb.dgesv_batched(dA, dB)
    torch.cuda.synchronize()
    assert int((info != 0).sum()) == 0
    # backward error on the GPU in FP64 with torch (checker arithmetic only)
    Ar = A0.transpose(1, 2).contiguous()
    Xr = dB.transpose(1, 2).contiguous()
    Br = B0.transpose(1, 2).contiguous()
    R = Br - torch.bmm(Ar, Xr)
    be = R.abs().sum(2).amax(1) / (n * Ar.abs().sum(2).amax(1) * Xr.abs().sum(2).amax(1))
    assert float(be.max()) < TESTER_TOL
    sample = rng.choice(batch, 64, replace=False)
    A_s = A0[sample].cpu().numpy()
    LUo, ipo, _ = oracle.dgetrf(A_s)
    tf = oracle.pivot_margin(A_s) > 1e-10
    assert np.array_equal(ipiv.cpu().numpy()[sample][tf], ipo[tf])
    fe = oracle.lu_forward_error(LUo[tf], dA[sample].cpu().numpy()[tf], n)
    assert fe.max() < LU_FWD_TOL
