"""GPU parity of the single-matrix (dense) path against the oracle:
magma_dgetrf_gpu / magma_dgetrf_native / magma_sgetrf_gpu (pivots bit-exact
on tie-free inputs, LU forward error, the tester's residual bound, first
zero pivot in info), magma_dgetrs_gpu (NoTrans and Trans, against the
oracle's dgetrs on the same factors), and the refinement driver
magma_dxgesv_gmres_gpu on one large system (src/dxgesv_gmres_gpu.cpp:160-832)
for BF16 and FP16 updates, classical and GMRES refinement, condition numbers
1e2 to 1e6, at n = 2,048 and at the full n = 24,576."""
import numpy as np
import pytest

from conftest import EPS, TESTER_TOL

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

LU_FWD_TOL = 1e3 * EPS


@pytest.fixture(scope="module")
def dn():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from magma_amd import dense
    return dense


def cm(rowmajor: np.ndarray) -> np.ndarray:
    """row-major (m, n) -> column storage (n, m)"""
    return np.ascontiguousarray(rowmajor.T)


def run_getrf(dn, Arow, native=False, dtype=np.float64):
    m, n = Arow.shape
    dA = torch.from_numpy(cm(Arow).astype(dtype)).cuda()
    ipiv, info = dn.getrf_gpu(dA, m=m, native=native)
    return dA.cpu().numpy(), ipiv, info


@pytest.mark.parametrize("m,n", [(300, 300), (1000, 1000), (2048, 2048), (1500, 700), (700, 1500), (2049, 2049)])
def test_dense_getrf_parity(dn, oracle, m, n):
    rng = np.random.default_rng(m * 7 + n)
    Arow = rng.uniform(-1, 1, (m, n))
    LU, ipiv, info = run_getrf(dn, Arow)
    A = cm(Arow)[None]
    LUo, ipo, info_o = oracle.dgetrf(A, m=m)
    assert info == int(info_o[0]) == 0
    res = oracle.lu_residual(A, LU[None], ipiv[None], m, n)[0]
    assert res < TESTER_TOL, res
    if m * n <= 1500 * 1500:
        tie_free = oracle.pivot_margin(A, m)[0] > 1e-10
    else:
        tie_free = True       # random U[-1,1]: ties at 1e-10 are far from likely at these orders
    if tie_free:
        assert np.array_equal(ipiv, ipo[0]), np.nonzero(ipiv != ipo[0])[0][:10]
        fe = oracle.lu_forward_error(LUo, LU[None], m)[0]
        assert fe < LU_FWD_TOL, fe


def test_dense_getrf_4096(dn, oracle):
    rng = np.random.default_rng(4096)
    n = 4096
    Arow = rng.uniform(-1, 1, (n, n))
    LU, ipiv, info = run_getrf(dn, Arow)
    A = cm(Arow)[None]
    LUo, ipo, info_o = oracle.dgetrf(A)
    assert info == 0
    assert np.array_equal(ipiv, ipo[0])
    assert oracle.lu_forward_error(LUo, LU[None], n)[0] < LU_FWD_TOL
    assert oracle.lu_residual(A, LU[None], ipiv[None], n, n)[0] < TESTER_TOL


@pytest.mark.parametrize("prec,m,n", [("d", 2049, 2049), ("d", 1800, 2300), ("s", 2600, 2600), ("bf16", 2100, 1900)])
def test_dense_lookahead_matches_serial(dn, monkeypatch, prec, m, n):
    """The two-stream schedule (the next block factored while the rest of the
    trailing matrix is updated) gives bit-identical factors and pivots to the
    one-stream schedule: every column receives the same updates in the same
    order. Forced at these orders with MAGMA_DENSE_LOOKAHEAD_MIN."""
    rng = np.random.default_rng(m + n)
    Arow = rng.uniform(-1, 1, (m, n))
    dt = torch.float64 if prec == "d" else torch.float32
    out = []
    for lam in ("1000000000", "0"):
        monkeypatch.setenv("MAGMA_DENSE_LOOKAHEAD_MIN", lam)
        dA = torch.from_numpy(cm(Arow)).to(dt).cuda()
        if prec == "d":
            ipiv, info = dn.getrf_gpu(dA, m=m)
        else:
            ipiv, info = dn.sgetrf_prec_gpu(dA, "fp32" if prec == "s" else prec, m=m)
        assert info == 0
        out.append((dA.cpu().numpy(), np.array(ipiv)))
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("prec,lookahead", [("d", False), ("d", True), ("s", True)])
def test_dense_exact_permuted_upper(dn, monkeypatch, prec, lookahead):
    """A = P U with U integer upper triangular: every multiplier is 0 and all
    arithmetic exact, so the dense LU must return U bit for bit and ipiv must
    sort A's rows back into U (the row interchanges checked exactly), with
    and without the two-stream schedule."""
    from test_gpu_lu import apply_ipiv, permuted_upper
    monkeypatch.setenv("MAGMA_DENSE_LOOKAHEAD_MIN", "0" if lookahead else "1000000000")
    n = 1300
    Arow, U = permuted_upper(np.random.default_rng(3), n)
    dt = np.float64 if prec == "d" else np.float32
    LU, ipiv, info = run_getrf(dn, Arow, dtype=dt)
    assert info == 0
    assert np.array_equal(LU.T.astype(np.float64), U)
    assert np.array_equal(apply_ipiv(Arow, ipiv), U)


def test_dense_native_matches_gpu(dn):
    rng = np.random.default_rng(3)
    Arow = rng.uniform(-1, 1, (777, 777))
    LU1, p1, i1 = run_getrf(dn, Arow)
    LU2, p2, i2 = run_getrf(dn, Arow, native=True)
    assert np.array_equal(p1, p2) and i1 == i2
    assert np.array_equal(LU1, LU2)


def test_dense_getrf_singular_and_ties(dn, oracle):
    rng = np.random.default_rng(11)
    n = 600
    Arow = rng.uniform(-1, 1, (n, n))
    Arow[:, 137] = 0.0                      # first zero pivot at column 138 (1-based)
    Arow[:, 301] = 0.0
    LU, ipiv, info = run_getrf(dn, Arow)
    LUo, ipo, info_o = oracle.dgetrf(cm(Arow)[None])
    assert info == int(info_o[0]) == 138
    assert np.array_equal(ipiv[:137], ipo[0][:137])
    # integer matrix with exact ties in every column: lowest index wins
    T = rng.integers(-3, 4, (300, 300)).astype(np.float64)
    LU, ipiv, info = run_getrf(dn, T)
    LUo, ipo, info_o = oracle.dgetrf(cm(T)[None])
    assert info == int(info_o[0])
    # the first column's pivot is decided before any rounding
    assert ipiv[0] == ipo[0][0]


@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("n,nrhs", [(1000, 1), (2048, 7), (333, 20)])
def test_dense_getrs(dn, oracle, n, nrhs, trans):
    rng = np.random.default_rng(n + nrhs + trans)
    Arow = rng.uniform(-1, 1, (n, n))
    Brow = rng.uniform(-1, 1, (n, nrhs))
    dA = torch.from_numpy(cm(Arow)).cuda()
    ipiv, info = dn.getrf_gpu(dA)
    assert info == 0
    dB = torch.from_numpy(cm(Brow)).cuda()
    info = dn.getrs_gpu(dA, ipiv, dB, trans=112 if trans else 111)
    assert info == 0
    X = dB.cpu().numpy().T
    Xo = oracle.dgetrs(trans, dA.cpu().numpy()[None], ipiv[None], cm(Brow)[None])[0].T
    Aop = Arow.T if trans else Arow
    be = np.abs(Brow - Aop @ X).max() / (np.abs(Aop).sum(1).max() * np.abs(X).max() * n)
    assert be < TESTER_TOL, be
    rel = np.abs(X - Xo).max() / np.abs(Xo).max()
    assert rel < 1e-10, rel


def test_dense_sgetrf(dn):
    rng = np.random.default_rng(5)
    m, n = 1500, 1200
    Arow = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    LU, ipiv, info = run_getrf(dn, Arow.astype(np.float64), dtype=np.float32)
    assert info == 0
    L = np.tril(LU.T[:, :n].astype(np.float64), -1) + np.eye(m, n)
    U = np.triu(LU.T[:n, :].astype(np.float64))
    PA = Arow.astype(np.float64).copy()
    for j, p in enumerate(ipiv):
        PA[[j, p - 1]] = PA[[p - 1, j]]
    res = np.linalg.norm(PA - L @ U) / (np.linalg.norm(PA) * n)
    assert res < 30 * np.finfo(np.float32).eps, res


@pytest.mark.parametrize("facto,u", [("bf16", 2.0 ** -8), ("fp16", 2.0 ** -11), ("bf16x3", 2.0 ** -16),
                                     ("fp8", 2.0 ** -4)])
def test_dense_sgetrf_prec_residual(dn, facto, u):
    """magma_sgetrf_prec_gpu: the FP32 LU with 16-bit (or, as an experiment,
    FP8 e4m3) trailing updates has finite factors and a residual at the level
    of the update format: ||PA - LU||_F / ||A||_F below 40 u, u its unit
    roundoff (measured at this order: 19 u for BF16 and FP16, 9 u for
    split-BF16, 21 u for FP8), with the pivots a valid permutation."""
    rng = np.random.default_rng(11)
    n = 1024
    Arow = rng.uniform(-1, 1, (n, n))
    dA = torch.from_numpy(cm(Arow).astype(np.float32)).cuda()
    ipiv, info = dn.sgetrf_prec_gpu(dA, facto)
    assert info == 0
    LU = dA.cpu().numpy().astype(np.float64).T
    assert np.isfinite(LU).all()
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = Arow.copy()
    for j, p in enumerate(ipiv):
        assert j + 1 <= p <= n
        PA[[j, p - 1]] = PA[[p - 1, j]]
    res = np.linalg.norm(PA - L @ U) / np.linalg.norm(PA)
    print(f"{facto}: residual {res:.3e} (u = {u:.3e})")
    assert res < 40 * u, res


def test_dense_fp8_refinement(dn, oracle):
    """The FP8 experiment inside magma_dxgesv_gmres_gpu: whatever the FP8
    factors are worth as a preconditioner, the answer is of FP64 quality
    (converged to MAGMA's test, or the FP64 fallback with iter < 0)."""
    from magma_amd import iteref
    n = 1024
    A = svd_geo_gpu(n, 10.0, seed=4)
    b = torch.from_numpy(np.random.default_rng(4).uniform(-1, 1, (1, n))).cuda()
    for solver in ("ir", "gmres"):
        X, ipiv, dipiv, it, info = iteref.dhgesv_iteref_gpu(A.clone(), b.clone(), facto="fp8", solver=solver)
        assert info == 0
        be = normwise_be(A, X, b)
        assert (it >= 0 and be <= 1e-14) or (it < 0 and be <= TESTER_TOL * n), (solver, it, be)
        print(f"fp8-{solver}: iter {it} be {be:.2e}")


def test_dense_argument_codes(dn):
    from magma_amd import _ffi
    lib = _ffi.load()
    import ctypes
    info = ctypes.c_int(0)
    piv = (ctypes.c_int * 4)()
    assert lib.magma_dgetrf_gpu(-1, 4, None, 4, piv, ctypes.byref(info)) == -1
    assert lib.magma_dgetrf_gpu(4, -1, None, 4, piv, ctypes.byref(info)) == -2
    assert lib.magma_dgetrf_gpu(4, 4, None, 3, piv, ctypes.byref(info)) == -4
    assert lib.magma_dgetrs_gpu(999, 4, 1, None, 4, piv, None, 4, ctypes.byref(info)) == -1
    assert lib.magma_dgetrs_gpu(111, 4, 1, None, 4, piv, None, 3, ctypes.byref(info)) == -8


# --------------------------------------------------------------------------
# Refinement on one large system
# --------------------------------------------------------------------------
def svd_geo_gpu(n, cond, seed):
    """A = U diag(sigma) V^T with U, V Haar orthogonal (QR of Gaussian
    matrices with the sign fix), sigma_i = cond^(-(i-1)/(n-1))
    (testing/magma_generate.cpp, svd_geo), built on the GPU in FP64.
    Returns the column storage (n, n) tensor."""
    g = torch.Generator(device="cuda").manual_seed(seed)

    def haar():
        Z = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
        Q, R = torch.linalg.qr(Z)
        return Q * torch.sign(torch.diagonal(R))[None, :]
    U, V = haar(), haar()
    sig = cond ** (-torch.arange(n, dtype=torch.float64, device="cuda") / max(n - 1, 1))
    A = (U * sig[None, :]) @ V.T.contiguous()
    return A.T.contiguous()      # column storage


def dense_refine(A_cols, b, facto, solver):
    from magma_amd import iteref
    dA = A_cols.clone()
    dB = b.clone()
    X, ipiv, dipiv, it, info = iteref.dhgesv_iteref_gpu(dA, dB, facto=facto, solver=solver)
    return X, it, info, dA


def normwise_be(A_cols, X, B):
    """max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf) on the GPU in FP64."""
    A = A_cols.T
    R = B.T - A @ X.T
    an = A.abs().sum(1).max()
    return float((R.abs().amax(0) / (an * X.T.abs().amax(0))).max())


@pytest.mark.parametrize("facto,solver", [("bf16", "gmres"), ("fp16", "ir"), ("bf16", "ir"), ("fp16", "gmres"),
                                          ("fp32", "ir")])
@pytest.mark.parametrize("cond", [1e2, 1e4, 1e6])
def test_dense_refinement_2048(dn, oracle, facto, solver, cond):
    n = 2048
    A = svd_geo_gpu(n, cond, seed=int(np.log10(cond)))
    rng = np.random.default_rng(int(cond) % 1000)
    b = torch.from_numpy(rng.uniform(-1, 1, (1, n))).cuda()
    X, it, info, dA = dense_refine(A, b, facto, solver)
    assert info == 0
    be = normwise_be(A, X, b)
    if it >= 0:
        assert be <= 1e-14, (it, be)
    else:
        # FP64 fallback (iter < 0): the tester's bound
        assert be <= TESTER_TOL * n, (it, be)
    # forward error against the oracle's FP64 solution
    Xo = oracle.dgesv(A.cpu().numpy()[None], b.cpu().numpy()[None])[3][0]
    fe = float(np.abs(X.cpu().numpy() - Xo).max() / np.abs(Xo).max())
    assert fe <= max(cond, 10.0) * 1e-13, fe
    print(f"n={n} {facto}-{solver} cond={cond:.0e}: iter {it} be {be:.2e} fe {fe:.2e}")


def MAGMA_BWD(n):
    """Backward error the refinement's stopping test guarantees:
    ||r||_inf <= ||x||_inf * ||A||_inf * eps * sqrt(n)
    (src/dxgesv_gmres_gpu.cpp:490-493, eps = dlamch('Epsilon') = 2^-53);
    below 1e-14 for n <= 8,192."""
    return 2.0 ** -53 * np.sqrt(n)


MODES = ["irstrs", "irdtrs", "irgmstrs", "irgmdtrs", "gmstrs", "gmdtrs", "gmgmstrs", "gmgmdtrs"]


@pytest.mark.parametrize("solver", MODES)
@pytest.mark.parametrize("n,facto,cond,trans", [(300, "fp32", 1e4, 0), (1500, "fp16", 1e4, 0), (777, "bf16", 1e2, 1)])
def test_dense_refinement_modes(dn, oracle, solver, n, facto, cond, trans):
    """Every solver type of src/dxgesv_gmres_gpu.cpp:328-359: the answer
    meets 1e-14 normwise backward error (or the tester's bound after the FP64
    fallback) and matches the oracle's FP64 solve; iterations are counted as
    the reference counts them (inner iterations for the nested modes)."""
    from magma_amd import iteref
    A = svd_geo_gpu(n, cond, seed=n)
    rng = np.random.default_rng(n + len(solver))
    b = torch.from_numpy(rng.uniform(-1, 1, (2, n))).cuda()
    dA = A.clone()
    from magma_amd import _ffi
    X, ipiv, dipiv, it, info = iteref.dhgesv_iteref_gpu(dA, b.clone(), facto=facto, solver=solver,
                                                        trans=_ffi.MagmaTrans if trans else _ffi.MagmaNoTrans)
    assert info == 0
    At = A if trans == 0 else A.T.contiguous()        # column storage of op(A)
    be = normwise_be(At, X, b)
    if it >= 0:
        assert be <= 1e-14, (it, be)
    else:
        assert be <= TESTER_TOL * n, (it, be)
    Xo = oracle.dgesv(At.cpu().numpy()[None], b.cpu().numpy()[None])[3][0]
    fe = float(np.abs(X.cpu().numpy() - Xo).max() / np.abs(Xo).max())
    assert fe <= max(cond, 10.0) * 1e-13, fe
    # the low-precision LU's pivots come back in ipiv and their row map in dipiv
    if it >= 0:
        perm = np.arange(n)
        for k, p in enumerate(ipiv):
            assert 1 <= p <= n
            perm[[k, p - 1]] = perm[[p - 1, k]]
        dp = dipiv.cpu().numpy()
        assert np.array_equal(dp, perm if trans == 0 else np.argsort(perm))
    print(f"n={n} {facto}-{solver} cond={cond:.0e} trans={trans}: iter {it} be {be:.2e} fe {fe:.2e}")


def test_dsgesv_iteref_gpu(dn, oracle):
    """magma_dsgesv_iteref_gpu (src/dxgesv_gmres_gpu.cpp:966): FP32 LU +
    FGMRES, at a size of the batched engine and one of the dense LU."""
    from magma_amd import iteref
    for n in (200, 1200):
        A = svd_geo_gpu(n, 1e5, seed=3)
        b = torch.from_numpy(np.random.default_rng(n).uniform(-1, 1, (3, n))).cuda()
        X, ipiv, dipiv, it, info = iteref.dhgesv_iteref_gpu(A.clone(), b.clone(), facto="fp32", solver="gmres")
        assert info == 0 and it >= 0, (n, it, info)
        assert normwise_be(A, X, b) <= 1e-14
        Xo = oracle.dgesv(A.cpu().numpy()[None], b.cpu().numpy()[None])[3][0]
        assert float(np.abs(X.cpu().numpy() - Xo).max() / np.abs(Xo).max()) <= 1e5 * 1e-13


@pytest.mark.timeout(900)
def test_dense_refinement_full_size(dn):
    """n = 24,576 (the dense half of the mixed-precision workload), condition
    1e2, 1e4, 1e6, BF16 and FP16 updates, classical and GMRES refinement:
    FP64-level backward error (or the FP64 fallback with the tester's bound)
    and forward error against the FP64 GPU solve (magma_dgetrf_gpu +
    magma_dgetrs_gpu)."""
    n = 24576
    g = torch.Generator(device="cuda")
    g.manual_seed(7)

    def haar():
        Z = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
        Q, R = torch.linalg.qr(Z)
        del Z
        return Q * torch.sign(torch.diagonal(R))[None, :]
    U = haar()
    V = haar()
    rng = np.random.default_rng(1)
    b = torch.from_numpy(rng.uniform(-1, 1, (1, n))).cuda()
    for cond in (1e2, 1e4, 1e6):
        sig = cond ** (-torch.arange(n, dtype=torch.float64, device="cuda") / (n - 1))
        A = ((U * sig[None, :]) @ V.T.contiguous()).T.contiguous()
        LU = A.clone()
        ipiv, info = dn.getrf_gpu(LU)
        assert info == 0
        X64 = b.clone()
        assert dn.getrs_gpu(LU, ipiv, X64) == 0
        del LU
        be64 = normwise_be(A, X64, b)
        assert be64 <= TESTER_TOL * n
        for facto in ("bf16", "fp16"):
            for solver in ("ir", "gmres"):
                X, it, info, dA = dense_refine(A, b, facto, solver)
                del dA
                assert info == 0
                be = normwise_be(A, X, b)
                # MAGMA stops at ||r||_inf <= ||x||_inf ||A||_inf eps sqrt(n): a
                # backward error of at most eps sqrt(n) = 1.74e-14 at this order
                assert (it >= 0 and be <= MAGMA_BWD(n)) or (it < 0 and be <= TESTER_TOL * n), \
                    (cond, facto, solver, it, be)
                fe = float((X - X64).abs().max() / X64.abs().max())
                assert fe <= cond * 1e-12, fe
                print(f"n={n} {facto}-{solver} cond={cond:.0e}: iter {it} be {be:.2e} fp64 be {be64:.2e} "
                      f"fe {fe:.2e}", flush=True)
                torch.cuda.empty_cache()
        del A
        torch.cuda.empty_cache()
