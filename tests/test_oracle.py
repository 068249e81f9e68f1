"""Pins the oracle (CPU only): golden host-LAPACK fixtures, closed forms,
brute force on tiny integer matrices, singular and tie cases, and the
testing/flops.h formulas."""
import itertools
from fractions import Fraction

import numpy as np
import pytest

from conftest import EPS, golden_cases, uniform_batch


def cm(A):
    """Row-major (m, n) -> column-major batch of one: (1, n, m)."""
    return np.ascontiguousarray(A.T)[None]


def test_golden_getrf(oracle, golden):
    for name in golden_cases(golden):
        A = golden[f"{name}/A"]
        m, n, _ = golden[f"{name}/shape"]
        LU, ipiv, info = oracle.dgetrf(cm(A), m=m)
        ref_piv = golden[f"{name}/ipiv"]
        margin = oracle.pivot_margin(cm(A), m)[0]
        assert int(info[0]) == int(golden[f"{name}/info"]), name
        if margin > 1e-8 or name.startswith(("tie", "int", "perm", "sing")):
            # tie-free or exactly-representable cases: pivots must match
            assert np.array_equal(ipiv[0], ref_piv), name
        ref_lu = cm(golden[f"{name}/LU"])
        err = oracle.lu_forward_error(ref_lu, LU, m)[0]
        assert err < 1e3 * EPS, (name, err)


def test_golden_getrs(oracle, golden):
    for name in golden_cases(golden):
        if f"{name}/XN" not in golden.files:
            continue
        A = golden[f"{name}/A"]
        B = golden[f"{name}/B"]
        LU, ipiv, info = oracle.dgetrf(cm(A))
        assert info[0] == 0
        for trans, key in ((0, "XN"), (1, "XT")):
            X = oracle.dgetrs(trans, LU, ipiv, np.ascontiguousarray(B.T)[None])
            ref = golden[f"{name}/{key}"]
            rel = np.abs(X[0].T - ref).max() / max(np.abs(ref).max(), 1e-300)
            cond = np.linalg.cond(A)
            assert rel < 50 * EPS * cond * A.shape[0], (name, key, rel)


def test_diag_dominant_has_no_interchanges(oracle):
    rng = np.random.default_rng(1)
    A = uniform_batch(rng, 8, 40)
    A += 100.0 * np.eye(40)[None]
    LU, ipiv, info = oracle.dgetrf(A)
    assert (ipiv == np.arange(1, 41)[None]).all()
    assert (info == 0).all()


def test_permutation_matrix(oracle):
    rng = np.random.default_rng(2)
    for n in (1, 5, 16, 33):
        p = rng.permutation(n)
        P = np.eye(n)[p]            # row i of P is e_{p[i]}
        LU, ipiv, info = oracle.dgetrf(cm(P))
        assert info[0] == 0
        # LU of a permutation matrix is the identity; the interchanges sort it.
        assert np.array_equal(LU[0], np.eye(n))
        perm = oracle.perm_from_ipiv(ipiv, n)[0]
        assert np.array_equal(P[perm], np.eye(n))


def test_upper_triangular_unchanged(oracle):
    rng = np.random.default_rng(3)
    n = 20
    U = np.triu(rng.uniform(-1, 1, (n, n))) + 3 * np.eye(n)
    LU, ipiv, info = oracle.dgetrf(cm(U))
    assert np.array_equal(ipiv[0], np.arange(1, n + 1))
    assert np.array_equal(LU[0], cm(U)[0])


def _exact_lu_pivots(A):
    """Partial pivoting in exact rational arithmetic; ties -> first row."""
    M = [[Fraction(int(v)) for v in row] for row in A]
    n = len(M)
    piv = []
    for k in range(n):
        best, p = -1, k
        for i in range(k, n):
            if abs(M[i][k]) > best:
                best, p = abs(M[i][k]), i
        piv.append(p + 1)
        M[k], M[p] = M[p], M[k]
        if M[k][k] != 0:
            for i in range(k + 1, n):
                f = M[i][k] / M[k][k]
                M[i][k] = f
                for j in range(k + 1, n):
                    M[i][j] -= f * M[k][j]
    return piv, M


def test_bruteforce_small_integer(oracle):
    """n <= 4 integer matrices: pivots equal exact-arithmetic partial pivoting
    wherever the exact candidates are not tied, and L U reproduces P A."""
    rng = np.random.default_rng(4)
    checked = 0
    for n in (1, 2, 3, 4):
        for _ in range(300):
            A = rng.integers(-6, 7, (n, n)).astype(np.float64)
            piv_exact, M = _exact_lu_pivots(A)
            LU, ipiv, info = oracle.dgetrf(cm(A))
            # exact ties can round either way in floating point; skip those
            Mx = [[Fraction(int(v)) for v in row] for row in A]
            tied = False
            for k in range(n):
                col = sorted((abs(Mx[i][k]) for i in range(k, n)), reverse=True)
                if len(col) > 1 and col[0] == col[1] and col[0] != 0:
                    tied = True
                p = piv_exact[k] - 1
                Mx[k], Mx[p] = Mx[p], Mx[k]
                if Mx[k][k] != 0:
                    for i in range(k + 1, n):
                        f = Mx[i][k] / Mx[k][k]
                        for j in range(k, n):
                            Mx[i][j] -= f * Mx[k][j]
            if not tied:
                assert list(ipiv[0]) == piv_exact
                checked += 1
            if info[0] == 0:
                r = oracle.lu_residual(cm(A), LU, ipiv, n, n)[0]
                assert r < 10 * EPS
    assert checked > 500


def test_bruteforce_all_permutations_n3(oracle):
    """For every row order of a fixed 3x3 integer matrix the factors satisfy
    P A = L U and the chosen pivots maximise |u_kk| at each step."""
    A0 = np.array([[2., -1., 3.], [4., 1., -2.], [-1., 5., 1.]])
    for perm in itertools.permutations(range(3)):
        A = A0[list(perm)]
        LU, ipiv, info = oracle.dgetrf(cm(A))
        piv_exact, _ = _exact_lu_pivots(A)
        assert list(ipiv[0]) == piv_exact
        assert oracle.lu_residual(cm(A), LU, ipiv, 3, 3)[0] < 4 * EPS


def test_singular_columns(oracle):
    rng = np.random.default_rng(5)
    A = rng.uniform(-1, 1, (12, 12))
    A[:, 4] = 0.0
    LU, ipiv, info = oracle.dgetrf(cm(A))
    assert info[0] == 5
    A[:, 2] = 0.0
    LU, ipiv, info = oracle.dgetrf(cm(A))
    assert info[0] == 3
    assert np.isfinite(LU).all()   # no NaN from a zero pivot


def test_ties_first_index(oracle):
    # column 0 = [1, -2, 2, -2]: |.| ties between rows 1, 2, 3 -> row 1 wins
    A = np.array([[1., 0, 0, 0], [-2., 1, 0, 0], [2., 0, 1, 0], [-2., 0, 0, 1]])
    LU, ipiv, info = oracle.dgetrf(cm(A))
    assert ipiv[0][0] == 2
    assert oracle.idamax(np.array([3.0, -3.0, 3.0])) == 0


def test_rectangular(oracle):
    rng = np.random.default_rng(6)
    for m, n in ((30, 12), (12, 30), (65, 64)):
        A = rng.uniform(-1, 1, (m, n))
        LU, ipiv, info = oracle.dgetrf(cm(A), m=m)
        assert ipiv.shape == (1, min(m, n))
        assert oracle.lu_residual(cm(A), LU, ipiv, m, n)[0] < 20 * EPS


def test_getrs_exact_on_integer_triangular(oracle):
    """Unit-lower L with integer entries and U = I: the solves are exact."""
    rng = np.random.default_rng(7)
    n = 10
    L = np.tril(rng.integers(-3, 4, (n, n)).astype(float), -1) + np.eye(n)
    LU = L.copy()                      # strictly lower part = L, diag/upper = I
    LU[np.triu_indices(n)] = 0.0
    LU += np.eye(n)
    ipiv = np.arange(1, n + 1, dtype=np.int32)[None]
    X = rng.integers(-5, 6, (n, 2)).astype(float)
    B = L @ X
    Xs = oracle.dgetrs(0, cm(LU), ipiv, np.ascontiguousarray(B.T)[None])
    assert np.array_equal(Xs[0].T, X)
    # transposed: L^T X = B'
    Bt = L.T @ X
    Xt = oracle.dgetrs(1, cm(LU), ipiv, np.ascontiguousarray(Bt.T)[None])
    assert np.array_equal(Xt[0].T, X)


def test_getrs_transpose_residual(oracle):
    rng = np.random.default_rng(8)
    A = uniform_batch(rng, 20, 50)
    B = uniform_batch(rng, 20, 50, cols=3)
    LU, ipiv, info = oracle.dgetrf(A)
    XT = oracle.dgetrs(1, LU, ipiv, B)
    At = np.ascontiguousarray(np.swapaxes(A, 1, 2))
    err = oracle.gesv_backward_error(At, XT, B, 50)
    assert (err < 30 * EPS).all()


def test_swp2pswp_hand_examples(oracle):
    # ipiv (1-based) [3, 3, 3]: swap(0,2) then swap(1,2) on [0,1,2]
    # -> [2,1,0] -> [2,0,1]; row i of P^T B is B[perm[i]].
    assert list(oracle.swp2pswp(0, np.array([3, 3, 3]))) == [2, 0, 1]
    assert list(oracle.swp2pswp(1, np.array([3, 3, 3]))) == [1, 2, 0]
    assert list(oracle.swp2pswp(0, np.array([1, 2, 3]))) == [0, 1, 2]
    assert list(oracle.swp2pswp(0, np.array([2, 2]))) == [1, 0]


def test_swp2pswp_is_inverse_pair(oracle):
    rng = np.random.default_rng(9)
    for n in (1, 5, 31):
        ipiv = np.array([rng.integers(k, n) + 1 for k in range(n)], dtype=np.int32)
        f = oracle.swp2pswp(0, ipiv)
        b = oracle.swp2pswp(1, ipiv)
        assert np.array_equal(f[b], np.arange(n))
        assert np.array_equal(f, oracle.perm_from_ipiv(ipiv[None], n)[0])


def test_flops_formulas(oracle):
    # values quoted in the baseline plan for these sizes
    assert oracle.flops_dgetrf(32, 32) == pytest.approx(21360.0)
    assert oracle.flops_dgetrs(32, 1) == pytest.approx(2016.0)
    assert oracle.flops_dgetrf(128, 128) == pytest.approx(1390016.0)
    assert oracle.flops_dgetrs(128, 4) == pytest.approx(130560.0)
    # 2n^3/3 - n^2/2 + 5n/6 for square matrices
    for n in (1, 7, 100):
        assert oracle.flops_dgetrf(n, n) == pytest.approx(2 * n**3 / 3 - n**2 / 2 + 5 * n / 6)


def test_host_lapack_cross_check(oracle):
    L = oracle.host_lapack()
    if L is None:
        pytest.skip("no host LAPACK library found")
    rng = np.random.default_rng(10)
    n = 96
    A = uniform_batch(rng, 6, n)
    B = uniform_batch(rng, 6, n, cols=2)
    LUh, ipivh, infoh, Xh, _ = oracle.host_dgesv_batch(A, B, 2)
    LU, ipiv, info = oracle.dgetrf(A)
    margin = oracle.pivot_margin(A)
    for s in range(6):
        if margin[s] > 1e-8:
            assert np.array_equal(ipiv[s], ipivh[s])
    assert (oracle.lu_forward_error(LUh, LU, n) < 1e3 * EPS).all()
    assert (oracle.gesv_backward_error(A, Xh, B, n) < 30 * EPS).all()
