"""Write the golden LU fixtures from host LAPACK (the reference testers'
oracle: testing/testing_zgetrf_batched.cpp:254-278 calls lapackf77_zgetrf,
testing/testing_zgesv_batched.cpp:158-180 lapackf77_zgesv).

MAGMA ships no golden vectors for this path, so these fixtures are produced
here by the host LAPACK that SciPy bundles (scipy.linalg.lapack dgetrf /
dgetrs / dgesv) on seeded inputs, and committed as data:

  lu_fixtures.npz   per case "<name>/A", "<name>/LU", "<name>/ipiv",
                    "<name>/info", "<name>/B", "<name>/XN", "<name>/XT",
                    plus "<name>/shape" = (m, n, nrhs).

Run: python tests/golden/make_golden.py   (writes next to this script)
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import scipy
import scipy.linalg.lapack as lapack

HERE = Path(__file__).resolve().parent


def cases():
    rng = np.random.default_rng(20240917)
    out = []
    for n in (1, 2, 3, 4, 7, 8, 15, 16, 17, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 200):
        for rep in range(2 if n <= 33 else 1):
            A = rng.uniform(-1.0, 1.0, (n, n))
            out.append((f"rand_n{n}_{rep}", A, 3))
    for m, n in ((40, 24), (24, 40), (128, 64), (64, 128), (33, 7), (7, 33)):
        out.append((f"rect_{m}x{n}", rng.uniform(-1.0, 1.0, (m, n)), 0))
    # singular: zero column 5 (info = 6) and zero columns 3 and 9 (info = 4)
    A = rng.uniform(-1.0, 1.0, (20, 20)); A[:, 5] = 0.0
    out.append(("sing_col5", A, 2))
    A = rng.uniform(-1.0, 1.0, (20, 20)); A[:, 3] = 0.0; A[:, 9] = 0.0
    out.append(("sing_col3_9", A, 2))
    # exact ties among candidates: small-integer entries with repeated values
    A = np.array([[1., 2., 0.], [-2., 1., 1.], [2., 0., 3.]])
    out.append(("tie_3", A, 1))
    A = rng.integers(-2, 3, (12, 12)).astype(np.float64)
    out.append(("int_12", A, 2))
    # diagonally dominant: no interchanges
    A = rng.uniform(-1.0, 1.0, (48, 48)) + 64.0 * np.eye(48)
    out.append(("diagdom_48", A, 2))
    # permutation matrix
    P = np.eye(16)[rng.permutation(16)]
    out.append(("perm_16", P, 1))
    return out, rng


def main() -> None:
    data = {}
    items, rng = cases()
    for name, A, nrhs in items:
        m, n = A.shape
        lu, piv, info = lapack.dgetrf(A)
        data[f"{name}/A"] = A
        data[f"{name}/LU"] = lu
        data[f"{name}/ipiv"] = (piv + 1).astype(np.int32)       # SciPy returns 0-based pivots
        data[f"{name}/info"] = np.int32(info)
        data[f"{name}/shape"] = np.array([m, n, nrhs], dtype=np.int32)
        if m == n and nrhs > 0:
            B = rng.uniform(-1.0, 1.0, (n, nrhs))
            data[f"{name}/B"] = B
            if info == 0:
                xn, inf1 = lapack.dgetrs(lu, piv, B, trans=0)
                xt, inf2 = lapack.dgetrs(lu, piv, B, trans=1)
                assert inf1 == 0 and inf2 == 0
                data[f"{name}/XN"] = xn
                data[f"{name}/XT"] = xt
    data["_meta/lapack"] = np.array(f"scipy {scipy.__version__} bundled OpenBLAS")
    np.savez_compressed(HERE / "lu_fixtures.npz", **data)
    print("wrote", HERE / "lu_fixtures.npz", len(items), "cases")


if __name__ == "__main__":
    main()
