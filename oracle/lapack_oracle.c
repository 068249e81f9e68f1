/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the LAPACK semantics that MAGMA's batched LU path
 * is checked against (its testers compare with host LAPACK dgetrf / dgetrs /
 * dgesv, testing/testing_zgetrf_batched.cpp:254-278,
 * testing/testing_zgesv_batched.cpp:158-180). Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load this library; the product path
 * never calls it.
 *
 * Pins (what checks each function against something other than itself):
 *   orc_dgetf2   pivots equal host LAPACK dgetrf on tie-free inputs and LU
 *                within rounding (tests/golden fixtures written by host
 *                LAPACK); closed forms (permutation, diagonally dominant,
 *                triangular inputs); brute force over all row orders for
 *                n <= 4 on small-integer matrices; singular-column cases.
 *   orc_dgetrs   host LAPACK dgetrs 'N' and 'T' (fixtures); exact on integer
 *                unit-triangular systems; residual bound.
 *   orc_swp2pswp hand-computed row maps for small interchange sequences
 *                (tests/test_oracle.py), and the identity P^T P = I.
 *   orc_flops_*  the closed forms of testing/flops.h:80-91 at known sizes
 *                (n=32 -> 21,360 + 2,016 flops; n=128 -> 1,390,016 + 130,560).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define SFMIN 2.2250738585072014e-308   /* dlamch('S') */

/* BLAS idamax: first index of max |x| (strict >), 0-based. */
int orc_idamax(int n, const double *x, int incx)
{
    if (n <= 0) return -1;
    int imax = 0;
    double dmax = fabs(x[0]);
    for (int i = 1; i < n; ++i) {
        double v = fabs(x[(size_t)i * incx]);
        if (v > dmax) { dmax = v; imax = i; }
    }
    return imax;
}

/*
 * Unblocked right-looking LU with partial pivoting, LAPACK dgetf2
 * (and the dgetrf2 leaf): pivot = idamax of the current column, interchange of
 * whole rows, scaling by the reciprocal pivot when |pivot| >= sfmin (else
 * division), rank-1 update a(i,j) = fma(-l(i), u(j), a(i,j)). A zero pivot
 * sets info (first occurrence, 1-based, plus gbstep) and the factorisation
 * continues (magmablas/zgetf2_devicefunc.cuh:254,268-278).
 * ipiv receives 1-based interchanges (+ gbstep). Returns info.
 */
int orc_dgetf2(int m, int n, double *A, int lda, int *ipiv, int gbstep)
{
    int info = 0;
    int minmn = m < n ? m : n;
    for (int j = 0; j < minmn; ++j) {
        int jp = j + orc_idamax(m - j, &A[j + (size_t)j * lda], 1);
        ipiv[j] = jp + 1 + gbstep;
        double piv = A[jp + (size_t)j * lda];
        if (piv != 0.0) {
            if (jp != j) {
                for (int c = 0; c < n; ++c) {
                    double t = A[j + (size_t)c * lda];
                    A[j + (size_t)c * lda] = A[jp + (size_t)c * lda];
                    A[jp + (size_t)c * lda] = t;
                }
            }
            if (fabs(piv) >= SFMIN) {
                double r = 1.0 / piv;
                for (int i = j + 1; i < m; ++i) A[i + (size_t)j * lda] *= r;
            } else {
                for (int i = j + 1; i < m; ++i) A[i + (size_t)j * lda] /= piv;
            }
        } else if (info == 0) {
            info = j + 1 + gbstep;
        }
        for (int c = j + 1; c < n; ++c) {
            double u = A[j + (size_t)c * lda];
            for (int i = j + 1; i < m; ++i)
                A[i + (size_t)c * lda] = fma(-A[i + (size_t)j * lda], u, A[i + (size_t)c * lda]);
        }
    }
    return info;
}

/* LAPACK dlaswp on ncols columns, interchanges k1..k2 (1-based, inclusive)
 * applied forward (incx > 0) or backward (incx < 0). */
void orc_dlaswp(int ncols, double *B, int ldb, int k1, int k2, const int *ipiv, int incx)
{
    if (incx > 0) {
        for (int k = k1; k <= k2; ++k) {
            int p = ipiv[k - 1];
            if (p != k)
                for (int c = 0; c < ncols; ++c) {
                    double t = B[(k - 1) + (size_t)c * ldb];
                    B[(k - 1) + (size_t)c * ldb] = B[(p - 1) + (size_t)c * ldb];
                    B[(p - 1) + (size_t)c * ldb] = t;
                }
        }
    } else {
        for (int k = k2; k >= k1; --k) {
            int p = ipiv[k - 1];
            if (p != k)
                for (int c = 0; c < ncols; ++c) {
                    double t = B[(k - 1) + (size_t)c * ldb];
                    B[(k - 1) + (size_t)c * ldb] = B[(p - 1) + (size_t)c * ldb];
                    B[(p - 1) + (size_t)c * ldb] = t;
                }
        }
    }
}

/*
 * LAPACK dgetrs with the factors of orc_dgetf2. trans = 0: X = U\(L\(P^T B));
 * trans = 1: X = P (L^T \ (U^T \ B)) (src/zgetrs_gpu.cpp:203-227 order).
 * Every triangular step is column oriented: solve one entry, then
 * b(i) = fma(-t(i,k), x(k), b(i)) for the rows still open.
 */
void orc_dgetrs(int trans, int n, int nrhs, const double *LU, int lda, const int *ipiv,
                double *B, int ldb)
{
    if (trans == 0) {
        orc_dlaswp(nrhs, B, ldb, 1, n, ipiv, 1);
        for (int c = 0; c < nrhs; ++c) {
            double *b = B + (size_t)c * ldb;
            for (int k = 0; k < n - 1; ++k)
                for (int i = k + 1; i < n; ++i)
                    b[i] = fma(-LU[i + (size_t)k * lda], b[k], b[i]);
            for (int k = n - 1; k >= 0; --k) {
                b[k] = b[k] / LU[k + (size_t)k * lda];
                for (int i = 0; i < k; ++i)
                    b[i] = fma(-LU[i + (size_t)k * lda], b[k], b[i]);
            }
        }
    } else {
        for (int c = 0; c < nrhs; ++c) {
            double *b = B + (size_t)c * ldb;
            for (int k = 0; k < n; ++k) {
                b[k] = b[k] / LU[k + (size_t)k * lda];
                for (int i = k + 1; i < n; ++i)
                    b[i] = fma(-LU[k + (size_t)i * lda], b[k], b[i]);
            }
            for (int k = n - 1; k > 0; --k)
                for (int i = 0; i < k; ++i)
                    b[i] = fma(-LU[k + (size_t)i * lda], b[k], b[i]);
        }
        orc_dlaswp(nrhs, B, ldb, 1, n, ipiv, -1);
    }
}

/* dgesv = dgetf2 + dgetrs('N') when the factorisation is nonsingular. */
int orc_dgesv(int n, int nrhs, double *A, int lda, int *ipiv, double *B, int ldb)
{
    int info = orc_dgetf2(n, n, A, lda, ipiv, 0);
    if (info == 0) orc_dgetrs(0, n, nrhs, A, lda, ipiv, B, ldb);
    return info;
}

/* Batched loops over contiguous storage (matrix s at A + s*strideA). */
void orc_dgetf2_batch(int m, int n, double *A, int lda, long strideA, int *ipiv, int *info, int batch)
{
    int minmn = m < n ? m : n;
    #pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < batch; ++s)
        info[s] = orc_dgetf2(m, n, A + (size_t)s * strideA, lda, ipiv + (size_t)s * minmn, 0);
}

void orc_dgetrs_batch(int trans, int n, int nrhs, const double *LU, int lda, long strideA,
                      const int *ipiv, double *B, int ldb, long strideB, int batch)
{
    #pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < batch; ++s)
        orc_dgetrs(trans, n, nrhs, LU + (size_t)s * strideA, lda, ipiv + (size_t)s * n,
                   B + (size_t)s * strideB, ldb);
}

/*
 * The row map behind magma_swp2pswp (control/auxiliary.cpp:81-139): LAPACK's
 * sequential interchanges applied to the identity give perm with
 * (P^T B)(i,:) = B(perm[i],:); that map (0-based) is returned for trans = 0,
 * and its inverse for trans != 0, which is what MAGMA's incremental loop
 * produces in each case.
 */
void orc_swp2pswp(int trans, int n, const int *ipiv, int *newipiv)
{
    int *perm = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
        int p = ipiv[k] - 1, t = perm[k];
        perm[k] = perm[p];
        perm[p] = t;
    }
    if (trans == 0)
        for (int i = 0; i < n; ++i) newipiv[i] = perm[i];
    else
        for (int i = 0; i < n; ++i) newipiv[perm[i]] = i;
    free(perm);
}

/* testing/flops.h:80-91 */
double orc_flops_dgetrf(double m, double n)
{
    double fmuls, fadds;
    if (m < n) {
        fmuls = 0.5 * m * (m * (n - (1. / 3.) * m - 1.) + n) + (2. / 3.) * m;
        fadds = 0.5 * m * (m * (n - (1. / 3.) * m) - n) + (1. / 6.) * m;
    } else {
        fmuls = 0.5 * n * (n * (m - (1. / 3.) * n - 1.) + m) + (2. / 3.) * n;
        fadds = 0.5 * n * (n * (m - (1. / 3.) * n) - m) + (1. / 6.) * n;
    }
    return fmuls + fadds;
}

double orc_flops_dgetrs(double n, double nrhs)
{
    return nrhs * n * n + nrhs * n * (n - 1.);
}
