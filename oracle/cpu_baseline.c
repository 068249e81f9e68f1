/*
 * ORACLE -- test infrastructure only (CPU baseline leg of bench.py and the
 * host-LAPACK cross-checks in tests/).
 *
 * The reference's CPU path for the batched solve: host LAPACK dgesv per
 * system, OpenMP parallel for schedule(dynamic) over the batch, LAPACK itself
 * single-threaded (testing/testing_zgesv_batched.cpp:158-180). The LAPACK
 * library is opened at run time (MKL's single dynamic library or any library
 * exporting the Fortran dgesv_/dgetrf_/dgetrs_ symbols) and its path recorded.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef void (*dgesv_t)(const int *, const int *, double *, const int *, int *, double *, const int *, int *);
typedef void (*dgetrf_t)(const int *, const int *, double *, const int *, int *, int *);
typedef void (*dgetrs_t)(const char *, const int *, const int *, const double *, const int *, const int *,
                         double *, const int *, int *, size_t);
typedef void (*set_threads_t)(int);

static void *g_lib = NULL;
static dgesv_t g_dgesv = NULL;
static dgetrf_t g_dgetrf = NULL;
static dgetrs_t g_dgetrs = NULL;
static char g_path[1024] = "";

/* Open the host LAPACK at `path`; returns 0 on success. */
int cb_open(const char *path)
{
    if (g_lib) return 0;
    /* MKL: pick the sequential layer so that each dgesv runs on one thread
     * inside the OpenMP loop, as the reference tester arranges. */
    setenv("MKL_THREADING_LAYER", "SEQUENTIAL", 0);
    setenv("MKL_INTERFACE_LAYER", "LP64", 0);
    setenv("OPENBLAS_NUM_THREADS", "1", 0);
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 1;
    /* Plain Fortran names first, then the prefixed names of SciPy's bundled
     * OpenBLAS. */
    static const char *prefixes[] = {"", "scipy_"};
    for (int i = 0; i < 2 && !g_dgesv; ++i) {
        char name[64];
        snprintf(name, sizeof(name), "%sdgesv_", prefixes[i]);
        g_dgesv = (dgesv_t)dlsym(h, name);
        snprintf(name, sizeof(name), "%sdgetrf_", prefixes[i]);
        g_dgetrf = (dgetrf_t)dlsym(h, name);
        snprintf(name, sizeof(name), "%sdgetrs_", prefixes[i]);
        g_dgetrs = (dgetrs_t)dlsym(h, name);
    }
    if (!g_dgesv || !g_dgetrf || !g_dgetrs) {
        g_dgesv = NULL; g_dgetrf = NULL; g_dgetrs = NULL;
        dlclose(h);
        return 2;
    }
    set_threads_t st = (set_threads_t)dlsym(h, "openblas_set_num_threads");
    if (!st) st = (set_threads_t)dlsym(h, "scipy_openblas_set_num_threads");
    if (st) st(1);
    set_threads_t mst = (set_threads_t)dlsym(h, "MKL_Set_Num_Threads");
    if (mst) mst(1);
    g_lib = h;
    strncpy(g_path, path, sizeof(g_path) - 1);
    return 0;
}

const char *cb_path(void) { return g_path; }

int cb_max_threads(void) { return omp_get_max_threads(); }

/* Solve `batch` systems stored contiguously (A_s at A + s*lda*n, B_s at
 * B + s*ldb*nrhs) with host dgesv on `nthreads` threads. Returns wall seconds
 * of the loop; info[s] gets LAPACK's info. */
double cb_dgesv_batch(int n, int nrhs, double *A, int lda, double *B, int ldb, int *ipiv, int *info,
                      int batch, int nthreads)
{
    if (!g_lib) return -1.0;
    double t0 = omp_get_wtime();
    #pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int s = 0; s < batch; ++s) {
        int li = 0;
        g_dgesv(&n, &nrhs, A + (size_t)s * lda * n, &lda, ipiv + (size_t)s * n,
                B + (size_t)s * ldb * nrhs, &ldb, &li);
        info[s] = li;
    }
    return omp_get_wtime() - t0;
}

/* Host dgetrf on one matrix (for cross-checks). */
int cb_dgetrf(int m, int n, double *A, int lda, int *ipiv)
{
    int info = 0;
    if (!g_lib) return -1000;
    g_dgetrf(&m, &n, A, &lda, ipiv, &info);
    return info;
}

/* Host dgetrs on one system; trans is 'N' or 'T'. */
int cb_dgetrs(char trans, int n, int nrhs, const double *LU, int lda, const int *ipiv, double *B, int ldb)
{
    int info = 0;
    if (!g_lib) return -1000;
    g_dgetrs(&trans, &n, &nrhs, LU, &lda, ipiv, B, &ldb, &info, 1);
    return info;
}
