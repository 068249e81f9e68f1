// Single-matrix drivers on the dense LU: magma_[ds]getrf_gpu,
// magma_dgetrf_native, magma_[ds]getrs_gpu (include/magma_z.h:612,638,713),
// and the refinement solver for one large system used by
// magma_dxgesv_gmres_gpu (src/dxgesv_gmres_gpu.cpp:160-832):
//   1. A is rounded to FP32 with an overflow check and ||A||_inf
//      (magmablas_dlag2s, dlange) -- overflow: iter = -2, FP64 fallback;
//   2. the FP32 copy is factored by the dense LU with FP32, BF16, FP16 or
//      split-BF16 trailing updates -- a zero pivot or a non-finite factor:
//      iter = -3, FP64 fallback;
//   3. refinement, with FP64 residuals r = b - A x and the test
//      ||r_j||_inf <= ||x_j||_inf * cte, cte = ||A||_inf * eps * sqrt(n)
//      (:490-493, :648, :755). The correction solve applies the factors
//      in FP32 (..STRS, magma_dsgetrs_gpu) or their FP64 copy (..DTRS,
//      magma_dgetrs_gpu on slag2d of them, :445-452):
//        IR*TRS    x = LU^{-1} b, then x += LU^{-1} r;
//        IRGM*TRS  x += d with d from an inner FGMRES on A d = r
//                  (50 iterations to 1e-4, preconditioned by the solve);
//        GM*TRS    FGMRES right-preconditioned by the solve, per column,
//                  restart = ITERMAX (:574-575), until ||r||_2 <= cte
//                  ||r0||_2 (src/dgmres_plu_gpu.cpp) or the test above
//                  holds for the true residual (MAGMA's test alone cannot
//                  be met in FP64 when ||x|| is large);
//        GMGM*TRS  the same with the inner FGMRES as preconditioner;
//      ITERMAX = min(n, 500); with an inner FGMRES the outer level takes
//      max(10, ITERMAX / 50) steps (:211-219, 328-359);
//   4. no convergence: FP64 LU and solve of the original system
//      (:801-810), iter = -steps. (MAGMA's FGMRES modes return without this
//      check; here every mode falls back, so the answer is always FP64
//      quality.)
// The host drives the iterations; vectors, residuals and the Krylov basis
// stay in device memory, and only norms and the Hessenberg column come back.
#include "magma_internal.h"
#include "lu_device.h"
#include "dense_kernels.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <type_traits>
#include <vector>

namespace mgi {

template <typename T>
magma_int_t dense_trsv_both(int trans, int n, int nrhs, const T* LU, int ld, T* Y, int ldy, const SolveWork& s,
                            hipStream_t st);
magma_int_t dense_solve_check_err(const SolveWork& s, hipStream_t stream);

namespace {

constexpr int kRT = 256;          // threads of the vector kernels
constexpr int kGmresInner = 50;   // inner_itermax of src/dxgesv_gmres_gpu.cpp:212

// --------------------------------------------------------------------------
// FP64 -> FP32 copy with the overflow flag and partial row sums of |a| over
// 256-column chunks: part[chunk * n + row].
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(kRT)
dn_lag2s_kernel(int n, const double* __restrict__ A, int lda, float* __restrict__ S, int lds,
                double* __restrict__ part, int* __restrict__ ovf)
{
    const int row = blockIdx.x * kRT + threadIdx.x;
    const int c0 = blockIdx.y * 256;
    if (row >= n) return;
    const int c1 = min(n, c0 + 256);
    double rs = 0.0;
    bool over = false;
    for (int j = c0; j < c1; ++j) {
        const double a = A[(size_t)row + (size_t)j * lda];
        rs += fabs(a);
        over = over || (fabs(a) > (double)FLT_MAX);
        S[(size_t)row + (size_t)j * lds] = (float)a;
    }
    part[(size_t)blockIdx.y * n + row] = rs;
    if (over) atomicOr(ovf, 1);
}

// Row sums over the chunks, maximum per workgroup -> bmax[blockIdx.x].
__global__ void __launch_bounds__(kRT)
dn_rowmax_kernel(int n, int nchunk, const double* __restrict__ part, double* __restrict__ bmax)
{
    __shared__ double red[kRT / kWave];
    const int row = blockIdx.x * kRT + threadIdx.x;
    double s = 0.0;
    if (row < n)
        for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * n + row];
    if (s != s) s = INFINITY;
    for (int off = 32; off >= 1; off >>= 1) s = fmax(s, shfl_xor_d(s, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < kRT / kWave; ++q) m = fmax(m, red[q]);
        bmax[blockIdx.x] = m;
    }
}

__global__ void __launch_bounds__(kRT)
dn_finite_kernel(size_t count, int n, const float* __restrict__ S, int lds, int* __restrict__ bad)
{
    for (size_t e = (size_t)blockIdx.x * kRT + threadIdx.x; e < count; e += (size_t)gridDim.x * kRT) {
        const size_t i = e % n, j = e / n;
        if (!isfinite(S[i + j * lds])) { atomicOr(bad, 1); return; }
    }
}

// --------------------------------------------------------------------------
// Residual R = B - op(A) X (FP64), partial sums over column (NoTrans) or row
// (Trans) chunks of A, then a reduction that also takes the per-column
// infinity norms of R and X for the stopping test.
// --------------------------------------------------------------------------
constexpr int kResChunks = 32;

// NoTrans: part[(cc * nrhs + q) * n + row] = sum_{j in chunk cc} A[row, j] X[j, q]
template <int NR>
__global__ void __launch_bounds__(kRT)
dn_gemv_n_kernel(int n, int nrhs, const double* __restrict__ A, int lda, const double* __restrict__ X, int ldx,
                 double* __restrict__ part)
{
    __shared__ double xs[256][NR];
    const int row = blockIdx.x * kRT + threadIdx.x;
    const int cc = blockIdx.y;
    const int chunk = (n + kResChunks - 1) / kResChunks;
    const int c0 = cc * chunk, c1 = min(n, c0 + chunk);
    const int q0 = blockIdx.z * NR;
    const int nq = min(NR, nrhs - q0);
    double acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = 0.0;
    for (int j0 = c0; j0 < c1; j0 += 256) {
        __syncthreads();
        for (int e = threadIdx.x; e < 256 * NR; e += kRT) {
            const int jj = e % 256, q = e / 256;
            xs[jj][q] = (j0 + jj < c1 && q < nq) ? X[(size_t)(j0 + jj) + (size_t)(q0 + q) * ldx] : 0.0;
        }
        __syncthreads();
        if (row < n) {
            const int je = min(256, c1 - j0);
            for (int jj = 0; jj < je; ++jj) {
                const double a = A[(size_t)row + (size_t)(j0 + jj) * lda];
#pragma unroll
                for (int q = 0; q < NR; ++q) acc[q] = fma(a, xs[jj][q], acc[q]);
            }
        }
    }
    if (row < n) {
#pragma unroll
        for (int q = 0; q < NR; ++q)
            if (q < nq) part[((size_t)cc * nrhs + q0 + q) * n + row] = acc[q];
    }
}

// Trans: part[(rc * nrhs + q) * n + col] = sum_{i in row chunk rc} A[i, col] X[i, q];
// one wave per column.
template <int NR>
__global__ void __launch_bounds__(kRT)
dn_gemv_t_kernel(int n, int nrhs, const double* __restrict__ A, int lda, const double* __restrict__ X, int ldx,
                 double* __restrict__ part)
{
    const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
    const int rc = blockIdx.y;
    const int chunk = (n + kResChunks - 1) / kResChunks;
    const int r0 = rc * chunk, r1 = min(n, r0 + chunk);
    const int q0 = blockIdx.z * NR;
    const int nq = min(NR, nrhs - q0);
    for (int col = blockIdx.x * 16 + wv; col < min(n, blockIdx.x * 16 + 16); col += 4) {
        double acc[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = 0.0;
        for (int i = r0 + ln; i < r1; i += kWave) {
            const double a = A[(size_t)i + (size_t)col * lda];
#pragma unroll
            for (int q = 0; q < NR; ++q)
                if (q < nq) acc[q] = fma(a, X[(size_t)i + (size_t)(q0 + q) * ldx], acc[q]);
        }
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            double v = acc[q];
            for (int off = 32; off >= 1; off >>= 1) v += shfl_xor_d(v, off);
            if (ln == 0 && q < nq) part[((size_t)rc * nrhs + q0 + q) * n + col] = v;
        }
    }
}

// R[:, q] = B[:, q] - sum_c part; nrm[(blockIdx.x * nrhs + q) * 2 + {0,1}] =
// block maxima of |R| and |X| (NaN counted as +inf).
__global__ void __launch_bounds__(kRT)
dn_resid_reduce_kernel(int n, int nrhs, int nchunk, const double* __restrict__ part, const double* __restrict__ B,
                       int ldb, const double* __restrict__ X, int ldx, double* __restrict__ R, int ldr,
                       double* __restrict__ nrm)
{
    __shared__ double red[2][kRT / kWave];
    const int row = blockIdx.x * kRT + threadIdx.x;
    const int q = blockIdx.y;
    double rmax = 0.0, xmax = 0.0;
    if (row < n) {
        double s = 0.0;
        for (int c = 0; c < nchunk; ++c) s += part[((size_t)c * nrhs + q) * n + row];
        const double r = B[(size_t)row + (size_t)q * ldb] - s;
        R[(size_t)row + (size_t)q * ldr] = r;
        const double ra = fabs(r), xa = fabs(X[(size_t)row + (size_t)q * ldx]);
        rmax = (ra == ra) ? ra : INFINITY;
        xmax = (xa == xa) ? xa : INFINITY;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        rmax = fmax(rmax, shfl_xor_d(rmax, off));
        xmax = fmax(xmax, shfl_xor_d(xmax, off));
    }
    if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x / 64] = rmax; red[1][threadIdx.x / 64] = xmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red[0][0], b = red[1][0];
        for (int w = 1; w < kRT / kWave; ++w) { a = fmax(a, red[0][w]); b = fmax(b, red[1][w]); }
        nrm[((size_t)blockIdx.x * nrhs + q) * 2 + 0] = a;
        nrm[((size_t)blockIdx.x * nrhs + q) * 2 + 1] = b;
    }
}

// out[row] = sum_c part[c * n + row]
__global__ void dn_sum_parts_kernel(int n, int nchunk, const double* __restrict__ part, double* __restrict__ out)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * n + row];
    out[row] = s;
}

// FP64 -> TD with the row map (NoTrans) or without (Trans)
template <typename TD>
__global__ void dn_gather_kernel(int n, int nrhs, const int* __restrict__ perm, const double* __restrict__ S, int lds,
                                 TD* __restrict__ D, int ldd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i < n && q < nrhs) D[(size_t)i + (size_t)q * ldd] = (TD)S[(size_t)(perm ? perm[i] : i) + (size_t)q * lds];
}

// X[dest(i)] (+)= Y[i]: dest = i (NoTrans) or perm[i] (Trans); acc = 0 stores.
template <typename TS>
__global__ void dn_apply_kernel(int n, int nrhs, const int* __restrict__ perm, const TS* __restrict__ Y, int ldy,
                                double* __restrict__ X, int ldx, int acc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i < n && q < nrhs) {
        double* p = X + (size_t)(perm ? perm[i] : i) + (size_t)q * ldx;
        const double y = (double)Y[(size_t)i + (size_t)q * ldy];
        *p = acc ? *p + y : y;
    }
}

// FP32 factors -> FP64 copy, ld n (magmablas_slag2d)
__global__ void dn_lag2d_kernel(int n, const float* __restrict__ S, int lds, double* __restrict__ D)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n) D[(size_t)i + (size_t)j * n] = (double)S[(size_t)i + (size_t)j * lds];
}

__global__ void dn_copy_d_kernel(int n, int nrhs, const double* __restrict__ S, int lds, double* __restrict__ D,
                                 int ldd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i < n && q < nrhs) D[(size_t)i + (size_t)q * ldd] = S[(size_t)i + (size_t)q * lds];
}

__global__ void dn_zero_d_kernel(int n, int nrhs, double* __restrict__ D, int ldd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i < n && q < nrhs) D[(size_t)i + (size_t)q * ldd] = 0.0;
}

// --------------------------------------------------------------------------
// FGMRES vector kernels (one right-hand side; V, Z are n x (m+1), ld n)
// --------------------------------------------------------------------------
constexpr int kDotChunks = 64;

// part[c * (cnt+1) + j] = sum_{rows of chunk c} V[:, j] * w   (j < cnt);
// part[c * (cnt+1) + cnt] = sum w^2 when with_norm.
__global__ void __launch_bounds__(kRT)
dn_dots_kernel(int n, int cnt, const double* __restrict__ V, const double* __restrict__ w, double* __restrict__ part,
               int with_norm)
{
    __shared__ double red[kRT / kWave];
    const int c = blockIdx.x;
    const int chunk = (n + kDotChunks - 1) / kDotChunks;
    const int r0 = c * chunk, r1 = min(n, r0 + chunk);
    const int total = cnt + (with_norm ? 1 : 0);
    for (int j = 0; j < total; ++j) {
        double s = 0.0;
        for (int i = r0 + threadIdx.x; i < r1; i += kRT) {
            const double wi = w[i];
            s = fma(j < cnt ? V[(size_t)i + (size_t)j * n] : wi, wi, s);
        }
        for (int off = 32; off >= 1; off >>= 1) s += shfl_xor_d(s, off);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0.0;
            for (int q = 0; q < kRT / kWave; ++q) a += red[q];
            part[(size_t)c * (cnt + 1) + j] = a;
        }
    }
}

// h[j] (+)= sum_c part[c * (cnt+1) + j] for j < total; accumulate into hacc
// too when hacc != null.
__global__ void dn_dots_reduce_kernel(int nchunk, int cnt, int total, const double* __restrict__ part,
                                      double* __restrict__ h, double* __restrict__ hacc)
{
    const int j = threadIdx.x;
    if (j >= total) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * (cnt + 1) + j];
    h[j] = s;
    if (hacc) hacc[j] += s;
}

// w -= V[:, 0..cnt) h
__global__ void dn_orth_update_kernel(int n, int cnt, const double* __restrict__ V, const double* __restrict__ h,
                                      double* __restrict__ w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = w[i];
    for (int j = 0; j < cnt; ++j) s = fma(-V[(size_t)i + (size_t)j * n], h[j], s);
    w[i] = s;
}

// dst = src * (1 / *nrm2) where *nrm2 holds a squared norm (a zero norm
// leaves dst = 0)
__global__ void dn_scale_kernel(int n, const double* __restrict__ src, const double* __restrict__ nrm2,
                                double* __restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double nv = sqrt(*nrm2);
    dst[i] = nv > 0.0 ? src[i] / nv : 0.0;
}

// x += Z[:, 0..cnt) y
__global__ void dn_xupdate_kernel(int n, int cnt, const double* __restrict__ Z, const double* __restrict__ y,
                                  double* __restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = x[i];
    for (int j = 0; j < cnt; ++j) s = fma(Z[(size_t)i + (size_t)j * n], y[j], s);
    x[i] = s;
}

__global__ void dn_zero_i_kernel(int* p, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

// --------------------------------------------------------------------------
// Host helpers
// --------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, std::max<size_t>(bytes, 256)); }
};

struct StreamHolder {
    hipStream_t s = nullptr;
    ~StreamHolder() { if (s) (void)hipStreamDestroy(s); }
};

unsigned gx(int n) { return (unsigned)ceildiv(std::max(n, 1), kRT); }

size_t a256_bytes(size_t b) { return (b + 255) & ~size_t(255); }

// FP64 LU of one matrix with its own scratch (host ipiv).
template <typename T>
magma_int_t getrf_one(int mode, int m, int n, T* dA, int ldda, int* h_ipiv, int* h_info, hipStream_t st)
{
    const int minmn = std::min(m, n);
    const int gmax = std::max(dense_panel_gmax(m), 1);
    if (gmax > device_cu_count()) return MAGMA_ERR_NOT_SUPPORTED;
    DevBuf work, small;
    MGI_HIP_CHECK(work.alloc(dense_work_bytes(m, n, (int)sizeof(T), gmax)));
    MGI_HIP_CHECK(small.alloc(sizeof(int) * (minmn + 64)));
    DenseWork w;
    dense_work_carve(work.p, m, n, (int)sizeof(T), gmax, w);
    int* d_info = static_cast<int*>(small.p);
    int* d_ipiv = d_info + 64;
    magma_int_t rc;
    StreamHolder side;
    side.s = dense_side_stream(m, (int)sizeof(T));
    if constexpr (std::is_same<T, double>::value)
        rc = dense_getrf(kUpdF64, m, n, dA, ldda, d_ipiv, d_info, w, st, 0, side.s);
    else rc = dense_sgetrf(mode, m, n, dA, ldda, d_ipiv, d_info, w, st, 0, side.s);
    if (rc != MAGMA_SUCCESS) return rc;
    if ((rc = dense_check_err(w, st)) != MAGMA_SUCCESS) return rc;
    MGI_HIP_CHECK(hipMemcpy(h_ipiv, d_ipiv, sizeof(int) * minmn, hipMemcpyDeviceToHost));
    MGI_HIP_CHECK(hipMemcpy(h_info, d_info, sizeof(int), hipMemcpyDeviceToHost));
    return MAGMA_SUCCESS;
}

template <typename T>
magma_int_t getrs_one(int trans, int n, int nrhs, const T* dA, int ldda, const int* h_ipiv, T* dB, int lddb,
                      hipStream_t st)
{
    const int nrhs_max = std::min(nrhs, 64);
    DevBuf sw;
    MGI_HIP_CHECK(sw.alloc(solve_work_bytes(n, nrhs_max, (int)sizeof(T))));
    SolveWork s;
    solve_work_carve(sw.p, n, nrhs_max, (int)sizeof(T), s);
    magma_int_t rc = dense_set_perm(n, h_ipiv, s, st);
    if (rc != MAGMA_SUCCESS) return rc;
    if ((rc = dense_diag_inverses<T>(n, dA, ldda, s, st)) != MAGMA_SUCCESS) return rc;
    if ((rc = dense_getrs<T>(trans, n, nrhs, dA, ldda, dB, lddb, s, st)) != MAGMA_SUCCESS) return rc;
    return dense_solve_check_err(s, st);
}

template <typename T>
magma_int_t getrf_entry(const char* name, int mode, magma_int_t m, magma_int_t n, T* dA, magma_int_t ldda,
                        magma_int_t* ipiv, magma_int_t* info)
{
    *info = 0;
    if (m < 0) *info = -1;
    else if (n < 0) *info = -2;
    else if (ldda < std::max(1, m)) *info = -4;
    if (*info != 0) {
        magma_xerbla(name, -(*info));
        return *info;
    }
    if (m == 0 || n == 0) return *info;
    StreamHolder sh;
    MGI_HIP_CHECK(hipStreamCreateWithFlags(&sh.s, hipStreamNonBlocking));
    magma_int_t rc = getrf_one<T>(mode, m, n, dA, ldda, ipiv, info, sh.s);
    if (rc != MAGMA_SUCCESS) return rc;
    return *info;
}

template <typename T>
magma_int_t getrs_entry(const char* name, magma_trans_t trans, magma_int_t n, magma_int_t nrhs, const T* dA,
                        magma_int_t ldda, const magma_int_t* ipiv, T* dB, magma_int_t lddb, magma_int_t* info)
{
    *info = 0;
    if (trans != MagmaNoTrans && trans != MagmaTrans && trans != MagmaConjTrans) *info = -1;
    else if (n < 0) *info = -2;
    else if (nrhs < 0) *info = -3;
    else if (ldda < std::max(1, n)) *info = -5;
    else if (lddb < std::max(1, n)) *info = -8;
    if (*info != 0) {
        magma_xerbla(name, -(*info));
        return *info;
    }
    if (n == 0 || nrhs == 0) return *info;
    StreamHolder sh;
    MGI_HIP_CHECK(hipStreamCreateWithFlags(&sh.s, hipStreamNonBlocking));
    magma_int_t rc = getrs_one<T>(trans == MagmaNoTrans ? 0 : 1, n, nrhs, dA, ldda, ipiv, dB, lddb, sh.s);
    if (rc != MAGMA_SUCCESS) return rc;
    return *info;
}

}  // namespace

// --------------------------------------------------------------------------
// Refinement driver for one system (see the file header).
// --------------------------------------------------------------------------
struct DenseIrCtx {
    int trans, n, nrhs;
    double* dA; int ldda;
    double* dB; int lddb;
    double* dX; int lddx;
    double* dR; int ldr;        // FP64 residual (caller's dworkd when given)
    float* S; int lds;          // FP32 LU
    float* Ys;                  // FP32 staging n x nrhs
    SolveWork sw;
    bool dtrs;                  // correction solves on the FP64 copy of the factors
    double* Dl;                 // that copy, ld n
    double* Yd;                 // FP64 staging n x nrhs
    SolveWork swd;
    double* part; double* nrm;
    double cte;
    hipStream_t st;
};

// R = B - op(A) X for nq columns; per-column ||R||_inf and ||X||_inf.
static magma_int_t ir_residual(DenseIrCtx& c, const double* B, int ldb, const double* X, int ldx, double* R, int ldr,
                               int nq, std::vector<double>& rn, std::vector<double>& xn)
{
    const int n = c.n;
    if (c.trans == 0) {
        dn_gemv_n_kernel<1><<<dim3(gx(n), kResChunks, nq), kRT, 0, c.st>>>(n, nq, c.dA, c.ldda, X, ldx, c.part);
    } else {
        dn_gemv_t_kernel<1><<<dim3((unsigned)ceildiv(n, 16), kResChunks, nq), kRT, 0, c.st>>>(n, nq, c.dA, c.ldda, X,
                                                                                              ldx, c.part);
    }
    MGI_LAUNCH_CHECK();
    dn_resid_reduce_kernel<<<dim3(gx(n), nq), kRT, 0, c.st>>>(n, nq, kResChunks, c.part, B, ldb, X, ldx, R, ldr,
                                                               c.nrm);
    MGI_LAUNCH_CHECK();
    std::vector<double> h((size_t)gx(n) * nq * 2);
    MGI_HIP_CHECK(hipMemcpyAsync(h.data(), c.nrm, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.st));
    MGI_HIP_CHECK(hipStreamSynchronize(c.st));
    rn.assign(nq, 0.0);
    xn.assign(nq, 0.0);
    for (unsigned b = 0; b < gx(n); ++b)
        for (int q = 0; q < nq; ++q) {
            rn[q] = std::max(rn[q], h[((size_t)b * nq + q) * 2 + 0]);
            xn[q] = std::max(xn[q], h[((size_t)b * nq + q) * 2 + 1]);
        }
    return MAGMA_SUCCESS;
}

// D (+)= op(LU)^{-1} S for nq columns (S, D FP64): S solves round S to FP32
// and use the FP32 factors (magma_dsgetrs_gpu), D solves use the FP64 copy
// of the same factors (magma_dgetrs_gpu on slag2d of them).
static magma_int_t ir_trs(DenseIrCtx& c, const double* Sv, int ldsv, int nq, double* D, int ldd, int acc)
{
    // at most the solve scratch's column count per pass
    const int cap = c.dtrs ? c.swd.nrhs_max : c.sw.nrhs_max;
    for (int q0 = cap; q0 < nq; q0 += cap) {
        const magma_int_t rc = ir_trs(c, Sv + (size_t)q0 * ldsv, ldsv, std::min(cap, nq - q0), D + (size_t)q0 * ldd,
                                      ldd, acc);
        if (rc != MAGMA_SUCCESS) return rc;
    }
    nq = std::min(nq, cap);
    const int n = c.n;
    dim3 g(gx(n), (unsigned)nq);
    magma_int_t rc;
    if (!c.dtrs) {
        dn_gather_kernel<float><<<g, kRT, 0, c.st>>>(n, nq, c.trans == 0 ? c.sw.perm : nullptr, Sv, ldsv, c.Ys, n);
        MGI_LAUNCH_CHECK();
        if ((rc = dense_trsv_both<float>(c.trans, n, nq, c.S, c.lds, c.Ys, n, c.sw, c.st)) != MAGMA_SUCCESS) return rc;
        dn_apply_kernel<float><<<g, kRT, 0, c.st>>>(n, nq, c.trans == 0 ? nullptr : c.sw.perm, c.Ys, n, D, ldd, acc);
    } else {
        dn_gather_kernel<double><<<g, kRT, 0, c.st>>>(n, nq, c.trans == 0 ? c.swd.perm : nullptr, Sv, ldsv, c.Yd, n);
        MGI_LAUNCH_CHECK();
        if ((rc = dense_trsv_both<double>(c.trans, n, nq, c.Dl, n, c.Yd, n, c.swd, c.st)) != MAGMA_SUCCESS) return rc;
        dn_apply_kernel<double><<<g, kRT, 0, c.st>>>(n, nq, c.trans == 0 ? nullptr : c.swd.perm, c.Yd, n, D, ldd,
                                                     acc);
    }
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

static bool col_ok(double r, double x, double cte)
{
    return std::isfinite(r) && std::isfinite(x) && r <= x * cte;
}

// Krylov workspace of one FGMRES level: V n x (m+1), Z n x m, r n, and the
// reduction buffers (pointers into the memory a GmOwner holds).
struct GmBuf {
    double* V = nullptr; double* Z = nullptr; double* r = nullptr;
    double* part = nullptr; double* hd = nullptr; double* hacc = nullptr; double* nrm2 = nullptr;
    int m = 0;
};

struct GmOwner {
    DevBuf vb, hb;
    GmBuf g;
    hipError_t alloc(int n, int restart)
    {
        const int m = restart;
        hipError_t e = vb.alloc(sizeof(double) * (size_t)n * (2 * m + 2));
        if (e != hipSuccess) return e;
        if ((e = hb.alloc(sizeof(double) * ((size_t)kDotChunks * (m + 2) + 4 * (m + 2)))) != hipSuccess) return e;
        g.m = m;
        g.V = static_cast<double*>(vb.p);
        g.Z = g.V + (size_t)n * (m + 1);
        g.r = g.Z + (size_t)n * m;
        g.part = static_cast<double*>(hb.p);             // kDotChunks x (m+2)
        g.hd = g.part + (size_t)kDotChunks * (m + 2);    // m+2: h of the current pass
        g.hacc = g.hd + (m + 2);                         // m+2: h summed over the two passes
        g.nrm2 = g.hacc + (m + 2);                       // [0] squared norm
        return hipSuccess;
    }
};

// One level of the FGMRES of src/dgmres_plu_gpu.cpp (Saad, Iterative
// Methods, Alg. 9.6) for one column: A x = b from the x given, restart g.m,
// at most maxit inner iterations. A cycle starts from the true residual
// r = b - A x; it stops when ||r||_2 <= tol ||r0||_2 and, for an outer
// solve, also when the refinement test ||r||_inf <= ||x||_inf cte holds
// (MAGMA's test alone cannot be met in FP64 once ||x|| is large). The
// preconditioner is the triangular solve, or (nested != null) an inner
// FGMRES of 50 iterations to 1e-4 preconditioned by it (GMGM, IRGM modes).
// *iters: iterations of this level; *inner: iterations of the nested level.
static magma_int_t fgmres_col(DenseIrCtx& c, const double* b, double* x, GmBuf& g, int maxit, double tol, bool outer,
                              GmBuf* nested, int* iters, long* inner, bool* converged)
{
    const int n = c.n, m = g.m;
    const unsigned g1 = gx(n);
    std::vector<double> rn, xn;
    magma_int_t rc;
    int total = 0;
    double eps1 = -1.0;
    *converged = false;
    for (;;) {
        // r = b - A x, beta = ||r||_2, v_0 = r / beta
        if ((rc = ir_residual(c, b, n, x, n, g.r, n, 1, rn, xn)) != MAGMA_SUCCESS) return rc;
        dn_dots_kernel<<<kDotChunks, kRT, 0, c.st>>>(n, 0, g.V, g.r, g.part, 1);
        dn_dots_reduce_kernel<<<1, 64, 0, c.st>>>(kDotChunks, 0, 1, g.part, g.nrm2, nullptr);
        dn_scale_kernel<<<g1, kRT, 0, c.st>>>(n, g.r, g.nrm2, g.V);
        MGI_LAUNCH_CHECK();
        double beta2 = 0.0;
        MGI_HIP_CHECK(hipMemcpyAsync(&beta2, g.nrm2, sizeof(double), hipMemcpyDeviceToHost, c.st));
        MGI_HIP_CHECK(hipStreamSynchronize(c.st));
        const double beta = std::sqrt(beta2);
        if (!std::isfinite(beta)) break;
        if (eps1 < 0.0) eps1 = tol * beta;
        if (beta <= eps1 || beta == 0.0 || (outer && col_ok(rn[0], xn[0], c.cte))) {
            *converged = true;
            break;
        }
        if (total >= maxit) break;
        // inner stop on the residual estimate; for an outer solve also at half
        // the refinement bound with ||x|| of the cycle start (the next cycle
        // checks the true residual)
        const double eps2 = outer ? std::max(eps1, 0.5 * c.cte * xn[0]) : eps1;
        std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m, 0.0), sn(m, 0.0), gv(m + 1, 0.0);
        gv[0] = beta;
        int k = 0;
        for (int i = 0; i < m && total < maxit; ++i) {
            double* vi = g.V + (size_t)i * n;
            double* zi = g.Z + (size_t)i * n;
            double* w = g.V + (size_t)(i + 1) * n;
            if (nested) {
                dn_zero_d_kernel<<<dim3(g1, 1), kRT, 0, c.st>>>(n, 1, zi, n);
                MGI_LAUNCH_CHECK();
                int it_in = 0;
                bool conv_in = false;
                if ((rc = fgmres_col(c, vi, zi, *nested, nested->m, 1e-4, false, nullptr, &it_in, nullptr,
                                     &conv_in)) != MAGMA_SUCCESS)
                    return rc;
                if (inner) *inner += it_in;
            } else if ((rc = ir_trs(c, vi, n, 1, zi, n, 0)) != MAGMA_SUCCESS) {
                return rc;
            }
            // w = A z_i
            if (c.trans == 0)
                dn_gemv_n_kernel<1><<<dim3(g1, kResChunks, 1), kRT, 0, c.st>>>(n, 1, c.dA, c.ldda, zi, n, c.part);
            else
                dn_gemv_t_kernel<1><<<dim3((unsigned)ceildiv(n, 16), kResChunks, 1), kRT, 0, c.st>>>(
                    n, 1, c.dA, c.ldda, zi, n, c.part);
            MGI_LAUNCH_CHECK();
            dn_sum_parts_kernel<<<g1, kRT, 0, c.st>>>(n, kResChunks, c.part, w);
            MGI_LAUNCH_CHECK();
            // two classical Gram-Schmidt passes against v_0..v_i, then ||w||
            dn_zero_d_kernel<<<dim3(1, 1), kRT, 0, c.st>>>(i + 2, 1, g.hacc, m + 2);
            for (int pass = 0; pass < 2; ++pass) {
                dn_dots_kernel<<<kDotChunks, kRT, 0, c.st>>>(n, i + 1, g.V, w, g.part, 0);
                dn_dots_reduce_kernel<<<1, 64, 0, c.st>>>(kDotChunks, i + 1, i + 1, g.part, g.hd, g.hacc);
                dn_orth_update_kernel<<<g1, kRT, 0, c.st>>>(n, i + 1, g.V, g.hd, w);
            }
            dn_dots_kernel<<<kDotChunks, kRT, 0, c.st>>>(n, 0, g.V, w, g.part, 1);
            dn_dots_reduce_kernel<<<1, 64, 0, c.st>>>(kDotChunks, 0, 1, g.part, g.nrm2, nullptr);
            dn_scale_kernel<<<g1, kRT, 0, c.st>>>(n, w, g.nrm2, w);
            MGI_LAUNCH_CHECK();
            std::vector<double> hcol(i + 2);
            MGI_HIP_CHECK(hipMemcpyAsync(hcol.data(), g.hacc, sizeof(double) * (i + 1), hipMemcpyDeviceToHost, c.st));
            MGI_HIP_CHECK(hipMemcpyAsync(&beta2, g.nrm2, sizeof(double), hipMemcpyDeviceToHost, c.st));
            MGI_HIP_CHECK(hipStreamSynchronize(c.st));
            hcol[i + 1] = std::sqrt(beta2);
            ++total;
            // Givens rotations (Saad, Iterative Methods, Alg. 6.9)
            for (int j = 0; j < i; ++j) {
                const double a = hcol[j], bb = hcol[j + 1];
                hcol[j] = cs[j] * a + sn[j] * bb;
                hcol[j + 1] = -sn[j] * a + cs[j] * bb;
            }
            const double a = hcol[i], bb = hcol[i + 1];
            const double den = std::hypot(a, bb);
            cs[i] = den > 0.0 ? a / den : 1.0;
            sn[i] = den > 0.0 ? bb / den : 0.0;
            hcol[i] = den;
            hcol[i + 1] = 0.0;
            gv[i + 1] = -sn[i] * gv[i];
            gv[i] = cs[i] * gv[i];
            for (int j = 0; j <= i; ++j) H[(size_t)j + (size_t)i * (m + 1)] = hcol[j];
            k = i + 1;
            if (!std::isfinite(gv[i + 1])) break;
            if (std::fabs(gv[i + 1]) <= eps2) break;
        }
        // y = H(0:k,0:k)^{-1} g, x += Z y
        std::vector<double> y(k, 0.0);
        for (int j = k - 1; j >= 0; --j) {
            double sum = gv[j];
            for (int l = j + 1; l < k; ++l) sum -= H[(size_t)j + (size_t)l * (m + 1)] * y[l];
            const double d = H[(size_t)j + (size_t)j * (m + 1)];
            y[j] = d != 0.0 ? sum / d : 0.0;
        }
        if (k > 0) {
            MGI_HIP_CHECK(hipMemcpyAsync(g.hd, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, c.st));
            dn_xupdate_kernel<<<g1, kRT, 0, c.st>>>(n, k, g.Z, g.hd, x);
            MGI_LAUNCH_CHECK();
            MGI_HIP_CHECK(hipStreamSynchronize(c.st));   // y is pageable host memory
        }
        // an inner solve is one cycle (restart = maxit = 50), ended on the
        // residual estimate as in src/dgmres_plu_gpu.cpp
        if (!outer) {
            *converged = k > 0 && std::fabs(gv[k]) <= eps2;
            break;
        }
    }
    *iters = total;
    return MAGMA_SUCCESS;
}

// Classical refinement (IRSTRS, IRDTRS, IRGMSTRS, IRGMDTRS), all columns
// together: x = LU^{-1} b, then x += d with d = LU^{-1} r, or d from an inner
// FGMRES on A d = r (nested != null), until every column passes the test.
// *steps: outer steps, or with an inner solver its iterations in total.
static magma_int_t ir_classical(DenseIrCtx& c, int itermax, GmBuf* nested, double* dz, int* steps, bool* converged)
{
    const int n = c.n, nrhs = c.nrhs;
    *converged = false;
    *steps = 0;
    long tot_inner = 0;
    magma_int_t rc = ir_trs(c, c.dB, c.lddb, nrhs, c.dX, c.lddx, 0);
    if (rc != MAGMA_SUCCESS) return rc;
    std::vector<double> rn, xn;
    for (int it = 0; it <= itermax; ++it) {
        if ((rc = ir_residual(c, c.dB, c.lddb, c.dX, c.lddx, c.dR, c.ldr, nrhs, rn, xn)) != MAGMA_SUCCESS) return rc;
        bool all = true, bad = false;
        for (int q = 0; q < nrhs; ++q) {
            if (!std::isfinite(rn[q]) || !std::isfinite(xn[q])) bad = true;
            if (!col_ok(rn[q], xn[q], c.cte)) all = false;
        }
        *steps = nested ? (int)std::min<long>(tot_inner, INT32_MAX) : it;
        if (all) { *converged = true; return MAGMA_SUCCESS; }
        if (bad || it == itermax) return MAGMA_SUCCESS;
        if (!nested) {
            if ((rc = ir_trs(c, c.dR, c.ldr, nrhs, c.dX, c.lddx, 1)) != MAGMA_SUCCESS) return rc;
            continue;
        }
        for (int q = 0; q < nrhs; ++q) {
            dn_zero_d_kernel<<<dim3(gx(n), 1), kRT, 0, c.st>>>(n, 1, dz, n);
            MGI_LAUNCH_CHECK();
            int it_in = 0;
            bool conv_in = false;
            if ((rc = fgmres_col(c, c.dR + (size_t)q * c.ldr, dz, *nested, nested->m, 1e-4, false, nullptr, &it_in,
                                 nullptr, &conv_in)) != MAGMA_SUCCESS)
                return rc;
            tot_inner += it_in;
            dn_apply_kernel<double><<<dim3(gx(n), 1), kRT, 0, c.st>>>(n, 1, nullptr, dz, n,
                                                                      c.dX + (size_t)q * c.lddx, c.lddx, 1);
            MGI_LAUNCH_CHECK();
        }
    }
    return MAGMA_SUCCESS;
}

// FGMRES refinement (GMSTRS, GMDTRS, GMGMSTRS, GMGMDTRS) per column from
// x = 0. *steps: outer iterations, or with an inner solver its iterations.
static magma_int_t ir_fgmres(DenseIrCtx& c, int itermax, GmBuf& outer, GmBuf* nested, int* steps, bool* converged)
{
    *steps = 0;
    *converged = true;
    int total = 0;
    long tot_inner = 0;
    for (int col = 0; col < c.nrhs; ++col) {
        double* x = c.dX + (size_t)col * c.lddx;
        double* b = c.dB + (size_t)col * c.lddb;
        dn_zero_d_kernel<<<dim3(gx(c.n), 1), kRT, 0, c.st>>>(c.n, 1, x, c.lddx);
        MGI_LAUNCH_CHECK();
        int it = 0;
        bool conv = false;
        magma_int_t rc = fgmres_col(c, b, x, outer, itermax, c.cte, true, nested, &it, &tot_inner, &conv);
        if (rc != MAGMA_SUCCESS) return rc;
        total += it;
        if (!conv) { *converged = false; break; }
    }
    *steps = nested ? (int)std::min<long>(tot_inner, INT32_MAX) : total;
    return MAGMA_SUCCESS;
}

// Row map of the interchanges as magma_swp2pswp (control/auxiliary.cpp:81)
static void swp2pswp(int trans, int n, const int* ipiv, std::vector<int>& out)
{
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) std::swap(perm[k], perm[ipiv[k] - 1]);
    out.assign(n, 0);
    if (trans == 0) out = perm;
    else for (int i = 0; i < n; ++i) out[perm[i]] = i;
}

magma_int_t dense_iteref(int trans, int n, int nrhs, double* dA, int ldda, int* ipiv, int* dipiv, double* dB,
                         int lddb, double* dX, int lddx, double* dworkd, float* dworks, int mode, int solver,
                         int* iter, int* info)
{
    *iter = 0;
    *info = 0;
    const int gmax = std::max(dense_panel_gmax(n), 1);
    if (gmax > device_cu_count()) return MAGMA_ERR_NOT_SUPPORTED;
    const bool gm_outer = (solver & kRefineOuterGM) != 0;
    const bool nested = (solver & kRefineInnerGM) != 0;
    StreamHolder sh;
    MGI_HIP_CHECK(hipStreamCreateWithFlags(&sh.s, hipStreamNonBlocking));
    hipStream_t st = sh.s;
    const int nchunk = (int)ceildiv(n, 256);
    const int nrhs_max = std::max(nrhs, 1);
    // the caller's FP32 workspace (n*(n+nrhs) floats) holds the FP32 factors
    // (ld n) and the FP32 right-hand sides, as MAGMA documents dworks
    const size_t lu_bytes = dworks ? 0 : a256_bytes(sizeof(float) * (size_t)n * n);
    DevBuf big;
    const size_t ys_bytes = dworks ? 0 : a256_bytes(sizeof(float) * (size_t)n * nrhs_max);
    const size_t dw_bytes = a256_bytes(dense_work_bytes(n, n, 4, gmax));
    const size_t sw_bytes = a256_bytes(solve_work_bytes(n, std::min(nrhs_max, 64), 4));
    const size_t part_elems = std::max<size_t>((size_t)kResChunks * nrhs_max * n, (size_t)nchunk * n);
    const size_t part_bytes = a256_bytes(sizeof(double) * part_elems);
    const size_t nrm_bytes = a256_bytes(sizeof(double) * ((size_t)gx(n) * nrhs_max * 2 + 64));
    const size_t r_bytes = dworkd ? 0 : a256_bytes(sizeof(double) * (size_t)n * nrhs_max);
    const size_t small_bytes = a256_bytes(sizeof(int) * ((size_t)n + 64));
    const size_t dz_bytes = a256_bytes(sizeof(double) * (size_t)n);
    MGI_HIP_CHECK(big.alloc(lu_bytes + ys_bytes + dw_bytes + sw_bytes + part_bytes + nrm_bytes + r_bytes +
                            small_bytes + dz_bytes));
    char* p = static_cast<char*>(big.p);
    DenseIrCtx c{};
    c.trans = trans; c.n = n; c.nrhs = nrhs;
    c.dA = dA; c.ldda = ldda; c.dB = dB; c.lddb = lddb; c.dX = dX; c.lddx = lddx;
    if (dworks) {
        c.S = dworks; c.lds = n;
        c.Ys = dworks + (size_t)n * n;
    } else {
        c.S = reinterpret_cast<float*>(p); c.lds = n;         p += lu_bytes;
        c.Ys = reinterpret_cast<float*>(p);                   p += ys_bytes;
    }
    DenseWork dw;
    dense_work_carve(p, n, n, 4, gmax, dw);                      p += dw_bytes;
    solve_work_carve(p, n, std::min(nrhs_max, 64), 4, c.sw);  p += sw_bytes;
    c.part = reinterpret_cast<double*>(p);                    p += part_bytes;
    c.nrm = reinterpret_cast<double*>(p);                     p += nrm_bytes;
    if (dworkd) { c.dR = dworkd; c.ldr = n; }
    else { c.dR = reinterpret_cast<double*>(p); c.ldr = n;    p += r_bytes; }
    int* d_small = reinterpret_cast<int*>(p);                 p += small_bytes;  // [0] ovf, [1] info, [2] bad, [64..] ipiv
    int* d_lpipiv = d_small + 64;
    double* dz = reinterpret_cast<double*>(p);
    c.st = st;
    c.dtrs = (solver & kRefineDTRS) != 0;

    dn_zero_i_kernel<<<1, 64, 0, st>>>(d_small, 64);
    // 1. down-cast with overflow check, ||A||_inf
    dn_lag2s_kernel<<<dim3(gx(n), nchunk), kRT, 0, st>>>(n, dA, ldda, c.S, c.lds, c.part, d_small);
    MGI_LAUNCH_CHECK();
    dn_rowmax_kernel<<<gx(n), kRT, 0, st>>>(n, nchunk, c.part, c.nrm);
    MGI_LAUNCH_CHECK();
    std::vector<double> bm(gx(n));
    int h_small[4] = {0, 0, 0, 0};
    MGI_HIP_CHECK(hipMemcpyAsync(bm.data(), c.nrm, sizeof(double) * bm.size(), hipMemcpyDeviceToHost, st));
    MGI_HIP_CHECK(hipMemcpyAsync(h_small, d_small, sizeof(h_small), hipMemcpyDeviceToHost, st));
    MGI_HIP_CHECK(hipStreamSynchronize(st));
    double anrm = 0.0;
    for (double v : bm) anrm = std::max(anrm, v);
    const double eps = DBL_EPSILON / 2;   // dlamch('Epsilon')
    c.cte = anrm * eps * std::sqrt((double)n);
    int it = 0;
    bool fallback = false;
    DevBuf dcopy, dsw;
    if (h_small[0] != 0) {
        it = -2;
        fallback = true;
    } else {
        // 2. low-precision LU
        StreamHolder side;
        side.s = dense_side_stream(n, 4);
        magma_int_t rc = dense_sgetrf(mode, n, n, c.S, c.lds, d_lpipiv, d_small + 1, dw, st, 0, side.s);
        if (rc != MAGMA_SUCCESS) return rc;
        if ((rc = dense_check_err(dw, st)) != MAGMA_SUCCESS) return rc;
        dn_finite_kernel<<<1024, kRT, 0, st>>>((size_t)n * n, n, c.S, c.lds, d_small + 2);
        MGI_LAUNCH_CHECK();
        MGI_HIP_CHECK(hipMemcpyAsync(h_small, d_small, sizeof(h_small), hipMemcpyDeviceToHost, st));
        MGI_HIP_CHECK(hipMemcpyAsync(ipiv, d_lpipiv, sizeof(int) * n, hipMemcpyDeviceToHost, st));
        MGI_HIP_CHECK(hipStreamSynchronize(st));
        if (h_small[1] != 0 || h_small[2] != 0) {
            it = -3;
            fallback = true;
        } else {
            std::vector<int> rowmap;
            swp2pswp(trans, n, ipiv, rowmap);
            if (dipiv) {
                MGI_HIP_CHECK(hipMemcpyAsync(dipiv, rowmap.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
                MGI_HIP_CHECK(hipStreamSynchronize(st));
            }
            if (c.dtrs) {
                // FP64 copy of the factors (slag2d) with its own solve state
                const int nr = std::min(nrhs_max, 64);
                MGI_HIP_CHECK(dcopy.alloc(sizeof(double) * ((size_t)n * n + (size_t)n * nrhs_max)));
                MGI_HIP_CHECK(dsw.alloc(solve_work_bytes(n, nr, 8)));
                c.Dl = static_cast<double*>(dcopy.p);
                c.Yd = c.Dl + (size_t)n * n;
                solve_work_carve(dsw.p, n, nr, 8, c.swd);
                dn_lag2d_kernel<<<dim3(gx(n), (unsigned)n), kRT, 0, st>>>(n, c.S, c.lds, c.Dl);
                MGI_LAUNCH_CHECK();
                if ((rc = dense_set_perm(n, ipiv, c.swd, st)) != MAGMA_SUCCESS) return rc;
                if ((rc = dense_diag_inverses<double>(n, c.Dl, n, c.swd, st)) != MAGMA_SUCCESS) return rc;
            } else {
                if ((rc = dense_set_perm(n, ipiv, c.sw, st)) != MAGMA_SUCCESS) return rc;
                if ((rc = dense_diag_inverses<float>(n, c.S, c.lds, c.sw, st)) != MAGMA_SUCCESS) return rc;
            }
            // 3. refinement; ITERMAX = min(n, 500), an outer level over an
            // inner FGMRES runs max(10, ITERMAX / 50) steps, and the outer
            // FGMRES restarts after as many iterations as it may take
            // (src/dxgesv_gmres_gpu.cpp:211-219, 328-359, 574-575)
            const int itermax_all = std::min(n, 500);
            const int itermax = nested ? std::max(10, itermax_all / 50) : itermax_all;
            GmOwner gin, gout;
            if (nested) MGI_HIP_CHECK(gin.alloc(n, std::min(n, kGmresInner)));
            if (gm_outer) MGI_HIP_CHECK(gout.alloc(n, itermax));
            int steps = 0;
            bool conv = false;
            rc = gm_outer ? ir_fgmres(c, itermax, gout.g, nested ? &gin.g : nullptr, &steps, &conv)
                          : ir_classical(c, itermax, nested ? &gin.g : nullptr, dz, &steps, &conv);
            if (rc != MAGMA_SUCCESS) return rc;
            if ((rc = dense_solve_check_err(c.dtrs ? c.swd : c.sw, st)) != MAGMA_SUCCESS) return rc;
            if (conv) it = steps;
            else { it = -std::max(1, steps); fallback = true; }
        }
    }
    if (fallback) {
        // 4. FP64 LU of A in place and X = A^{-1} B (src/dxgesv_gmres_gpu.cpp:801-810)
        magma_int_t rc = getrf_one<double>(kUpdF64, n, n, dA, ldda, ipiv, info, st);
        if (rc != MAGMA_SUCCESS) return rc;
        if (*info == 0) {
            dn_copy_d_kernel<<<dim3(gx(n), (unsigned)nrhs), kRT, 0, st>>>(n, nrhs, dB, lddb, dX, lddx);
            MGI_LAUNCH_CHECK();
            if ((rc = getrs_one<double>(trans, n, nrhs, dA, ldda, ipiv, dX, lddx, st)) != MAGMA_SUCCESS) return rc;
        }
        if (dipiv) {
            std::vector<int> rowmap;
            swp2pswp(trans, n, ipiv, rowmap);
            MGI_HIP_CHECK(hipMemcpyAsync(dipiv, rowmap.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
        }
    }
    MGI_HIP_CHECK(hipStreamSynchronize(st));
    *iter = it;
    return MAGMA_SUCCESS;
}

}  // namespace mgi

using namespace mgi;

extern "C" {

// include/magma_z.h:612 (src/zgetrf_gpu.cpp:542): LU of dA in place, host ipiv.
magma_int_t magma_dgetrf_gpu(magma_int_t m, magma_int_t n, magmaDouble_ptr dA, magma_int_t ldda, magma_int_t* ipiv,
                             magma_int_t* info)
{
    return getrf_entry<double>(__func__, kUpdF64, m, n, dA, ldda, ipiv, info);
}

// include/magma_z.h:638 (src/zgetrf_gpu.cpp:607): the all-GPU variant; here
// the same code path as magma_dgetrf_gpu.
magma_int_t magma_dgetrf_native(magma_int_t m, magma_int_t n, magmaDouble_ptr dA, magma_int_t ldda,
                                magma_int_t* ipiv, magma_int_t* info)
{
    return getrf_entry<double>(__func__, kUpdF64, m, n, dA, ldda, ipiv, info);
}

// include/magma_s.h (s form of magma_z.h:612)
magma_int_t magma_sgetrf_gpu(magma_int_t m, magma_int_t n, magmaFloat_ptr dA, magma_int_t ldda, magma_int_t* ipiv,
                             magma_int_t* info)
{
    return getrf_entry<float>(__func__, kUpdF32, m, n, dA, ldda, ipiv, info);
}

// include/magma_z.h:713 (src/zgetrs_gpu.cpp:292)
magma_int_t magma_dgetrs_gpu(magma_trans_t trans, magma_int_t n, magma_int_t nrhs, magmaDouble_ptr dA,
                             magma_int_t ldda, magma_int_t* ipiv, magmaDouble_ptr dB, magma_int_t lddb,
                             magma_int_t* info)
{
    return getrs_entry<double>(__func__, trans, n, nrhs, dA, ldda, ipiv, dB, lddb, info);
}

magma_int_t magma_sgetrs_gpu(magma_trans_t trans, magma_int_t n, magma_int_t nrhs, magmaFloat_ptr dA,
                             magma_int_t ldda, magma_int_t* ipiv, magmaFloat_ptr dB, magma_int_t lddb,
                             magma_int_t* info)
{
    return getrs_entry<float>(__func__, trans, n, nrhs, dA, ldda, ipiv, dB, lddb, info);
}

// include/magma_htc.h:119 (src/xshgetrf_gpu.cpp:86): FP32 LU whose trailing
// updates take FP16 inputs with FP32 accumulation on the matrix cores
// (enable_tc = Magma_MP_ENABLE_TC_MATH) or FP32 inputs otherwise
// (Magma_MP_ENABLE_DFLT_MATH, or mp_algo_type Magma_MP_SGEMM /
// Magma_MP_GEMEX_I32_O32_C32). The 16-bit output forms of cublasGemmEx
// (I16_O16_*) are served by FP16 inputs with FP32 accumulation.
magma_int_t magma_xshgetrf_gpu(magma_int_t m, magma_int_t n, magmaFloat_ptr dA, magma_int_t ldda, magma_int_t* ipiv,
                               magma_int_t* info, magma_mp_type_t enable_tc, magma_mp_type_t mp_algo_type)
{
    const bool fp32 = enable_tc != Magma_MP_ENABLE_TC_MATH || mp_algo_type == Magma_MP_SGEMM ||
                      mp_algo_type == Magma_MP_TC_SGEMM || mp_algo_type == Magma_MP_GEMEX_I32_O32_C32 ||
                      mp_algo_type == Magma_MP_TC_GEMEX_I32_O32_C32;
    return getrf_entry<float>(__func__, fp32 ? kUpdF32 : kUpdFP16, m, n, dA, ldda, ipiv, info);
}

// No MAGMA counterpart: the FP32 LU with the trailing-update format named by
// a refinement factorisation type (Magma_PREC_SS FP32, _SHT FP16, _SBT BF16,
// _SB3T split BF16, _SF8T FP8 e4m3), as used inside magma_dxgesv_gmres_gpu.
magma_int_t magma_sgetrf_prec_gpu(magma_int_t m, magma_int_t n, magmaFloat_ptr dA, magma_int_t ldda,
                                  magma_int_t* ipiv, magma_int_t* info, magma_refinement_t facto_type)
{
    int mode = -1;
    switch (facto_type) {
        case Magma_PREC_SS: case Magma_PREC_XSH_S: mode = kUpdF32; break;
        case Magma_PREC_SHT: case Magma_PREC_XSH_163232TC: mode = kUpdFP16; break;
        case Magma_PREC_SBT: mode = kUpdBF16; break;
        case Magma_PREC_SB3T: mode = kUpdBF16x3; break;
        case Magma_PREC_SF8T: mode = kUpdFP8; break;
        default: break;
    }
    if (mode < 0) {
        *info = -7;
        magma_xerbla(__func__, 7);
        return *info;
    }
    return getrf_entry<float>(__func__, mode, m, n, dA, ldda, ipiv, info);
}

}  // extern "C"
