// Batched LU without pivoting, random butterfly transforms and the batched
// inverse, on the kernels of the pivoted path:
//   magma_dgetrf_nopiv_batched (src/zgetrf_nopiv_batched.cpp): the blocked
//     LU with the diagonal entry as every pivot (a null pivot array), info =
//     first exactly-zero pivot;
//   magma_dgetrs_nopiv_batched / magma_dgesv_nopiv_batched: the solves with
//     no interchanges;
//   magma_dgetrf_nopiv_vbatched (src/zgetrf_nopiv_vbatched.cpp): per-matrix
//     sizes;
//   magma_dgesv_rbt_batched (src/zgesv_rbt_batched.cpp, src/zgerbt_batched.cpp):
//     A := U^T A V and B := U^T B with two-level random butterflies
//     (magmablas/zgerbt_kernels.cu), the LU without pivoting, then X := V Y;
//   magma_dgetri_outofplace_batched (src/zgetri_outofplace_batched.cpp):
//     inv(A) from the LU and pivots, as the solve of A X = I.
#include "magma_internal.h"
#include "getrf_kernels.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

namespace mgi {

namespace {

// One butterfly level on the ar x ac block at (ai, aj) of every matrix
// (magmablas/zgerbt_kernels.cu, elementary multiplication): with the block
// cut into quadrants at ar1 = ceil(ar/2), ac1 = ceil(ac/2),
//   A00 := u1 v1 (a00 + a01 + a10 + a11)   A01 := u1 v2 (a00 - a01 + a10 - a11)
//   A10 := u2 v1 (a00 + a01 - a10 - a11)   A11 := u2 v2 (a00 - a01 - a10 + a11)
// i.e. D_u H A H^T D_v with H = [I I; I -I]; entries beyond the block count 0.
__global__ void rbt_matrix_kernel(int ar, int ac, double** __restrict__ dA_array, int ai, int aj, int ldda,
                                  const double* __restrict__ u, const double* __restrict__ v)
{
    const int ar1 = (ar + 1) / 2, ar2 = ar - ar1;
    const int ac1 = (ac + 1) / 2, ac2 = ac - ac1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= ar1 || j >= ac1) return;
    double* A = dA_array[blockIdx.z] + ai + (size_t)aj * ldda;
    const bool r2 = i < ar2, c2 = j < ac2;
    const double a00 = A[i + (size_t)j * ldda];
    const double a01 = c2 ? A[i + (size_t)(j + ac1) * ldda] : 0.0;
    const double a10 = r2 ? A[i + ar1 + (size_t)j * ldda] : 0.0;
    const double a11 = (r2 && c2) ? A[i + ar1 + (size_t)(j + ac1) * ldda] : 0.0;
    const double b1 = a00 + a01, b2 = a10 + a11, b3 = a00 - a01, b4 = a10 - a11;
    const double u1 = u[i], v1 = v[j];
    A[i + (size_t)j * ldda] = u1 * v1 * (b1 + b2);
    if (c2) A[i + (size_t)(j + ac1) * ldda] = u1 * v[ac1 + j] * (b3 + b4);
    if (r2) A[i + ar1 + (size_t)j * ldda] = u[ar1 + i] * v1 * (b1 - b2);
    if (r2 && c2) A[i + ar1 + (size_t)(j + ac1) * ldda] = u[ar1 + i] * v[ac1 + j] * (b3 - b4);
}

// b := D_u H b on the len rows at offset off (U^T b for one level), or
// b := H^T D_u b (V b for one level), for every column of every B.
__global__ void rbt_vector_kernel(int len, int nrhs, double** __restrict__ dB_array, int off, int lddb,
                                  const double* __restrict__ u, int transpose)
{
    const int n1 = (len + 1) / 2, n2 = len - n1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n1 || c >= nrhs) return;
    double* b = dB_array[blockIdx.z] + off + (size_t)c * lddb;
    const bool h2 = i < n2;
    if (transpose) {
        const double x0 = b[i], x1 = h2 ? b[n1 + i] : 0.0;
        b[i] = u[i] * (x0 + x1);
        if (h2) b[n1 + i] = u[n1 + i] * (x0 - x1);
    } else {
        const double y0 = u[i] * b[i], y1 = h2 ? u[n1 + i] * b[n1 + i] : 0.0;
        b[i] = y0 + y1;
        if (h2) b[n1 + i] = y0 - y1;
    }
}

__global__ void identity_kernel(int n, double** __restrict__ dA_array, int ldda)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n && j < n) dA_array[blockIdx.z][i + (size_t)j * ldda] = (i == j) ? 1.0 : 0.0;
}

__global__ void nopiv_fill_kernel(int* p, int v, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

// Butterfly entries: exp((r - 1/2) / 10), r uniform in [0, 1]
// (src/zgerbt_gpu.cpp, init_butterfly), from a seeded generator.
void make_butterfly(int n2, std::vector<double>& u, std::vector<double>& v, unsigned seed)
{
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    u.resize(n2);
    v.resize(n2);
    for (int i = 0; i < n2; ++i) {
        u[i] = std::exp((uni(gen) - 0.5) / 10.0);
        v[i] = std::exp((uni(gen) - 0.5) / 10.0);
    }
}

magma_int_t nopiv_getrf_async(int m, int n, double** dA_array, int ldda, const int* m_arr, const int* n_arr,
                              const int* ld_arr, int max_m, int max_n, int* info_array, int batch, hipStream_t st)
{
    if (!getrf_blocked_supported(max_m)) return MAGMA_ERR_NOT_SUPPORTED;
    return getrf_blocked_launch(m, n, ldda, m_arr, n_arr, ld_arr, max_m, max_n, dA_array, nullptr, info_array,
                                batch, st);
}

// split a batch over grid z
template <typename F>
magma_int_t over_z(int batch, F&& f)
{
    for (int c0 = 0; c0 < batch; c0 += kMaxGridYZ) {
        magma_int_t rc = f(c0, std::min(kMaxGridYZ, batch - c0));
        if (rc != MAGMA_SUCCESS) return rc;
    }
    return MAGMA_SUCCESS;
}

}  // namespace

}  // namespace mgi

using namespace mgi;

extern "C" {

// include/magma_zbatched.h:977
magma_int_t magma_dgetrf_nopiv_batched(magma_int_t m, magma_int_t n, double** dA_array, magma_int_t ldda,
                                       magma_int_t* info_array, magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t arginfo = 0;
    if (m < 0) arginfo = -1;
    else if (n < 0) arginfo = -2;
    else if (ldda < std::max(1, m)) arginfo = -4;
    else if (batchCount < 0) arginfo = -6;
    if (arginfo != 0) {
        magma_xerbla(__func__, -arginfo);
        return arginfo;
    }
    if (m == 0 || n == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    MGI_HIP_CHECK(hipSetDevice(queue->device));
    magma_int_t rc = nopiv_getrf_async(m, n, dA_array, ldda, nullptr, nullptr, nullptr, m, n, info_array, batchCount,
                                       queue->stream);
    if (rc != MAGMA_SUCCESS) return rc;
    MGI_HIP_CHECK(hipStreamSynchronize(queue->stream));
    return 0;
}

// include/magma_zbatched.h:985 (magma_zgetrs_nopiv_batched)
magma_int_t magma_dgetrs_nopiv_batched(magma_trans_t trans, magma_int_t n, magma_int_t nrhs, double** dA_array,
                                       magma_int_t ldda, double** dB_array, magma_int_t lddb,
                                       magma_int_t* info_array, magma_int_t batchCount, magma_queue_t queue)
{
    (void)info_array;
    magma_int_t info = 0;
    if (trans != MagmaNoTrans && trans != MagmaTrans && trans != MagmaConjTrans) info = -1;
    else if (n < 0) info = -2;
    else if (nrhs < 0) info = -3;
    else if (ldda < std::max(1, n)) info = -5;
    else if (lddb < std::max(1, n)) info = -7;
    else if (batchCount < 0) info = -9;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    MGI_HIP_CHECK(hipSetDevice(queue->device));
    if (getrs_generic_bytes(n) > max_dynamic_lds()) return MAGMA_ERR_NOT_SUPPORTED;
    return getrs_generic_launch(trans == MagmaNoTrans ? 0 : 1, n, nrhs, dA_array, ldda, nullptr, dB_array, lddb,
                                batchCount, queue->stream);
}

// include/magma_zbatched.h:993 (magma_zgesv_nopiv_batched)
magma_int_t magma_dgesv_nopiv_batched(magma_int_t n, magma_int_t nrhs, double** dA_array, magma_int_t ldda,
                                      double** dB_array, magma_int_t lddb, magma_int_t* info_array,
                                      magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (n < 0) info = -1;
    else if (nrhs < 0) info = -2;
    else if (ldda < std::max(1, n)) info = -4;
    else if (lddb < std::max(1, n)) info = -6;
    else if (batchCount < 0) info = -8;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    MGI_HIP_CHECK(hipSetDevice(queue->device));
    magma_int_t rc = nopiv_getrf_async(n, n, dA_array, ldda, nullptr, nullptr, nullptr, n, n, info_array, batchCount,
                                       queue->stream);
    if (rc != MAGMA_SUCCESS) return rc;
    if (getrs_generic_bytes(n) > max_dynamic_lds()) return MAGMA_ERR_NOT_SUPPORTED;
    return getrs_generic_launch(0, n, nrhs, dA_array, ldda, nullptr, dB_array, lddb, batchCount, queue->stream);
}

// include/magma_zvbatched.h:218: m, n, ldda are device arrays
magma_int_t magma_dgetrf_nopiv_vbatched(magma_int_t* m, magma_int_t* n, double** dA_array, magma_int_t* ldda,
                                        magma_int_t* info_array, magma_int_t batchCount, magma_queue_t queue)
{
    if (batchCount < 0) {
        magma_xerbla(__func__, 6);
        return -6;
    }
    if (batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    MGI_HIP_CHECK(hipSetDevice(queue->device));
    std::vector<int> hm(batchCount), hn(batchCount), hl(batchCount);
    MGI_HIP_CHECK(hipMemcpyAsync(hm.data(), m, sizeof(int) * batchCount, hipMemcpyDeviceToHost, queue->stream));
    MGI_HIP_CHECK(hipMemcpyAsync(hn.data(), n, sizeof(int) * batchCount, hipMemcpyDeviceToHost, queue->stream));
    MGI_HIP_CHECK(hipMemcpyAsync(hl.data(), ldda, sizeof(int) * batchCount, hipMemcpyDeviceToHost, queue->stream));
    MGI_HIP_CHECK(hipStreamSynchronize(queue->stream));
    int max_m = 0, max_n = 0;
    magma_int_t arginfo = 0;
    for (int i = 0; i < batchCount; ++i) {
        if (hm[i] < 0) arginfo = -1;
        else if (hn[i] < 0) arginfo = -2;
        else if (hl[i] < std::max(1, hm[i])) arginfo = -4;
        if (arginfo) break;
        max_m = std::max(max_m, hm[i]);
        max_n = std::max(max_n, hn[i]);
    }
    if (arginfo != 0) {
        magma_xerbla(__func__, -arginfo);
        return arginfo;
    }
    if (max_m == 0 || max_n == 0) {
        nopiv_fill_kernel<<<(unsigned)ceildiv(batchCount, 256), 256, 0, queue->stream>>>(info_array, 0, batchCount);
        MGI_LAUNCH_CHECK();
        MGI_HIP_CHECK(hipStreamSynchronize(queue->stream));
        return 0;
    }
    magma_int_t rc = nopiv_getrf_async(max_m, max_n, dA_array, max_m, m, n, ldda, max_m, max_n, info_array,
                                       batchCount, queue->stream);
    if (rc != MAGMA_SUCCESS) return rc;
    MGI_HIP_CHECK(hipStreamSynchronize(queue->stream));
    return 0;
}

// include/magma_zbatched.h:1001: two-level random butterflies, LU without
// pivoting, solve, then X := V Y. The butterflies come from a fixed seed, so
// a call is reproducible.
magma_int_t magma_dgesv_rbt_batched(magma_int_t n, magma_int_t nrhs, double** dA_array, magma_int_t ldda,
                                    double** dB_array, magma_int_t lddb, magma_int_t* dinfo_array,
                                    magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (n < 0) info = -1;
    else if (nrhs < 0) info = -2;
    else if (ldda < std::max(1, n)) info = -4;
    else if (lddb < std::max(1, n)) info = -6;
    else if (batchCount < 0) info = -8;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    MGI_HIP_CHECK(hipSetDevice(queue->device));
    hipStream_t st = queue->stream;
    std::vector<double> hu, hv;
    make_butterfly(2 * n, hu, hv, 20240613u + (unsigned)n);
    double* d = static_cast<double*>(queue_scratch(queue, sizeof(double) * 4 * (size_t)n));
    if (d == nullptr) return MAGMA_ERR_DEVICE_ALLOC;
    double* du = d;
    double* dv = d + 2 * (size_t)n;
    MGI_HIP_CHECK(hipMemcpyAsync(du, hu.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, st));
    MGI_HIP_CHECK(hipMemcpyAsync(dv, hv.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, st));
    const int n1 = (n + 1) / 2, n2 = n - n1;
    // A := U^T A V: the four quadrants (second level, entries n.. of u and v), then the whole
    // matrix (first level, entries 0..n-1)
    magma_int_t rc = over_z(batchCount, [&](int c0, int cb) -> magma_int_t {
        double** pa = dA_array + c0;
        auto level = [&](int ar, int ac, int ai, int aj, const double* u, const double* v) {
            if (ar <= 0 || ac <= 0) return;
            dim3 g((unsigned)ceildiv((ar + 1) / 2, 128), (unsigned)((ac + 1) / 2), (unsigned)cb);
            rbt_matrix_kernel<<<g, 128, 0, st>>>(ar, ac, pa, ai, aj, ldda, u, v);
        };
        level(n1, n1, 0, 0, du + n, dv + n);
        level(n1, n2, 0, n1, du + n, dv + n + n1);
        level(n2, n1, n1, 0, du + n + n1, dv + n);
        level(n2, n2, n1, n1, du + n + n1, dv + n + n1);
        level(n, n, 0, 0, du, dv);
        MGI_LAUNCH_CHECK();
        // B := U^T B: halves, then the whole
        double** pb = dB_array + c0;
        auto vec = [&](int len, int off, const double* u, int transpose) {
            if (len <= 0) return;
            dim3 g((unsigned)ceildiv((len + 1) / 2, 128), (unsigned)nrhs, (unsigned)cb);
            rbt_vector_kernel<<<g, 128, 0, st>>>(len, nrhs, pb, off, lddb, u, transpose);
        };
        vec(n1, 0, du + n, 1);
        vec(n2, n1, du + n + n1, 1);
        vec(n, 0, du, 1);
        MGI_LAUNCH_CHECK();
        return MAGMA_SUCCESS;
    });
    if (rc != MAGMA_SUCCESS) return rc;
    rc = nopiv_getrf_async(n, n, dA_array, ldda, nullptr, nullptr, nullptr, n, n, dinfo_array, batchCount, st);
    if (rc != MAGMA_SUCCESS) return rc;
    if (getrs_generic_bytes(n) > max_dynamic_lds()) return MAGMA_ERR_NOT_SUPPORTED;
    rc = getrs_generic_launch(0, n, nrhs, dA_array, ldda, nullptr, dB_array, lddb, batchCount, st);
    if (rc != MAGMA_SUCCESS) return rc;
    // X := V Y: the whole, then the halves
    rc = over_z(batchCount, [&](int c0, int cb) -> magma_int_t {
        double** pb = dB_array + c0;
        auto vec = [&](int len, int off, const double* v) {
            if (len <= 0) return;
            dim3 g((unsigned)ceildiv((len + 1) / 2, 128), (unsigned)nrhs, (unsigned)cb);
            rbt_vector_kernel<<<g, 128, 0, st>>>(len, nrhs, pb, off, lddb, v, 0);
        };
        vec(n, 0, dv);
        vec(n1, 0, dv + n);
        vec(n2, n1, dv + n + n1);
        MGI_LAUNCH_CHECK();
        return MAGMA_SUCCESS;
    });
    if (rc != MAGMA_SUCCESS) return rc;
    // the butterflies (pageable host vectors) are read by the queued copies
    MGI_HIP_CHECK(hipStreamSynchronize(st));
    return 0;
}

// include/magma_zbatched.h:872: inv(A) from the factors of magma_dgetrf_batched,
// out of place: dinvA := I, then dinvA := A^{-1} dinvA (= U^{-1} L^{-1} P).
// Host-synchronous, like the reference.
magma_int_t magma_dgetri_outofplace_batched(magma_int_t n, double** dA_array, magma_int_t ldda,
                                            magma_int_t** dipiv_array, double** dinvA_array, magma_int_t lddia,
                                            magma_int_t* info_array, magma_int_t batchCount, magma_queue_t queue)
{
    (void)info_array;
    magma_int_t info = 0;
    if (n < 0) info = -1;
    else if (ldda < std::max(1, n)) info = -3;
    else if (lddia < std::max(1, n)) info = -6;
    else if (batchCount < 0) info = -8;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    MGI_HIP_CHECK(hipSetDevice(queue->device));
    hipStream_t st = queue->stream;
    magma_int_t rc = over_z(batchCount, [&](int c0, int cb) -> magma_int_t {
        dim3 g((unsigned)ceildiv(n, 128), (unsigned)n, (unsigned)cb);
        identity_kernel<<<g, 128, 0, st>>>(n, dinvA_array + c0, lddia);
        MGI_LAUNCH_CHECK();
        return MAGMA_SUCCESS;
    });
    if (rc != MAGMA_SUCCESS) return rc;
    if (getrs_generic_bytes(n) > max_dynamic_lds()) return MAGMA_ERR_NOT_SUPPORTED;
    rc = getrs_generic_launch(0, n, n, dA_array, ldda, dipiv_array, dinvA_array, lddia, batchCount, st);
    if (rc != MAGMA_SUCCESS) return rc;
    MGI_HIP_CHECK(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
