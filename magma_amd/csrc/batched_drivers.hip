// Public batched LU drivers: argument checking, quick returns and kernel
// selection. Argument codes and return conventions follow
// src/zgetrf_batched.cpp:82-105, src/zgetrs_batched.cpp:86-117 and
// src/zgesv_batched.cpp:93-154.
#include "magma_internal.h"
#include "getrf_kernels.h"

#include <algorithm>

namespace mgi {

// Asynchronous factorisation; the public entry point adds the host sync.
magma_int_t getrf_batched_async(int m, int n, double** dA_array, int ldda, int** ipiv_array,
                                int* info_array, int batch, hipStream_t stream)
{
    if (m == n && getrf_tile_supported(n))
        return getrf_tile_launch(n, dA_array, ldda, ipiv_array, info_array, batch, stream);
    if (getrf_lds_bytes(m, n) <= max_dynamic_lds())
        return getrf_lds_launch(m, n, dA_array, ldda, ipiv_array, info_array, batch, stream);
    if (getrf_blocked_supported(m))
        return getrf_blocked_launch(m, n, ldda, nullptr, nullptr, nullptr, m, n, dA_array, ipiv_array,
                                    info_array, batch, stream);
    return MAGMA_ERR_NOT_SUPPORTED;
}

magma_int_t getrs_batched_async(int trans, int n, int nrhs, double** dA_array, int ldda, int** ipiv_array,
                                double** dB_array, int lddb, int batch, hipStream_t stream)
{
    if (getrs_generic_bytes(n) <= max_dynamic_lds())
        return getrs_generic_launch(trans, n, nrhs, dA_array, ldda, ipiv_array, dB_array, lddb, batch, stream);
    return MAGMA_ERR_NOT_SUPPORTED;
}

}  // namespace mgi

extern "C" {

magma_int_t magma_dgetrf_batched(magma_int_t m, magma_int_t n, double** dA_array, magma_int_t ldda,
                                 magma_int_t** ipiv_array, magma_int_t* info_array,
                                 magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t arginfo = 0;
    if (m < 0)
        arginfo = -1;
    else if (n < 0)
        arginfo = -2;
    else if (ldda < std::max(1, m))
        arginfo = -4;
    else if (batchCount < 0)
        arginfo = -7;
    if (arginfo != 0) {
        magma_xerbla(__func__, -arginfo);
        return arginfo;
    }
    if (m == 0 || n == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    hipSetDevice(queue->device);
    magma_int_t rc = mgi::getrf_batched_async(m, n, dA_array, ldda, ipiv_array, info_array,
                                              batchCount, queue->stream);
    if (rc != MAGMA_SUCCESS) return rc;
    MGI_HIP_CHECK(hipStreamSynchronize(queue->stream));
    return 0;
}

magma_int_t magma_dgetrs_batched(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
                                 double** dA_array, magma_int_t ldda, magma_int_t** dipiv_array,
                                 double** dB_array, magma_int_t lddb,
                                 magma_int_t batchCount, magma_queue_t queue)
{
    const bool notran = (trans == MagmaNoTrans);
    magma_int_t info = 0;
    if (!notran && trans != MagmaTrans && trans != MagmaConjTrans)
        info = -1;
    else if (n < 0)
        info = -2;
    else if (nrhs < 0)
        info = -3;
    else if (ldda < std::max(1, n))
        info = -5;
    else if (lddb < std::max(1, n))
        info = -8;
    else if (batchCount < 0)
        info = -9;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    hipSetDevice(queue->device);
    return mgi::getrs_batched_async(notran ? 0 : 1, n, nrhs, dA_array, ldda, dipiv_array,
                                    dB_array, lddb, batchCount, queue->stream);
}

magma_int_t magma_dgesv_batched(magma_int_t n, magma_int_t nrhs, double** dA_array, magma_int_t ldda,
                                magma_int_t** dipiv_array, double** dB_array, magma_int_t lddb,
                                magma_int_t* dinfo_array, magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (n < 0)
        info = -1;
    else if (nrhs < 0)
        info = -2;
    else if (ldda < std::max(1, n))
        info = -4;
    else if (lddb < std::max(1, n))
        info = -6;
    else if (batchCount < 0)
        info = -9;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    hipSetDevice(queue->device);
    if (mgi::gesv_tile_supported(n, nrhs))
        return mgi::gesv_tile_launch(n, nrhs, dA_array, ldda, dipiv_array, dinfo_array, dB_array, lddb,
                                     batchCount, queue->stream);
    info = mgi::getrf_batched_async(n, n, dA_array, ldda, dipiv_array, dinfo_array, batchCount, queue->stream);
    if (info != MAGMA_SUCCESS) return info;
    return mgi::getrs_batched_async(0, n, nrhs, dA_array, ldda, dipiv_array, dB_array, lddb,
                                    batchCount, queue->stream);
}

void magma_get_dgetrf_batched_nbparam(magma_int_t n, magma_int_t* nb, magma_int_t* recnb)
{
    // The blocking the square n x n batched LU uses: the register-tile kernel
    // factors 16-column panels; the LDS kernel factors the whole matrix one
    // column at a time; the blocked LU factors nb-wide blocks of 32-wide
    // panels.
    int b = 16, r = 16;
    if (!mgi::getrf_tile_supported(n)) {
        if (mgi::getrf_lds_bytes(n, n) <= mgi::max_dynamic_lds()) {
            b = std::max(1, n);
            r = 1;
        } else {
            b = std::max(mgi::getrf_blocked_nb(n), 1);
            r = std::min(b, 32);
        }
    }
    if (nb) *nb = b;
    if (recnb) *recnb = r;
}

void magma_get_dgetrf_vbatched_nbparam(magma_int_t max_m, magma_int_t max_n, magma_int_t* nb, magma_int_t* recnb)
{
    // The vbatched LU puts systems up to 128 on the register tile (16-column
    // panels) and larger ones on the blocked LU sized by the largest row count.
    int b = 16, r = 16;
    if (max_m > 128 || max_n > 128) {
        b = std::max(mgi::getrf_blocked_nb(max_m), 1);
        r = std::min(b, 32);
    }
    if (nb) *nb = b;
    if (recnb) *recnb = r;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Single precision (the "s" forms of the same drivers).
// ---------------------------------------------------------------------------
namespace mgi {

magma_int_t sgetrf_batched_async(int m, int n, float** dA_array, int ldda, int** ipiv_array, int* info_array,
                                 int batch, hipStream_t stream)
{
    if (getrf_lds_bytes(m, n, sizeof(float)) <= max_dynamic_lds())
        return sgetrf_lds_launch(m, n, dA_array, ldda, ipiv_array, info_array, batch, stream);
    if (getrf_blocked_supported(m))
        return sgetrf_blocked_launch(kUpdF32, m, n, ldda, nullptr, nullptr, nullptr, m, n, dA_array, ipiv_array,
                                     info_array, batch, stream);
    return MAGMA_ERR_NOT_SUPPORTED;
}

magma_int_t sgetrs_batched_async(int trans, int n, int nrhs, float** dA_array, int ldda, int** ipiv_array,
                                 float** dB_array, int lddb, int batch, hipStream_t stream)
{
    RsSizes rs{n, nrhs, ldda, lddb, n, nullptr, nullptr, nullptr, nullptr, nullptr};
    return getrs_generic_launch_t<float>(trans, rs, nrhs, dA_array, ipiv_array, dB_array, batch, stream);
}

}  // namespace mgi

extern "C" {

magma_int_t magma_sgetrf_batched(magma_int_t m, magma_int_t n, float** dA_array, magma_int_t ldda,
                                 magma_int_t** ipiv_array, magma_int_t* info_array,
                                 magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t arginfo = 0;
    if (m < 0)
        arginfo = -1;
    else if (n < 0)
        arginfo = -2;
    else if (ldda < std::max(1, m))
        arginfo = -4;
    else if (batchCount < 0)
        arginfo = -7;
    if (arginfo != 0) {
        magma_xerbla(__func__, -arginfo);
        return arginfo;
    }
    if (m == 0 || n == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    hipSetDevice(queue->device);
    magma_int_t rc = mgi::sgetrf_batched_async(m, n, dA_array, ldda, ipiv_array, info_array, batchCount,
                                               queue->stream);
    if (rc != MAGMA_SUCCESS) return rc;
    MGI_HIP_CHECK(hipStreamSynchronize(queue->stream));
    return 0;
}

magma_int_t magma_sgetrs_batched(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
                                 float** dA_array, magma_int_t ldda, magma_int_t** dipiv_array,
                                 float** dB_array, magma_int_t lddb,
                                 magma_int_t batchCount, magma_queue_t queue)
{
    const bool notran = (trans == MagmaNoTrans);
    magma_int_t info = 0;
    if (!notran && trans != MagmaTrans && trans != MagmaConjTrans)
        info = -1;
    else if (n < 0)
        info = -2;
    else if (nrhs < 0)
        info = -3;
    else if (ldda < std::max(1, n))
        info = -5;
    else if (lddb < std::max(1, n))
        info = -8;
    else if (batchCount < 0)
        info = -9;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    hipSetDevice(queue->device);
    return mgi::sgetrs_batched_async(notran ? 0 : 1, n, nrhs, dA_array, ldda, dipiv_array, dB_array, lddb,
                                     batchCount, queue->stream);
}

magma_int_t magma_sgesv_batched(magma_int_t n, magma_int_t nrhs, float** dA_array, magma_int_t ldda,
                                magma_int_t** dipiv_array, float** dB_array, magma_int_t lddb,
                                magma_int_t* dinfo_array, magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (n < 0)
        info = -1;
    else if (nrhs < 0)
        info = -2;
    else if (ldda < std::max(1, n))
        info = -4;
    else if (lddb < std::max(1, n))
        info = -6;
    else if (batchCount < 0)
        info = -9;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    hipSetDevice(queue->device);
    info = mgi::sgetrf_batched_async(n, n, dA_array, ldda, dipiv_array, dinfo_array, batchCount, queue->stream);
    if (info != MAGMA_SUCCESS) return info;
    return mgi::sgetrs_batched_async(0, n, nrhs, dA_array, ldda, dipiv_array, dB_array, lddb, batchCount,
                                     queue->stream);
}

}  // extern "C"
