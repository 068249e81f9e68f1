// Launchers of the batched LU kernels (internal to libmagma_mi355x.so).
// Every launcher enqueues on `stream` and returns MAGMA_SUCCESS,
// MAGMA_ERR_NOT_SUPPORTED when the shape is outside what the kernel handles,
// or a launch error code.
#pragma once

#include "magma_internal.h"

namespace mgi {

size_t max_dynamic_lds();

// getrf_lds.hip: any m x n with (m+1)*n doubles <= LDS.
size_t getrf_lds_bytes(int m, int n);
magma_int_t getrf_lds_launch(int m, int n, double** dA_array, int ldda, int** ipiv_array,
                             int* info_array, int batch, hipStream_t stream);

// getrs_generic.hip: trans 0 = NoTrans, 1 = Trans.
size_t getrs_generic_bytes(int n);
magma_int_t getrs_generic_launch(int trans, int n, int nrhs, double** dA_array, int ldda, int** ipiv_array,
                                 double** dB_array, int lddb, int batch, hipStream_t stream);

// getrf_blocked.hip: blocked panel/laswp/trsm/gemm steps, fixed size (m_arr
// == nullptr) or per-matrix sizes in device arrays; max_m <= 2048.
bool getrf_blocked_supported(int max_m);
magma_int_t getrf_blocked_launch(int m, int n, int ldda, const int* m_arr, const int* n_arr, const int* ld_arr,
                                 int max_m, int max_n, double** dA_array, int** ipiv_array, int* info_array,
                                 int batch, hipStream_t stream);

// getrf_tile.hip: square n <= 128, register-tiled, MFMA FP64 trailing update.
bool getrf_tile_supported(int n);
magma_int_t getrf_tile_launch(int n, double** dA_array, int ldda, int** ipiv_array, int* info_array,
                              int batch, hipStream_t stream);
// Fused LU + solve (A X = B) for n <= 128, nrhs <= 16.
bool gesv_tile_supported(int n, int nrhs);
magma_int_t gesv_tile_launch(int n, int nrhs, double** dA_array, int ldda, int** ipiv_array, int* info_array,
                             double** dB_array, int lddb, int batch, hipStream_t stream);

}  // namespace mgi
