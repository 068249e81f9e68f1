// Launchers of the batched LU kernels (internal to libmagma_mi355x.so).
// Every launcher enqueues on `stream` and returns MAGMA_SUCCESS,
// MAGMA_ERR_NOT_SUPPORTED when the shape is outside what the kernel handles,
// or a launch error code.
#pragma once

#include "magma_internal.h"

namespace mgi {

// Input format of the trailing-update products in the blocked LU.
enum UpdateMode {
    kUpdF64 = 0,     // FP64 storage, FP64 MFMA
    kUpdF32 = 1,     // FP32 storage, FP32 MFMA
    kUpdBF16 = 2,    // FP32 storage, BF16 inputs, FP32 accumulation
    kUpdFP16 = 3,    // FP32 storage, FP16 inputs, FP32 accumulation
    kUpdBF16x3 = 4,  // FP32 storage, hi/lo BF16 split (three products)
    kUpdFP8 = 5,     // FP32 storage, FP8 (e4m3) inputs, FP32 accumulation (dense LU only)
};

// Matrix sizes of a batch: fixed (arrays null) or per matrix (device arrays).
struct SizeArgs {
    int m, n, ld;
    const int* m_arr;
    const int* n_arr;
    const int* ld_arr;
    const int* map;          // optional: workgroup index -> matrix index
    int base = 0;            // first workgroup index of this launch (a batch split over launches)
    __device__ __forceinline__ int mat(int bi) const { return map ? map[base + bi] : base + bi; }
    // each per-matrix array is optional on its own; a null one means the
    // uniform value
    __device__ __forceinline__ void get(int b, int& mm, int& nn, int& ll) const
    {
        mm = m_arr ? m_arr[b] : m;
        nn = n_arr ? n_arr[b] : n;
        ll = ld_arr ? ld_arr[b] : ld;
    }
};

size_t max_dynamic_lds();

// getrf_lds.hip: any m x n with (m+1)*n doubles <= LDS.
size_t getrf_lds_bytes(int m, int n);
size_t getrf_lds_bytes(int m, int n, int elem);
magma_int_t sgetrf_lds_launch(int m, int n, float** dA_array, int ldda, int** ipiv_array,
                              int* info_array, int batch, hipStream_t stream);
magma_int_t getrf_lds_launch(int m, int n, double** dA_array, int ldda, int** ipiv_array,
                             int* info_array, int batch, hipStream_t stream);

// Sizes for the solve kernels: fixed (arrays null) or per matrix; `map`
// (optional) sends workgroup b to matrix map[b] of the pointer arrays.
struct RsSizes {
    int n, nrhs, ldda, lddb;
    int max_n;
    const int* n_arr;
    const int* nrhs_arr;
    const int* ldda_arr;
    const int* lddb_arr;
    const int* map;
    __device__ __forceinline__ void get(int b, int& nn, int& rr, int& la, int& lb) const
    {
        nn = n_arr ? n_arr[b] : n;
        rr = nrhs_arr ? nrhs_arr[b] : nrhs;
        la = ldda_arr ? ldda_arr[b] : ldda;
        lb = lddb_arr ? lddb_arr[b] : lddb;
    }
};

// getrs_generic.hip: trans 0 = NoTrans, 1 = Trans.
size_t getrs_generic_bytes(int n, int elem);
template <typename T>
magma_int_t getrs_generic_launch_t(int trans, const RsSizes& sz, int max_nrhs, T** dA_array, int** ipiv_array,
                                   T** dB_array, int batch, hipStream_t stream);
size_t getrs_generic_bytes(int n);
magma_int_t getrs_generic_launch(int trans, int n, int nrhs, double** dA_array, int ldda, int** ipiv_array,
                                 double** dB_array, int lddb, int batch, hipStream_t stream);

// getrf_blocked.hip: blocked panel/laswp/trsm/gemm steps, fixed size (m_arr
// == nullptr) or per-matrix sizes in device arrays; max_m <= 2048.
bool getrf_blocked_supported(int max_m);
int getrf_blocked_nb(int max_m);   // panel width the blocked LU uses (0: unsupported)
magma_int_t getrf_blocked_launch(int m, int n, int ldda, const int* m_arr, const int* n_arr, const int* ld_arr,
                                 int max_m, int max_n, double** dA_array, int** ipiv_array, int* info_array,
                                 int batch, hipStream_t stream);
// FP32 storage; mode is one of kUpdF32, kUpdBF16, kUpdFP16, kUpdBF16x3.
magma_int_t sgetrf_blocked_launch(int mode, int m, int n, int ldda, const int* m_arr, const int* n_arr,
                                  const int* ld_arr, int max_m, int max_n, float** dA_array, int** ipiv_array,
                                  int* info_array, int batch, hipStream_t stream);

// Per-matrix arguments of the tile kernels (fixed size when arrays are null;
// `map` sends workgroup b to matrix map[b]).
struct TileArgs {
    int n, nrhs, ldda, lddb;
    const int* map;
    const int* n_arr;
    const int* nrhs_arr;
    const int* ldda_arr;
    const int* lddb_arr;
};

// getrf_tile.hip: square n <= 128, register-tiled, MFMA FP64 trailing update.
// The *_args forms take per-matrix sizes; nt = ceil(max n / 16) selects the
// kernel instance.
magma_int_t getrf_tile_launch_args(int nt, const TileArgs& a, double** dA_array, int** ipiv_array, int* info_array,
                                   int batch, hipStream_t stream);
magma_int_t gesv_tile_launch_args(int nt, const TileArgs& a, double** dA_array, int** ipiv_array, int* info_array,
                                  double** dB_array, int batch, hipStream_t stream);
// getrf_blocked.hip, general form (any SizeArgs incl. map)
magma_int_t getrf_blocked_launch_sz(const SizeArgs& sz, int max_m, int max_n, double** dA_array, int** ipiv_array,
                                    int* info_array, int batch, hipStream_t stream, int force_nb = 0);
bool blocked_nb_valid(int max_m, int nb);
bool getrf_tile_supported(int n);
magma_int_t getrf_tile_launch(int n, double** dA_array, int ldda, int** ipiv_array, int* info_array,
                              int batch, hipStream_t stream);
// Fused LU + solve (A X = B) for n <= 128, nrhs <= 16.
bool gesv_tile_supported(int n, int nrhs);
magma_int_t gesv_tile_launch(int n, int nrhs, double** dA_array, int ldda, int** ipiv_array, int* info_array,
                             double** dB_array, int lddb, int batch, hipStream_t stream);

}  // namespace mgi
