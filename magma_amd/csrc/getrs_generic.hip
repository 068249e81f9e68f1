// Generic batched triangular solve with LU factors (LAPACK dgetrs semantics),
// FP64 or FP32, fixed or per-matrix sizes, for any n whose right-hand-side
// block fits in LDS.
//
// NoTrans: B := U \ (L \ (P^T B))          (src/zgetrs_batched.cpp:119-145)
// Trans:   B := P (L^T \ (U^T \ B))        (src/zgetrs_gpu.cpp:203-227; the
//          batched reference's transposed branch uses the wrong triangles and
//          pivot order, so the non-batched LAPACK order is followed here).
//
// Each step is column-oriented (an axpy of the solved entry into the rows not
// yet solved, with fused multiply-add), the order reference BLAS dtrsm uses
// for the NoTrans case.
#include "magma_internal.h"
#include "lu_device.h"
#include "getrf_kernels.h"

namespace mgi {

constexpr int kRsThreads = 256;
constexpr int kRsCols = 8;     // right-hand sides per workgroup

template <typename T>
__global__ void __launch_bounds__(kRsThreads)
getrs_generic_kernel(int trans, RsSizes sz, T** __restrict__ dA_array, int** __restrict__ ipiv_array,
                     T** __restrict__ dB_array)
{
    extern __shared__ __align__(16) unsigned char rs_raw[];
    const int b = sz.map ? sz.map[blockIdx.x] : blockIdx.x;
    int n, nrhs, ldda, lddb;
    sz.get(b, n, nrhs, ldda, lddb);
    const int c0 = blockIdx.y * kRsCols;
    if (n <= 0 || c0 >= nrhs) return;
    T* sB = reinterpret_cast<T*>(rs_raw);                       // n x kRsCols, ld = n
    int* perm = reinterpret_cast<int*>(rs_raw + (size_t)sz.max_n * kRsCols * sizeof(T));
    const int tid = threadIdx.x;
    const int nr = min(kRsCols, nrhs - c0);
    const T* A = dA_array[b];
    const int* ipiv = ipiv_array[b];
    T* B = dB_array[b] + (size_t)c0 * lddb;

    for (int i = tid; i < n; i += kRsThreads) perm[i] = i;
    __syncthreads();
    if (tid == 0) {
        // Compose the LAPACK interchange sequence into a row map:
        // (P^T B)(i,:) = B(perm[i],:). An interchange outside [k, n) cannot
        // come from a valid LU; it is ignored rather than allowed to index
        // outside the map.
        for (int k = 0; k < n; ++k) {
            int p = ipiv[k] - 1;
            if (p < k || p >= n) continue;
            int t = perm[k]; perm[k] = perm[p]; perm[p] = t;
        }
    }
    __syncthreads();

    if (trans == 0) {
        for (int idx = tid; idx < n * nr; idx += kRsThreads) {
            int i = idx % n, c = idx / n;
            sB[i + c * n] = B[perm[i] + (size_t)c * lddb];
        }
        __syncthreads();
        for (int k = 0; k < n - 1; ++k) {                  // unit lower, forward
            const int rows = n - k - 1;
            for (int idx = tid; idx < rows * nr; idx += kRsThreads) {
                int i = k + 1 + idx % rows, c = idx / rows;
                sB[i + c * n] = fma(-A[i + (size_t)k * ldda], sB[k + c * n], sB[i + c * n]);
            }
            __syncthreads();
        }
        for (int k = n - 1; k >= 0; --k) {                 // non-unit upper, backward
            if (tid < nr) sB[k + tid * n] = sB[k + tid * n] / A[k + (size_t)k * ldda];
            __syncthreads();
            for (int idx = tid; idx < k * nr; idx += kRsThreads) {
                int i = idx % k, c = idx / k;
                sB[i + c * n] = fma(-A[i + (size_t)k * ldda], sB[k + c * n], sB[i + c * n]);
            }
            __syncthreads();
        }
        for (int idx = tid; idx < n * nr; idx += kRsThreads) {
            int i = idx % n, c = idx / n;
            B[i + (size_t)c * lddb] = sB[i + c * n];
        }
    } else {
        for (int idx = tid; idx < n * nr; idx += kRsThreads) {
            int i = idx % n, c = idx / n;
            sB[i + c * n] = B[i + (size_t)c * lddb];
        }
        __syncthreads();
        for (int k = 0; k < n; ++k) {                      // U^T: non-unit lower, forward
            if (tid < nr) sB[k + tid * n] = sB[k + tid * n] / A[k + (size_t)k * ldda];
            __syncthreads();
            const int rows = n - k - 1;
            for (int idx = tid; idx < rows * nr; idx += kRsThreads) {
                int i = k + 1 + idx % rows, c = idx / rows;
                sB[i + c * n] = fma(-A[k + (size_t)i * ldda], sB[k + c * n], sB[i + c * n]);
            }
            __syncthreads();
        }
        for (int k = n - 1; k > 0; --k) {                  // L^T: unit upper, backward
            for (int idx = tid; idx < k * nr; idx += kRsThreads) {
                int i = idx % k, c = idx / k;
                sB[i + c * n] = fma(-A[k + (size_t)i * ldda], sB[k + c * n], sB[i + c * n]);
            }
            __syncthreads();
        }
        for (int idx = tid; idx < n * nr; idx += kRsThreads) {
            int i = idx % n, c = idx / n;
            B[perm[i] + (size_t)c * lddb] = sB[i + c * n];
        }
    }
}

size_t getrs_generic_bytes(int n, int elem) { return (size_t)n * kRsCols * elem + (size_t)n * sizeof(int); }
size_t getrs_generic_bytes(int n) { return getrs_generic_bytes(n, sizeof(double)); }

template <typename T>
magma_int_t getrs_generic_launch_t(int trans, const RsSizes& sz, int max_nrhs, T** dA_array, int** ipiv_array,
                                   T** dB_array, int batch, hipStream_t stream)
{
    const size_t lds = getrs_generic_bytes(sz.max_n, sizeof(T));
    if (lds > max_dynamic_lds()) return MAGMA_ERR_NOT_SUPPORTED;
    MGI_HIP_CHECK(hipFuncSetAttribute((const void*)getrs_generic_kernel<T>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    dim3 grid(batch, ceildiv(max_nrhs, kRsCols));
    getrs_generic_kernel<T><<<grid, kRsThreads, lds, stream>>>(trans, sz, dA_array, ipiv_array, dB_array);
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

template magma_int_t getrs_generic_launch_t<double>(int, const RsSizes&, int, double**, int**, double**, int,
                                                    hipStream_t);
template magma_int_t getrs_generic_launch_t<float>(int, const RsSizes&, int, float**, int**, float**, int,
                                                   hipStream_t);

magma_int_t getrs_generic_launch(int trans, int n, int nrhs, double** dA_array, int ldda, int** ipiv_array,
                                 double** dB_array, int lddb, int batch, hipStream_t stream)
{
    RsSizes sz{n, nrhs, ldda, lddb, n, nullptr, nullptr, nullptr, nullptr, nullptr};
    return getrs_generic_launch_t<double>(trans, sz, nrhs, dA_array, ipiv_array, dB_array, batch, stream);
}

}  // namespace mgi
