// Generic batched LU for any m-by-n matrix whose padded copy fits in one
// workgroup's LDS, FP64 or FP32. It follows LAPACK dgetf2 step by step
// (idamax pivot with first-index ties, whole-row interchange, reciprocal
// scaling when |pivot| >= sfmin, rank-1 update with fused multiply-add), so
// its factors are bitwise those of the same recurrence on the host. It serves
// rectangular and odd shapes; the square fast path lives in getrf_tile.hip.
//
// Reference behaviour: magmablas/zgetf2_devicefunc.cuh:218-294 (panel step),
// magmablas/magma_templates.h:44-60 (ties), zgetf2_devicefunc.cuh:254,268-278
// (zero pivot: info set, no swap, factorisation continues).
#include "magma_internal.h"
#include "lu_device.h"
#include "getrf_kernels.h"

namespace mgi {

constexpr int kLdsThreads = 256;

__device__ __forceinline__ void lds_wave_argmax(double& v, int& i) { wave_argmax(v, i); }
__device__ __forceinline__ void lds_wave_argmax(float& v, int& i) { wave_argmax_f(v, i); }
__device__ __forceinline__ double sfmin_of(double) { return kSfminD; }
__device__ __forceinline__ float sfmin_of(float) { return kSfminF; }

template <typename T>
__global__ void __launch_bounds__(kLdsThreads)
getrf_lds_kernel(int m, int n, T** __restrict__ dA_array, int ldda, int** __restrict__ ipiv_array,
                 int* __restrict__ info_array, int gbstep)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* sA = reinterpret_cast<T*>(lds_raw);
    __shared__ int s_piv;
    __shared__ T s_pv;
    const int tid = threadIdx.x;
    const int ld = m + 1;                  // odd stride: row walks spread over banks
    auto A = gptr(dA_array[blockIdx.x]);
    auto ipiv = gptr(ipiv_array[blockIdx.x]);

    for (int idx = tid; idx < m * n; idx += kLdsThreads) {
        int i = idx % m, j = idx / m;
        sA[i + j * ld] = A[i + (size_t)j * ldda];
    }
    __syncthreads();

    int info = 0;
    const int minmn = m < n ? m : n;
    for (int k = 0; k < minmn; ++k) {
        if (tid < kWave) {
            T best = T(-1);
            int bi = k;
            for (int i = k + tid; i < m; i += kWave) {
                T v = fabs(sA[i + k * ld]);
                if (i == k && v != v) v = T(INFINITY);     // idamax keeps a NaN first entry
                if (v > best) { best = v; bi = i; }
            }
            lds_wave_argmax(best, bi);
            // the pivot value is published with its index: the row swap below
            // may move it before every thread has read it from sA
            if (tid == 0) { s_piv = bi; s_pv = sA[bi + k * ld]; }
        }
        __syncthreads();
        const int p = s_piv;
        const T pv = s_pv;
        if (tid == 0) ipiv[k] = p + 1 + gbstep;
        if (pv != T(0)) {
            if (p != k) {
                for (int j = tid; j < n; j += kLdsThreads) {
                    T t = sA[k + j * ld];
                    sA[k + j * ld] = sA[p + j * ld];
                    sA[p + j * ld] = t;
                }
            }
            __syncthreads();
            if (fabs(pv) >= sfmin_of(pv)) {
                const T r = T(1) / pv;
                for (int i = k + 1 + tid; i < m; i += kLdsThreads) sA[i + k * ld] *= r;
            } else {
                for (int i = k + 1 + tid; i < m; i += kLdsThreads) sA[i + k * ld] /= pv;
            }
        } else if (info == 0) {
            info = k + 1 + gbstep;
        }
        __syncthreads();
        const int rows = m - k - 1, cols = n - k - 1;
        for (int idx = tid; idx < rows * cols; idx += kLdsThreads) {
            int i = k + 1 + idx % rows, j = k + 1 + idx / rows;
            sA[i + j * ld] = fma(-sA[i + k * ld], sA[k + j * ld], sA[i + j * ld]);
        }
        __syncthreads();
    }

    for (int idx = tid; idx < m * n; idx += kLdsThreads) {
        int i = idx % m, j = idx / m;
        A[i + (size_t)j * ldda] = sA[i + j * ld];
    }
    if (tid == 0) info_array[blockIdx.x] = info;
}

size_t getrf_lds_bytes(int m, int n, int elem) { return size_t(m + 1) * size_t(n) * size_t(elem); }
size_t getrf_lds_bytes(int m, int n) { return getrf_lds_bytes(m, n, sizeof(double)); }

template <typename T>
static magma_int_t lds_launch(int m, int n, T** dA_array, int ldda, int** ipiv_array, int* info_array, int batch,
                              hipStream_t stream)
{
    const size_t lds = getrf_lds_bytes(m, n, sizeof(T));
    if (lds > max_dynamic_lds()) return MAGMA_ERR_NOT_SUPPORTED;
    MGI_HIP_CHECK(hipFuncSetAttribute((const void*)getrf_lds_kernel<T>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    getrf_lds_kernel<T><<<batch, kLdsThreads, lds, stream>>>(m, n, dA_array, ldda, ipiv_array, info_array, 0);
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

magma_int_t getrf_lds_launch(int m, int n, double** dA_array, int ldda, int** ipiv_array,
                             int* info_array, int batch, hipStream_t stream)
{
    return lds_launch<double>(m, n, dA_array, ldda, ipiv_array, info_array, batch, stream);
}

magma_int_t sgetrf_lds_launch(int m, int n, float** dA_array, int ldda, int** ipiv_array,
                              int* info_array, int batch, hipStream_t stream)
{
    return lds_launch<float>(m, n, dA_array, ldda, ipiv_array, info_array, batch, stream);
}

size_t max_dynamic_lds()
{
    static size_t cached = 0;
    if (cached == 0) {
        int dev = 0, v = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 65536;
        // Leave room for the kernels' static LDS.
        cached = size_t(v) - 1024;
    }
    return cached;
}

}  // namespace mgi
