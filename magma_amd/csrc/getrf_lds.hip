// Generic batched LU for any m-by-n matrix whose padded copy fits in one
// workgroup's LDS. It follows LAPACK dgetf2 step by step (idamax pivot with
// first-index ties, whole-row interchange, reciprocal scaling when
// |pivot| >= sfmin, rank-1 update with fused multiply-add), so its factors are
// bitwise those of the same recurrence on the host. It serves rectangular and
// odd shapes; the square fast paths live in getrf_tile.hip.
//
// Reference behaviour: magmablas/zgetf2_devicefunc.cuh:218-294 (panel step),
// magmablas/magma_templates.h:44-60 (ties), zgetf2_devicefunc.cuh:254,268-278
// (zero pivot: info set, no swap, factorisation continues).
#include "magma_internal.h"
#include "lu_device.h"
#include "getrf_kernels.h"

namespace mgi {

constexpr int kLdsThreads = 256;

__global__ void __launch_bounds__(kLdsThreads)
getrf_lds_kernel(int m, int n, double** __restrict__ dA_array, int ldda,
                 int** __restrict__ ipiv_array, int* __restrict__ info_array, int gbstep)
{
    extern __shared__ double sA[];
    __shared__ int s_piv;
    const int tid = threadIdx.x;
    const int ld = m + 1;                  // odd stride: row walks spread over banks
    double* A = dA_array[blockIdx.x];
    int* ipiv = ipiv_array[blockIdx.x];

    for (int idx = tid; idx < m * n; idx += kLdsThreads) {
        int i = idx % m, j = idx / m;
        sA[i + j * ld] = A[i + (size_t)j * ldda];
    }
    __syncthreads();

    int info = 0;
    const int minmn = m < n ? m : n;
    for (int k = 0; k < minmn; ++k) {
        if (tid < kWave) {
            double best = -1.0;
            int bi = k;
            for (int i = k + tid; i < m; i += kWave) {
                double v = fabs(sA[i + k * ld]);
                if (v > best) { best = v; bi = i; }
            }
            wave_argmax(best, bi);
            if (tid == 0) s_piv = bi;
        }
        __syncthreads();
        const int p = s_piv;
        const double pv = sA[p + k * ld];
        if (tid == 0) ipiv[k] = p + 1 + gbstep;
        if (pv != 0.0) {
            if (p != k) {
                for (int j = tid; j < n; j += kLdsThreads) {
                    double t = sA[k + j * ld];
                    sA[k + j * ld] = sA[p + j * ld];
                    sA[p + j * ld] = t;
                }
            }
            __syncthreads();
            if (fabs(pv) >= kSfminD) {
                const double r = 1.0 / pv;
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

size_t getrf_lds_bytes(int m, int n) { return size_t(m + 1) * size_t(n) * sizeof(double); }

magma_int_t getrf_lds_launch(int m, int n, double** dA_array, int ldda, int** ipiv_array,
                             int* info_array, int batch, hipStream_t stream)
{
    const size_t lds = getrf_lds_bytes(m, n);
    if (lds > max_dynamic_lds()) return MAGMA_ERR_NOT_SUPPORTED;
    MGI_HIP_CHECK(hipFuncSetAttribute((const void*)getrf_lds_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    getrf_lds_kernel<<<batch, kLdsThreads, lds, stream>>>(m, n, dA_array, ldda, ipiv_array, info_array, 0);
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

size_t max_dynamic_lds()
{
    static size_t cached = 0;
    if (cached == 0) {
        int dev = 0, v = 0;
        hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 65536;
        // Leave room for the kernels' static LDS.
        cached = size_t(v) - 1024;
    }
    return cached;
}

}  // namespace mgi
