// Adjoint of the batched solve X = A^{-1} B (the backward pass of the
// differentiable solve; no MAGMA counterpart): grad_A = -grad_B X^T as one
// rank-nrhs outer product per system, read straight from the column-major
// n x nrhs blocks the solves produce and written as the caller's row-major
// n x n gradient. One workgroup covers a 16-row by 256-column tile of one
// system: thread t owns column j = tile + t, so the X loads and the grad_A
// stores are contiguous across the wave, and the 16 x nrhs slice of grad_B
// it needs is staged in LDS once per workgroup.
#include "magma_internal.h"

#include <algorithm>

namespace mgi {

namespace {

constexpr int kOutRows = 16;
constexpr int kOutCols = 256;
constexpr int kOutMaxR = 32;        // right-hand sides staged per pass

template <typename T>
__global__ void __launch_bounds__(kOutCols)
adjoint_outer_kernel(int n, int nrhs, const T* __restrict__ gB, int ldb, long long sb, const T* __restrict__ X,
                     int ldx, long long sx, T* __restrict__ gA, int lda, long long sa)
{
    __shared__ T g[kOutMaxR][kOutRows];
    const int s = blockIdx.z;
    const int i0 = blockIdx.y * kOutRows;
    const int j = blockIdx.x * kOutCols + threadIdx.x;
    const T* gBs = gB + (size_t)s * sb;
    const T* Xs = X + (size_t)s * sx;
    T acc[kOutRows];
#pragma unroll
    for (int r = 0; r < kOutRows; ++r) acc[r] = T(0);
    for (int q0 = 0; q0 < nrhs; q0 += kOutMaxR) {
        const int nq = min(kOutMaxR, nrhs - q0);
        __syncthreads();
        for (int e = threadIdx.x; e < kOutMaxR * kOutRows; e += kOutCols) {
            const int q = e / kOutRows, r = e % kOutRows;
            g[q][r] = (q < nq && i0 + r < n) ? gBs[(size_t)(q0 + q) * ldb + i0 + r] : T(0);
        }
        __syncthreads();
        if (j < n) {
            for (int q = 0; q < nq; ++q) {
                const T xj = Xs[(size_t)(q0 + q) * ldx + j];
#pragma unroll
                for (int r = 0; r < kOutRows; ++r) acc[r] = fma(g[q][r], xj, acc[r]);
            }
        }
    }
    if (j < n) {
        T* out = gA + (size_t)s * sa;
#pragma unroll
        for (int r = 0; r < kOutRows; ++r)
            if (i0 + r < n) out[(size_t)(i0 + r) * lda + j] = -acc[r];
    }
}

template <typename T>
magma_int_t outer_launch(int n, int nrhs, const void* gB, int ldb, long long sb, const void* X, int ldx, long long sx,
                         void* gA, int lda, long long sa, int batch, hipStream_t st)
{
    // the batch goes in grid z: at most 65,535 systems per launch
    for (int b0 = 0; b0 < batch; b0 += kMaxGridYZ) {
        const int nb = std::min(batch - b0, kMaxGridYZ);
        dim3 grid((unsigned)ceildiv(n, kOutCols), (unsigned)ceildiv(n, kOutRows), (unsigned)nb);
        adjoint_outer_kernel<T><<<grid, kOutCols, 0, st>>>(
            n, nrhs, static_cast<const T*>(gB) + (size_t)b0 * sb, ldb, sb, static_cast<const T*>(X) + (size_t)b0 * sx,
            ldx, sx, static_cast<T*>(gA) + (size_t)b0 * sa, lda, sa);
        MGI_LAUNCH_CHECK();
    }
    return MAGMA_SUCCESS;
}

}  // namespace
}  // namespace mgi

using namespace mgi;

extern "C" {

// grad_A[s] = -grad_B[s] X[s]^T for contiguous batches: grad_B and X are
// column-major n x nrhs blocks (leading dimensions ldb, ldx, strides sb, sx
// elements between systems), grad_A row-major n x n (row stride lda, system
// stride sa). elem 8 = FP64, 4 = FP32. Asynchronous on the queue.
magma_int_t magma_adjoint_outer_batched(magma_int_t elem, magma_int_t n, magma_int_t nrhs, const void* dgB,
                                        magma_int_t ldb, long long strideB, const void* dX, magma_int_t ldx,
                                        long long strideX, void* dgA, magma_int_t lda, long long strideA,
                                        magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (elem != 8 && elem != 4) info = -1;
    else if (n < 0) info = -2;
    else if (nrhs < 0) info = -3;
    else if (ldb < std::max(1, n)) info = -5;
    else if (ldx < std::max(1, n)) info = -8;
    else if (lda < std::max(1, n)) info = -11;
    else if (batchCount < 0) info = -13;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    hipSetDevice(queue->device);
    if (elem == 8)
        return outer_launch<double>(n, nrhs, dgB, ldb, strideB, dX, ldx, strideX, dgA, lda, strideA, batchCount,
                                    queue->stream);
    return outer_launch<float>(n, nrhs, dgB, ldb, strideB, dX, ldx, strideX, dgA, lda, strideA, batchCount,
                               queue->stream);
}

}  // extern "C"
