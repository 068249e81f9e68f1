// magma_dgetrf_native_noalloc: the dense FP64 LU of magma_dgetrf_native for
// callers that cannot allocate (code under graph capture, serving loops,
// harnesses that forbid hipMalloc). Its scratch -- the panel's candidate
// slots, one block's row pairs and one diagonal block's inverse -- lives in
// the file-scope device array `scratch` below; the caller passes the matrix,
// the pivot and info arrays (device memory) and a queue.
#include "magma_internal.h"
#include "dense_kernels.h"

#include <mutex>

namespace mgi {

namespace {

// Panel workgroups the scratch holds: m <= 65,536 rows.
constexpr int kNaMaxRows = 65536;
constexpr int kNaGmax = kNaMaxRows / kDensePanelRows;
constexpr size_t kNaBytes = 1024 * 1024;

// 1 MB: the panel's slot words (528 KB), one block inverse (128 KB), pairs.
// Every device holds its own copy until the process ends. Calls that overlap
// on one device share it and race.
__device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaMaxDevices = 64;
void* g_scratch[kNaMaxDevices];
hipError_t g_scratch_err[kNaMaxDevices];
std::once_flag g_scratch_once[kNaMaxDevices];

// Device address of `scratch` on device dev (current), looked up once per
// device. The array's host-side name is not a device address.
void* scratch_on(int dev)
{
    if (dev < 0 || dev >= kNaMaxDevices) return nullptr;
    std::call_once(g_scratch_once[dev], [dev]() {
        void* p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch[dev] = g_scratch_err[dev] == hipSuccess ? p : nullptr;
    });
    if (g_scratch_err[dev] != hipSuccess) {
        hip_report(g_scratch_err[dev], "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        return nullptr;
    }
    return g_scratch[dev];
}

__global__ void na_err_to_info_kernel(const int* err, int* info)
{
    if (*err != 0) *info = MAGMA_ERR_UNKNOWN;
}

}  // namespace

}  // namespace mgi

using namespace mgi;

extern "C" {

magma_int_t magma_dgetrf_native_noalloc(magma_int_t m, magma_int_t n, magmaDouble_ptr dA, magma_int_t ldda,
                                        magmaInt_ptr d_ipiv, magmaInt_ptr d_info, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (m < 0) info = -1;
    else if (n < 0) info = -2;
    else if (ldda < std::max(1, m)) info = -4;
    else if (queue == nullptr) info = -7;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (m == 0 || n == 0) return MAGMA_SUCCESS;
    const int g = dense_panel_gmax(m);
    if (g > kNaGmax || g > device_cu_count()) {
        std::fprintf(stderr, "magma_dgetrf_native_noalloc: m = %d rows exceed its scratch limit of %d rows "
                     "(%d per compute unit, at most %d)\n", (int)m,
                     kDensePanelRows * std::min(kNaGmax, device_cu_count()), kDensePanelRows, kNaMaxRows);
        return MAGMA_ERR_NOT_SUPPORTED;
    }
    if (dense_work_bytes(m, n, (int)sizeof(double), kNaGmax) > kNaBytes) return MAGMA_ERR_NOT_SUPPORTED;
    MGI_HIP_CHECK(hipSetDevice(queue->device));
    void* base = scratch_on(queue->device);
    if (base == nullptr) return MAGMA_ERR_UNKNOWN;
    DenseWork w;
    dense_work_carve(base, m, n, (int)sizeof(double), kNaGmax, w);
    magma_int_t rc = dense_getrf(kUpdF64, m, n, dA, ldda, d_ipiv, d_info, w, queue->stream);
    if (rc != MAGMA_SUCCESS) return rc;
    na_err_to_info_kernel<<<1, 1, 0, queue->stream>>>(w.err, d_info);
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

}  // extern "C"
