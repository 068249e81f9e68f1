// Runtime boundary of libmagma_mi355x.so: queues, device memory, copies,
// error reporting and pointer-array helpers. Semantics follow MAGMA's
// interface_cuda/ (interface.cpp:931-999 queues, alloc.cpp, copy_v2.cpp,
// error.cpp) and control/xerbla.cpp, re-implemented over the HIP runtime.
#include "magma_internal.h"

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace mgi {

void hip_report(hipError_t err, const char* what, const char* file, int line)
{
    std::fprintf(stderr, "magma_mi355x: HIP error %d (%s) in %s at %s:%d\n",
                 (int)err, hipGetErrorString(err), what, file, line);
}

void* queue_scratch(magma_queue_t queue, size_t bytes)
{
    if (queue == nullptr) return nullptr;
    if (bytes <= queue->scratch_bytes) return queue->scratch;
    if (queue->scratch != nullptr) {
        // The old arena may still be read by queued kernels.
        hipStreamSynchronize(queue->stream);
        hipFree(queue->scratch);
        queue->scratch = nullptr;
        queue->scratch_bytes = 0;
    }
    size_t want = bytes < (size_t(1) << 20) ? (size_t(1) << 20) : bytes;
    want = (want + 255) & ~size_t(255);
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) return nullptr;
    queue->scratch = p;
    queue->scratch_bytes = want;
    return p;
}

}  // namespace mgi

namespace {
std::once_flag g_init_once;
bool g_initialized = false;
}

extern "C" {

// interface_cuda/interface.cpp:147 -- one-time process setup.
magma_int_t magma_init(void)
{
    magma_int_t rc = MAGMA_SUCCESS;
    std::call_once(g_init_once, [&]() {
        int ndev = 0;
        hipError_t e = hipGetDeviceCount(&ndev);
        if (e != hipSuccess || ndev <= 0) {
            rc = MAGMA_ERR_UNKNOWN;
            return;
        }
        g_initialized = true;
    });
    if (!g_initialized) return MAGMA_ERR_UNKNOWN;
    return rc;
}

magma_int_t magma_finalize(void) { return MAGMA_SUCCESS; }

real_Double_t magma_wtime(void)
{
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

real_Double_t magma_sync_wtime(magma_queue_t queue)
{
    if (queue) hipStreamSynchronize(queue->stream);
    else hipDeviceSynchronize();
    return magma_wtime();
}

magma_int_t magma_malloc(magma_ptr* ptr_ptr, size_t bytes)
{
    if (ptr_ptr == nullptr) return MAGMA_ERR_INVALID_PTR;
    if (bytes == 0) bytes = 1;   // MAGMA allocates a minimal buffer for 0 bytes
    if (hipMalloc(ptr_ptr, bytes) != hipSuccess) {
        *ptr_ptr = nullptr;
        return MAGMA_ERR_DEVICE_ALLOC;
    }
    return MAGMA_SUCCESS;
}

magma_int_t magma_free_internal(magma_ptr ptr, const char* func, const char* file, int line)
{
    if (ptr == nullptr) return MAGMA_SUCCESS;
    hipError_t e = hipFree(ptr);
    if (e != hipSuccess) {
        std::fprintf(stderr, "magma_free: invalid pointer %p in %s (%s:%d)\n", ptr, func, file, line);
        return MAGMA_ERR_INVALID_PTR;
    }
    return MAGMA_SUCCESS;
}

magma_int_t magma_malloc_cpu(void** ptr_ptr, size_t bytes)
{
    if (ptr_ptr == nullptr) return MAGMA_ERR_INVALID_PTR;
    if (bytes == 0) bytes = 1;
    if (posix_memalign(ptr_ptr, 64, bytes) != 0) {
        *ptr_ptr = nullptr;
        return MAGMA_ERR_HOST_ALLOC;
    }
    return MAGMA_SUCCESS;
}

magma_int_t magma_free_cpu(void* ptr)
{
    std::free(ptr);
    return MAGMA_SUCCESS;
}

void magma_getdevice(magma_device_t* dev)
{
    int d = 0;
    hipGetDevice(&d);
    if (dev) *dev = d;
}

void magma_setdevice(magma_device_t dev) { hipSetDevice(dev); }

void magma_queue_create_internal(magma_device_t device, magma_queue_t* queue_ptr,
                                 const char* func, const char* file, int line)
{
    if (queue_ptr == nullptr) return;
    *queue_ptr = nullptr;
    auto* q = new (std::nothrow) magma_queue();
    if (q == nullptr) return;
    q->device = device;
    q->own_stream = true;
    q->scratch = nullptr;
    q->scratch_bytes = 0;
    hipSetDevice(device);
    // Default priority: streams of mixed priority are avoided on purpose.
    if (hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) != hipSuccess) {
        std::fprintf(stderr, "magma_queue_create: stream creation failed in %s (%s:%d)\n", func, file, line);
        delete q;
        return;
    }
    *queue_ptr = q;
}

void magma_queue_create_from_hip_internal(magma_device_t device, void* hip_stream,
                                          void* /*hipblas_handle*/, void* /*hipsparse_handle*/,
                                          magma_queue_t* queue_ptr,
                                          const char* /*func*/, const char* /*file*/, int /*line*/)
{
    if (queue_ptr == nullptr) return;
    auto* q = new (std::nothrow) magma_queue();
    *queue_ptr = q;
    if (q == nullptr) return;
    q->device = device;
    q->stream = (hipStream_t)hip_stream;
    q->own_stream = false;
    q->scratch = nullptr;
    q->scratch_bytes = 0;
}

void magma_queue_destroy_internal(magma_queue_t queue, const char* /*func*/, const char* /*file*/, int /*line*/)
{
    if (queue == nullptr) return;
    hipStreamSynchronize(queue->stream);
    if (queue->scratch) hipFree(queue->scratch);
    if (queue->own_stream) hipStreamDestroy(queue->stream);
    delete queue;
}

void magma_queue_sync_internal(magma_queue_t queue, const char* func, const char* file, int line)
{
    if (queue == nullptr) return;
    hipError_t e = hipStreamSynchronize(queue->stream);
    if (e != hipSuccess) mgi::hip_report(e, func, file, line);
}

void* magma_queue_get_hip_stream(magma_queue_t queue) { return queue ? (void*)queue->stream : nullptr; }
magma_device_t magma_queue_get_device(magma_queue_t queue) { return queue ? queue->device : -1; }

// interface_cuda/copy_v2.cpp: the non-async forms synchronize the queue.
void magma_setvector_internal(magma_int_t n, magma_int_t elemSize, const void* hx_src, magma_int_t incx,
                              magma_ptr dy_dst, magma_int_t incy, magma_queue_t queue,
                              const char* func, const char* file, int line)
{
    if (n <= 0) return;
    hipStream_t s = queue ? queue->stream : nullptr;
    hipError_t e = hipMemcpy2DAsync(dy_dst, size_t(incy) * elemSize, hx_src, size_t(incx) * elemSize,
                                    elemSize, n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) mgi::hip_report(e, func, file, line);
}

void magma_getvector_internal(magma_int_t n, magma_int_t elemSize, const void* dx_src, magma_int_t incx,
                              void* hy_dst, magma_int_t incy, magma_queue_t queue,
                              const char* func, const char* file, int line)
{
    if (n <= 0) return;
    hipStream_t s = queue ? queue->stream : nullptr;
    hipError_t e = hipMemcpy2DAsync(hy_dst, size_t(incy) * elemSize, dx_src, size_t(incx) * elemSize,
                                    elemSize, n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) mgi::hip_report(e, func, file, line);
}

void magma_setmatrix_internal(magma_int_t m, magma_int_t n, magma_int_t elemSize,
                              const void* hA_src, magma_int_t lda, magma_ptr dB_dst, magma_int_t lddb,
                              magma_queue_t queue, const char* func, const char* file, int line)
{
    if (m <= 0 || n <= 0) return;
    hipStream_t s = queue ? queue->stream : nullptr;
    hipError_t e = hipMemcpy2DAsync(dB_dst, size_t(lddb) * elemSize, hA_src, size_t(lda) * elemSize,
                                    size_t(m) * elemSize, n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) mgi::hip_report(e, func, file, line);
}

void magma_getmatrix_internal(magma_int_t m, magma_int_t n, magma_int_t elemSize,
                              const void* dA_src, magma_int_t ldda, void* hB_dst, magma_int_t ldb,
                              magma_queue_t queue, const char* func, const char* file, int line)
{
    if (m <= 0 || n <= 0) return;
    hipStream_t s = queue ? queue->stream : nullptr;
    hipError_t e = hipMemcpy2DAsync(hB_dst, size_t(ldb) * elemSize, dA_src, size_t(ldda) * elemSize,
                                    size_t(m) * elemSize, n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) mgi::hip_report(e, func, file, line);
}

// control/xerbla.cpp:51 -- error report; neg_info is minus the info code, so
// a positive value below -MAGMA_ERR names the offending argument.
void magma_xerbla(const char* srname, magma_int_t neg_info)
{
    if (neg_info < 0) {
        std::fprintf(stderr, "Error in %s, function-specific error (info = %d)\n", srname, (int)-neg_info);
    } else if (neg_info == 0) {
        std::fprintf(stderr, "No error, why is %s calling xerbla? (info = %d)\n", srname, (int)-neg_info);
    } else if (neg_info >= -MAGMA_ERR) {
        std::fprintf(stderr, "Error in %s, %s (info = %d)\n", srname, magma_strerror(-neg_info), (int)-neg_info);
    } else {
        std::fprintf(stderr, "On entry to %s, parameter %d had an illegal value (info = %d)\n",
                     srname, (int)neg_info, (int)-neg_info);
    }
}

// interface_cuda/error.cpp -- message text per MAGMA error code.
const char* magma_strerror(magma_int_t error)
{
    if (error > 0) return "function-specific error, see documentation";
    if (error < 0 && error > MAGMA_ERR) return "invalid argument";
    switch (error) {
        case MAGMA_SUCCESS:             return "success";
        case MAGMA_ERR:                 return "unknown error";
        case MAGMA_ERR_NOT_INITIALIZED: return "not initialized";
        case MAGMA_ERR_NOT_SUPPORTED:   return "not supported";
        case MAGMA_ERR_HOST_ALLOC:      return "could not allocate CPU host memory";
        case MAGMA_ERR_DEVICE_ALLOC:    return "could not allocate GPU device memory";
        case MAGMA_ERR_INVALID_PTR:     return "invalid pointer";
        case MAGMA_ERR_UNKNOWN:         return "unknown error";
        case MAGMA_ERR_NOT_IMPLEMENTED: return "not implemented";
        default:                        return "unknown MAGMA error code";
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Pointer arrays (magmablas/zset_pointer.cu:17-40 semantics).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void set_pointer_kernel(T** out, T* input, long lda, long row, long column,
                                   long batch_offset, int batchCount)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < batchCount) out[i] = input + (long)i * batch_offset + row + column * lda;
}

extern "C" void magma_dset_pointer(double** output_array, double* input, magma_int_t lda,
                                   magma_int_t row, magma_int_t column, magma_int_t batch_offset,
                                   magma_int_t batchCount, magma_queue_t queue)
{
    if (batchCount <= 0) return;
    hipStream_t s = queue ? queue->stream : nullptr;
    set_pointer_kernel<double><<<mgi::ceildiv(batchCount, 256), 256, 0, s>>>(
        output_array, input, lda, row, column, batch_offset, batchCount);
}

extern "C" void magma_iset_pointer(magma_int_t** output_array, magma_int_t* input, magma_int_t lda,
                                   magma_int_t row, magma_int_t column, magma_int_t batchSize,
                                   magma_int_t batchCount, magma_queue_t queue)
{
    if (batchCount <= 0) return;
    hipStream_t s = queue ? queue->stream : nullptr;
    set_pointer_kernel<magma_int_t><<<mgi::ceildiv(batchCount, 256), 256, 0, s>>>(
        output_array, input, lda, row, column, batchSize, batchCount);
}

extern "C" void magma_sset_pointer(float** output_array, float* input, magma_int_t lda, magma_int_t row,
                                   magma_int_t column, magma_int_t batch_offset, magma_int_t batchCount,
                                   magma_queue_t queue)
{
    if (batchCount <= 0) return;
    hipStream_t s = queue ? queue->stream : nullptr;
    set_pointer_kernel<float><<<mgi::ceildiv(batchCount, 256), 256, 0, s>>>(
        output_array, input, lda, row, column, batch_offset, batchCount);
}
