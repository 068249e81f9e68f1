// Internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <cstddef>
#include "../../include/magma_mi355x.h"

// One queue = one HIP stream on one device, plus a device scratch arena that
// grows on demand and is reused across calls on the same queue (MAGMA's queue
// likewise owns pointer-array workspace, control/magma_internal.h:112-120).
// A queue must not be used from two host threads at once.
struct magma_queue {
    int          device;
    hipStream_t  stream;
    bool         own_stream;
    void*        scratch;
    size_t       scratch_bytes;
};

namespace mgi {

inline magma_int_t ceildiv(magma_int_t a, magma_int_t b) { return (a + b - 1) / b; }
inline magma_int_t roundup(magma_int_t a, magma_int_t b) { return ceildiv(a, b) * b; }

// Returns a device buffer of at least `bytes` owned by the queue. The previous
// contents are not preserved when it grows. Not valid across concurrent calls
// on one queue.
void* queue_scratch(magma_queue_t queue, size_t bytes);

// Report a failing HIP call with the caller location; returns MAGMA_ERR_UNKNOWN
// style codes to the caller.
void hip_report(hipError_t err, const char* what, const char* file, int line);

// Largest grid extent in y and z: launches that put the batch there split it
// into launches of at most this many matrices.
constexpr int kMaxGridYZ = 65535;

// A matrix pointer read from a device array of pointers carries no address
// space, so accesses through it compile to flat instructions, which count
// against the LDS/scalar wait counter as well: every wait for an LDS read
// then also waits for the outstanding global loads. The batch's matrices
// always live in global memory; gptr() states that to the compiler.
#ifndef MGI_GLOBAL_AS
#define MGI_GLOBAL_AS __attribute__((address_space(1)))
#endif
template <typename T>
__device__ __forceinline__ MGI_GLOBAL_AS T* gptr(T* p)
{
    return (MGI_GLOBAL_AS T*)p;
}

}  // namespace mgi

#define MGI_HIP_CHECK(call)                                                    \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            mgi::hip_report(e_, #call, __FILE__, __LINE__);                    \
            return MAGMA_ERR_UNKNOWN;                                          \
        }                                                                      \
    } while (0)

#define MGI_LAUNCH_CHECK()                                                     \
    do {                                                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess) {                                                \
            mgi::hip_report(e_, "kernel launch", __FILE__, __LINE__);          \
            return MAGMA_ERR_UNKNOWN;                                          \
        }                                                                      \
    } while (0)
