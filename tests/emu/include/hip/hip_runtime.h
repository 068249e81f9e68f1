// CPU emulation of the HIP subset the library uses, for out-of-bounds hunting
// under AddressSanitizer. One OS thread per lane; wave intrinsics rendezvous
// the 64 lanes of a wave; blocks of a grid run one after another.
#pragma once
#include <atomic>
#include <barrier>
#include <cfloat>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct EmuIdx { unsigned x = 0, y = 0, z = 0; };
extern EmuIdx threadIdx, blockIdx;
extern EmuIdx blockDim, gridDim;

typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
typedef void* hipEvent_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
                     hipMemcpyDefault };
enum { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
enum { hipDeviceAttributeMaxSharedMemoryPerBlock = 74, hipDeviceAttributeMultiprocessorCount = 63 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
#define hipStreamNonBlocking 1

namespace emu {
struct Wave {
    uint64_t buf[64][8];
    bool alive[64];
};
struct Block {
    std::vector<Wave> waves;
    unsigned char* dyn = nullptr;
    std::map<int, std::vector<unsigned char>> shvars;   // static LDS of this workgroup
};
extern Block* cur_block;
extern int cur_lane_global;
void block_sync();
void wave_sync();
void yield();
unsigned long long realtime();
void* shared_var(int id, size_t bytes);
// a __shared__ declaration: storage of this workgroup (the build rewrites
// `__shared__ T name[..];` into `auto& name = emu::shvar<T[..]>(id);`)
template <class A>
inline A& shvar(int id) { return *reinterpret_cast<A*>(shared_var(id, sizeof(A))); }
void launch_impl(dim3 grid, dim3 block, size_t shm, const std::function<void()>& fn);
inline unsigned char* dyn_shared() { return cur_block->dyn; }
inline Wave& wave() { return cur_block->waves[cur_lane_global / 64]; }
inline int lane() { return cur_lane_global % 64; }
template <class T>
inline void put(const T& v)
{
    static_assert(sizeof(T) <= 64, "wave exchange value too large");
    std::memcpy(wave().buf[lane()], &v, sizeof(T));
}
template <class T>
inline T get(int l)
{
    T v;
    std::memcpy(&v, wave().buf[l], sizeof(T));
    return v;
}
inline void wsync() { wave_sync(); }
}  // namespace emu

inline void __syncthreads() { emu::block_sync(); }
// a wave executes in lockstep: its scheduling barrier is a rendezvous here
#define __builtin_amdgcn_wave_barrier() emu::wave_sync()
#define __builtin_amdgcn_fence(...) ((void)0)
#define __builtin_amdgcn_sched_barrier(...) ((void)0)
#define __builtin_amdgcn_s_waitcnt(...) ((void)0)
#define MGI_GLOBAL_AS

template <class T>
inline T emu_shfl(T v, int src)
{
    emu::put(v);
    emu::wsync();
    T r = emu::get<T>(src & 63);
    emu::wsync();
    return r;
}
inline int __shfl_xor(int v, int mask, int = 64) { return emu_shfl(v, emu::lane() ^ mask); }
inline float __shfl_xor(float v, int mask, int = 64) { return emu_shfl(v, emu::lane() ^ mask); }
inline int __shfl(int v, int src, int = 64) { return emu_shfl(v, src); }
inline float __shfl(float v, int src, int = 64) { return emu_shfl(v, src); }
inline unsigned long long __ballot(int pred)
{
    emu::put<int>(pred ? 1 : 0);
    emu::wave().alive[emu::lane()] = true;
    emu::wsync();
    unsigned long long m = 0;
    for (int l = 0; l < 64; ++l)
        if (emu::wave().alive[l] && emu::get<int>(l)) m |= 1ull << l;
    emu::wsync();
    return m;
}
inline int __builtin_amdgcn_readlane(int v, int l) { return emu_shfl(v, l); }
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound)
{
    (void)bank_mask;
    (void)bound;
    const int l = emu::lane();
    int s;
    if (ctrl < 0x100) s = (l & ~3) + ((ctrl >> (2 * (l & 3))) & 3);
    else if (ctrl == 0x140) s = (l & ~15) + (15 - (l & 15));
    else if (ctrl == 0x141) s = (l & ~7) + (7 - (l & 7));
    else if (ctrl == 0x142) s = (l & ~15) - 1;     // lane 15 of the previous row
    else if (ctrl == 0x143) s = 31;
    else {
        std::fprintf(stderr, "emu: unsupported dpp ctrl %#x\n", ctrl);
        std::abort();
    }
    const int r = emu_shfl(src, s < 0 ? l : s);
    const int row = l >> 4;
    if (!((row_mask >> row) & 1)) return old;
    if (s < 0) return old;
    return r;
}
inline int __popcll(unsigned long long m) { return __builtin_popcountll(m); }
inline int __ffsll(long long m) { return __builtin_ffsll(m); }
inline int __double2loint(double d) { int64_t b; std::memcpy(&b, &d, 8); return (int)(uint32_t)b; }
inline int __double2hiint(double d) { int64_t b; std::memcpy(&b, &d, 8); return (int)(b >> 32); }
inline double __hiloint2double(int hi, int lo)
{
    uint64_t b = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}
inline float __int_as_float(int i) { float f; std::memcpy(&f, &i, 4); return f; }
inline int __float_as_int(float f) { int i; std::memcpy(&i, &f, 4); return i; }
inline float __uint_as_float(unsigned i) { float f; std::memcpy(&f, &i, 4); return f; }
inline unsigned __float_as_uint(float f) { unsigned i; std::memcpy(&i, &f, 4); return i; }
inline double __longlong_as_double(long long i) { double f; std::memcpy(&f, &i, 8); return f; }
inline long long __double_as_longlong(double f) { long long i; std::memcpy(&i, &f, 8); return i; }
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline int atomicCAS(int* p, int cmp, int v)
{
    __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return cmp;
}
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), __ATOMIC_SEQ_CST)
inline void __threadfence() {}
// a wait yields to the other lanes and workgroups in flight
inline void __builtin_amdgcn_s_sleep(int) { emu::yield(); }
inline void __builtin_amdgcn_s_setprio(int) {}
inline unsigned long long __builtin_amdgcn_s_memrealtime() { return emu::realtime(); }
inline int atomicSub(int* p, int v) { return __atomic_fetch_sub(p, v, __ATOMIC_SEQ_CST); }

template <class A, class B>
inline auto min(A a, B b) -> typename std::common_type<A, B>::type { return a < b ? a : b; }
template <class A, class B>
inline auto max(A a, B b) -> typename std::common_type<A, B>::type { return a > b ? a : b; }
using std::fabs;
using std::fma;
using std::fmax;
using std::fmin;
using std::isfinite;
using std::sqrt;

// MFMA: operands exchanged across the wave; layouts of gfx950.
typedef double emu_d4 __attribute__((ext_vector_type(4)));
typedef float emu_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 emu_bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 emu_h8 __attribute__((ext_vector_type(8)));

inline emu_d4 __builtin_amdgcn_mfma_f64_16x16x4f64(double a, double b, emu_d4 c, int, int, int)
{
    struct AB { double a, b; };
    emu::put(AB{a, b});
    emu::wsync();
    const int l = emu::lane();
    emu_d4 out = c;
    for (int r = 0; r < 4; ++r) {
        const int row = 4 * r + l / 16, col = l % 16;
        double acc = c[r];
        for (int k = 0; k < 4; ++k) acc += emu::get<AB>(k * 16 + row).a * emu::get<AB>(k * 16 + col).b;
        out[r] = acc;
    }
    emu::wsync();
    return out;
}
inline emu_f4 __builtin_amdgcn_mfma_f32_16x16x4f32(float a, float b, emu_f4 c, int, int, int)
{
    struct AB { float a, b; };
    emu::put(AB{a, b});
    emu::wsync();
    const int l = emu::lane();
    emu_f4 out = c;
    for (int r = 0; r < 4; ++r) {
        const int row = 4 * (l / 16) + r, col = l % 16;
        float acc = c[r];
        for (int k = 0; k < 4; ++k) acc += emu::get<AB>(k * 16 + row).a * emu::get<AB>(k * 16 + col).b;
        out[r] = acc;
    }
    emu::wsync();
    return out;
}
template <class V8>
inline emu_f4 emu_mfma_x32(V8 a, V8 b, emu_f4 c)
{
    struct AB { float a[8], b[8]; };
    AB ab;
    for (int v = 0; v < 8; ++v) { ab.a[v] = (float)a[v]; ab.b[v] = (float)b[v]; }
    emu::put(ab);
    emu::wsync();
    const int l = emu::lane();
    emu_f4 out = c;
    for (int r = 0; r < 4; ++r) {
        const int row = 4 * (l / 16) + r, col = l % 16;
        float acc = c[r];
        for (int q = 0; q < 4; ++q)
            for (int v = 0; v < 8; ++v) acc += emu::get<AB>(q * 16 + row).a[v] * emu::get<AB>(q * 16 + col).b[v];
        out[r] = acc;
    }
    emu::wsync();
    return out;
}
inline emu_f4 __builtin_amdgcn_mfma_f32_16x16x32_bf16(emu_bf8 a, emu_bf8 b, emu_f4 c, int, int, int)
{
    return emu_mfma_x32(a, b, c);
}
inline emu_f4 __builtin_amdgcn_mfma_f32_16x16x32_f16(emu_h8 a, emu_h8 b, emu_f4 c, int, int, int)
{
    return emu_mfma_x32(a, b, c);
}
// FP8 (OCP e4m3fn on gfx950): round to nearest even, no infinity; beyond the
// largest finite value the result is NaN (the conversion runs without clamp).
inline float emu_fp8_dec(unsigned b)
{
    b &= 0xffu;
    const float sg = (b & 0x80u) ? -1.f : 1.f;
    const unsigned e = (b >> 3) & 0xfu, m = b & 7u;
    if (e == 15u && m == 7u) return NAN;
    if (e == 0u) return sg * std::ldexp((float)m, -9);
    return sg * std::ldexp(1.f + (float)m / 8.f, (int)e - 7);
}
inline unsigned emu_fp8_enc(float x)
{
    if (x != x) return 0x7fu;
    const unsigned sg = std::signbit(x) ? 0x80u : 0u;
    const float a = std::fabs(x);
    if (a >= 464.f) return sg | 0x7fu;
    if (a < 0.015625f) {
        const float q = std::nearbyint(a * 512.f);          // units of 2^-9, ties to even
        return sg | (unsigned)q;                            // q == 8 is the smallest normal
    }
    int e;
    const float fr = std::frexp(a, &e);                     // a = fr * 2^e, fr in [0.5, 1)
    int ex = e - 1;                                         // a = (2 fr) * 2^ex
    float q = std::nearbyint((2.f * fr - 1.f) * 8.f);
    if (q >= 8.f) { q = 0.f; ++ex; }
    if (ex + 7 > 15 || (ex + 7 == 15 && q > 6.f)) return sg | 0x7eu;   // 448
    return sg | ((unsigned)(ex + 7) << 3) | (unsigned)q;
}
inline int __builtin_amdgcn_cvt_pk_fp8_f32(float a, float b, int old, bool hi)
{
    const unsigned v = emu_fp8_enc(a) | (emu_fp8_enc(b) << 8);
    const unsigned o = (unsigned)old;
    return (int)(hi ? ((o & 0x0000ffffu) | (v << 16)) : ((o & 0xffff0000u) | v));
}
inline float __builtin_amdgcn_cvt_f32_fp8(int src, int byte_sel)
{
    return emu_fp8_dec((unsigned)src >> (8 * byte_sel));
}
struct emu_f8x8 {
    float v[8];
    float operator[](int i) const { return v[i]; }
};
inline emu_f4 __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(long a, long b, emu_f4 c, int, int, int)
{
    emu_f8x8 fa, fb;
    for (int v = 0; v < 8; ++v) {
        fa.v[v] = emu_fp8_dec((unsigned)((unsigned long)a >> (8 * v)));
        fb.v[v] = emu_fp8_dec((unsigned)((unsigned long)b >> (8 * v)));
    }
    return emu_mfma_x32(fa, fb, c);
}

// Runtime API (host memory stands in for device memory).
hipError_t hipMalloc(void** p, size_t bytes);
template <class T>
inline hipError_t hipMalloc(T** p, size_t bytes) { return hipMalloc((void**)p, bytes); }
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned = 0);
hipError_t hipHostFree(void* p);
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return 0; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t = nullptr)
{
    std::memmove(d, s, n);
    return 0;
}
inline hipError_t hipMemcpy2DAsync(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h,
                                   hipMemcpyKind, hipStream_t = nullptr)
{
    for (size_t r = 0; r < h; ++r) std::memmove((char*)d + r * dp, (const char*)s + r * sp, w);
    return 0;
}
inline hipError_t hipMemset(void* d, int v, size_t n) { std::memset(d, v, n); return 0; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t = nullptr) { std::memset(d, v, n); return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
inline hipError_t hipDeviceSynchronize() { return 0; }
inline hipError_t hipGetLastError() { return 0; }
inline hipError_t hipPeekAtLastError() { return 0; }
inline const char* hipGetErrorString(hipError_t) { return "emu"; }
inline hipError_t hipSetDevice(int) { return 0; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return 0; }
inline hipError_t hipGetDeviceCount(int* c) { *c = 1; return 0; }
inline hipError_t hipDeviceGetAttribute(int* v, int attr, int)
{
    *v = (attr == hipDeviceAttributeMaxSharedMemoryPerBlock) ? 163840 : 256;
    return 0;
}
inline hipError_t hipFuncSetAttribute(const void*, int, int) { return 0; }
// a __device__ array is a host array here: its address is the symbol's
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return 0; }
inline hipError_t hipStreamCreate(hipStream_t* s) { *s = (void*)0x1; return 0; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)0x1; return 0; }
inline hipError_t hipStreamDestroy(hipStream_t) { return 0; }
// launches run when issued, so every stream is the one host order, which
// respects any event dependencies the host expressed
inline hipError_t hipExtStreamCreateWithCUMask(hipStream_t* s, uint32_t, const uint32_t*) { *s = (void*)0x2; return 0; }
#define hipEventDisableTiming 2
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = (void*)0x3; return 0; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return 0; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return 0; }
inline hipError_t hipEventDestroy(hipEvent_t) { return 0; }
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return 1; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return 1; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return 1; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return 1; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return 0; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return 0; }

#define EMU_LAUNCH(GRID, BLOCK, ...) emu::Launch(dim3(GRID), dim3(BLOCK), ##__VA_ARGS__)
namespace emu {
struct Launch {
    dim3 g, b;
    size_t shm;
    Launch(dim3 g_, dim3 b_, size_t shm_ = 0, hipStream_t = nullptr) : g(g_), b(b_), shm(shm_) {}
    void operator<<(const std::function<void()>& fn) const { launch_impl(g, b, shm, fn); }
};
}  // namespace emu
