// Solves with the factors of one large LU (MAGMA's magma_dgetrs_gpu,
// src/zgetrs_gpu.cpp:203-227: NoTrans applies the interchanges, then L
// (unit lower) and U; Trans solves U^T, L^T, then applies the interchanges
// in reverse order).
//
// The interchanges are composed once into a row map (perm), so applying them
// is one gather (NoTrans) or scatter (Trans). Each triangular solve is one
// launch: the rows are cut into blocks of 128; the workgroup that takes
// block b (in the order the workgroups start, through a ticket counter)
// subtracts the products of the off-diagonal blocks with every solved block
// as soon as that block's values are published, then multiplies by the
// precomputed inverse of its diagonal block and publishes its own rows.
// A solved value is published as tagged 64-bit words ((epoch << 32) |
// 32 bits of the value; two words for FP64), so a reader that sees the
// solve's epoch in every word it needs has the values themselves: one
// memory round trip per block, no flag and no fence. A workgroup only waits
// for blocks taken by workgroups that started before it, so the solve cannot
// deadlock whatever the residency.
#include "magma_internal.h"
#include "lu_device.h"
#include "dense_kernels.h"

#include <algorithm>
#include <vector>

namespace mgi {

template <typename T>
magma_int_t dense_trtri_launch(bool upper, int n, int k0, int bs, int nblk, const T* A, int lda, T* X,
                               hipStream_t st);

namespace {

constexpr long long kSolveWaitTicks = 200000000LL;   // 2 s at the 100 MHz constant clock
constexpr int BS = kSolveBS;

__device__ __forceinline__ int ld_agent_i(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The low 32 bits of the word at p once its high half is epoch (bounded by
// kSolveWaitTicks from t0; a timeout, or another workgroup's, raises err and
// returns 0).
__device__ __forceinline__ unsigned wait_word(const unsigned long long* p, unsigned epoch, int* err,
                                              unsigned long long t0)
{
    for (;;) {
        const unsigned long long v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((unsigned)(v >> 32) == epoch) return (unsigned)v;
        if (ld_agent_i(err) != 0) return 0u;
        if ((long long)(__builtin_amdgcn_s_memrealtime() - t0) > kSolveWaitTicks) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return 0u;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

__device__ __forceinline__ void put_word(unsigned long long* p, unsigned epoch, unsigned payload)
{
    __hip_atomic_store(p, ((unsigned long long)epoch << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T> struct XWords;
template <> struct XWords<float> {
    static constexpr int W = 1;
    __device__ static float get(const unsigned long long* p, unsigned e, int* err, unsigned long long t0)
    {
        return __uint_as_float(wait_word(p, e, err, t0));
    }
    __device__ static void put(unsigned long long* p, unsigned e, float v) { put_word(p, e, __float_as_uint(v)); }
};
template <> struct XWords<double> {
    static constexpr int W = 2;
    __device__ static double get(const unsigned long long* p, unsigned e, int* err, unsigned long long t0)
    {
        const unsigned lo = wait_word(p, e, err, t0);
        const unsigned hi = wait_word(p + 1, e, err, t0);
        return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    }
    __device__ static void put(unsigned long long* p, unsigned e, double v)
    {
        const unsigned long long b = (unsigned long long)__double_as_longlong(v);
        put_word(p, e, (unsigned)b);
        put_word(p + 1, e, (unsigned)(b >> 32));
    }
};

}  // namespace

// KIND 0: L (unit lower), forward; 1: U, backward; 2: U^T (lower), forward;
// 3: L^T (upper, unit), backward. NR right-hand sides per workgroup.
// Thread t works on row t % 128 of its block and on half t / 128 of the
// columns of each off-diagonal block. Its 64 matrix values of a block do not
// depend on the solution, so they are loaded before the block's flag is
// awaited (one memory latency per block instead of one per column).
// xw: [nchunk][nblk][NR][BS][W] tagged words of the solved values.
template <typename T, int KIND, int NR>
__global__ void __launch_bounds__(256)
dense_trsv_kernel(int n, int nrhs, const T* __restrict__ LU, int ld, const T* __restrict__ invD,
                  T* __restrict__ X, int ldx, unsigned long long* __restrict__ xw, int* __restrict__ ticket,
                  int* __restrict__ err, unsigned epoch)
{
    constexpr int W = XWords<T>::W;
    constexpr bool kFwd = (KIND == 0 || KIND == 2);
    constexpr bool kTr = (KIND == 2 || KIND == 3);
    constexpr int H = BS / 2;
    __shared__ T xs[BS][NR];
    __shared__ T red[BS][NR];
    __shared__ int s_tk;
    const int t = threadIdx.x;
    const int rr = t % BS, h = t / BS;
    const int nblk = (n + BS - 1) / BS;
    if (t == 0) s_tk = atomicAdd(ticket, 1);
    __syncthreads();
    const int tk = s_tk;
    const int chunk = tk / nblk;
    const int bi = tk % nblk;
    const int b = kFwd ? bi : nblk - 1 - bi;
    const int r0 = b * BS;
    const int nb = min(BS, n - r0);
    const int q0 = chunk * NR;
    const int nq = min(NR, nrhs - q0);
    unsigned long long* xwc = xw + (size_t)chunk * nblk * (NR * BS * W);

    T acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q)
        acc[q] = (h == 0 && rr < nb && q < nq) ? X[(size_t)(r0 + rr) + (size_t)(q0 + q) * ldx] : T(0);

    const int ndep = kFwd ? b : nblk - 1 - b;
    for (int d = 0; d < ndep; ++d) {
        const int c = kFwd ? d : nblk - 1 - d;
        const int c0 = c * BS;
        const int cn = min(BS, n - c0);
        // this thread's part of block (b, c), independent of x_c
        T mv[H];
#pragma unroll
        for (int jj = 0; jj < H; ++jj) {
            const int col = h * H + jj;
            mv[jj] = (rr < nb && col < cn) ? (kTr ? LU[(size_t)(c0 + col) + (size_t)(r0 + rr) * ld]
                                                  : LU[(size_t)(r0 + rr) + (size_t)(c0 + col) * ld])
                                           : T(0);
        }
        // the values of block c, published by its workgroup
        __syncthreads();      // xs of the previous block consumed
        const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
        const unsigned long long* xc = xwc + (size_t)c * (NR * BS * W);
        for (int e = t; e < BS * NR; e += 256) {
            const int i = e % BS, q = e / BS;
            xs[i][q] = (i < cn && q < nq) ? XWords<T>::get(xc + (size_t)(q * BS + i) * W, epoch, err, t0) : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int jj = 0; jj < H; ++jj) {
#pragma unroll
            for (int q = 0; q < NR; ++q) acc[q] = fma(-mv[jj], xs[h * H + jj][q], acc[q]);
        }
    }
    // y = sum of the two halves; x_b = D^{-1} y
    const T* D = invD + (size_t)b * BS * BS;
    T dv[H];
#pragma unroll
    for (int jj = 0; jj < H; ++jj) {
        const int col = h * H + jj;
        dv[jj] = (rr < nb && col < nb) ? (kTr ? D[(size_t)col + (size_t)rr * BS] : D[(size_t)rr + (size_t)col * BS])
                                       : T(0);
    }
    __syncthreads();
    if (h == 1) {
#pragma unroll
        for (int q = 0; q < NR; ++q) red[rr][q] = acc[q];
    }
    __syncthreads();
    if (h == 0) {
#pragma unroll
        for (int q = 0; q < NR; ++q) xs[rr][q] = (rr < nb) ? acc[q] + red[rr][q] : T(0);
    }
    __syncthreads();
    T out[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) out[q] = T(0);
#pragma unroll
    for (int jj = 0; jj < H; ++jj) {
#pragma unroll
        for (int q = 0; q < NR; ++q) out[q] = fma(dv[jj], xs[h * H + jj][q], out[q]);
    }
    if (h == 1) {
#pragma unroll
        for (int q = 0; q < NR; ++q) red[rr][q] = out[q];
    }
    __syncthreads();
    if (h == 0 && rr < nb) {
        unsigned long long* xb = xwc + (size_t)b * (NR * BS * W);
#pragma unroll
        for (int q = 0; q < NR; ++q)
            if (q < nq) {
                const T v = out[q] + red[rr][q];
                XWords<T>::put(xb + (size_t)(q * BS + rr) * W, epoch, v);
                X[(size_t)(r0 + rr) + (size_t)(q0 + q) * ldx] = v;
            }
    }
}

// Y (ldy) = rows perm[i] of S (NoTrans gather), or X[perm[i]] = Y[i] (scatter),
// with conversion between the two element types.
template <typename TS, typename TD>
__global__ void dense_gather_kernel(int n, int nrhs, const int* __restrict__ perm, const TS* __restrict__ S, int lds,
                                    TD* __restrict__ D, int ldd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i < n && q < nrhs) D[(size_t)i + (size_t)q * ldd] = (TD)S[(size_t)(perm ? perm[i] : i) + (size_t)q * lds];
}

template <typename TS, typename TD>
__global__ void dense_scatter_kernel(int n, int nrhs, const int* __restrict__ perm, const TS* __restrict__ S, int lds,
                                     TD* __restrict__ D, int ldd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i < n && q < nrhs) D[(size_t)(perm ? perm[i] : i) + (size_t)q * ldd] = (TD)S[(size_t)i + (size_t)q * lds];
}

__global__ void dense_solve_reset_kernel(unsigned long long* xw, size_t nwords, int* ticket, int* err)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (size_t)gridDim.x * blockDim.x)
        xw[i] = 0ull;                       // epoch 0 is never a solve's
    if (blockIdx.x == 0 && threadIdx.x == 0) { *ticket = 0; *err = 0; }
}

__global__ void dense_ticket_reset_kernel(int* ticket) { *ticket = 0; }

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static size_t a256s(size_t b) { return (b + 255) & ~size_t(255); }

size_t solve_work_bytes(int n, int nrhs_max, int elem)
{
    const int nblk = (int)ceildiv(std::max(n, 1), BS);
    const int nchunk = (int)ceildiv(std::max(nrhs_max, 1), 8);
    size_t b = 0;
    b += 2 * a256s((size_t)elem * nblk * BS * BS);        // invL, invU
    b += a256s(sizeof(int) * std::max(n, 1));             // perm
    b += a256s(sizeof(unsigned long long) * 2 * (size_t)nblk * BS * nchunk * 8);   // tagged values
    b += a256s(sizeof(int) * 4);                          // ticket
    b += a256s(sizeof(int) * 4);                          // err
    b += a256s((size_t)elem * std::max(n, 1) * std::max(nrhs_max, 1));   // Y
    return b;
}

void solve_work_carve(void* base, int n, int nrhs_max, int elem, SolveWork& s)
{
    const int nblk = (int)ceildiv(std::max(n, 1), BS);
    const int nchunk = (int)ceildiv(std::max(nrhs_max, 1), 8);
    char* p = static_cast<char*>(base);
    s.invL = p;                            p += a256s((size_t)elem * nblk * BS * BS);
    s.invU = p;                            p += a256s((size_t)elem * nblk * BS * BS);
    s.perm = reinterpret_cast<int*>(p);    p += a256s(sizeof(int) * std::max(n, 1));
    s.xw = reinterpret_cast<unsigned long long*>(p);
    p += a256s(sizeof(unsigned long long) * 2 * (size_t)nblk * BS * nchunk * 8);
    s.ticket = reinterpret_cast<int*>(p);  p += a256s(sizeof(int) * 4);
    s.err = reinterpret_cast<int*>(p);     p += a256s(sizeof(int) * 4);
    s.Y = p;
    s.nrhs_max = std::max(nrhs_max, 1);
}

template <typename T>
magma_int_t dense_diag_inverses(int n, const T* dLU, int ldlu, const SolveWork& s, hipStream_t stream)
{
    if (n <= 0) return MAGMA_SUCCESS;
    const int nblk = (int)ceildiv(n, BS);
    magma_int_t rc = dense_trtri_launch<T>(false, n, 0, BS, nblk, dLU, ldlu, reinterpret_cast<T*>(s.invL), stream);
    if (rc != MAGMA_SUCCESS) return rc;
    return dense_trtri_launch<T>(true, n, 0, BS, nblk, dLU, ldlu, reinterpret_cast<T*>(s.invU), stream);
}

magma_int_t dense_set_perm(int n, const int* h_ipiv, const SolveWork& s, hipStream_t stream)
{
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = 0; i < n; ++i) {
        const int p = h_ipiv[i] - 1;
        if (p < 0 || p >= n) return MAGMA_ERR_ILLEGAL_VALUE;
        std::swap(perm[i], perm[p]);
    }
    MGI_HIP_CHECK(hipMemcpyAsync(s.perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream));
    // perm is pageable host memory read by the queued copy
    MGI_HIP_CHECK(hipStreamSynchronize(stream));
    return MAGMA_SUCCESS;
}

namespace {

template <typename T, int NR>
magma_int_t trsv_pair(int trans, int n, int nrhs, const T* LU, int ld, T* Y, int ldy, const SolveWork& s,
                      hipStream_t st)
{
    const int nblk = (int)ceildiv(n, BS);
    const int nchunk = (int)ceildiv(nrhs, NR);
    const int nwg = nblk * nchunk;
    const T* iL = reinterpret_cast<const T*>(s.invL);
    const T* iU = reinterpret_cast<const T*>(s.invU);
    // the tagged words this pair uses, cleared (capacity: nrhs_max <= 64)
    const size_t nwords = (size_t)nchunk * nblk * NR * BS * XWords<T>::W;
    if (nwords > (size_t)2 * nblk * BS * ceildiv(s.nrhs_max, 8) * 8) return MAGMA_ERR_NOT_SUPPORTED;
    dense_solve_reset_kernel<<<(unsigned)std::min<size_t>(ceildiv((long)nwords, 256L), 1024), 256, 0, st>>>(
        s.xw, nwords, s.ticket, s.err);
    MGI_LAUNCH_CHECK();
    const unsigned e1 = 1, e2 = 2;     // the words are cleared for every pair of solves
    if (trans == 0) {
        dense_trsv_kernel<T, 0, NR><<<nwg, 256, 0, st>>>(n, nrhs, LU, ld, iL, Y, ldy, s.xw, s.ticket, s.err, e1);
        MGI_LAUNCH_CHECK();
        dense_ticket_reset_kernel<<<1, 1, 0, st>>>(s.ticket);
        dense_trsv_kernel<T, 1, NR><<<nwg, 256, 0, st>>>(n, nrhs, LU, ld, iU, Y, ldy, s.xw, s.ticket, s.err, e2);
    } else {
        dense_trsv_kernel<T, 2, NR><<<nwg, 256, 0, st>>>(n, nrhs, LU, ld, iU, Y, ldy, s.xw, s.ticket, s.err, e1);
        MGI_LAUNCH_CHECK();
        dense_ticket_reset_kernel<<<1, 1, 0, st>>>(s.ticket);
        dense_trsv_kernel<T, 3, NR><<<nwg, 256, 0, st>>>(n, nrhs, LU, ld, iL, Y, ldy, s.xw, s.ticket, s.err, e2);
    }
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

}  // namespace

// Solve in place on Y (n x nrhs, ldy, already permuted for NoTrans).
template <typename T>
magma_int_t dense_trsv_both(int trans, int n, int nrhs, const T* LU, int ld, T* Y, int ldy, const SolveWork& s,
                            hipStream_t st)
{
    if (nrhs == 1) return trsv_pair<T, 1>(trans, n, nrhs, LU, ld, Y, ldy, s, st);
    return trsv_pair<T, 8>(trans, n, nrhs, LU, ld, Y, ldy, s, st);
}
template magma_int_t dense_trsv_both<double>(int, int, int, const double*, int, double*, int, const SolveWork&,
                                             hipStream_t);
template magma_int_t dense_trsv_both<float>(int, int, int, const float*, int, float*, int, const SolveWork&,
                                            hipStream_t);

template <typename T>
magma_int_t dense_getrs(int trans, int n, int nrhs, const T* dLU, int ldlu, T* dX, int ldx, const SolveWork& s,
                        hipStream_t stream)
{
    if (n <= 0 || nrhs <= 0) return MAGMA_SUCCESS;
    T* Y = reinterpret_cast<T*>(s.Y);
    for (int q0 = 0; q0 < nrhs; q0 += s.nrhs_max) {
        const int nq = std::min(s.nrhs_max, nrhs - q0);
        T* Xq = dX + (size_t)q0 * ldx;
        dim3 g((unsigned)ceildiv(n, 256), (unsigned)nq);
        if (trans == 0) dense_gather_kernel<T, T><<<g, 256, 0, stream>>>(n, nq, s.perm, Xq, ldx, Y, n);
        else dense_gather_kernel<T, T><<<g, 256, 0, stream>>>(n, nq, nullptr, Xq, ldx, Y, n);
        MGI_LAUNCH_CHECK();
        magma_int_t rc = dense_trsv_both<T>(trans, n, nq, dLU, ldlu, Y, n, s, stream);
        if (rc != MAGMA_SUCCESS) return rc;
        if (trans == 0) dense_scatter_kernel<T, T><<<g, 256, 0, stream>>>(n, nq, nullptr, Y, n, Xq, ldx);
        else dense_scatter_kernel<T, T><<<g, 256, 0, stream>>>(n, nq, s.perm, Y, n, Xq, ldx);
        MGI_LAUNCH_CHECK();
    }
    return MAGMA_SUCCESS;
}

template magma_int_t dense_diag_inverses<double>(int, const double*, int, const SolveWork&, hipStream_t);
template magma_int_t dense_diag_inverses<float>(int, const float*, int, const SolveWork&, hipStream_t);
template magma_int_t dense_getrs<double>(int, int, int, const double*, int, double*, int, const SolveWork&,
                                         hipStream_t);
template magma_int_t dense_getrs<float>(int, int, int, const float*, int, float*, int, const SolveWork&,
                                        hipStream_t);

magma_int_t dense_solve_check_err(const SolveWork& s, hipStream_t stream)
{
    int h = 0;
    MGI_HIP_CHECK(hipMemcpyAsync(&h, s.err, sizeof(int), hipMemcpyDeviceToHost, stream));
    MGI_HIP_CHECK(hipStreamSynchronize(stream));
    if (h != 0) {
        std::fprintf(stderr, "magma_mi355x: dense triangular solve wait timed out\n");
        return MAGMA_ERR_UNKNOWN;
    }
    return MAGMA_SUCCESS;
}

}  // namespace mgi
