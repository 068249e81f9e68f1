// Blocked right-looking batched LU for matrices too large for one workgroup's
// on-chip storage, fixed or variable size, FP64 or FP32 storage. Per block
// column of width jb:
//   1. panel:  LU with partial pivoting of the (m-k) x jb panel, rows spread
//              over the workgroup's threads in registers;
//   2. laswp:  the panel's interchanges applied to the columns left and right
//              of the panel;
//   3. trsm:   A12 := L11^{-1} A12 (unit lower);
//   4. gemm:   A22 := A22 - A21 * A12 on MFMA. FP64 storage uses
//              v_mfma_f64_16x16x4f64. FP32 storage uses, by update format:
//              FP32 inputs (v_mfma_f32_16x16x4f32), BF16 or FP16 inputs with
//              FP32 accumulation (v_mfma_f32_16x16x32_{bf16,f16}), or a
//              three-term BF16 split (hi*hi + hi*lo + lo*hi).
// Matrices whose size is exhausted exit at once, so one launch per step
// serves a batch of mixed sizes.
//
// Reference structure: src/zgetrf_batched.cpp:149-203 (blocked loop),
// src/zgetrf_panel_batched.cpp:102-196 (panel, laswp, trsm, gemm per step),
// src/zgetrf_vbatched.cpp:20-124 (variable sizes, per-kernel early exit),
// src/xshgetrf_gpu.cpp:214-350 (FP32 panel, FP16-input / FP32-accumulate
// trailing update). Pivot, scaling and zero-pivot semantics follow LAPACK
// dgetf2 as in getrf_lds.hip.
#include "magma_internal.h"
#include "lu_device.h"
#include "getrf_kernels.h"

#include <type_traits>
#include <utility>

namespace mgi {

constexpr int kPanelThreads = 256;

template <int... Js, typename F>
__device__ __forceinline__ void static_for(std::integer_sequence<int, Js...>, F&& f)
{
    (f(std::integral_constant<int, Js>{}), ...);
}

template <typename T> struct Limits;
template <> struct Limits<double> { static constexpr double sfmin = kSfminD; };
template <> struct Limits<float>  { static constexpr float sfmin = kSfminF; };

__device__ __forceinline__ void wave_argmax_t(double& v, int& idx) { wave_argmax(v, idx); }
__device__ __forceinline__ void wave_argmax_t(float& v, int& idx) { wave_argmax_f(v, idx); }

// ---------------------------------------------------------------------------
// 1. Panel: thread t owns panel rows t + 256*i (i < R); the jb <= NB columns
//    of those rows sit in registers.
// ---------------------------------------------------------------------------
template <typename T, int R, int NB>
__global__ void __launch_bounds__(kPanelThreads)
panel_kernel(SizeArgs sz, int k, T** __restrict__ dA_array, int** __restrict__ ipiv_array,
             int* __restrict__ info_array)
{
    constexpr int kW = kPanelThreads / kWave;
    // per column step (double-buffered by step parity, so one barrier per
    // step suffices): each wave's best candidate with its whole panel row,
    // and row j as it was before the interchange
    __shared__ T s_val[2][kW];
    __shared__ int s_idx[2][kW];
    __shared__ T s_cand[2][kW][NB];
    __shared__ T s_rowj[2][NB];

    const int b = sz.mat(blockIdx.x);
    int m, n, ld;
    sz.get(b, m, n, ld);
    const int minmn = min(m, n);
    if (k >= minmn) return;
    const int jb = min(NB, minmn - k);
    const int mloc = m - k;
    const int t = threadIdx.x;
    const int w = t / kWave;
    auto A = gptr(dA_array[b]) + k + (size_t)k * ld;     // panel origin
    // a null pivot array asks for the factorisation without pivoting
    // (src/zgetrf_nopiv_batched.cpp): the diagonal entry is the pivot
    const bool nopiv = ipiv_array == nullptr;
    auto ipiv = gptr(nopiv ? nullptr : ipiv_array[b]);

    T a[R][NB];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int r = t + kPanelThreads * i;
#pragma unroll
        for (int j = 0; j < NB; ++j)
            a[i][j] = (r < mloc && j < jb) ? A[r + (size_t)j * ld] : T(0);
    }

    int info = 0;
    // every column step is instantiated with its index as a constant (a
    // run-time column index would move the register panel to scratch memory)
    static_for(std::make_integer_sequence<int, NB>{}, [&](auto jc) {
        constexpr int j = decltype(jc)::value;
        constexpr int par = j & 1;
        if (j < jb) {
            T best = T(-1);
            int bi = 0x7fffffff, bs = 0;
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int r = t + kPanelThreads * i;
                if (r >= j && r < mloc && (!nopiv || r == j)) {
                    // idamax: a NaN never wins a comparison, except in the
                    // first candidate, which keeps the pivot (counted as +inf)
                    T v = fabs(a[i][j]);
                    if (r == j && (v != v || nopiv)) v = T(INFINITY);
                    if (v > best) { best = v; bi = r; bs = i; }
                }
            }
            T wb = best;
            int wi = bi;
            wave_argmax_t(wb, wi);
            // the holder of the wave's best publishes it with its row; the
            // holder of row j (thread j, slot 0) publishes row j
            if (bi == wi && best == wb && wi != 0x7fffffff) {
                s_val[par][w] = wb;
                s_idx[par][w] = wi;
#pragma unroll
                for (int i = 0; i < R; ++i)
                    if (bs == i) {
#pragma unroll
                        for (int c = 0; c < NB; ++c) s_cand[par][w][c] = a[i][c];
                    }
            } else if ((t & (kWave - 1)) == 0 && wi == 0x7fffffff) {
                s_val[par][w] = T(-1);
                s_idx[par][w] = 0x7fffffff;
            }
            if (t == j) {
#pragma unroll
                for (int c = 0; c < NB; ++c) s_rowj[par][c] = a[0][c];
            }
            __syncthreads();
            T pbest = s_val[par][0];
            int p = s_idx[par][0], pw = 0;
#pragma unroll
            for (int q = 1; q < kW; ++q) {
                const T ov = s_val[par][q];
                const int oi = s_idx[par][q];
                if (ov > pbest || (ov == pbest && oi < p)) { pbest = ov; p = oi; pw = q; }
            }
            const T* prow = s_cand[par][pw];
            const T* jrow = s_rowj[par];
            const T pv = prow[j];
            if (t == 0 && !nopiv) ipiv[k + j] = k + p + 1;
            T rinv = T(1);
            bool recip = true;
            if (pv != T(0)) {
                if (p != j) {
                    if (t == j) {
#pragma unroll
                        for (int c = 0; c < NB; ++c) a[0][c] = prow[c];
                    }
#pragma unroll
                    for (int i = 0; i < R; ++i)
                        if (t + kPanelThreads * i == p) {
#pragma unroll
                            for (int c = 0; c < NB; ++c) a[i][c] = jrow[c];
                        }
                }
                recip = fabs(pv) >= Limits<T>::sfmin;
                rinv = T(1) / pv;
            } else if (info == 0) {
                info = k + j + 1;
            }
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int r = t + kPanelThreads * i;
                if (r > j && r < mloc) {
                    T l = a[i][j];
                    if (pv != T(0)) l = recip ? l * rinv : l / pv;
                    a[i][j] = l;
#pragma unroll
                    for (int c = 0; c < NB; ++c)
                        if (c > j && c < jb) a[i][c] = fma(-l, prow[c], a[i][c]);
                }
            }
        }
    });

#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int r = t + kPanelThreads * i;
        if (r < mloc) {
#pragma unroll
            for (int j = 0; j < NB; ++j)
                if (j < jb) A[r + (size_t)j * ld] = a[i][j];
        }
    }
    if (t == 0 && info != 0 && info_array[b] == 0) info_array[b] = info;
}

// ---------------------------------------------------------------------------
// 2. Row interchanges of panel k..k+jb-1 on every column outside the panel.
//    The jb <= 64 interchanges touch at most 2*jb rows: the panel rows
//    k..k+jb-1 and the pivot rows below them. Each lane takes up to two
//    touched destination rows and traces each back through the interchanges
//    (last to first) to the row whose content ends there; the wave then moves
//    one column at a time, every lane loading its sources before any lane
//    stores (the composition equals applying the interchanges in order, as
//    dlaswp).
//    Interchanges outside [k+j, m) cannot come from a valid panel and are
//    treated as none.
// ---------------------------------------------------------------------------
constexpr int kLaswpCols = 32;     // columns per workgroup
// Column ranges of the two-level blocked LU (64-wide blocks factored as two
// 32-wide panels): see blocked_steps2.
constexpr int kInPanelRight = 1;
constexpr int kInPanelLeft = 2;
constexpr int kRightOnly = 3;     // only the columns right of the panel

template <typename T>
__global__ void __launch_bounds__(256)
laswp_kernel(SizeArgs sz, int k, int nb, T** __restrict__ dA_array, int** __restrict__ ipiv_array, int mode)
{
    __shared__ int s_piv[64];
    const int b = sz.mat(blockIdx.y);
    int m, n, ld;
    sz.get(b, m, n, ld);
    const int minmn = min(m, n);
    if (k >= minmn) return;
    const int jb = min(nb, minmn - k);
    // mode 0: every column outside the panel; kRightOnly: the columns right
    // of the panel (the left ones get the deferred pass of
    // laswp_deferred_kernel); kInPanelRight: the columns of
    // the second half of a 64-wide block (panel = its first half); kInPanelLeft:
    // the 32 columns left of the panel (panel = the block's second half)
    int span = n - jb, lo = 0;
    if (mode == kInPanelRight) {
        const int W = min(2 * nb, minmn - k);
        if (W <= nb) return;
        lo = k + nb;
        span = W - nb;
    } else if (mode == kInPanelLeft) {
        lo = k - nb;
        span = nb;
    } else if (mode == kRightOnly) {
        lo = k + jb;
        span = n - k - jb;
    }
    const int cbeg = blockIdx.x * kLaswpCols;
    if (cbeg >= span) return;
    const int t = threadIdx.x;
    if (t < jb) {
        const int p = gptr(ipiv_array[b])[k + t] - 1;
        s_piv[t] = (p > k + t && p < m) ? p : k + t;
    }
    __syncthreads();
    // lane l covers panel row k + l and the pivot row of step l (when that
    // lies below the panel; a pivot row inside the panel has its own lane)
    const int l = t % kWave, w = t / kWave;
    int dst0 = -1, dst1 = -1;
    if (l < jb) {
        dst0 = k + l;
        dst1 = s_piv[l];
        if (dst1 < k + jb) dst1 = -1;
    }
    int src0 = dst0, src1 = dst1;
    for (int j = jb - 1; j >= 0; --j) {
        const int p = s_piv[j];
        src0 = (src0 == k + j) ? p : (src0 == p ? k + j : src0);
        src1 = (src1 == k + j) ? p : (src1 == p ? k + j : src1);
    }
    const bool mv0 = dst0 >= 0 && src0 != dst0, mv1 = dst1 >= 0 && src1 != dst1;
    const int cend = min(span, cbeg + kLaswpCols);
    // each wave moves its kLaswpCols / 4 columns at once: every source is
    // read before any store (the columns are disjoint, so the loads of all
    // of them are in flight together)
    constexpr int kPer = kLaswpCols / (256 / kWave);
    T v0[kPer], v1[kPer];
    auto base = gptr(dA_array[b]);
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int c = cbeg + w + u * (256 / kWave);
        v0[u] = T(0);
        v1[u] = T(0);
        if (c < cend) {
            const int col = mode == 0 ? (c < k ? c : c + jb) : lo + c;
            auto A = base + (size_t)col * ld;
            if (mv0) v0[u] = A[src0];
            if (mv1) v1[u] = A[src1];
        }
    }
    // every load of the wave has returned before the first store issues
    __builtin_amdgcn_s_waitcnt(0x0F70);            // vmcnt(0); expcnt, lgkmcnt unconstrained
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int c = cbeg + w + u * (256 / kWave);
        if (c < cend) {
            const int col = mode == 0 ? (c < k ? c : c + jb) : lo + c;
            auto A = base + (size_t)col * ld;
            if (mv0) A[dst0] = v0[u];
            if (mv1) A[dst1] = v1[u];
        }
    }
}

// ---------------------------------------------------------------------------
// 2b. Deferred interchanges of the factored columns. The L columns of block
//     bk (columns bk*nbk .. bk*nbk + nbk - 1) are read by no later step, so
//     the interchanges of the later steps (rows t >= T0 = (bk+1)*nbk) are
//     applied to them once, after the last step, instead of one scattered
//     pass per step. Every row r >= T0 is traced back through those
//     interchanges (last to first) to the row whose content ends there, as
//     in laswp_kernel; then each column moves its rows, all loads of a batch
//     of columns completing before any store. The composition equals applying
//     the interchanges in order, so the factors are identical to those of a
//     step-by-step dlaswp.
// ---------------------------------------------------------------------------
constexpr int kDeferMaxSteps = 2048;   // later steps one workgroup traces (max_m of the blocked LU)

template <typename T, int R>
__global__ void __launch_bounds__(256)
laswp_deferred_kernel(SizeArgs sz, int nbk, T** __restrict__ dA_array, int** __restrict__ ipiv_array)
{
    __shared__ int s_piv[kDeferMaxSteps];
    const int b = sz.mat(blockIdx.y);
    int m, n, ld;
    sz.get(b, m, n, ld);
    const int minmn = min(m, n);
    const int c0 = blockIdx.x * nbk;
    const int T0 = c0 + nbk;
    if (T0 >= minmn) return;                     // no later step for this block
    const int ncols = nbk;                       // c0 + nbk = T0 < minmn <= n
    const int tid = threadIdx.x;
    auto ip = gptr(ipiv_array[b]);
    for (int t = T0 + tid; t < minmn; t += 256) {
        const int p = ip[t] - 1;
        s_piv[t - T0] = (p > t && p < m) ? p : t;
    }
    __syncthreads();
    int src[R];
#pragma unroll
    for (int i = 0; i < R; ++i) src[i] = T0 + tid + 256 * i;
    for (int t = minmn - 1; t >= T0; --t) {
        const int p = s_piv[t - T0];
#pragma unroll
        for (int i = 0; i < R; ++i) src[i] = (src[i] == t) ? p : (src[i] == p ? t : src[i]);
    }
    bool mv[R];
#pragma unroll
    for (int i = 0; i < R; ++i) mv[i] = T0 + tid + 256 * i < m && src[i] != T0 + tid + 256 * i;
    auto A = gptr(dA_array[b]) + (size_t)c0 * ld;
    constexpr int CB = R <= 2 ? 8 : (R <= 4 ? 4 : 2);   // columns per batch
    for (int cb = 0; cb < ncols; cb += CB) {
        T v[CB][R];
#pragma unroll
        for (int u = 0; u < CB; ++u)
#pragma unroll
            for (int i = 0; i < R; ++i)
                v[u][i] = (cb + u < ncols && mv[i]) ? A[(size_t)(cb + u) * ld + src[i]] : T(0);
        __builtin_amdgcn_s_waitcnt(0x0F70);       // vmcnt(0): every load has returned
        __syncthreads();
#pragma unroll
        for (int u = 0; u < CB; ++u)
#pragma unroll
            for (int i = 0; i < R; ++i)
                if (cb + u < ncols && mv[i]) A[(size_t)(cb + u) * ld + T0 + tid + 256 * i] = v[u][i];
        __syncthreads();
    }
}

template <typename T>
static magma_int_t deferred_launch(int R, const SizeArgs& sz, int nbk, int max_minmn, T** dA_array,
                                   int** ipiv_array, int batch, hipStream_t stream)
{
    const int nblk = ceildiv(max_minmn, nbk) - 1;   // blocks followed by a later step
    if (ipiv_array == nullptr || nblk <= 0) return MAGMA_SUCCESS;
    const dim3 grid(nblk, batch);
    switch (R) {
        case 1: laswp_deferred_kernel<T, 1><<<grid, 256, 0, stream>>>(sz, nbk, dA_array, ipiv_array); break;
        case 2: laswp_deferred_kernel<T, 2><<<grid, 256, 0, stream>>>(sz, nbk, dA_array, ipiv_array); break;
        case 4: laswp_deferred_kernel<T, 4><<<grid, 256, 0, stream>>>(sz, nbk, dA_array, ipiv_array); break;
        case 8: laswp_deferred_kernel<T, 8><<<grid, 256, 0, stream>>>(sz, nbk, dA_array, ipiv_array); break;
        default: return MAGMA_ERR_NOT_SUPPORTED;
    }
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

// ---------------------------------------------------------------------------
// 3. A12 := L11^{-1} A12 (unit lower). NB lanes run down the rows of one
//    column (64/NB columns per wave instruction), so loads and stores are
//    contiguous; the solved x_j reaches the rows below through scalar lane
//    reads (no LDS round trip in the dependent chain). Same operation order
//    as a column-oriented dtrsm.
// ---------------------------------------------------------------------------
constexpr int kTrsmCols = 64;      // columns per workgroup

template <typename T, int NB>
__global__ void __launch_bounds__(256)
trsm_kernel(SizeArgs sz, int k, T** __restrict__ dA_array, int mode)
{
    constexpr int CPW = kWave / NB;                 // columns per wave pass
    __shared__ T sL[NB * NB];
    const int b = sz.mat(blockIdx.y);
    int m, n, ld;
    sz.get(b, m, n, ld);
    const int minmn = min(m, n);
    if (k >= minmn) return;
    const int jb = min(NB, minmn - k);
    int ncols = n - k - jb;
    if (mode == kInPanelRight) {                    // only the block's second half
        const int W = min(2 * NB, minmn - k);
        if (W <= NB) return;
        ncols = W - NB;
    }
    if (ncols <= 0) return;
    const int c0 = blockIdx.x * kTrsmCols;
    if (c0 >= ncols) return;
    const int t = threadIdx.x;
    auto L = gptr(dA_array[b]) + k + (size_t)k * ld;
    for (int idx = t; idx < NB * NB; idx += 256) {
        const int i = idx % NB, j = idx / NB;
        sL[idx] = (i < jb && j < jb) ? L[i + (size_t)j * ld] : T(0);
    }
    __syncthreads();
    const int l = t % kWave, w = t / kWave;
    const int i = l % NB, cq = l / NB;
    const int cend = min(ncols, c0 + kTrsmCols);
    // Q column groups per wave pass: their loads are in flight together and
    // their chains interleave (each column's operation order is unchanged)
    constexpr int Q = 4;
    constexpr int kStride = (256 / kWave) * CPW;    // columns between a wave's groups
    auto A0 = gptr(dA_array[b]) + k + (size_t)(k + jb) * ld;
    for (int cb = c0 + w * CPW; cb < cend; cb += kStride * Q) {
        T x[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int c = cb + q * kStride + cq;
            x[q] = (c < cend && i < jb) ? A0[(size_t)c * ld + i] : T(0);
        }
#pragma unroll
        for (int j = 0; j < NB - 1; ++j) {
            const T lij = sL[i + j * NB];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                // x_j of this lane's column: one scalar read per column group
                T xj = readlane_t(x[q], j);
#pragma unroll
                for (int g = 1; g < CPW; ++g) {
                    const T v = readlane_t(x[q], g * NB + j);
                    xj = cq == g ? v : xj;
                }
                if (i > j) x[q] = fma(-lij, xj, x[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int c = cb + q * kStride + cq;
            if (c < cend && i < jb) A0[(size_t)c * ld + i] = x[q];
        }
    }
}

// ---------------------------------------------------------------------------
// 4. A22 := A22 - A21 * A12. A workgroup of 4 waves computes a 64x64 tile of
//    A22 (each wave a 32x32 quarter = 2x2 MFMA tiles); K = jb <= NB is staged
//    in LDS.
//
//    MFMA operand layouts (lane l of 64):
//      16x16x4 (f64, f32):  A[l%16][l/16], B[l/16][l%16]
//      16x16x32 (bf16, f16): A[l%16][8*(l/16)+v], B[8*(l/16)+v][l%16], v < 8
//    accumulators (4 per lane, r = 0..3):
//      f64:  C[4*r + l/16][l%16]
//      f32:  C[4*(l/16) + r][l%16]
//    The product is formed transposed (A12^T as the first operand, A21^T as
//    the second), so lane l holds row l%16 of the tile: each store writes 16
//    consecutive rows of one column (one contiguous segment per column)
//    instead of one element in each of 16 columns.
// ---------------------------------------------------------------------------
typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));

template <typename T, int NB, int MODE>
__global__ void __launch_bounds__(256)
gemm_update_kernel(SizeArgs sz, int k, int nb, T** __restrict__ dA_array, int mode)
{
    constexpr bool kWide = (MODE == kUpdBF16 || MODE == kUpdFP16 || MODE == kUpdBF16x3);
    constexpr int KP = kWide ? (NB + 31) / 32 * 32 : NB;   // K staged (zero padded to whole MFMA steps)
    __shared__ T sA[KP][64 + 1];                     // A21 tile, [kk][row]
    __shared__ T sB[KP][64 + 1];                     // A12 tile, [kk][col]
    const int b = sz.mat(blockIdx.z);
    int m, n, ld;
    sz.get(b, m, n, ld);
    const int minmn = min(m, n);
    if (k >= minmn) return;
    const int jb = min(nb, minmn - k);
    const int mr = m - k - jb;
    int nc = n - k - jb;
    if (mode == kInPanelRight) {                    // only the block's second half
        const int W = min(2 * nb, minmn - k);
        if (W <= nb) return;
        nc = W - nb;
    }
    const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
    if (r0 >= mr || c0 >= nc) return;
    auto A = gptr(dA_array[b]);
    auto A21 = A + (k + jb) + (size_t)k * ld;
    auto A12 = A + k + (size_t)(k + jb) * ld;
    auto C = A + (k + jb) + (size_t)(k + jb) * ld;
    const int t = threadIdx.x;
    for (int idx = t; idx < KP * 64; idx += 256) {
        const int r = idx % 64, kk = idx / 64;           // A21: walk down columns
        sA[kk][r] = (kk < jb && r0 + r < mr) ? A21[r0 + r + (size_t)kk * ld] : T(0);
        const int kb = idx % KP, cc = idx / KP;          // A12: walk down its short columns
        sB[kb][cc] = (kb < jb && c0 + cc < nc) ? A12[kb + (size_t)(c0 + cc) * ld] : T(0);
    }
    __syncthreads();
    const int w = t / kWave, l = t % kWave;
    const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
    const int lr = l % 16, lq = l / 16;

    if constexpr (MODE == kUpdF64) {
        double4_t acc[2][2];
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kb = 0; kb < NB; kb += 4) {
            if (kb < jb) {
                double af[2], bf[2];
#pragma unroll
                for (int ti = 0; ti < 2; ++ti) af[ti] = sA[kb + lq][wr + 16 * ti + lr];
#pragma unroll
                for (int tj = 0; tj < 2; ++tj) bf[tj] = sB[kb + lq][wc + 16 * tj + lr];
#pragma unroll
                for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                    for (int tj = 0; tj < 2; ++tj)
                        acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[tj], af[ti], acc[ti][tj], 0, 0, 0);
            }
        }
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) {
                const int row = r0 + wr + 16 * ti + lr;
                if (row >= mr) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = c0 + wc + 16 * tj + 4 * r + lq;
                    if (col < nc) {
                        auto cp = C + row + (size_t)col * ld;
                        *cp = *cp - acc[ti][tj][r];
                    }
                }
            }
    } else {
        float4_t acc[2][2];
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = float4_t{0.f, 0.f, 0.f, 0.f};
        if constexpr (MODE == kUpdF32) {
#pragma unroll
            for (int kb = 0; kb < NB; kb += 4) {
                if (kb < jb) {
                    float af[2], bf[2];
#pragma unroll
                    for (int ti = 0; ti < 2; ++ti) af[ti] = sA[kb + lq][wr + 16 * ti + lr];
#pragma unroll
                    for (int tj = 0; tj < 2; ++tj) bf[tj] = sB[kb + lq][wc + 16 * tj + lr];
#pragma unroll
                    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                        for (int tj = 0; tj < 2; ++tj)
                            acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[tj], af[ti], acc[ti][tj], 0, 0, 0);
                }
            }
        } else {
            // K = 32 per MFMA step: lane holds k = kb + 8*(l/16) + v, v < 8
#pragma unroll
            for (int kb = 0; kb < KP; kb += 32) {
                float av[2][8], bv[2][8];
#pragma unroll
                for (int v = 0; v < 8; ++v) {
#pragma unroll
                    for (int ti = 0; ti < 2; ++ti) av[ti][v] = sA[kb + 8 * lq + v][wr + 16 * ti + lr];
#pragma unroll
                    for (int tj = 0; tj < 2; ++tj) bv[tj][v] = sB[kb + 8 * lq + v][wc + 16 * tj + lr];
                }
                if constexpr (MODE == kUpdFP16) {
                    f16x8_t ah[2], bh[2];
#pragma unroll
                    for (int v = 0; v < 8; ++v) {
#pragma unroll
                        for (int q = 0; q < 2; ++q) {
                            ah[q][v] = (_Float16)av[q][v];
                            bh[q][v] = (_Float16)bv[q][v];
                        }
                    }
#pragma unroll
                    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                        for (int tj = 0; tj < 2; ++tj)
                            acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bh[tj], ah[ti], acc[ti][tj], 0, 0, 0);
                } else {
                    bf16x8_t ah[2], bh[2], al[2], bl[2];
#pragma unroll
                    for (int v = 0; v < 8; ++v) {
#pragma unroll
                        for (int q = 0; q < 2; ++q) {
                            ah[q][v] = (__bf16)av[q][v];
                            bh[q][v] = (__bf16)bv[q][v];
                            al[q][v] = (__bf16)(av[q][v] - (float)ah[q][v]);
                            bl[q][v] = (__bf16)(bv[q][v] - (float)bh[q][v]);
                        }
                    }
#pragma unroll
                    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                        for (int tj = 0; tj < 2; ++tj) {
                            if constexpr (MODE == kUpdBF16x3) {
                                acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[tj], al[ti], acc[ti][tj], 0, 0, 0);
                                acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[tj], ah[ti], acc[ti][tj], 0, 0, 0);
                            }
                            acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[tj], ah[ti], acc[ti][tj], 0, 0, 0);
                        }
                }
            }
        }
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) {
                const int row = r0 + wr + 16 * ti + lr;
                if (row >= mr) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = c0 + wc + 16 * tj + 4 * lq + r;
                    if (col < nc) {
                        auto cp = C + row + (size_t)col * ld;
                        *cp = *cp - acc[ti][tj][r];
                    }
                }
            }
    }
}

__global__ void fill_int_kernel(int* p, int v, int count, const int* map)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[map ? map[i] : i] = v;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct BlockedPlan {
    int rows_per_thread;   // R
    int nb;                // panel width
};

static bool plan_for(int max_m, BlockedPlan& p)
{
    if (max_m <= 256)       { p.rows_per_thread = 1; p.nb = 32; }
    else if (max_m <= 512)  { p.rows_per_thread = 2; p.nb = 64; }   // two-level: 2 x 32 per 64-wide block
    else if (max_m <= 1024) { p.rows_per_thread = 4; p.nb = 16; }
    else if (max_m <= 2048) { p.rows_per_thread = 8; p.nb = 8; }
    else return false;
    return true;
}

template <typename T, int NB>
static magma_int_t panel_launch(int R, const SizeArgs& sz, int k, T** dA_array, int** ipiv_array, int* info_array,
                                int batch, hipStream_t stream)
{
    // only panels of R * NB <= 64 register values per thread exist
    // (blocked_nb_valid)
    auto panel = [&](auto rc) -> magma_int_t {
        constexpr int RR = decltype(rc)::value;
        if constexpr (RR * NB <= 64) {
            panel_kernel<T, RR, NB><<<batch, kPanelThreads, 0, stream>>>(sz, k, dA_array, ipiv_array, info_array);
            return MAGMA_SUCCESS;
        } else {
            return MAGMA_ERR_NOT_SUPPORTED;
        }
    };
    magma_int_t prc = MAGMA_ERR_NOT_SUPPORTED;
    switch (R) {
        case 1: prc = panel(std::integral_constant<int, 1>{}); break;
        case 2: prc = panel(std::integral_constant<int, 2>{}); break;
        case 4: prc = panel(std::integral_constant<int, 4>{}); break;
        case 8: prc = panel(std::integral_constant<int, 8>{}); break;
        default: break;
    }
    if (prc != MAGMA_SUCCESS) return prc;
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

// The full trailing update in FP64 or FP32: TM x TM output tiles (TM = 128
// for trailing matrices of at least 256 x 256, so each A21 / U12 strip is
// read half as often; TM = 64 below that). Four waves of TM/2 x TM/2
// (TM/32 x TM/32 MFMA tiles each); K staged through LDS 16 at a time,
// double-buffered, one barrier per stage, so the next stage's loads overlap
// the current products; the product is formed transposed (the same store
// pattern as above). Grid: x = tile of the system (column-major over the
// tile grid), y = system.
template <typename T, int TM>
__global__ void __launch_bounds__(256, 2)
gemm_update_db_kernel(SizeArgs sz, int k, int nb, T** __restrict__ dA_array, int mode)
{
    constexpr int BK = 16, LDP = TM + 4;
    constexpr int WT = TM / 2;              // each wave's square of the tile
    constexpr int MT = WT / 16;             // MFMA tiles per wave and dimension
    constexpr int AS = 256 / TM;            // k stride between a thread's A elements
    constexpr int AQ = TM * BK / 256;       // A elements per thread and stage
    constexpr int BQ = TM / 16;             // B elements per thread and stage
    const int b = sz.mat(blockIdx.y);
    int m, n, ld;
    sz.get(b, m, n, ld);
    const int minmn = min(m, n);
    if (k >= minmn) return;
    const int jb = min(nb, minmn - k);
    int N = n - k - jb;
    if (mode == kInPanelRight) {                    // only the block's second half
        const int W = min(2 * nb, minmn - k);
        if (W <= nb) return;
        N = W - nb;
    }
    const int M = m - k - jb, K = jb;
    if (M <= 0 || N <= 0) return;
    const int tiles_m = (M + TM - 1) / TM;
    const int tid = blockIdx.x;
    if (tid >= tiles_m * ((N + TM - 1) / TM)) return;
    const int m0 = (tid % tiles_m) * TM, n0 = (tid / tiles_m) * TM;
    auto Ab = gptr(dA_array[b]);
    auto A = Ab + (k + jb) + (size_t)k * ld;              // A21
    auto B = Ab + k + (size_t)(k + jb) * ld;              // U12
    auto C = Ab + (k + jb) + (size_t)(k + jb) * ld;
    const int t = threadIdx.x;
    const int w = t / kWave, l = t % kWave;
    const int wr = (w & 1) * WT, wc = (w >> 1) * WT;
    const int lr = l % 16, lq = l / 16;
    __shared__ T sA[2][BK][LDP];
    __shared__ T sB[2][BK][LDP];
    using acc_t = typename std::conditional<std::is_same<T, double>::value, double4_t, float4_t>::type;
    acc_t acc[MT][MT];
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int c = 0; c < MT; ++c) acc[a][c] = acc_t{0, 0, 0, 0};
    T ra[AQ], rb[BQ];
    const int ai = t % TM, ak = t / TM;        // A tile: row ai, k = ak + AS q
    const int bk = t & 15, bj = t >> 4;        // B tile: k = bk, column bj + 16q
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < AQ; ++q) {
            const int kk = ak + AS * q;
            ra[q] = (m0 + ai < M && k0 + kk < K) ? A[(size_t)(m0 + ai) + (size_t)(k0 + kk) * ld] : T(0);
        }
#pragma unroll
        for (int q = 0; q < BQ; ++q) {
            const int jj = bj + 16 * q;
            rb[q] = (k0 + bk < K && n0 + jj < N) ? B[(size_t)(k0 + bk) + (size_t)(n0 + jj) * ld] : T(0);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < AQ; ++q) sA[buf][ak + AS * q][ai] = ra[q];
#pragma unroll
        for (int q = 0; q < BQ; ++q) sB[buf][bk][bj + 16 * q] = rb[q];
    };
    const int nst = (K + BK - 1) / BK;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        if (st + 1 < nst) gload((st + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T pf[MT], qf[MT];
#pragma unroll
            for (int x = 0; x < MT; ++x) {
                pf[x] = sB[buf][kk + lq][wc + 16 * x + lr];
                qf[x] = sA[buf][kk + lq][wr + 16 * x + lr];
            }
#pragma unroll
            for (int tj = 0; tj < MT; ++tj)
#pragma unroll
                for (int ti = 0; ti < MT; ++ti) {
                    if constexpr (std::is_same<T, double>::value)
                        acc[tj][ti] = __builtin_amdgcn_mfma_f64_16x16x4f64(pf[tj], qf[ti], acc[tj][ti], 0, 0, 0);
                    else
                        acc[tj][ti] = __builtin_amdgcn_mfma_f32_16x16x4f32(pf[tj], qf[ti], acc[tj][ti], 0, 0, 0);
                }
        }
        if (st + 1 < nst) lstore(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int tj = 0; tj < MT; ++tj)
#pragma unroll
        for (int ti = 0; ti < MT; ++ti) {
            const int row = m0 + wr + 16 * ti + lr;
            if (row >= M) continue;
#pragma unroll
            for (int rg = 0; rg < 4; ++rg) {
                const int cc = std::is_same<T, double>::value ? (4 * rg + lq) : (4 * lq + rg);
                const int col = n0 + wc + 16 * tj + cc;
                if (col < N) {
                    auto cp = C + (size_t)row + (size_t)col * ld;
                    *cp = *cp - acc[tj][ti][rg];
                }
            }
        }
}

// The full trailing update after a block of NB columns: 128-wide tiles for
// the FP64 / FP32 updates while the trailing matrix is at least 256 x 256
// (below that the 64-wide tiles give each system enough workgroups: at
// n = 256 the 128-wide tiles were 14 % slower), the 64-wide kernel otherwise
// and for the 16-bit formats.
template <typename T, int NB, int MODE>
static magma_int_t trailing_update(const SizeArgs& sz, int k, int rest_m, int rest_n, T** dA_array, int batch,
                                   hipStream_t stream)
{
    if constexpr (MODE == kUpdF64 || MODE == kUpdF32) {
        if (rest_m >= 256 && rest_n >= 256) {
            const int tiles = ceildiv(rest_m, 128) * ceildiv(rest_n, 128);
            gemm_update_db_kernel<T, 128><<<dim3(tiles, batch), 256, 0, stream>>>(sz, k, NB, dA_array, 0);
            MGI_LAUNCH_CHECK();
            return MAGMA_SUCCESS;
        }
        const int tiles = ceildiv(rest_m, 64) * ceildiv(rest_n, 64);
        gemm_update_db_kernel<T, 64><<<dim3(tiles, batch), 256, 0, stream>>>(sz, k, NB, dA_array, 0);
        MGI_LAUNCH_CHECK();
        return MAGMA_SUCCESS;
    }
    dim3 grid(ceildiv(rest_m, 64), ceildiv(rest_n, 64), batch);
    gemm_update_kernel<T, NB, MODE><<<grid, 256, 0, stream>>>(sz, k, NB, dA_array, 0);
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

template <typename T, int NB, int MODE>
static magma_int_t blocked_steps(int R, const SizeArgs& sz, int max_m, int max_n, int max_minmn,
                                 T** dA_array, int** ipiv_array, int* info_array, int batch,
                                 hipStream_t stream)
{
    for (int k = 0; k < max_minmn; k += NB) {
        magma_int_t prc = panel_launch<T, NB>(R, sz, k, dA_array, ipiv_array, info_array, batch, stream);
        if (prc != MAGMA_SUCCESS) return prc;
        if (ipiv_array && max_n - k - 1 > 0) {        // columns right of the panel
            laswp_kernel<T><<<dim3(ceildiv(max_n - k - 1, kLaswpCols), batch), 256, 0, stream>>>(
                sz, k, NB, dA_array, ipiv_array, kRightOnly);
            MGI_LAUNCH_CHECK();
        }
        const int rest_n = max_n - k - 1;
        const int rest_m = max_m - k - 1;
        if (rest_n > 0) {
            trsm_kernel<T, NB><<<dim3(ceildiv(rest_n, kTrsmCols), batch), 256, 0, stream>>>(sz, k, dA_array, 0);
            MGI_LAUNCH_CHECK();
            if (rest_m > 0) {
                const magma_int_t urc = trailing_update<T, NB, MODE>(sz, k, rest_m, rest_n, dA_array, batch, stream);
                if (urc != MAGMA_SUCCESS) return urc;
            }
        }
    }
    // the interchanges of later steps on the factored columns
    return deferred_launch<T>(R, sz, NB, max_minmn, dA_array, ipiv_array, batch, stream);
}

// Two-level blocking: each 64-wide block is factored as two 32-wide panels
// (the second half updated from the first inside the block), then the
// interchanges, L11^{-1} and the trailing update are applied once for the
// whole block (K = 64). Same arithmetic as the 32-wide blocked loop, with
// half as many passes over the trailing matrix and over the row swaps.
template <typename T, int MODE>
static magma_int_t blocked_steps2(int R, const SizeArgs& sz, int max_m, int max_n, int max_minmn,
                                  T** dA_array, int** ipiv_array, int* info_array, int batch,
                                  hipStream_t stream)
{
    constexpr int H = 32;
    for (int k = 0; k < max_minmn; k += 2 * H) {
        // first half: panel, then its interchanges, U12 and update on the
        // block's second half only
        magma_int_t prc = panel_launch<T, H>(R, sz, k, dA_array, ipiv_array, info_array, batch, stream);
        if (prc != MAGMA_SUCCESS) return prc;
        if (ipiv_array) {
            laswp_kernel<T><<<dim3(1, batch), 256, 0, stream>>>(sz, k, H, dA_array, ipiv_array, kInPanelRight);
            MGI_LAUNCH_CHECK();
        }
        trsm_kernel<T, H><<<dim3(1, batch), 256, 0, stream>>>(sz, k, dA_array, kInPanelRight);
        MGI_LAUNCH_CHECK();
        if (max_m - k - H > 0) {
            if constexpr (MODE == kUpdF64 || MODE == kUpdF32) {
                // at most H columns: one 64-wide column tile per row tile
                gemm_update_db_kernel<T, 64><<<dim3(ceildiv(max_m - k - H, 64), batch), 256, 0, stream>>>(
                    sz, k, H, dA_array, kInPanelRight);
            } else {
                gemm_update_kernel<T, H, MODE><<<dim3(ceildiv(max_m - k - H, 64), 1, batch), 256, 0, stream>>>(
                    sz, k, H, dA_array, kInPanelRight);
            }
            MGI_LAUNCH_CHECK();
        }
        // second half: panel, then its interchanges on the first half
        if (k + H < max_minmn) {
            prc = panel_launch<T, H>(R, sz, k + H, dA_array, ipiv_array, info_array, batch, stream);
            if (prc != MAGMA_SUCCESS) return prc;
            if (ipiv_array) {
                laswp_kernel<T><<<dim3(1, batch), 256, 0, stream>>>(sz, k + H, H, dA_array, ipiv_array,
                                                                    kInPanelLeft);
                MGI_LAUNCH_CHECK();
            }
        }
        // the whole block's interchanges outside it, U12 and the trailing update
        if (ipiv_array && max_n - k - 1 > 0) {        // columns right of the block
            laswp_kernel<T><<<dim3(ceildiv(max_n - k - 1, kLaswpCols), batch), 256, 0, stream>>>(
                sz, k, 2 * H, dA_array, ipiv_array, kRightOnly);
            MGI_LAUNCH_CHECK();
        }
        const int rest_n = max_n - k - 1;
        const int rest_m = max_m - k - 1;
        if (rest_n > 0) {
            trsm_kernel<T, 2 * H><<<dim3(ceildiv(rest_n, kTrsmCols), batch), 256, 0, stream>>>(sz, k, dA_array, 0);
            MGI_LAUNCH_CHECK();
            if (rest_m > 0) {
                const magma_int_t urc =
                    trailing_update<T, 2 * H, MODE>(sz, k, rest_m, rest_n, dA_array, batch, stream);
                if (urc != MAGMA_SUCCESS) return urc;
            }
        }
    }
    // the interchanges of later blocks on the factored columns
    return deferred_launch<T>(R, sz, 2 * H, max_minmn, dA_array, ipiv_array, batch, stream);
}

bool getrf_blocked_supported(int max_m) { BlockedPlan p; return plan_for(max_m, p); }
int getrf_blocked_nb(int max_m) { BlockedPlan p; return plan_for(max_m, p) ? p.nb : 0; }

// Panel widths a given row count supports: the panel keeps R x NB values per
// thread in registers, R = ceil(max_m / 256), and R * NB <= 64. Width 64
// means the two-level form (two 32-wide panels per 64-wide block), R <= 2.
bool blocked_nb_valid(int max_m, int nb)
{
    BlockedPlan p;
    if (!plan_for(max_m, p)) return false;
    if (nb == 64) return p.rows_per_thread <= 2;     // two-level: 64-wide blocks of two 32-wide panels
    return (nb == 8 || nb == 16 || nb == 32) && p.rows_per_thread * nb <= 64;
}

template <typename T, int MODE>
static magma_int_t blocked_launch(const SizeArgs& sz, int max_m, int max_n, T** dA_array, int** ipiv_array,
                                  int* info_array, int batch, hipStream_t stream, int force_nb = 0)
{
    BlockedPlan plan;
    if (!plan_for(max_m, plan)) return MAGMA_ERR_NOT_SUPPORTED;
    if (force_nb != 0) {
        if (!blocked_nb_valid(max_m, force_nb)) return MAGMA_ERR_NOT_SUPPORTED;
        plan.nb = force_nb;
    }
    fill_int_kernel<<<ceildiv(batch, 256), 256, 0, stream>>>(info_array, 0, batch, sz.map);
    MGI_LAUNCH_CHECK();
    const int max_minmn = std::min(max_m, max_n);
    // the row-interchange, triangular-solve and update kernels put the batch
    // in grid y or z: split larger batches (src/zgetrf_batched.cpp splits at
    // the queue's maximum batch the same way)
    for (int c0 = 0; c0 < batch; c0 += kMaxGridYZ) {
        SizeArgs part = sz;
        part.base = sz.base + c0;
        const int cb = std::min(kMaxGridYZ, batch - c0);
        magma_int_t rc;
        switch (plan.nb) {
            case 64: rc = blocked_steps2<T, MODE>(plan.rows_per_thread, part, max_m, max_n, max_minmn, dA_array, ipiv_array, info_array, cb, stream); break;
            case 32: rc = blocked_steps<T, 32, MODE>(plan.rows_per_thread, part, max_m, max_n, max_minmn, dA_array, ipiv_array, info_array, cb, stream); break;
            case 16: rc = blocked_steps<T, 16, MODE>(plan.rows_per_thread, part, max_m, max_n, max_minmn, dA_array, ipiv_array, info_array, cb, stream); break;
            case 8:  rc = blocked_steps<T, 8, MODE>(plan.rows_per_thread, part, max_m, max_n, max_minmn, dA_array, ipiv_array, info_array, cb, stream); break;
            default: rc = MAGMA_ERR_NOT_SUPPORTED; break;
        }
        if (rc != MAGMA_SUCCESS) return rc;
    }
    return MAGMA_SUCCESS;
}

magma_int_t getrf_blocked_launch(int m, int n, int ldda, const int* m_arr, const int* n_arr, const int* ld_arr,
                                 int max_m, int max_n, double** dA_array, int** ipiv_array, int* info_array,
                                 int batch, hipStream_t stream)
{
    SizeArgs sz{m, n, ldda, m_arr, n_arr, ld_arr, nullptr};
    return blocked_launch<double, kUpdF64>(sz, max_m, max_n, dA_array, ipiv_array, info_array, batch, stream);
}

magma_int_t getrf_blocked_launch_sz(const SizeArgs& sz, int max_m, int max_n, double** dA_array, int** ipiv_array,
                                    int* info_array, int batch, hipStream_t stream, int force_nb)
{
    return blocked_launch<double, kUpdF64>(sz, max_m, max_n, dA_array, ipiv_array, info_array, batch, stream,
                                           force_nb);
}

magma_int_t sgetrf_blocked_launch(int mode, int m, int n, int ldda, const int* m_arr, const int* n_arr,
                                  const int* ld_arr, int max_m, int max_n, float** dA_array, int** ipiv_array,
                                  int* info_array, int batch, hipStream_t stream)
{
    SizeArgs sz{m, n, ldda, m_arr, n_arr, ld_arr, nullptr};
    switch (mode) {
        case kUpdF32:   return blocked_launch<float, kUpdF32>(sz, max_m, max_n, dA_array, ipiv_array, info_array, batch, stream);
        case kUpdBF16:  return blocked_launch<float, kUpdBF16>(sz, max_m, max_n, dA_array, ipiv_array, info_array, batch, stream);
        case kUpdFP16:  return blocked_launch<float, kUpdFP16>(sz, max_m, max_n, dA_array, ipiv_array, info_array, batch, stream);
        case kUpdBF16x3: return blocked_launch<float, kUpdBF16x3>(sz, max_m, max_n, dA_array, ipiv_array, info_array, batch, stream);
        default: return MAGMA_ERR_NOT_SUPPORTED;
    }
}

}  // namespace mgi
