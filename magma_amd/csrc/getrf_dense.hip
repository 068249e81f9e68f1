// Blocked right-looking LU with partial pivoting of one large column-major
// matrix on the whole GPU (MAGMA's magma_dgetrf_gpu / magma_dgetrf_native,
// src/zgetrf_gpu.cpp:213-452 and :542-663, and the FP32 LU with low-precision
// trailing updates of src/xshgetrf_gpu.cpp:86-419).
//
// Blocks of nb columns (FP64 256, FP32 512), each factored as panels of PW
// columns (FP64 64, FP32 128):
//   panel    dense_panel_kernel: the (m-k) x PW panel is spread over G
//            workgroups, one row per thread in registers. Per column every
//            workgroup publishes its best candidate row; after all G have
//            published, each workgroup picks the same pivot (largest |a|,
//            lowest LAPACK position, NaN only at the first position), and
//            applies the rank-1 update to its rows. Rows are not moved while
//            the panel runs: each row carries its LAPACK position, and the
//            final rows are written to their positions at the end.
//   swaps    dense_pivmap_kernel composes a block's interchanges into
//            (source, destination) row pairs; dense_laswp_kernel moves those
//            rows in every other column, one wave per column.
//   U12      L11^{-1} (dense_trtri_kernel) times A12 on MFMA, in the storage
//            precision (FP64, or FP32 as MAGMA's strsm).
//   update   A22 -= A21 U12 on MFMA, 128 x 128 tiles: FP64 or FP32
//            (dense_gemm_kernel); or FP32 storage with BF16, FP16,
//            split-BF16 or (as an experiment) FP8 e4m3 inputs and FP32
//            accumulation, A21 and U12 converted once per update to 16- or
//            8-bit rows / columns with k contiguous (dense_cvt_a/b_kernel,
//            dense_gemm_h_kernel).
// Inside a block, each panel's interchanges go to the block's other columns
// and its U rows and update to the block's columns on its right before the
// next panel; the whole block's interchanges, U12 and update (K = nb) are
// then applied once to the rest. The triangular solves for U run in strips
// of 128 rows (the inverse of each strip's diagonal block, then a GEMM).
#include "magma_internal.h"
#include "lu_device.h"
#include "dense_kernels.h"

#include <cstdlib>
#include <vector>

#include <algorithm>
#include <climits>
#include <type_traits>
#include <utility>

namespace mgi {

namespace {

template <int... Js, typename F>
__device__ __forceinline__ void dense_for(std::integer_sequence<int, Js...>, F&& f)
{
    (f(std::integral_constant<int, Js>{}), ...);
}

template <typename T> struct DLim;
template <> struct DLim<double> { static constexpr double sfmin = kSfminD; };
template <> struct DLim<float>  { static constexpr float sfmin = kSfminF; };

__device__ __forceinline__ void argmax_t(double& v, int& idx) { wave_argmax(v, idx); }
__device__ __forceinline__ void argmax_t(float& v, int& idx) { wave_argmax_f(v, idx); }

// A wait gives up after this many ticks of the 100 MHz constant clock (2 s)
// and raises the error word instead of spinning on.
constexpr long long kDenseWaitTicks = 200000000LL;

__device__ __forceinline__ int ld_agent(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));

}  // namespace

// ---------------------------------------------------------------------------
// Panel: G workgroups of 256 threads; thread t of workgroup g holds row
// k + 256 g + t (columns k .. k+jb-1) in registers.
//
// Exchange between the workgroups, per column: workgroup g writes its best
// candidate into slot (column parity, g) as 64-bit words that each carry the
// column's tag in their upper half: word 0 = tag | position, then the row's
// values from the current column on, 32 bits per word (FP64 values take two
// words). A reader accepts a word only when its tag is the current column's,
// so every word validates itself: no fence, flag or ordering between words
// is needed, only the single-copy atomicity of aligned 64-bit accesses.
// Wave 0 reads every slot's position and current value, picks the pivot, then
// reads the winning row; both reads repeat until all their words carry the
// tag (with a time limit that raises the error word instead of hanging).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ unsigned piece_of(T v, int q)
{
    if constexpr (sizeof(T) == 8) return (unsigned)(q == 0 ? __double2hiint(v) : __double2loint(v));
    else return (unsigned)__float_as_int(v);
}

template <typename T>
__device__ __forceinline__ T value_of(unsigned hi, unsigned lo)
{
    if constexpr (sizeof(T) == 8) return __hiloint2double((int)hi, (int)lo);
    else return __int_as_float((int)hi);
}

__device__ __forceinline__ unsigned long long ld_word(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_word(unsigned long long* p, unsigned tag, unsigned payload)
{
    __hip_atomic_store(p, ((unsigned long long)tag << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T, int PW>
__global__ void __launch_bounds__(kDensePanelRows)
dense_panel_kernel(int m, int k, int jb, T* __restrict__ A, int lda, int* __restrict__ ipiv, int* __restrict__ info,
                   DenseWork w)
{
    constexpr int kW = kDensePanelRows / kWave;
    constexpr int NW = sizeof(T) / 4;              // words per value
    constexpr int SW = 1 + PW * NW;                 // words per slot
    __shared__ T s_wv[2][kW];       // per wave: best |a| and its position (parity-buffered)
    __shared__ int s_wp[2][kW];
    __shared__ T s_cand[2][kW][PW]; // per wave: its best row
    __shared__ T s_u[2][PW];        // pivot row
    __shared__ int s_pp[2];         // pivot position

    const int G = gridDim.x;
    const int g = blockIdx.x;
    const int t = threadIdx.x;
    const int wv = t / kWave, ln = t % kWave;
    const int mloc = m - k;
    const int r = g * kDensePanelRows + t;
    const bool valid = r < mloc;
    int pos = k + r;                 // LAPACK position of the row this thread holds
    bool active = valid;             // not yet chosen as a pivot
    T* P = A + k + (size_t)k * lda;
    const int gmax = w.gmax;
    unsigned long long* words = reinterpret_cast<unsigned long long*>(w.crow);
    // the panel is the critical path of the factorisation: its waves go
    // first where they share a SIMD with the trailing update's
    __builtin_amdgcn_s_setprio(3);

    T a[PW];
#pragma unroll
    for (int c = 0; c < PW; ++c) a[c] = (valid && c < jb) ? P[r + (size_t)c * lda] : T(0);

    dense_for(std::make_integer_sequence<int, PW>{}, [&](auto jc) {
        constexpr int j = decltype(jc)::value;
        constexpr int par = j & 1;
        if (j < jb) {
            const int diag = k + j;
            const unsigned tag = (unsigned)diag;
            // 1. best candidate of each wave (LAPACK idamax: largest |a|,
            //    lowest position; a NaN wins only at the first position)
            T v = T(-1);
            int vp = INT_MAX;
            if (active) {
                T x = fabs(a[j]);
                if (x != x) x = (pos == diag) ? T(INFINITY) : T(-1);
                if (x >= T(0)) { v = x; vp = pos; }
            }
            argmax_t(v, vp);
            if (ln == 0) { s_wv[par][wv] = v; s_wp[par][wv] = vp; }
            if (vp != INT_MAX && active && pos == vp) {
#pragma unroll
                for (int c = j; c < PW; ++c) s_cand[par][wv][c] = a[c];
            }
            __syncthreads();
            // 2. wave 0: publish the workgroup's best, gather every
            //    workgroup's, pick the pivot and stage its row
            if (wv == 0) {
                T bv = s_wv[par][0];
                int bp = s_wp[par][0], bw = 0;
#pragma unroll
                for (int q = 1; q < kW; ++q) {
                    const T ov = s_wv[par][q];
                    const int op = s_wp[par][q];
                    if (ov > bv || (ov == bv && op < bp)) { bv = ov; bp = op; bw = q; }
                }
                const bool has = bp != INT_MAX;
                unsigned long long* mine = words + ((size_t)par * gmax + g) * SW;
                if (has) {
                    for (int e = ln; e < (PW - j) * NW; e += kWave) {
                        const int c = j + e / NW, q = e % NW;
                        st_word(mine + 1 + c * NW + q, tag, piece_of(s_cand[par][bw][c], q));
                    }
                }
                if (ln == 0) st_word(mine, tag, has ? (unsigned)bp : 0xFFFFFFFFu);
                // every slot's position and current value
                const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
                T cv = T(-1);
                int cp = INT_MAX, cg = 0;
                for (;;) {
                    bool ok = true;
                    cv = T(-1);
                    cp = INT_MAX;
                    cg = 0;
                    for (int q = ln; q < G; q += kWave) {
                        const unsigned long long* sl = words + ((size_t)par * gmax + q) * SW;
                        // header and value together (one round trip): each
                        // word carries its own tag, a stale one is discarded
                        const unsigned long long hw = ld_word(sl);
                        const unsigned long long w0 = ld_word(sl + 1 + j * NW);
                        const unsigned long long w1 = NW == 2 ? ld_word(sl + 1 + j * NW + 1) : w0;
                        if ((unsigned)(hw >> 32) != tag) { ok = false; continue; }
                        const int p = (int)(unsigned)hw;
                        if (p < 0) continue;                        // no candidate in that workgroup
                        if ((unsigned)(w0 >> 32) != tag || (unsigned)(w1 >> 32) != tag) { ok = false; continue; }
                        T x = fabs(value_of<T>((unsigned)w0, (unsigned)w1));
                        if (x != x) x = (p == diag) ? T(INFINITY) : T(-1);
                        if (x >= T(0) && (x > cv || (x == cv && p < cp))) { cv = x; cp = p; cg = q; }
                    }
                    if (__ballot(!ok) == 0ull) break;
                    if (ld_agent(w.err) != 0) break;
                    if ((long long)(__builtin_amdgcn_s_memrealtime() - t0) > kDenseWaitTicks) {
                        if (ln == 0) __hip_atomic_store(w.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                const int mycp = cp;
                argmax_t(cv, cp);
                const unsigned long long hit = __ballot(mycp == cp && cp != INT_MAX);
                int gw = 0;
                if (hit != 0ull) gw = __shfl(cg, __ffsll((long long)hit) - 1, kWave);
                // only a timed-out wait can leave no candidate; keep the
                // position in range then
                int pp = cp;
                if (pp < diag || pp >= m) pp = diag;
                // the winning row, from the current column on
                const unsigned long long* wsl = words + ((size_t)par * gmax + gw) * SW;
                const unsigned long long t1 = __builtin_amdgcn_s_memrealtime();
                for (;;) {
                    bool ok = true;
                    for (int c = j + ln; c < PW; c += kWave) {
                        const unsigned long long w0 = ld_word(wsl + 1 + c * NW);
                        const unsigned long long w1 = NW == 2 ? ld_word(wsl + 1 + c * NW + 1) : w0;
                        const bool good = hit != 0ull && (unsigned)(w0 >> 32) == tag && (unsigned)(w1 >> 32) == tag;
                        ok = ok && (good || hit == 0ull);
                        s_u[par][c] = good ? value_of<T>((unsigned)w0, (unsigned)w1) : T(0);
                    }
                    if (__ballot(!ok) == 0ull) break;
                    if (ld_agent(w.err) != 0) break;
                    if ((long long)(__builtin_amdgcn_s_memrealtime() - t1) > kDenseWaitTicks) {
                        if (ln == 0) __hip_atomic_store(w.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                if (ln == 0) s_pp[par] = pp;
            }
            __syncthreads();
            // 3. interchange bookkeeping, scaling and rank-1 update
            const int pp = s_pp[par];
            const T piv = s_u[par][j];
            if (g == 0 && t == 0) {
                ipiv[diag] = pp + 1;
                if (piv == T(0)) atomicCAS(info, 0, diag + 1);
            }
            const bool win = active && pos == pp;
            if (active && pos == diag) pos = pp;
            if (win) { pos = diag; active = false; }
            if (active && piv != T(0)) {
                T l;
                if (fabs(piv) >= DLim<T>::sfmin) l = a[j] * (T(1) / piv);
                else l = a[j] / piv;
                a[j] = l;
#pragma unroll
                for (int c = j + 1; c < PW; ++c) a[c] = fma(-l, s_u[par][c], a[c]);
            }
        }
    });
    // rows go to their final positions (all reads of the panel happened
    // before the first wait, every write happens after the last one)
    if (valid) {
#pragma unroll
        for (int c = 0; c < PW; ++c)
            if (c < jb) A[pos + (size_t)(k + c) * lda] = a[c];
    }
}

// ---------------------------------------------------------------------------
// Row pairs of the interchanges ipiv[s .. s+cnt) (LAPACK, 1-based, applied in
// order): destination d receives the row found by tracing d back through
// the interchanges. Destinations are the rows s .. s+cnt-1 and every
// interchange partner below them (counted once). cnt <= kDenseMaxNB, one
// thread per candidate destination (2 * kDenseMaxNB threads).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(2 * kDenseMaxNB)
dense_pivmap_kernel(const int* __restrict__ ipiv, int s, int cnt, int* __restrict__ pairs, int* __restrict__ npairs,
                    int maxp)
{
    __shared__ int piv[kDenseMaxNB];
    __shared__ int count;
    const int t = threadIdx.x;
    if (t == 0) count = 0;
    for (int i = t; i < cnt; i += blockDim.x) piv[i] = ipiv[s + i] - 1;
    __syncthreads();
    int d = -1;
    if (t < cnt) {
        d = s + t;
    } else if (t < 2 * cnt) {
        const int i = t - cnt;
        const int p = piv[i];
        if (p >= s + cnt) {
            bool first = true;
            for (int q = 0; q < i; ++q)
                if (piv[q] == p) { first = false; break; }
            if (first) d = p;
        }
    }
    if (d >= 0) {
        int row = d;
        for (int i = cnt - 1; i >= 0; --i) {
            const int p = piv[i];
            if (row == s + i) row = p;
            else if (row == p) row = s + i;
        }
        if (row != d) {
            const int idx = atomicAdd(&count, 1);
            if (idx < maxp) {
                pairs[idx] = row;
                pairs[maxp + idx] = d;
            }
        }
    }
    __syncthreads();
    if (t == 0) *npairs = min(count, maxp);
}

// Row moves of one block in the columns [c0, c1) and [c2, c3): one wave per
// column reads every source row into registers; after the workgroup barrier
// (which waits for every load) it writes every destination row.
template <typename T, int MAXPL>
__global__ void __launch_bounds__(256)
dense_laswp_kernel(int m, T* __restrict__ A, int lda, const int* __restrict__ pairs, const int* __restrict__ npairs,
                   int maxp, int c0, int c1, int c2, int c3)
{
    const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
    const int q = blockIdx.x * (256 / kWave) + wv;
    const int n1 = c1 - c0;
    const int ncol = n1 + (c3 - c2);
    const bool live = q < ncol;
    const int col = !live ? 0 : (q < n1 ? c0 + q : c2 + (q - n1));
    const int np = *npairs;
    T* Ac = A + (size_t)col * lda;
    T v[MAXPL];
    int dst[MAXPL];
#pragma unroll
    for (int i = 0; i < MAXPL; ++i) {
        const int e = ln + kWave * i;
        dst[i] = -1;
        v[i] = T(0);
        if (live && e < np) {
            const int s = pairs[e], d = pairs[maxp + e];
            if (s >= 0 && s < m && d >= 0 && d < m) { v[i] = Ac[s]; dst[i] = d; }
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXPL; ++i)
        if (dst[i] >= 0) Ac[dst[i]] = v[i];
}

// ---------------------------------------------------------------------------
// Inverses of diagonal blocks. Block b (blockIdx.y) starts at row and column
// k0 + b*bs and has order nb = min(bs, n - k0 - b*bs); its inverse goes to
// X + b*bs*bs (column-major, ld bs). UPPER = false: unit lower triangle;
// UPPER = true: upper triangle with its diagonal. One wave per column of the
// inverse (forward or back substitution on a unit vector); the triangle the
// workgroup needs is staged in LDS, packed by columns. bs <= 256.
// ---------------------------------------------------------------------------
__host__ __device__ inline size_t trtri_lds_elems(int bs) { return (size_t)bs * (bs + 1) / 2; }

template <typename T, bool UPPER>
__global__ void __launch_bounds__(256)
dense_trtri_kernel(int n, int k0, int bs, const T* __restrict__ A, int lda, T* __restrict__ X)
{
    extern __shared__ __align__(16) unsigned char tr_raw[];
    T* S = reinterpret_cast<T*>(tr_raw);
    const int b = blockIdx.y;
    const int o = k0 + b * bs;
    const int nb = min(bs, n - o);
    const int j0 = blockIdx.x * 4;
    if (nb <= 0 || j0 >= nb) return;
    const T* D = A + o + (size_t)o * lda;
    const int t = threadIdx.x, wv = t / kWave, ln = t % kWave;
    // lower: columns p >= j0, rows p .. nb-1, column p at offset
    //   sum_{q=j0}^{p-1} (nb - q);  upper: columns p < min(nb, j0+4), rows 0..p at p(p+1)/2
    auto loff = [&](int p) -> int { return (p - j0) * nb - (p * (p - 1) / 2 - j0 * (j0 - 1) / 2); };
    if (!UPPER) {
        for (int p = j0 + wv; p < nb; p += 4) {
            T* dstc = S + loff(p) - p;
            for (int rr = p + ln; rr < nb; rr += kWave) dstc[rr] = D[rr + (size_t)p * lda];
        }
    } else {
        const int pe = min(nb, j0 + 4);
        for (int p = wv; p < pe; p += 4) {
            T* dstc = S + p * (p + 1) / 2;
            for (int rr = ln; rr <= p; rr += kWave) dstc[rr] = D[rr + (size_t)p * lda];
        }
    }
    __syncthreads();
    const int j = j0 + wv;
    if (j >= nb) return;
    T x[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) x[s] = (ln + kWave * s == j) ? T(1) : T(0);
    if (!UPPER) {
        for (int p = j; p < nb - 1; ++p) {
            const int sp = p >> 6;
            T xs = x[0];
            if (sp == 1) xs = x[1];
            else if (sp == 2) xs = x[2];
            else if (sp == 3) xs = x[3];
            const T xp = readlane_t(xs, p & 63);
            const T* col = S + loff(p) - p;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int rr = ln + kWave * s;
                if (rr > p && rr < nb) x[s] = fma(-col[rr], xp, x[s]);
            }
        }
    } else {
        for (int p = j; p >= 0; --p) {
            const int sp = p >> 6;
            T xs = x[0];
            if (sp == 1) xs = x[1];
            else if (sp == 2) xs = x[2];
            else if (sp == 3) xs = x[3];
            const T* col = S + p * (p + 1) / 2;
            const T xp = readlane_t(xs, p & 63) / col[p];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int rr = ln + kWave * s;
                if (rr == p) x[s] = xp;
                else if (rr < p) x[s] = fma(-col[rr], xp, x[s]);
            }
        }
    }
    T* Xo = X + (size_t)b * bs * bs + (size_t)j * bs;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int rr = ln + kWave * s;
        if (rr < nb) Xo[rr] = x[s];
    }
}

// ---------------------------------------------------------------------------
// C := A B (SET) or C := C - A B, all column-major, 128 x 128 tiles per
// workgroup (4 waves of 64 x 64 = 4 x 4 MFMA tiles). The product is formed
// transposed (B^T as the first MFMA operand, A^T as the second), so each
// accumulator register holds 16 consecutive rows of one column of C and the
// stores are whole 64-byte (FP32) or 128-byte (FP64) column segments.
// Workgroup ids are spread so that each XCD works on a contiguous run of
// tiles (ids go round-robin over the 8 XCDs).
//   MFMA operand layouts (lane l):  16x16x4 f64/f32: P[l%16][l/16], Q[l/16][l%16]
//                                   16x16x32 bf16/f16: P[l%16][8(l/16)+v], Q[8(l/16)+v][l%16]
//   accumulators: f64 D[4r + l/16][l%16], f32 D[4(l/16) + r][l%16]
// ---------------------------------------------------------------------------
constexpr int kGBM = 128, kGBN = 128;

__device__ __forceinline__ int gemm_tile_id(int tiles)
{
    const int pid = blockIdx.x;
    const int G = gridDim.x;
    if (G >= 8 && (G & 7) == 0) {
        const int per = G >> 3;
        return (pid & 7) * per + (pid >> 3);
    }
    return pid;
}

template <typename T, int MODE, bool SET>
__global__ void __launch_bounds__(256, 2)
dense_gemm_kernel(int M, int N, int K, const T* __restrict__ A, int lda, const T* B, int ldb,
                  T* C, int ldc, int tiles_m, int tiles_n)
{
    constexpr bool kWide = (MODE == kUpdBF16 || MODE == kUpdFP16 || MODE == kUpdBF16x3);
    const int tid = gemm_tile_id(tiles_m * tiles_n);
    if (tid >= tiles_m * tiles_n) return;
    const int m0 = (tid % tiles_m) * kGBM;
    const int n0 = (tid / tiles_m) * kGBN;
    const int t = threadIdx.x;
    const int w = t / kWave, l = t % kWave;
    const int wr = (w & 1) * 64, wc = (w >> 1) * 64;
    const int lr = l % 16, lq = l / 16;

    if constexpr (!kWide) {
        constexpr int BK = 16;
        constexpr int LDP = kGBM + 4;
        __shared__ T sA[2][BK][LDP];
        __shared__ T sB[2][BK][LDP];
        using acc_t = typename std::conditional<std::is_same<T, double>::value, d4_t, f4_t>::type;
        acc_t acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
        T ra[8], rb[8];
        const int ai = t & 127, ak = t >> 7;      // A tile: row ai, k = ak + 2q
        const int bk = t & 15, bj = t >> 4;       // B tile: k = bk, column bj + 16q
        auto gload = [&](int k0) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int kk = ak + 2 * q;
                ra[q] = (m0 + ai < M && k0 + kk < K) ? A[(size_t)(m0 + ai) + (size_t)(k0 + kk) * lda] : T(0);
                const int jj = bj + 16 * q;
                rb[q] = (k0 + bk < K && n0 + jj < N) ? B[(size_t)(k0 + bk) + (size_t)(n0 + jj) * ldb] : T(0);
            }
        };
        auto lstore = [&](int buf) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                sA[buf][ak + 2 * q][ai] = ra[q];
                sB[buf][bk][bj + 16 * q] = rb[q];
            }
        };
        const int nst = (K + BK - 1) / BK;
        gload(0);
        lstore(0);
        __syncthreads();
        for (int s = 0; s < nst; ++s) {
            const int buf = s & 1;
            if (s + 1 < nst) gload((s + 1) * BK);
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4) {
                T pf[4], qf[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    pf[x] = sB[buf][kk + lq][wc + 16 * x + lr];
                    qf[x] = sA[buf][kk + lq][wr + 16 * x + lr];
                }
#pragma unroll
                for (int tj = 0; tj < 4; ++tj)
#pragma unroll
                    for (int ti = 0; ti < 4; ++ti) {
                        if constexpr (std::is_same<T, double>::value)
                            acc[tj][ti] = __builtin_amdgcn_mfma_f64_16x16x4f64(pf[tj], qf[ti], acc[tj][ti], 0, 0, 0);
                        else
                            acc[tj][ti] = __builtin_amdgcn_mfma_f32_16x16x4f32(pf[tj], qf[ti], acc[tj][ti], 0, 0, 0);
                    }
            }
            if (s + 1 < nst) lstore(buf ^ 1);
            __syncthreads();
        }
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
#pragma unroll
            for (int ti = 0; ti < 4; ++ti) {
                const int row = m0 + wr + 16 * ti + lr;
                if (row >= M) continue;
#pragma unroll
                for (int rg = 0; rg < 4; ++rg) {
                    const int cc = std::is_same<T, double>::value ? (4 * rg + lq) : (4 * lq + rg);
                    const int col = n0 + wc + 16 * tj + cc;
                    if (col < N) {
                        T* cp = C + (size_t)row + (size_t)col * ldc;
                        if constexpr (SET) *cp = acc[tj][ti][rg];
                        else *cp = *cp - acc[tj][ti][rg];
                    }
                }
            }
    } else {
        // FP32 storage, 16-bit MFMA inputs: tiles staged as 16-bit values
        // with k contiguous (one 16-byte LDS read per operand fragment)
        constexpr int BK = 64;
        constexpr int LDK = BK + 8;
        using h_t = typename std::conditional<MODE == kUpdFP16, _Float16, __bf16>::type;
        using v8_t = typename std::conditional<MODE == kUpdFP16, h8_t, bf8_t>::type;
        constexpr int kSplit = (MODE == kUpdBF16x3) ? 2 : 1;   // hi (and lo) parts
        __shared__ __align__(16) h_t sA[kSplit][2][kGBM][LDK];
        __shared__ __align__(16) h_t sB[kSplit][2][kGBN][LDK];
        f4_t acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = f4_t{0.f, 0.f, 0.f, 0.f};
        // A tile: thread owns row ai and k-groups ag + 2q (8 consecutive k each)
        const int ai = t & 127, ag = t >> 7;
        // B tile: 4-vectors e = t + 256 q: column e/16, k = 4 (e%16) .. +3
        float ra[4][8], rb[8][4];
        auto gload = [&](int k0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int kg = ag + 2 * q;
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    const int kk = k0 + 8 * kg + v;
                    ra[q][v] = (m0 + ai < M && kk < K) ? (float)A[(size_t)(m0 + ai) + (size_t)kk * lda] : 0.f;
                }
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = t + 256 * q;
                const int jj = e >> 4, kv = e & 15;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int kk = k0 + 4 * kv + v;
                    rb[q][v] = (n0 + jj < N && kk < K) ? (float)B[(size_t)kk + (size_t)(n0 + jj) * ldb] : 0.f;
                }
            }
        };
        typedef h_t v4h_t __attribute__((ext_vector_type(4)));
        auto lstore = [&](int buf) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int kg = ag + 2 * q;
                v8_t hi, lo;
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    hi[v] = (h_t)ra[q][v];
                    lo[v] = (h_t)(ra[q][v] - (float)hi[v]);
                }
                *reinterpret_cast<v8_t*>(&sA[0][buf][ai][8 * kg]) = hi;
                if constexpr (kSplit == 2) *reinterpret_cast<v8_t*>(&sA[kSplit - 1][buf][ai][8 * kg]) = lo;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = t + 256 * q;
                const int jj = e >> 4, kv = e & 15;
                v4h_t hi, lo;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    hi[v] = (h_t)rb[q][v];
                    lo[v] = (h_t)(rb[q][v] - (float)hi[v]);
                }
                *reinterpret_cast<v4h_t*>(&sB[0][buf][jj][4 * kv]) = hi;
                if constexpr (kSplit == 2) *reinterpret_cast<v4h_t*>(&sB[kSplit - 1][buf][jj][4 * kv]) = lo;
            }
        };
        const int nst = (K + BK - 1) / BK;
        gload(0);
        lstore(0);
        __syncthreads();
        for (int s = 0; s < nst; ++s) {
            const int buf = s & 1;
            if (s + 1 < nst) gload((s + 1) * BK);
#pragma unroll
            for (int kb = 0; kb < BK; kb += 32) {
                v8_t pf[4], qf[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    pf[x] = *reinterpret_cast<const v8_t*>(&sB[0][buf][wc + 16 * x + lr][kb + 8 * lq]);
                    qf[x] = *reinterpret_cast<const v8_t*>(&sA[0][buf][wr + 16 * x + lr][kb + 8 * lq]);
                }
                if constexpr (MODE == kUpdBF16x3) {
                    v8_t pl[4], ql[4];
#pragma unroll
                    for (int x = 0; x < 4; ++x) {
                        pl[x] = *reinterpret_cast<const v8_t*>(&sB[1][buf][wc + 16 * x + lr][kb + 8 * lq]);
                        ql[x] = *reinterpret_cast<const v8_t*>(&sA[1][buf][wr + 16 * x + lr][kb + 8 * lq]);
                    }
#pragma unroll
                    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
                        for (int ti = 0; ti < 4; ++ti) {
                            acc[tj][ti] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pl[tj], qf[ti], acc[tj][ti], 0, 0, 0);
                            acc[tj][ti] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pf[tj], ql[ti], acc[tj][ti], 0, 0, 0);
                        }
                }
#pragma unroll
                for (int tj = 0; tj < 4; ++tj)
#pragma unroll
                    for (int ti = 0; ti < 4; ++ti) {
                        if constexpr (MODE == kUpdFP16)
                            acc[tj][ti] = __builtin_amdgcn_mfma_f32_16x16x32_f16(pf[tj], qf[ti], acc[tj][ti], 0, 0, 0);
                        else
                            acc[tj][ti] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pf[tj], qf[ti], acc[tj][ti], 0, 0, 0);
                    }
            }
            if (s + 1 < nst) lstore(buf ^ 1);
            __syncthreads();
        }
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
#pragma unroll
            for (int ti = 0; ti < 4; ++ti) {
                const int row = m0 + wr + 16 * ti + lr;
                if (row >= M) continue;
#pragma unroll
                for (int rg = 0; rg < 4; ++rg) {
                    const int col = n0 + wc + 16 * tj + 4 * lq + rg;
                    if (col < N) {
                        T* cp = C + (size_t)row + (size_t)col * ldc;
                        if constexpr (SET) *cp = acc[tj][ti][rg];
                        else *cp = *cp - acc[tj][ti][rg];
                    }
                }
            }
    }
}

// ---------------------------------------------------------------------------
// 16-bit operands of the trailing update, converted once per update:
//   Ah[i * ldk + k] = h(A21[i, k])   (rows with k contiguous, via an LDS tile)
//   Bh[j * ldk + k] = h(U12[k, j])   (columns with k contiguous, as stored)
// for k < ldk (K rounded up to 64; the padding is zero). With a split
// (BF16x3) the second array of each holds the remainders lo = h(x - hi).
// ---------------------------------------------------------------------------
// Rounding to the update format and back: BF16 / FP16 by the conversions of
// the types; FP8 (OCP e4m3, one byte) by v_cvt_pk_fp8_f32 (round to nearest
// even, no scaling: multipliers are at most 1 in magnitude, and a U12 entry
// beyond 448 becomes NaN, which the non-finite-factor check reports).
template <typename h_t> struct HConv {
    __device__ static h_t to(float x) { return (h_t)x; }
    __device__ static float back(h_t v) { return (float)v; }
};
template <> struct HConv<unsigned char> {
    __device__ static unsigned char to(float x)
    {
        return (unsigned char)(__builtin_amdgcn_cvt_pk_fp8_f32(x, 0.f, 0, false) & 0xff);
    }
    __device__ static float back(unsigned char v) { return __builtin_amdgcn_cvt_f32_fp8((int)v, 0); }
};

template <typename h_t>
__global__ void __launch_bounds__(256)
dense_cvt_a_kernel(int M, int K, int ldk, const float* __restrict__ A, int lda, h_t* __restrict__ hi,
                   h_t* __restrict__ lo)
{
    __shared__ float tile[64][65];
    const int i0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
    const int t = threadIdx.x;
    for (int e = t; e < 64 * 64; e += 256) {
        const int i = e % 64, kk = e / 64;
        tile[kk][i] = (i0 + i < M && k0 + kk < K) ? A[(size_t)(i0 + i) + (size_t)(k0 + kk) * lda] : 0.f;
    }
    __syncthreads();
    typedef h_t v8h __attribute__((ext_vector_type(8)));
    for (int e = t; e < 64 * 8; e += 256) {
        const int i = e / 8, g = e % 8;
        if (i0 + i >= M) continue;
        v8h vh, vl;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            const float x = tile[8 * g + v][i];
            vh[v] = HConv<h_t>::to(x);
            vl[v] = HConv<h_t>::to(x - HConv<h_t>::back(vh[v]));
        }
        *reinterpret_cast<v8h*>(hi + (size_t)(i0 + i) * ldk + k0 + 8 * g) = vh;
        if (lo) *reinterpret_cast<v8h*>(lo + (size_t)(i0 + i) * ldk + k0 + 8 * g) = vl;
    }
}

template <typename h_t>
__global__ void __launch_bounds__(256)
dense_cvt_b_kernel(int N, int K, int ldk, const float* __restrict__ B, int ldb, h_t* __restrict__ hi,
                   h_t* __restrict__ lo)
{
    const int e = blockIdx.x * 256 + threadIdx.x;          // (column, group of 8 k)
    const int groups = ldk / 8;
    const int j = e / groups, g = e % groups;
    if (j >= N) return;
    typedef h_t v8h __attribute__((ext_vector_type(8)));
    v8h vh, vl;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const int kk = 8 * g + v;
        const float x = kk < K ? B[(size_t)kk + (size_t)j * ldb] : 0.f;
        vh[v] = HConv<h_t>::to(x);
        vl[v] = HConv<h_t>::to(x - HConv<h_t>::back(vh[v]));
    }
    *reinterpret_cast<v8h*>(hi + (size_t)j * ldk + 8 * g) = vh;
    if (lo) *reinterpret_cast<v8h*>(lo + (size_t)j * ldk + 8 * g) = vl;
}

// C := C - A B on 128 x 128 tiles from the converted operands (Ah rows and
// Bh columns with k contiguous, ld ldk, zero padded to a multiple of 64):
// each stage moves 16-byte vectors straight into the LDS tiles.
template <int MODE>
__global__ void __launch_bounds__(256, 2)
dense_gemm_h_kernel(int M, int N, int ldk, const void* __restrict__ Ahv, const void* __restrict__ Alv,
                    const void* __restrict__ Bhv, const void* __restrict__ Blv, float* C, int ldc, int tiles_m,
                    int tiles_n)
{
    // element type and the 8 elements one lane feeds an MFMA (16 bytes, or
    // 8 for FP8)
    using h_t = typename std::conditional<
        MODE == kUpdFP16, _Float16,
        typename std::conditional<MODE == kUpdFP8, unsigned char, __bf16>::type>::type;
    using v8_t = typename std::conditional<
        MODE == kUpdFP16, h8_t, typename std::conditional<MODE == kUpdFP8, long, bf8_t>::type>::type;
    constexpr int kSplit = (MODE == kUpdBF16x3) ? 2 : 1;
    constexpr int BK = 64;
    constexpr int LDK = BK + 8;
    const h_t* Ah = static_cast<const h_t*>(Ahv);
    const h_t* Bh = static_cast<const h_t*>(Bhv);
    const h_t* Al = static_cast<const h_t*>(Alv);
    const h_t* Bl = static_cast<const h_t*>(Blv);
    const int tid = gemm_tile_id(tiles_m * tiles_n);
    if (tid >= tiles_m * tiles_n) return;
    const int m0 = (tid % tiles_m) * kGBM;
    const int n0 = (tid / tiles_m) * kGBN;
    const int t = threadIdx.x;
    const int w = t / kWave, l = t % kWave;
    const int wr = (w & 1) * 64, wc = (w >> 1) * 64;
    const int lr = l % 16, lq = l / 16;
    __shared__ __align__(16) h_t sA[kSplit][2][kGBM][LDK];
    __shared__ __align__(16) h_t sB[kSplit][2][kGBN][LDK];
    f4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f4_t{0.f, 0.f, 0.f, 0.f};
    // 128 rows x 8 groups of 8 k per operand tile: 4 vectors per thread
    v8_t ra[kSplit][4], rb[kSplit][4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            const int r = e / 8, g = e % 8;
            const bool va = m0 + r < M, vb = n0 + r < N;
            const size_t oa = (size_t)(m0 + r) * ldk + k0 + 8 * g;
            const size_t ob = (size_t)(n0 + r) * ldk + k0 + 8 * g;
            ra[0][q] = va ? *reinterpret_cast<const v8_t*>(Ah + oa) : v8_t{};
            rb[0][q] = vb ? *reinterpret_cast<const v8_t*>(Bh + ob) : v8_t{};
            if constexpr (kSplit == 2) {
                ra[1][q] = va ? *reinterpret_cast<const v8_t*>(Al + oa) : v8_t{};
                rb[1][q] = vb ? *reinterpret_cast<const v8_t*>(Bl + ob) : v8_t{};
            }
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            const int r = e / 8, g = e % 8;
#pragma unroll
            for (int h = 0; h < kSplit; ++h) {
                *reinterpret_cast<v8_t*>(&sA[h][buf][r][8 * g]) = ra[h][q];
                *reinterpret_cast<v8_t*>(&sB[h][buf][r][8 * g]) = rb[h][q];
            }
        }
    };
    const int nst = ldk / BK;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int s = 0; s < nst; ++s) {
        const int buf = s & 1;
        if (s + 1 < nst) gload((s + 1) * BK);
#pragma unroll
        for (int kb = 0; kb < BK; kb += 32) {
            v8_t pf[4], qf[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                pf[x] = *reinterpret_cast<const v8_t*>(&sB[0][buf][wc + 16 * x + lr][kb + 8 * lq]);
                qf[x] = *reinterpret_cast<const v8_t*>(&sA[0][buf][wr + 16 * x + lr][kb + 8 * lq]);
            }
            if constexpr (MODE == kUpdBF16x3) {
                v8_t pl[4], ql[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    pl[x] = *reinterpret_cast<const v8_t*>(&sB[1][buf][wc + 16 * x + lr][kb + 8 * lq]);
                    ql[x] = *reinterpret_cast<const v8_t*>(&sA[1][buf][wr + 16 * x + lr][kb + 8 * lq]);
                }
#pragma unroll
                for (int tj = 0; tj < 4; ++tj)
#pragma unroll
                    for (int ti = 0; ti < 4; ++ti) {
                        acc[tj][ti] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pl[tj], qf[ti], acc[tj][ti], 0, 0, 0);
                        acc[tj][ti] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pf[tj], ql[ti], acc[tj][ti], 0, 0, 0);
                    }
            }
#pragma unroll
            for (int tj = 0; tj < 4; ++tj)
#pragma unroll
                for (int ti = 0; ti < 4; ++ti) {
                    if constexpr (MODE == kUpdFP16)
                        acc[tj][ti] = __builtin_amdgcn_mfma_f32_16x16x32_f16(pf[tj], qf[ti], acc[tj][ti], 0, 0, 0);
                    else if constexpr (MODE == kUpdFP8)
                        acc[tj][ti] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(pf[tj], qf[ti], acc[tj][ti], 0, 0, 0);
                    else
                        acc[tj][ti] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pf[tj], qf[ti], acc[tj][ti], 0, 0, 0);
                }
        }
        if (s + 1 < nst) lstore(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
        for (int ti = 0; ti < 4; ++ti) {
            const int row = m0 + wr + 16 * ti + lr;
            if (row >= M) continue;
#pragma unroll
            for (int rg = 0; rg < 4; ++rg) {
                const int col = n0 + wc + 16 * tj + 4 * lq + rg;
                if (col < N) {
                    float* cp = C + (size_t)row + (size_t)col * ldc;
                    *cp = *cp - acc[tj][ti][rg];
                }
            }
        }
}

template <typename T>
__global__ void dense_copy_kernel(int rows, int cols, const T* __restrict__ S, int lds, T* __restrict__ D, int ldd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < rows && j < cols) D[(size_t)i + (size_t)j * ldd] = S[(size_t)i + (size_t)j * lds];
}

__global__ void dense_fill_int_kernel(int* p, int v, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
int device_cu_count()
{
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 1;
    return v;
}

int dense_panel_gmax(int m) { return (int)ceildiv(std::max(m, 1), kDensePanelRows); }

static size_t a256(size_t b) { return (b + 255) & ~size_t(255); }

size_t dense_work_bytes(int m, int n, int elem, int gmax)
{
    const int PW = elem == 8 ? DenseShape<double>::PW : DenseShape<float>::PW;
    const int NB = kDenseMaxNB;
    size_t b = 0;
    b += a256(sizeof(unsigned long long) * 2 * gmax * (1 + PW * elem / 4));   // slot words
    b += a256(sizeof(int) * 4);                         // err
    b += a256(sizeof(int) * 2 * 2 * NB);                // pairs
    b += a256(sizeof(int) * 4);                         // npairs
    for (int lane = 0; lane < 2; ++lane) {                    // main stream, then the look-ahead's
        if (lane == 1) {
            b += a256(sizeof(int) * 2 * 2 * NB);            // pairs
            b += a256(sizeof(int) * 4);                     // npairs
        }
        b += a256((size_t)elem * kDenseStrip * kDenseStrip);   // X
        if (elem == 4) {                                          // Ah, Bh (hi and lo halves)
            const size_t h = (size_t)std::max(std::max(m, n), 1) * kDenseMaxNB;
            b += 2 * a256(2 * 2 * h);
        }
    }
    return b;
}

void dense_work_carve(void* base, int m, int n, int elem, int gmax, DenseWork& w)
{
    const int PW = elem == 8 ? DenseShape<double>::PW : DenseShape<float>::PW;
    const int NB = kDenseMaxNB;
    char* p = static_cast<char*>(base);
    w.crow = p;                          p += a256(sizeof(unsigned long long) * 2 * gmax * (1 + PW * elem / 4));
    w.nwords = 2 * gmax * (1 + PW * elem / 4);
    w.err = reinterpret_cast<int*>(p);   p += a256(sizeof(int) * 4);
    w.pairs = reinterpret_cast<int*>(p); p += a256(sizeof(int) * 2 * 2 * NB);
    w.npairs = reinterpret_cast<int*>(p); p += a256(sizeof(int) * 4);
    w.X = p;                             p += a256((size_t)elem * kDenseStrip * kDenseStrip);
    w.Ah = w.Bh = w.Ah2 = w.Bh2 = nullptr;
    w.h_elems = 0;
    const size_t h = (size_t)std::max(std::max(m, n), 1) * kDenseMaxNB;
    if (elem == 4) {
        w.Ah = p;                        p += a256(2 * 2 * h);
        w.Bh = p;                        p += a256(2 * 2 * h);
        w.h_elems = h;
    }
    w.pairs2 = reinterpret_cast<int*>(p);  p += a256(sizeof(int) * 2 * 2 * NB);
    w.npairs2 = reinterpret_cast<int*>(p); p += a256(sizeof(int) * 4);
    w.X2 = p;                              p += a256((size_t)elem * kDenseStrip * kDenseStrip);
    if (elem == 4) {
        w.Ah2 = p;                       p += a256(2 * 2 * h);
        w.Bh2 = p;
    }
    w.gmax = gmax;
}

template <typename T>
magma_int_t dense_trtri_launch(bool upper, int n, int k0, int bs, int nblk, const T* A, int lda, T* X, hipStream_t st)
{
    const size_t lds = sizeof(T) * trtri_lds_elems(bs);
    dim3 grid((unsigned)ceildiv(bs, 4), (unsigned)nblk);
    if (upper) {
        MGI_HIP_CHECK(hipFuncSetAttribute((const void*)dense_trtri_kernel<T, true>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        dense_trtri_kernel<T, true><<<grid, 256, lds, st>>>(n, k0, bs, A, lda, X);
    } else {
        MGI_HIP_CHECK(hipFuncSetAttribute((const void*)dense_trtri_kernel<T, false>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        dense_trtri_kernel<T, false><<<grid, 256, lds, st>>>(n, k0, bs, A, lda, X);
    }
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

template magma_int_t dense_trtri_launch<double>(bool, int, int, int, int, const double*, int, double*, hipStream_t);
template magma_int_t dense_trtri_launch<float>(bool, int, int, int, int, const float*, int, float*, hipStream_t);

namespace {

template <typename T, int MODE, bool SET>
magma_int_t gemm_launch(int M, int N, int K, const T* A, int lda, const T* B, int ldb, T* C, int ldc,
                        hipStream_t st)
{
    if (M <= 0 || N <= 0) return MAGMA_SUCCESS;
    const int tm = (int)ceildiv(M, kGBM), tn = (int)ceildiv(N, kGBN);
    const int total = tm * tn;
    const int grid = total >= 8 ? (int)roundup(total, 8) : total;
    dense_gemm_kernel<T, MODE, SET><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, tm, tn);
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

template <int MODE>
magma_int_t update_h_launch(int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
                            const DenseWork& w, hipStream_t st)
{
    using h_t = typename std::conditional<
        MODE == kUpdFP16, _Float16,
        typename std::conditional<MODE == kUpdFP8, unsigned char, __bf16>::type>::type;
    if (M <= 0 || N <= 0) return MAGMA_SUCCESS;
    const int ldk = (int)roundup(K, 64);
    if (ldk > kDenseMaxNB || (size_t)M * ldk > w.h_elems || (size_t)N * ldk > w.h_elems) return MAGMA_ERR_NOT_SUPPORTED;
    constexpr bool split = MODE == kUpdBF16x3;
    h_t* Ah = static_cast<h_t*>(w.Ah);
    h_t* Bh = static_cast<h_t*>(w.Bh);
    h_t* Al = split ? Ah + w.h_elems : nullptr;
    h_t* Bl = split ? Bh + w.h_elems : nullptr;
    dense_cvt_a_kernel<h_t><<<dim3((unsigned)ceildiv(M, 64), (unsigned)(ldk / 64)), 256, 0, st>>>(M, K, ldk, A, lda,
                                                                                                  Ah, Al);
    dense_cvt_b_kernel<h_t><<<(unsigned)ceildiv((long)N * (ldk / 8), 256L), 256, 0, st>>>(N, K, ldk, B, ldb, Bh, Bl);
    MGI_LAUNCH_CHECK();
    const int tm = (int)ceildiv(M, kGBM), tn = (int)ceildiv(N, kGBN);
    const int total = tm * tn;
    const int grid = total >= 8 ? (int)roundup(total, 8) : total;
    dense_gemm_h_kernel<MODE><<<grid, 256, 0, st>>>(M, N, ldk, Ah, Al, Bh, Bl, C, ldc, tm, tn);
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

template <typename T>
magma_int_t update_launch(int mode, int M, int N, int K, const T* A, int lda, const T* B, int ldb, T* C, int ldc,
                          const DenseWork& w, hipStream_t st)
{
    if constexpr (std::is_same<T, double>::value) {
        (void)w;
        return gemm_launch<double, kUpdF64, false>(M, N, K, A, lda, B, ldb, C, ldc, st);
    } else {
        switch (mode) {
            case kUpdF32:    return gemm_launch<float, kUpdF32, false>(M, N, K, A, lda, B, ldb, C, ldc, st);
            case kUpdBF16:   return update_h_launch<kUpdBF16>(M, N, K, A, lda, B, ldb, C, ldc, w, st);
            case kUpdFP16:   return update_h_launch<kUpdFP16>(M, N, K, A, lda, B, ldb, C, ldc, w, st);
            case kUpdBF16x3: return update_h_launch<kUpdBF16x3>(M, N, K, A, lda, B, ldb, C, ldc, w, st);
            case kUpdFP8:    return update_h_launch<kUpdFP8>(M, N, K, A, lda, B, ldb, C, ldc, w, st);
            default:         return MAGMA_ERR_NOT_SUPPORTED;
        }
    }
}

template <typename T>
magma_int_t panel_launch(int m, int k, int jb, T* A, int lda, int* ipiv, int* info, const DenseWork& w,
                         hipStream_t st)
{
    constexpr int PW = DenseShape<T>::PW;
    const int G = (int)ceildiv(m - k, kDensePanelRows);
    if (G > w.gmax) return MAGMA_ERR_NOT_SUPPORTED;
    dense_panel_kernel<T, PW><<<G, kDensePanelRows, 0, st>>>(m, k, jb, A, lda, ipiv, info, w);
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

// Interchanges ipiv[s .. s+cnt) applied to the columns [c0, c1) and [c2, c3).
template <typename T>
magma_int_t swaps_launch(int m, const int* ipiv, int s, int cnt, T* A, int lda, const DenseWork& w, int c0, int c1,
                         int c2, int c3, hipStream_t st)
{
    constexpr int MP = 2 * kDenseMaxNB;
    const int ncol = (c1 - c0) + (c3 - c2);
    if (ncol <= 0 || cnt <= 0) return MAGMA_SUCCESS;
    if (cnt > kDenseMaxNB) return MAGMA_ERR_NOT_SUPPORTED;
    dense_pivmap_kernel<<<1, 2 * kDenseMaxNB, 0, st>>>(ipiv, s, cnt, w.pairs, w.npairs, MP);
    MGI_LAUNCH_CHECK();
    // at most 2 cnt pairs: the row-move kernel holds that many per wave
    if (cnt <= 128)
        dense_laswp_kernel<T, 4><<<(unsigned)ceildiv(ncol, 4), 256, 0, st>>>(m, A, lda, w.pairs, w.npairs, MP, c0, c1,
                                                                            c2, c3);
    else
        dense_laswp_kernel<T, MP / kWave><<<(unsigned)ceildiv(ncol, 4), 256, 0, st>>>(m, A, lda, w.pairs, w.npairs, MP,
                                                                                     c0, c1, c2, c3);
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

// U12 := L11^{-1} A[r0:r0+kw, cb:ce) for the unit lower L11 at (r0, r0),
// then A[r0+kw:m, cb:ce) -= A[r0+kw:m, r0:r0+kw) U12. The triangular solve
// runs in strips of up to 128 rows: strip s is multiplied by the inverse of
// its diagonal block (one GEMM row tile: each workgroup owns all rows of its
// columns and writes them after reading them, so the product runs in place),
// then removed from the strips below (a GEMM update). The triangular part
// runs in the storage precision (FP64, or FP32 as MAGMA's strsm); the
// trailing update in the requested format.
template <typename T>
magma_int_t trsm_update(int mode, int m, int r0, int kw, int cb, int ce, T* A, int lda, const DenseWork& w,
                        hipStream_t st)
{
    const int nc = ce - cb;
    if (nc <= 0 || kw <= 0) return MAGMA_SUCCESS;
    T* X = reinterpret_cast<T*>(w.X);
    constexpr int S = kDenseStrip;
    static_assert(S <= kGBM, "in-place strips need one GEMM row tile");
    constexpr int kTri = std::is_same<T, double>::value ? (int)kUpdF64 : (int)kUpdF32;
    magma_int_t rc = MAGMA_SUCCESS;
    for (int s0 = 0; s0 < kw; s0 += S) {
        const int ws = std::min(S, kw - s0);
        const int rs = r0 + s0;
        if ((rc = dense_trtri_launch<T>(false, rs + ws, rs, S, 1, A, lda, X, st)) != MAGMA_SUCCESS) return rc;
        T* Us = A + rs + (size_t)cb * lda;
        if constexpr (kTri == kUpdF64) rc = gemm_launch<double, kUpdF64, true>(ws, nc, ws, X, S, Us, lda, Us, lda, st);
        else rc = gemm_launch<float, kUpdF32, true>(ws, nc, ws, X, S, Us, lda, Us, lda, st);
        if (rc != MAGMA_SUCCESS) return rc;
        const int below = kw - s0 - ws;
        if (below > 0) {
            if constexpr (kTri == kUpdF64)
                rc = gemm_launch<double, kUpdF64, false>(below, nc, ws, A + rs + ws + (size_t)rs * lda, lda, Us, lda,
                                                         A + rs + ws + (size_t)cb * lda, lda, st);
            else
                rc = gemm_launch<float, kUpdF32, false>(below, nc, ws, A + rs + ws + (size_t)rs * lda, lda, Us, lda,
                                                        A + rs + ws + (size_t)cb * lda, lda, st);
            if (rc != MAGMA_SUCCESS) return rc;
        }
    }
    if (m - r0 - kw > 0)
        rc = update_launch<T>(mode, m - r0 - kw, nc, kw, A + (r0 + kw) + (size_t)r0 * lda, lda, A + r0 + (size_t)cb * lda,
                              lda, A + (r0 + kw) + (size_t)cb * lda, lda, w, st);
    return rc;
}

// Factors block [k, k+jb) in place: nb / PW panels; after panel i, its
// interchanges go to the block's other columns and its U rows and update to
// the block's columns on its right.
template <typename T>
magma_int_t factor_block(int mode, int m, int k, int jb, T* A, int lda, int* ipiv, int* info, const DenseWork& w,
                         hipStream_t st)
{
    constexpr int PW = DenseShape<T>::PW;
    magma_int_t rc;
    for (int i0 = 0; i0 < jb; i0 += PW) {
        const int wdt = std::min(PW, jb - i0);
        const int kp = k + i0;
        if ((rc = panel_launch<T>(m, kp, wdt, A, lda, ipiv, info, w, st)) != MAGMA_SUCCESS) return rc;
        const int right = kp + wdt;
        if (i0 > 0 || right < k + jb) {
            if ((rc = swaps_launch<T>(m, ipiv, kp, wdt, A, lda, w, k, kp, right, k + jb, st)) != MAGMA_SUCCESS)
                return rc;
        }
        if (right < k + jb) {
            if ((rc = trsm_update<T>(mode, m, kp, wdt, right, k + jb, A, lda, w, st)) != MAGMA_SUCCESS) return rc;
        }
    }
    return MAGMA_SUCCESS;
}

struct EventPair {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~EventPair()
    {
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

// Blocks of nb columns (a multiple of the panel width PW). After block k is
// factored, its interchanges go to every column outside it and its U12 and
// trailing update (K = nb) to the columns on its right.
//
// With a second stream (look-ahead of one block): the stream in order
// updates only the next block's columns and goes on to factor that block,
// while the second stream applies the interchanges to the columns on the
// left and updates every column beyond the next block. Before the first
// stream updates the block after next, it waits for the second stream's
// update of the previous step, which covered those columns. The two streams
// touch disjoint columns between the events, and each has its own row-move,
// inverse and 16-bit operand buffers.
template <typename T>
magma_int_t dense_getrf_t(int mode, int nb, int m, int n, T* A, int lda, int* ipiv, int* info, const DenseWork& w,
                          hipStream_t st, hipStream_t side)
{
    constexpr int PW = DenseShape<T>::PW;
    if (nb < PW || nb % PW != 0) return MAGMA_ERR_NOT_SUPPORTED;
    if (dense_panel_gmax(m) > w.gmax) return MAGMA_ERR_NOT_SUPPORTED;
    // slot words with an impossible tag: nothing of an earlier call is accepted
    dense_fill_int_kernel<<<(unsigned)ceildiv(2 * w.nwords, 256), 256, 0, st>>>(static_cast<int*>(w.crow), -1,
                                                                               2 * w.nwords);
    dense_fill_int_kernel<<<1, 32, 0, st>>>(w.err, 0, 1);
    dense_fill_int_kernel<<<1, 32, 0, st>>>(info, 0, 1);
    MGI_LAUNCH_CHECK();
    const int minmn = std::min(m, n);
    magma_int_t rc = MAGMA_SUCCESS;
    if (side == nullptr || minmn <= nb) {
        for (int k = 0; k < minmn; k += nb) {
            const int jb = std::min(nb, minmn - k);
            if ((rc = factor_block<T>(mode, m, k, jb, A, lda, ipiv, info, w, st)) != MAGMA_SUCCESS) return rc;
            if ((rc = swaps_launch<T>(m, ipiv, k, jb, A, lda, w, 0, k, k + jb, n, st)) != MAGMA_SUCCESS) return rc;
            if ((rc = trsm_update<T>(mode, m, k, jb, k + jb, n, A, lda, w, st)) != MAGMA_SUCCESS) return rc;
        }
        return MAGMA_SUCCESS;
    }
    const DenseWork w2 = w.side();
    EventPair ev;     // e[0]: the first stream's block is ready; e[1]: the second stream's update is done
    for (hipEvent_t& x : ev.e) MGI_HIP_CHECK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    // the second stream starts after what the caller queued on the first
    MGI_HIP_CHECK(hipEventRecord(ev.e[0], st));
    MGI_HIP_CHECK(hipStreamWaitEvent(side, ev.e[0], 0));
    bool side_pending = false;
    for (int k = 0; k < minmn; k += nb) {
        const int jb = std::min(nb, minmn - k);
        if ((rc = factor_block<T>(mode, m, k, jb, A, lda, ipiv, info, w, st)) != MAGMA_SUCCESS) return rc;
        const int c0 = k + jb;                       // next block [c0, c1), beyond it [c1, n)
        const int c1 = std::min(n, c0 + nb);
        // the next block's columns were last written by the second stream's
        // update of the previous step
        if (side_pending) MGI_HIP_CHECK(hipStreamWaitEvent(st, ev.e[1], 0));
        if (c1 > c0) {
            if ((rc = swaps_launch<T>(m, ipiv, k, jb, A, lda, w, c0, c1, 0, 0, st)) != MAGMA_SUCCESS) return rc;
            if ((rc = trsm_update<T>(mode, m, k, jb, c0, c1, A, lda, w, st)) != MAGMA_SUCCESS) return rc;
        }
        MGI_HIP_CHECK(hipEventRecord(ev.e[0], st));
        MGI_HIP_CHECK(hipStreamWaitEvent(side, ev.e[0], 0));
        if ((rc = swaps_launch<T>(m, ipiv, k, jb, A, lda, w2, 0, k, c1, n, side)) != MAGMA_SUCCESS) return rc;
        if (c1 < n && (rc = trsm_update<T>(mode, m, k, jb, c1, n, A, lda, w2, side)) != MAGMA_SUCCESS) return rc;
        MGI_HIP_CHECK(hipEventRecord(ev.e[1], side));
        side_pending = true;
    }
    MGI_HIP_CHECK(hipStreamWaitEvent(st, ev.e[1], 0));
    return MAGMA_SUCCESS;
}

}  // namespace

magma_int_t dense_getrf(int mode, int m, int n, double* dA, int ldda, int* d_ipiv, int* d_info, const DenseWork& w,
                        hipStream_t stream, int nb, hipStream_t side)
{
    (void)mode;
    return dense_getrf_t<double>(kUpdF64, nb > 0 ? nb : DenseShape<double>::NB, m, n, dA, ldda, d_ipiv, d_info, w,
                                 stream, side);
}

magma_int_t dense_sgetrf(int mode, int m, int n, float* dA, int ldda, int* d_ipiv, int* d_info, const DenseWork& w,
                         hipStream_t stream, int nb, hipStream_t side)
{
    if (mode != kUpdF32 && mode != kUpdBF16 && mode != kUpdFP16 && mode != kUpdBF16x3 && mode != kUpdFP8)
        return MAGMA_ERR_NOT_SUPPORTED;
    return dense_getrf_t<float>(mode, nb > 0 ? nb : DenseShape<float>::NB, m, n, dA, ldda, d_ipiv, d_info, w, stream,
                                side);
}

hipStream_t dense_side_stream(int m, int elem)
{
    // below this order a block step is too short for two streams to gain;
    // MAGMA_DENSE_LOOKAHEAD_MIN overrides the order (tests run the two-stream
    // schedule at small orders with it)
    int min_m = 4096;
    if (const char* e = std::getenv("MAGMA_DENSE_LOOKAHEAD_MIN")) min_m = std::atoi(e);
    if (m < min_m) return nullptr;
    (void)elem;
    // an ordinary stream: a stream with a CU mask needs a hardware queue of
    // its own, and a process has few (creating them per call fails)
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    return s;
}

magma_int_t dense_check_err(const DenseWork& w, hipStream_t stream)
{
    int h = 0;
    MGI_HIP_CHECK(hipMemcpyAsync(&h, w.err, sizeof(int), hipMemcpyDeviceToHost, stream));
    MGI_HIP_CHECK(hipStreamSynchronize(stream));
    if (h != 0) {
        std::fprintf(stderr, "magma_mi355x: dense LU panel wait timed out\n");
        return MAGMA_ERR_UNKNOWN;
    }
    return MAGMA_SUCCESS;
}

}  // namespace mgi
