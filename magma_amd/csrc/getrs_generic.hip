// Generic batched triangular solve with LU factors (LAPACK dgetrs semantics),
// FP64 or FP32, fixed or per-matrix sizes, for any n whose right-hand-side
// block fits in LDS.
//
// NoTrans: B := U \ (L \ (P^T B))          (src/zgetrs_batched.cpp:119-145)
// Trans:   B := P (L^T \ (U^T \ B))        (src/zgetrs_gpu.cpp:203-227; the
//          batched reference's transposed branch uses the wrong triangles and
//          pivot order, so the non-batched LAPACK order is followed here).
//
// Each step is column-oriented (an axpy of the solved entry into the rows not
// yet solved, with fused multiply-add), the order reference BLAS dtrsm uses
// for the NoTrans case. The NoTrans sweeps are blocked by 64 rows: one wave
// per right-hand side solves the diagonal block with cross-lane reads (no
// workgroup barrier inside it), then all threads apply the block's 64
// columns to the remaining rows, reading the factor down its columns. The
// per-row operation order is unchanged by the blocking.
#include "magma_internal.h"
#include "lu_device.h"
#include "getrf_kernels.h"

namespace mgi {

constexpr int kRsThreads = 256;
constexpr int kRsCols = 8;     // right-hand sides per workgroup (16 when the batch has more than 8)

// Q columns of one 64-row diagonal block of the unit lower factor, solved
// as interleaved chains by one wave: lane l holds row l of the block
// (lrow = its strictly lower part) and element l of columns col,
// col + cstride, ... Row j's solved value reaches the rows below through a
// cross-lane read.
template <typename T, int Q>
__device__ __forceinline__ void diag_lower(const T (&lrow)[kWave], T* col, int cstride, bool in)
{
    const int l = threadIdx.x % kWave;
    T x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) x[q] = in ? col[q * cstride] : T(0);
#pragma unroll
    for (int j = 0; j < kWave - 1; ++j) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const T xj = readlane_t(x[q], j);
            if (l > j) x[q] = fma(-lrow[j], xj, x[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (in) col[q * cstride] = x[q];
}

// The same for the non-unit upper factor, backward: urow = the strictly upper
// part of row l, rd = 1 / u_ll, nbk = rows in the block.
template <typename T, int Q>
__device__ __forceinline__ void diag_upper(const T (&urow)[kWave], T rd, int nbk, T* col, int cstride, bool in)
{
    const int l = threadIdx.x % kWave;
    T x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) x[q] = in ? col[q * cstride] : T(0);
#pragma unroll
    for (int j = kWave - 1; j >= 0; --j) {
        if (j < nbk) {
            const T dj = readlane_t(rd, j);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const T xj = readlane_t(x[q], j) * dj;
                if (l == j) x[q] = xj;
                else if (l < j) x[q] = fma(-urow[j], xj, x[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (in) col[q * cstride] = x[q];
}

// Rows [i0, i1) of the NC right-hand sides in sB (ld n) minus the 64-column
// block kb of the factor times the block's solved rows: the block's steps in
// ascending order (UP false) or descending order (UP true), the order of the
// column-oriented sweep. TR reads the factor transposed (the Trans solves'
// U^T and L^T). One row per thread; NC is the column count rounded up to 4
// (the extra columns hold zeros and are not stored).
template <typename T, int NC, bool UP, bool TR = false>
__device__ __forceinline__ void apply_block(const MGI_GLOBAL_AS T* __restrict__ A, int ldda, T* sB, int n, int kb, int i0, int i1,
                                            int nr)
{
    for (int i = i0 + (int)threadIdx.x; i < i1; i += kRsThreads) {
        T acc[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] = sB[i + c * n];
        // step j's factor entry of row i: A(i, kb + j), or A(kb + j, i) for
        // the transposed factor (TR; one contiguous column segment per row)
        const size_t sj = TR ? 1 : (size_t)ldda;
        auto Ai = TR ? A + kb + (size_t)i * ldda : A + i + (size_t)kb * ldda;
        // the row's factor entries are loaded CH at a time, each chunk
        // whole before its first use
        constexpr int CH = 32;
        for (int j0 = 0; j0 < kWave; j0 += CH) {
            T a[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int j = UP ? kWave - 1 - (j0 + u) : j0 + u;
                a[u] = Ai[(size_t)j * sj];
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int j = UP ? kWave - 1 - (j0 + u) : j0 + u;
#pragma unroll
                for (int c = 0; c < NC; ++c) acc[c] = fma(-a[u], sB[kb + j + c * n], acc[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if (c < nr) sB[i + c * n] = acc[c];
    }
}

// Q columns of one 64-row diagonal block of a non-unit lower factor (U^T in
// the Trans solve), forward, as interleaved chains: lrow = the strictly lower
// part of row l, rd = 1 / diagonal of row l, nbk = rows in the block.
template <typename T, int Q>
__device__ __forceinline__ void diag_lower_nonunit(const T (&lrow)[kWave], T rd, int nbk, T* col, int cstride, bool in)
{
    const int l = threadIdx.x % kWave;
    T x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) x[q] = in ? col[q * cstride] : T(0);
#pragma unroll
    for (int j = 0; j < kWave; ++j) {
        if (j < nbk) {
            const T dj = readlane_t(rd, j);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const T xj = readlane_t(x[q], j) * dj;
                if (l == j) x[q] = xj;
                else if (l > j) x[q] = fma(-lrow[j], xj, x[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (in) col[q * cstride] = x[q];
}

template <typename T, int NCOLS, bool TRANS>
__global__ void __launch_bounds__(kRsThreads)
getrs_generic_kernel(RsSizes sz, T** __restrict__ dA_array, int** __restrict__ ipiv_array,
                     T** __restrict__ dB_array)
{
    extern __shared__ __align__(16) unsigned char rs_raw[];
    const int b = sz.map ? sz.map[blockIdx.x] : blockIdx.x;
    int n, nrhs, ldda, lddb;
    sz.get(b, n, nrhs, ldda, lddb);
    const int c0 = blockIdx.y * NCOLS;
    if (n <= 0 || c0 >= nrhs) return;
    T* sB = reinterpret_cast<T*>(rs_raw);                       // n x NCOLS, ld = n
    int* perm = reinterpret_cast<int*>(rs_raw + (size_t)sz.max_n * NCOLS * sizeof(T));
    const int tid = threadIdx.x;
    const int nr = min(NCOLS, nrhs - c0);
    auto A = gptr(dA_array[b]);
    auto ipiv = gptr(ipiv_array ? ipiv_array[b] : nullptr);   // null: factors without pivoting
    auto B = gptr(dB_array[b]) + (size_t)c0 * lddb;

    // (P^T B)(i,:) = B(perm[i],:), P the product of the interchanges
    // k <-> ipiv[k]-1 in order. Each row is traced back through the
    // interchanges (last to first) to its source; an interchange outside
    // [k, n) cannot come from a valid LU and is treated as none.
    int* spiv = perm + sz.max_n;
    for (int i = tid; i < n; i += kRsThreads) {
        const int p = ipiv ? ipiv[i] - 1 : i;
        spiv[i] = (p > i && p < n) ? p : i;
    }
    __syncthreads();
    for (int r0 = 0; r0 < n; r0 += kRsThreads) {
        const int r = r0 + tid;
        int src = r;
        for (int k = n - 1; k >= 0; --k) {
            const int p = spiv[k];
            src = (src == k) ? p : (src == p ? k : src);
        }
        if (r < n) perm[r] = src;
    }
    __syncthreads();

    if constexpr (!TRANS) {
        // columns nr .. nc4 * 4 - 1 are zero: the block updates below run on
        // whole groups of 4 columns
        const int nc4 = (nr + 3) / 4;
        for (int idx = tid; idx < n * nc4 * 4; idx += kRsThreads) {
            int i = idx % n, c = idx / n;
            sB[i + c * n] = c < nr ? B[perm[i] + (size_t)c * lddb] : T(0);
        }
        __syncthreads();
        const int l = tid % kWave, w = tid / kWave;
        constexpr int kWaves = kRsThreads / kWave;
        constexpr int kChains = NCOLS / kWaves;            // columns per wave in the diagonal-block solves
        for (int kb = 0; kb < n; kb += kWave) {            // unit lower, forward
            const int nbk = min(kWave, n - kb);
            const int r = kb + l;
            if (w < nr) {
                // lane l holds row r of the diagonal block (strictly lower
                // part); wave w solves columns w, w + 4, ... as interleaved
                // chains
                T lrow[kWave];
                auto Lc = A + r + (size_t)kb * ldda;
#pragma unroll
                for (int j = 0; j < kWave; ++j) lrow[j] = (l > j && l < nbk) ? Lc[(size_t)j * ldda] : T(0);
                const int nq = (nr - w + kWaves - 1) / kWaves;
                T* col = sB + r + w * n;
                const bool in = l < nbk;
                switch (nq) {
                    case 1: diag_lower<T, 1>(lrow, col, kWaves * n, in); break;
                    case 2: diag_lower<T, 2>(lrow, col, kWaves * n, in); break;
                    case 3: diag_lower<T, 3>(lrow, col, kWaves * n, in); break;
                    default: diag_lower<T, kChains>(lrow, col, kWaves * n, in); break;
                }
            }
            __syncthreads();
            // rows below exist only after a full block
            switch (nc4) {
                case 1: apply_block<T, 4, false>(A, ldda, sB, n, kb, kb + kWave, n, nr); break;
                case 2: apply_block<T, 8, false>(A, ldda, sB, n, kb, kb + kWave, n, nr); break;
                case 3: apply_block<T, 12, false>(A, ldda, sB, n, kb, kb + kWave, n, nr); break;
                default: apply_block<T, NCOLS, false>(A, ldda, sB, n, kb, kb + kWave, n, nr); break;
            }
            __syncthreads();
        }
        for (int kend = n; kend > 0; kend -= kWave) {      // non-unit upper, backward
            const int kb = max(0, kend - kWave);
            const int nbk = kend - kb;
            const int r = kb + l;
            if (w < nr) {
                // lane l holds row r of the diagonal block (strictly upper
                // part); columns w, w + 4, ... run as interleaved chains
                T urow[kWave];
                auto Uc = A + r + (size_t)kb * ldda;
#pragma unroll
                for (int j = 0; j < kWave; ++j) urow[j] = (l < j && j < nbk) ? Uc[(size_t)j * ldda] : T(0);
                // each lane's reciprocal pivot, formed in parallel, keeps the
                // division off the dependent chain
                const T rd = l < nbk ? T(1) / A[r + (size_t)r * ldda] : T(1);
                const int nq = (nr - w + kWaves - 1) / kWaves;
                T* col = sB + r + w * n;
                const bool in = l < nbk;
                switch (nq) {
                    case 1: diag_upper<T, 1>(urow, rd, nbk, col, kWaves * n, in); break;
                    case 2: diag_upper<T, 2>(urow, rd, nbk, col, kWaves * n, in); break;
                    case 3: diag_upper<T, 3>(urow, rd, nbk, col, kWaves * n, in); break;
                    default: diag_upper<T, kChains>(urow, rd, nbk, col, kWaves * n, in); break;
                }
            }
            __syncthreads();
            // rows above exist only above a full block
            switch (nc4) {
                case 1: apply_block<T, 4, true>(A, ldda, sB, n, kb, 0, kb, nr); break;
                case 2: apply_block<T, 8, true>(A, ldda, sB, n, kb, 0, kb, nr); break;
                case 3: apply_block<T, 12, true>(A, ldda, sB, n, kb, 0, kb, nr); break;
                default: apply_block<T, NCOLS, true>(A, ldda, sB, n, kb, 0, kb, nr); break;
            }
            __syncthreads();
        }
        for (int idx = tid; idx < n * nr; idx += kRsThreads) {
            int i = idx % n, c = idx / n;
            B[i + (size_t)c * lddb] = sB[i + c * n];
        }
    } else {
        // Blocked by 64 rows like the NoTrans sweeps: a wave per column group
        // solves the diagonal block, then all threads apply the block to the
        // remaining rows. Per element, the contributions arrive in the order
        // of the column-oriented sweeps (ascending for U^T, descending for
        // L^T).
        const int nc4 = (nr + 3) / 4;
        for (int idx = tid; idx < n * nc4 * 4; idx += kRsThreads) {
            int i = idx % n, c = idx / n;
            sB[i + c * n] = c < nr ? B[i + (size_t)c * lddb] : T(0);
        }
        __syncthreads();
        const int l = tid % kWave, w = tid / kWave;
        constexpr int kWaves = kRsThreads / kWave;
        constexpr int kChains = NCOLS / kWaves;
        for (int kb = 0; kb < n; kb += kWave) {            // U^T: non-unit lower, forward
            const int nbk = min(kWave, n - kb);
            const int r = kb + l;
            if (w < nr) {
                // lane l: row r of U^T in the block, U(kb + j, r) for j < l
                T lrow[kWave];
                auto Uc = A + kb + (size_t)r * ldda;
#pragma unroll
                for (int j = 0; j < kWave; ++j) lrow[j] = (j < l && l < nbk) ? Uc[j] : T(0);
                const T rd = l < nbk ? T(1) / A[r + (size_t)r * ldda] : T(1);
                const int nq = (nr - w + kWaves - 1) / kWaves;
                T* col = sB + r + w * n;
                const bool in = l < nbk;
                switch (nq) {
                    case 1: diag_lower_nonunit<T, 1>(lrow, rd, nbk, col, kWaves * n, in); break;
                    case 2: diag_lower_nonunit<T, 2>(lrow, rd, nbk, col, kWaves * n, in); break;
                    case 3: diag_lower_nonunit<T, 3>(lrow, rd, nbk, col, kWaves * n, in); break;
                    default: diag_lower_nonunit<T, kChains>(lrow, rd, nbk, col, kWaves * n, in); break;
                }
            }
            __syncthreads();
            switch (nc4) {
                case 1: apply_block<T, 4, false, true>(A, ldda, sB, n, kb, kb + kWave, n, nr); break;
                case 2: apply_block<T, 8, false, true>(A, ldda, sB, n, kb, kb + kWave, n, nr); break;
                case 3: apply_block<T, 12, false, true>(A, ldda, sB, n, kb, kb + kWave, n, nr); break;
                default: apply_block<T, NCOLS, false, true>(A, ldda, sB, n, kb, kb + kWave, n, nr); break;
            }
            __syncthreads();
        }
        for (int kend = n; kend > 0; kend -= kWave) {      // L^T: unit upper, backward
            const int kb = max(0, kend - kWave);
            const int nbk = kend - kb;
            const int r = kb + l;
            if (w < nr) {
                // lane l: row r of L^T in the block, L(kb + j, r) for j > l
                T urow[kWave];
                auto Lc = A + kb + (size_t)r * ldda;
#pragma unroll
                for (int j = 0; j < kWave; ++j) urow[j] = (l < j && j < nbk) ? Lc[j] : T(0);
                const int nq = (nr - w + kWaves - 1) / kWaves;
                T* col = sB + r + w * n;
                const bool in = l < nbk;
                switch (nq) {                                  // unit diagonal: reciprocal 1
                    case 1: diag_upper<T, 1>(urow, T(1), nbk, col, kWaves * n, in); break;
                    case 2: diag_upper<T, 2>(urow, T(1), nbk, col, kWaves * n, in); break;
                    case 3: diag_upper<T, 3>(urow, T(1), nbk, col, kWaves * n, in); break;
                    default: diag_upper<T, kChains>(urow, T(1), nbk, col, kWaves * n, in); break;
                }
            }
            __syncthreads();
            switch (nc4) {
                case 1: apply_block<T, 4, true, true>(A, ldda, sB, n, kb, 0, kb, nr); break;
                case 2: apply_block<T, 8, true, true>(A, ldda, sB, n, kb, 0, kb, nr); break;
                case 3: apply_block<T, 12, true, true>(A, ldda, sB, n, kb, 0, kb, nr); break;
                default: apply_block<T, NCOLS, true, true>(A, ldda, sB, n, kb, 0, kb, nr); break;
            }
            __syncthreads();
        }
        for (int idx = tid; idx < n * nr; idx += kRsThreads) {
            int i = idx % n, c = idx / n;
            B[perm[i] + (size_t)c * lddb] = sB[i + c * n];
        }
    }
}

// ---------------------------------------------------------------------------
// NoTrans for n <= 64 R (n <= 256 in FP64, 512 in FP32), one wave per
// (system, group of up to kWaveCols
// right-hand sides): lane l keeps rows l + 64 s (s < R) of x in registers,
// so the sweeps need neither LDS nor a workgroup barrier. Row k's solved
// value reaches the other lanes through a cross-lane read, and the factor is
// streamed down its columns (64 contiguous elements per slot, loads issued
// ahead of the dependent updates). The interchanges are applied by tracing
// every row back through them, as in the kernel above; the per-row operation
// order is that of the column-oriented sweeps above.
// ---------------------------------------------------------------------------
constexpr int kWaveCols = 4;

template <typename T, int R>
__global__ void __launch_bounds__(kRsThreads)
getrs_wave_kernel(RsSizes sz, int groups, int nsys, T** __restrict__ dA_array, int** __restrict__ ipiv_array,
                  T** __restrict__ dB_array)
{
    const int gw = blockIdx.x * (kRsThreads / kWave) + threadIdx.x / kWave;
    const int sys = gw / groups;
    if (sys >= nsys) return;
    const int b = sz.map ? sz.map[sys] : sys;
    int n, nrhs, ldda, lddb;
    sz.get(b, n, nrhs, ldda, lddb);
    const int c0 = (gw % groups) * kWaveCols;
    if (n <= 0 || c0 >= nrhs) return;
    const int nr = min(kWaveCols, nrhs - c0);
    const int l = threadIdx.x % kWave;
    auto A = gptr(dA_array[b]);
    auto ipiv = gptr(ipiv_array ? ipiv_array[b] : nullptr);   // null: factors without pivoting
    auto B = gptr(dB_array[b]) + (size_t)c0 * lddb;

    // (P^T B)(i,:) = B(src(i),:)
    int src[R];
#pragma unroll
    for (int s = 0; s < R; ++s) src[s] = l + 64 * s;
    for (int k = ipiv ? n - 1 : -1; k >= 0; --k) {
        int p = ipiv[k] - 1;
        p = (p > k && p < n) ? p : k;
#pragma unroll
        for (int s = 0; s < R; ++s) src[s] = (src[s] == k) ? p : (src[s] == p ? k : src[s]);
    }
    T x[R][kWaveCols];
#pragma unroll
    for (int s = 0; s < R; ++s)
#pragma unroll
        for (int c = 0; c < kWaveCols; ++c)
            x[s][c] = (l + 64 * s < n && c < nr) ? B[src[s] + (size_t)c * lddb] : T(0);

    // Columns are processed in chunks of CH; the factor entries of the next
    // chunk are loaded while the current one is applied.
    constexpr int CH = R <= 2 ? 8 : (R <= 4 ? 4 : 2);
    // unit lower, forward: step k = 64 s + kk
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int kend = min(64, n - 1 - 64 * s);
        if (kend <= 0) continue;
        T cur[CH][R], nxt[CH][R];
        auto load_l = [&](T (&buf)[CH][R], int kk0) {
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const int k = 64 * s + kk0 + j;
                const bool kin = kk0 + j < kend;
                auto Lk = A + (size_t)(kin ? k : 0) * ldda;
#pragma unroll
                for (int s2 = s; s2 < R; ++s2) {
                    const int i = l + 64 * s2;
                    buf[j][s2] = (kin && i > k && i < n) ? Lk[i] : T(0);
                }
            }
        };
        load_l(cur, 0);
        for (int kk0 = 0; kk0 < kend; kk0 += CH) {
            if (kk0 + CH < kend) load_l(nxt, kk0 + CH);
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const int kk = kk0 + j;
                if (kk < kend) {
                    const int k = 64 * s + kk;
#pragma unroll
                    for (int c = 0; c < kWaveCols; ++c) {
                        if (c < nr) {
                            const T xk = readlane_t(x[s][c], kk);
#pragma unroll
                            for (int s2 = s; s2 < R; ++s2)
                                if (l + 64 * s2 > k) x[s2][c] = fma(-cur[j][s2], xk, x[s2][c]);
                        }
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < CH; ++j)
#pragma unroll
                for (int s2 = s; s2 < R; ++s2) cur[j][s2] = nxt[j][s2];
        }
    }
    // non-unit upper, backward: step k = 64 s + kk, kk descending
#pragma unroll
    for (int s = R - 1; s >= 0; --s) {
        if (64 * s >= n) continue;
        const int ktop = min(63, n - 1 - 64 * s);        // first kk processed
        T cur[CH][R], nxt[CH][R], dcur[CH], dnxt[CH];
        auto load_u = [&](T (&buf)[CH][R], T (&dg)[CH], int kk0) {
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const int kk = kk0 - j;
                const bool kin = kk >= 0;
                const int k = 64 * s + (kin ? kk : 0);
                auto Uk = A + (size_t)k * ldda;
                dg[j] = Uk[k];
#pragma unroll
                for (int s2 = 0; s2 <= s; ++s2) {
                    const int i = l + 64 * s2;
                    buf[j][s2] = (kin && i < k) ? Uk[i] : T(0);
                }
            }
        };
        load_u(cur, dcur, ktop);
        for (int kk0 = ktop; kk0 >= 0; kk0 -= CH) {
            if (kk0 - CH >= 0) load_u(nxt, dnxt, kk0 - CH);
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const int kk = kk0 - j;
                if (kk >= 0) {
                    const int k = 64 * s + kk;
                    // one reciprocal per step, independent of the chain
                    const T rk = T(1) / dcur[j];
#pragma unroll
                    for (int c = 0; c < kWaveCols; ++c) {
                        if (c < nr) {
                            const T xk = readlane_t(x[s][c], kk) * rk;
                            if (l == kk) x[s][c] = xk;
#pragma unroll
                            for (int s2 = 0; s2 <= s; ++s2)
                                if (l + 64 * s2 < k) x[s2][c] = fma(-cur[j][s2], xk, x[s2][c]);
                        }
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                dcur[j] = dnxt[j];
#pragma unroll
                for (int s2 = 0; s2 <= s; ++s2) cur[j][s2] = nxt[j][s2];
            }
        }
    }
#pragma unroll
    for (int s = 0; s < R; ++s)
#pragma unroll
        for (int c = 0; c < kWaveCols; ++c)
            if (l + 64 * s < n && c < nr) B[l + 64 * s + (size_t)c * lddb] = x[s][c];
}

template <typename T, int R>
static magma_int_t getrs_wave_launch(const RsSizes& sz, int max_nrhs, T** dA_array, int** ipiv_array, T** dB_array,
                                     int batch, hipStream_t stream)
{
    const int groups = ceildiv(max_nrhs, kWaveCols);
    const long waves = (long)batch * groups;
    const int wpb = kRsThreads / kWave;
    getrs_wave_kernel<T, R><<<(int)ceildiv(waves, (long)wpb), kRsThreads, 0, stream>>>(sz, groups, batch, dA_array,
                                                                                     ipiv_array, dB_array);
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

static size_t generic_bytes(int n, int elem, int ncols) { return (size_t)n * ncols * elem + 2 * (size_t)n * sizeof(int); }
size_t getrs_generic_bytes(int n, int elem) { return generic_bytes(n, elem, kRsCols); }
size_t getrs_generic_bytes(int n) { return getrs_generic_bytes(n, sizeof(double)); }

template <typename T, int NC, bool TRANS>
static magma_int_t generic_launch(const RsSizes& sz, int max_nrhs, size_t lds, T** dA_array, int** ipiv_array,
                                  T** dB_array, int batch, hipStream_t stream)
{
    MGI_HIP_CHECK(hipFuncSetAttribute((const void*)getrs_generic_kernel<T, NC, TRANS>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    dim3 grid(batch, ceildiv(max_nrhs, NC));
    getrs_generic_kernel<T, NC, TRANS><<<grid, kRsThreads, lds, stream>>>(sz, dA_array, ipiv_array, dB_array);
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

template <typename T>
magma_int_t getrs_generic_launch_t(int trans, const RsSizes& sz, int max_nrhs, T** dA_array, int** ipiv_array,
                                   T** dB_array, int batch, hipStream_t stream)
{
    if (batch <= 0 || max_nrhs <= 0) return MAGMA_SUCCESS;
    if (trans == 0 && sz.max_n <= (sizeof(T) == 8 ? 256 : 512)) {
        if (sz.max_n <= 64) return getrs_wave_launch<T, 1>(sz, max_nrhs, dA_array, ipiv_array, dB_array, batch, stream);
        if (sz.max_n <= 128) return getrs_wave_launch<T, 2>(sz, max_nrhs, dA_array, ipiv_array, dB_array, batch, stream);
        if (sz.max_n <= 256) return getrs_wave_launch<T, 4>(sz, max_nrhs, dA_array, ipiv_array, dB_array, batch, stream);
        if constexpr (sizeof(T) == 4)
            return getrs_wave_launch<T, 8>(sz, max_nrhs, dA_array, ipiv_array, dB_array, batch, stream);
    }
    // more than 8 right-hand sides: 16 per workgroup, so each system's
    // factors are streamed once for up to 16 columns
    const size_t lds16 = generic_bytes(sz.max_n, sizeof(T), 2 * kRsCols);
    if (max_nrhs > kRsCols && lds16 <= max_dynamic_lds())
        return trans ? generic_launch<T, 2 * kRsCols, true>(sz, max_nrhs, lds16, dA_array, ipiv_array, dB_array, batch, stream)
                     : generic_launch<T, 2 * kRsCols, false>(sz, max_nrhs, lds16, dA_array, ipiv_array, dB_array, batch, stream);
    const size_t lds = generic_bytes(sz.max_n, sizeof(T), kRsCols);
    if (lds > max_dynamic_lds()) return MAGMA_ERR_NOT_SUPPORTED;
    return trans ? generic_launch<T, kRsCols, true>(sz, max_nrhs, lds, dA_array, ipiv_array, dB_array, batch, stream)
                 : generic_launch<T, kRsCols, false>(sz, max_nrhs, lds, dA_array, ipiv_array, dB_array, batch, stream);
}

template magma_int_t getrs_generic_launch_t<double>(int, const RsSizes&, int, double**, int**, double**, int,
                                                    hipStream_t);
template magma_int_t getrs_generic_launch_t<float>(int, const RsSizes&, int, float**, int**, float**, int,
                                                   hipStream_t);

magma_int_t getrs_generic_launch(int trans, int n, int nrhs, double** dA_array, int ldda, int** ipiv_array,
                                 double** dB_array, int lddb, int batch, hipStream_t stream)
{
    RsSizes sz{n, nrhs, ldda, lddb, n, nullptr, nullptr, nullptr, nullptr, nullptr};
    return getrs_generic_launch_t<double>(trans, sz, nrhs, dA_array, ipiv_array, dB_array, batch, stream);
}

}  // namespace mgi
