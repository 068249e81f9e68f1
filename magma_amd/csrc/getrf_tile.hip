// Square batched LU (and LU + solve) for n <= 128 with the whole matrix held
// in registers as 16x16 FP64 tiles, one workgroup (4 waves) per matrix.
//
// Roles. Waves 0..2 ("tile waves") hold the tiles: tile (I,J) sits in wave
// (I*NT + J) % 3, so every block row, block column and trailing submatrix is
// spread over the three. Wave 3 (the "panel wave") holds no tiles: it factors
// each 16-column panel, keeps the right-hand sides and solves the diagonal
// blocks of the back substitution. The two roles run separate programs with
// the same sequence of workgroup barriers, so the tile registers are not live
// in the panel wave's code (no register spills at n = 128).
//
// Tile layout. Within a tile, lane l holds
//     element (16I + l%16, 16J + 4r + l/16) in register r = 0..3,
// which is the accumulator layout of v_mfma_f64_16x16x4f64 for the transposed
// tile (so an update C^T += U^T L^T lands in place) and lets 16 lanes read 128
// contiguous bytes of one column when loading from column-major memory.
//
// Per 16-column panel P (rows k0 = 16P .. n-1):
//   S1 owners of tile column P stage it in LDS;
//   S2 the panel wave factors the panel: LAPACK pivoting (largest |a|, ties
//      to the lowest current position), the pivot row read out of its lane
//      into scalar registers, reciprocal scaling when |pivot| >= sfmin,
//      rank-1 updates; rows are not moved during the panel -- their LAPACK
//      positions are tracked and the panel is written back in position order;
//      then L11^{-1};
//   S3/S4 the panel's row permutation is applied to all other tile columns
//      through an LDS staging area that holds only the rows that move;
//   S5 U12 = L11^{-1} A12 with MFMA;
//   S6 A22 -= L21 U12 with v_mfma_f64_16x16x4f64 on the register tiles.
// With a right-hand side (gesv), B sits in LDS as extra columns: it receives
// the same interchanges, TRSM and updates (giving L^{-1} P^T B), then a blocked
// back substitution with U produces X. A, B are read once; LU, X written once.
//
// Matrices with n not a multiple of 16 are padded to [A 0; 0 I] on chip; the
// padding never becomes a pivot for a real column and is not written back.
//
// Reference behaviour: magmablas/zgetf2_devicefunc.cuh:218-294 (panel step),
// magmablas/magma_templates.h:44-60 (ties), magmablas/zgetrf_batched_smallsq_noshfl.cu
// (register LU with row tracking instead of physical swaps),
// src/zgetrf_panel_batched.cpp:102-196 (panel / laswp / trsm / gemm order).
#include "magma_internal.h"
#include "lu_device.h"
#include "getrf_kernels.h"

#include <climits>

namespace mgi {

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int kTileThreads = 256;
constexpr int kTileWaves = 4;
constexpr int kTileOwners = 3;      // waves 0..2 hold tiles, wave 3 factors panels

// The matrices are held through global-address-space pointers (gptr), so the
// compiler uses global loads with a scalar base and a 32-bit lane offset
// instead of 64-bit flat addresses.
typedef MGI_GLOBAL_AS double gdouble;

__device__ __forceinline__ double wave_max_d(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, shfl_xor_d(v, off));
    return v;
}

__device__ __forceinline__ int wave_min_i(int v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
    return v;
}

__device__ __forceinline__ int first_lane(unsigned long long mask) { return __ffsll((long long)mask) - 1; }

template <int NT, bool WITH_B>
struct TileSmem {
    static constexpr int NP = 16 * NT;
    static constexpr int LD16 = 17;                 // padded row length of 16-wide buffers
    static constexpr int UCOLS = NP + (WITH_B ? 16 : 0);
    static constexpr int ULD = UCOLS + 1;
    static constexpr int SLD = NP + (WITH_B ? 16 : 0) + 1;   // staged rows: A's columns, then B's
    static constexpr int S_DOUBLES = 33 * SLD;                    // 32 staged rows + a trash row
    static constexpr int U_DOUBLES = 16 * ULD;
    static constexpr int T_DOUBLES = (kTileWaves + 1) * 16 * LD16;
    static constexpr int UNION_DOUBLES =
        S_DOUBLES > U_DOUBLES ? (S_DOUBLES > T_DOUBLES ? S_DOUBLES : T_DOUBLES)
                              : (U_DOUBLES > T_DOUBLES ? U_DOUBLES : T_DOUBLES);
    double L[NP * LD16];                            // panel, position order
    double U[UNION_DOUBLES];                        // laswp staging / U12 / solve scratch
    double B[WITH_B ? NP * LD16 : 1];               // right-hand sides
    double Li[16 * LD16];                           // L11^{-1} of the current panel
    int rowmap[NP];                                 // position -> source row (panel-relative)
    int slotof[NP];                                 // moved row -> staging slot, or -1
    int info;
};

// Tile slots of a tile wave: slot k of wave w holds tile q = 3k + w (row-major
// tile index), i.e. tile (q / NT, q % NT). Consecutive tiles rotate over the
// three tile waves, so every block row, block column and trailing submatrix
// is shared about evenly.
template <int NT>
struct TileMap {
    static constexpr int KT = (NT * NT + kTileOwners - 1) / kTileOwners;
    __device__ static constexpr bool valid(int w, int k) { return kTileOwners * k + w < NT * NT; }
    __device__ static constexpr int row(int w, int k) { return (kTileOwners * k + w) / NT; }
    __device__ static constexpr int col(int w, int k) { return (kTileOwners * k + w) % NT; }
};

// One 16-column panel on one wave (LAPACK dgetf2 semantics, lazy pivoting):
// lane l holds the panel rows l (x0) and, with HAS1, l + 64 (x1); pos0/pos1
// are their current LAPACK positions and alive0/alive1 whether they are still
// candidates. All 16 columns are processed: in the last panel of a matrix
// whose order is not a multiple of 16 the extra columns belong to the
// identity padding, whose pivot is the padding row itself (so nothing
// changes); only the ipiv and info writes are limited to real columns.
//
// The loop is software pipelined: step c updates column c+1 first, then
// runs the pivot search of column c+1 (DPP maximum, ballots) in the same
// branch-free block as the remaining rank-1 updates, so the reduction's
// latency is hidden behind the FMAs.
template <bool HAS1>
__device__ __forceinline__ void panel_columns(double (&x0)[16], double (&x1)[16], int& pos0, int& pos1,
                                              bool& alive0, bool& alive1, int& info_local, MGI_GLOBAL_AS int* __restrict__ ipiv,
                                              int k0, int n, int l)
{
    // idamax candidates of column c: NaN never wins, except at the first
    // candidate position c, which keeps the pivot (counted as +inf)
    double v0, v1 = -1.0, m;
    unsigned long long b0, b1 = 0ull;
    auto search = [&](int c) {
        v0 = alive0 ? fabs(x0[c]) : -1.0;
        if (pos0 == c && v0 != v0) v0 = INFINITY;
        if (HAS1) {
            v1 = alive1 ? fabs(x1[c]) : -1.0;
            if (pos1 == c && v1 != v1) v1 = INFINITY;
        }
        m = wave_max_dpp(HAS1 ? fmax(v0, v1) : v0);
        b0 = __ballot(v0 == m);
        if (HAS1) b1 = __ballot(v1 == m);
    };
    search(0);
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        if (__popcll(b0) + __popcll(b1) != 1) {
            // exact tie: LAPACK keeps the lowest current position
            const int cand = min(v0 == m ? pos0 : INT_MAX, (HAS1 && v1 == m) ? pos1 : INT_MAX);
            const int mpos = wave_min_dpp(cand);
            b0 = __ballot(v0 == m && pos0 == mpos);
            if (HAS1) b1 = __ballot(v1 == m && pos1 == mpos);
        }
        const bool in1 = HAS1 && b0 == 0ull;
        const int pl = in1 ? first_lane(b1) : first_lane(b0);
        // pivot and next-column entry of the pivot row (scalar registers)
        const double pv = readlane_d(in1 ? x1[c] : x0[c], pl);
        const int cn = c + 1 < 16 ? c + 1 : c;      // next column (unused at c = 15)
        const double un = readlane_d(in1 ? x1[cn] : x0[cn], pl);
        const int pos_p = __builtin_amdgcn_readlane(in1 ? pos1 : pos0, pl);
        if (l == 0 && k0 + c < n) ipiv[k0 + c] = k0 + pos_p + 1;
        // LAPACK interchange of positions c and pos_p
        const bool me0 = !in1 && l == pl, me1 = in1 && l == pl;
        pos0 = me0 ? c : (pos0 == c ? pos_p : pos0);
        alive0 = alive0 && !me0;
        if (HAS1) {
            pos1 = me1 ? c : (pos1 == c ? pos_p : pos1);
            alive1 = alive1 && !me1;
        }
        // multipliers: reciprocal scaling, or division when |pivot| < sfmin
        const double rinv = 1.0 / pv;
        double l0 = x0[c] * rinv, l1 = HAS1 ? x1[c] * rinv : 0.0;
        if (!(pv != 0.0 && fabs(pv) >= kSfminD)) {
            if (pv != 0.0) {
                l0 = x0[c] / pv;
                if (HAS1) l1 = x1[c] / pv;
            } else {
                l0 = x0[c];
                l1 = HAS1 ? x1[c] : 0.0;
                if (info_local == 0 && k0 + c < n) info_local = k0 + c + 1;
            }
        }
        // rows still alive take the multiplier; others stay as they are
        const double g0 = alive0 ? l0 : 0.0;
        x0[c] = alive0 ? l0 : x0[c];
        double g1 = 0.0;
        if (HAS1) {
            g1 = alive1 ? l1 : 0.0;
            x1[c] = alive1 ? l1 : x1[c];
        }
        if (c + 1 < 16) {
            x0[cn] = fma(-g0, un, x0[cn]);
            if (HAS1) x1[cn] = fma(-g1, un, x1[cn]);
            search(cn);
            // the rest of the pivot row and the remaining rank-1 updates
#pragma unroll
            for (int c2 = 0; c2 < 16; ++c2) {
                if (c2 <= cn) continue;
                const double u = readlane_d(in1 ? x1[c2] : x0[c2], pl);
                x0[c2] = fma(-g0, u, x0[c2]);
                if (HAS1) x1[c2] = fma(-g1, u, x1[c2]);
            }
        }
    }
}

// Barrier schedule shared by the two wave programs below (each __syncthreads
// in one program has a partner in the other, in the same order):
//   load | per panel: S1 | S2 | S3 | S4a | S4b | S5 | S6 | back substitution:
//   2 per block row.
template <int NT, bool WITH_B>
struct TileCtx {
    TileSmem<NT, WITH_B>& sm;
    int n, nrhs, ldda, lddb;
    MGI_GLOBAL_AS double* A;
    MGI_GLOBAL_AS int* ipiv;
    MGI_GLOBAL_AS double* B;
    int l, lr, lq;
};

// ---------------------------------------------------------------------------
// Tile waves (0..2): the matrix in MFMA accumulator layout, row interchanges,
// U12 and the trailing updates, the factors' write-back, back-substitution
// partial products.
// ---------------------------------------------------------------------------
template <int NT, bool WITH_B, int W>
__device__ __forceinline__ void tile_wave_main(const TileCtx<NT, WITH_B>& cx)
{
    constexpr int w = W;                  // tile coordinates are compile-time constants
    constexpr int NP = 16 * NT;
    constexpr int KT = TileMap<NT>::KT;
    constexpr int LD16 = TileSmem<NT, WITH_B>::LD16;
    constexpr int ULD = TileSmem<NT, WITH_B>::ULD;
    constexpr int SLD = TileSmem<NT, WITH_B>::SLD;
    auto& sm = cx.sm;
    const int n = cx.n, ldda = cx.ldda, lr = cx.lr, lq = cx.lq;
    auto A = cx.A;

    // A tile is addressed as a wave-uniform base plus one of four 32-bit lane
    // offsets shared by all tiles (ldda < 2^27 keeps them in range).
    int off[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) off[r] = lr + (4 * r + lq) * ldda;
    double4_t acc[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        if (!TileMap<NT>::valid(w, k)) continue;
        const int I = TileMap<NT>::row(w, k), J = TileMap<NT>::col(w, k);
        const gdouble* At = A + 16 * I + (size_t)(16 * J) * ldda;
        if (16 * I + 16 <= n && 16 * J + 16 <= n) {
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[k][r] = At[off[r]];
        } else {
            const int row = 16 * I + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = 16 * J + 4 * r + lq;
                acc[k][r] = (row < n && col < n) ? At[off[r]] : (row == col ? 1.0 : 0.0);
            }
        }
    }
    __syncthreads();                                              // load

    const int npan = (n + 15) / 16;
    for (int P = 0; P < npan; ++P) {
        // S1: the owners of tile column P stage it (rows >= 16P) in sm.L
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            if (!TileMap<NT>::valid(w, k)) continue;
            const int I = TileMap<NT>::row(w, k), J = TileMap<NT>::col(w, k);
            if (J != P || I < P) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) sm.L[(16 * (I - P) + lr) * LD16 + 4 * r + lq] = acc[k][r];
        }
        __syncthreads();                                          // S1
        __syncthreads();                                          // S2 (panel wave)

        // S3: stage the rows that move; reload the factored panel. A row
        // that stays is written to the trash row, so no lane branches; each
        // row block's slot is read once.
        constexpr int kTrash = 32;
        int slot_of_block[NT];
#pragma unroll
        for (int I = 0; I < NT; ++I) {
            const int s = (I >= P) ? sm.slotof[16 * (I - P) + lr] : -1;
            slot_of_block[I] = s >= 0 ? s : kTrash;
        }
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            if (!TileMap<NT>::valid(w, k)) continue;
            const int I = TileMap<NT>::row(w, k), J = TileMap<NT>::col(w, k);
            if (I < P) continue;
            if (J == P) {
                const int rr = 16 * (I - P) + lr;
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[k][r] = sm.L[rr * LD16 + 4 * r + lq];
            } else {
                const int base = slot_of_block[I] * SLD + 16 * J + lq;
#pragma unroll
                for (int r = 0; r < 4; ++r) sm.U[base + 4 * r] = acc[k][r];
            }
        }
        __syncthreads();                                          // S3

        // S4a: apply the row permutation from the staging area (rows that
        // stay read the trash row and keep their value)
        bool moved_of_block[NT];
#pragma unroll
        for (int I = 0; I < NT; ++I) {
            bool mv = false;
            int s = kTrash;
            if (I >= P) {
                const int rr = 16 * (I - P) + lr;
                const int src = sm.rowmap[rr];
                mv = src != rr;
                if (mv) s = sm.slotof[src];
            }
            moved_of_block[I] = mv;
            slot_of_block[I] = s;
        }
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            if (!TileMap<NT>::valid(w, k)) continue;
            const int I = TileMap<NT>::row(w, k), J = TileMap<NT>::col(w, k);
            if (I < P || J == P) continue;
            const int base = slot_of_block[I] * SLD + 16 * J + lq;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double v = sm.U[base + 4 * r];
                acc[k][r] = moved_of_block[I] ? v : acc[k][r];
            }
        }
        __syncthreads();                                          // S4a

        // S4b: block row P right of the panel into sm.U
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            if (!TileMap<NT>::valid(w, k)) continue;
            const int I = TileMap<NT>::row(w, k), J = TileMap<NT>::col(w, k);
            if (I != P || J <= P) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) sm.U[lr * ULD + 16 * J + 4 * r + lq] = acc[k][r];
        }
        __syncthreads();                                          // S4b

        // S5: U12 = L11^{-1} A12 on the owners of block row P
        // (U12^T = A12^T L11^{-T}: the result lands in the tile layout)
        {
            double lb[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) lb[kk] = sm.Li[lr * LD16 + 4 * kk + lq];
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                if (!TileMap<NT>::valid(w, k)) continue;
                const int I = TileMap<NT>::row(w, k), J = TileMap<NT>::col(w, k);
                if (I != P || J <= P) continue;
                double4_t uacc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const double a12 = sm.U[(4 * kk + lq) * ULD + 16 * J + lr];
                    uacc = __builtin_amdgcn_mfma_f64_16x16x4f64(a12, lb[kk], uacc, 0, 0, 0);
                }
                acc[k] = uacc;
#pragma unroll
                for (int r = 0; r < 4; ++r) sm.U[lr * ULD + 16 * J + 4 * r + lq] = uacc[r];
            }
        }
        __syncthreads();                                          // S5

        // S6: trailing update; lane l uses -L(16I + l%16, 4kk + l/16) and
        // U12(4kk + l/16, 16J + l%16)
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            if (!TileMap<NT>::valid(w, k)) continue;
            const int I = TileMap<NT>::row(w, k), J = TileMap<NT>::col(w, k);
            if (I <= P || J <= P) continue;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const double uf = sm.U[(4 * kk + lq) * ULD + 16 * J + lr];
                const double lf = -sm.L[(16 * (I - P) + lr) * LD16 + 4 * kk + lq];
                acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(uf, lf, acc[k], 0, 0, 0);
            }
        }
        __syncthreads();                                          // S6
    }

    // the factors
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        if (!TileMap<NT>::valid(w, k)) continue;
        const int I = TileMap<NT>::row(w, k), J = TileMap<NT>::col(w, k);
        gdouble* At = A + 16 * I + (size_t)(16 * J) * ldda;
        if (16 * I + 16 <= n && 16 * J + 16 <= n) {
#pragma unroll
            for (int r = 0; r < 4; ++r) At[off[r]] = acc[k][r];
        } else {
            const int row = 16 * I + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = 16 * J + 4 * r + lq;
                if (row < n && col < n) At[off[r]] = acc[k][r];
            }
        }
    }

    if constexpr (WITH_B) {
        // back substitution, block rows bottom-up: each tile wave sums its
        // tiles' share of U_IJ X_J (J > I) into its own 16x16 slice of T; the
        // owner of U_II copies it to slice 3; the panel wave solves.
        double* T = sm.U;                         // [4][16][LD16]
        const int nb_real = (n + 15) / 16;
        for (int I = nb_real - 1; I >= 0; --I) {
            double4_t part = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                if (!TileMap<NT>::valid(w, k)) continue;
                const int Ik = TileMap<NT>::row(w, k), J = TileMap<NT>::col(w, k);
                if (Ik != I) continue;
                if (J > I) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const double xb = sm.B[(16 * J + 4 * r + lq) * LD16 + lr];
                        part = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[k][r], xb, part, 0, 0, 0);
                    }
                } else if (J == I) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) T[(3 * 16 + lr) * LD16 + 4 * r + lq] = acc[k][r];
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) T[(w * 16 + 4 * r + lq) * LD16 + lr] = part[r];
            __syncthreads();                                      // partials ready
            __syncthreads();                                      // block row solved
        }
    }
    (void)NP;
}

// ---------------------------------------------------------------------------
// Panel wave (3): the right-hand sides, the panel factorisation with LAPACK
// pivoting, L11^{-1}, B's interchanges / TRSM / updates and the 16x16
// triangular solves of the back substitution.
// ---------------------------------------------------------------------------
template <int NT, bool WITH_B>
__device__ __forceinline__ void panel_wave_main(const TileCtx<NT, WITH_B>& cx)
{
    constexpr int NP = 16 * NT;
    constexpr int LD16 = TileSmem<NT, WITH_B>::LD16;
    constexpr int ULD = TileSmem<NT, WITH_B>::ULD;
    auto& sm = cx.sm;
    const int n = cx.n, l = cx.l, lr = cx.lr, lq = cx.lq;
    auto ipiv = cx.ipiv;

    if constexpr (WITH_B) {
        for (int idx = l; idx < NP * 16; idx += kWave) {
            const int c = idx / NP, row = idx % NP;
            sm.B[row * LD16 + c] = (row < n && c < cx.nrhs) ? cx.B[row + (size_t)c * cx.lddb] : 0.0;
        }
    }
    if (l == 0) sm.info = 0;
    __syncthreads();                                              // load

    const int npan = (n + 15) / 16;
    for (int P = 0; P < npan; ++P) {
        const int k0 = 16 * P;
        const int mp = NP - k0;                       // panel rows (incl. padding)
        __syncthreads();                                          // S1 (tile waves)

        // S2: lane l holds panel rows l and l+64 (the second only while the
        // panel is taller than 64 rows); a row leaves the candidates once chosen
        {
            const bool has1 = mp > 64;
            double x0[16], x1[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                x0[c] = (l < mp) ? sm.L[l * LD16 + c] : 0.0;
                x1[c] = (has1 && l + 64 < mp) ? sm.L[(l + 64) * LD16 + c] : 0.0;
            }
            int pos0 = l, pos1 = l + 64;
            bool alive0 = l < mp, alive1 = has1 && (l + 64 < mp);
            int info_local = 0;
            if (has1) panel_columns<true>(x0, x1, pos0, pos1, alive0, alive1, info_local, ipiv, k0, n, l);
            else panel_columns<false>(x0, x1, pos0, pos1, alive0, alive1, info_local, ipiv, k0, n, l);
            // every panel row (padding included) to its LAPACK position
            if (l < mp) {
                sm.rowmap[pos0] = l;
#pragma unroll
                for (int c = 0; c < 16; ++c) sm.L[pos0 * LD16 + c] = x0[c];
            }
            if (has1 && l + 64 < mp) {
                sm.rowmap[pos1] = l + 64;
#pragma unroll
                for (int c = 0; c < 16; ++c) sm.L[pos1 * LD16 + c] = x1[c];
            }
            const bool mv0 = (l < mp) && pos0 != l;
            const bool mv1 = (has1 && l + 64 < mp) && pos1 != l + 64;
            const unsigned long long m0 = __ballot(mv0), m1 = __ballot(mv1);
            const unsigned long long below = (l == 0) ? 0ull : (~0ull >> (64 - l));
            if (l < mp) sm.slotof[l] = mv0 ? __popcll(m0 & below) : -1;
            if (has1 && l + 64 < mp) sm.slotof[l + 64] = mv1 ? (__popcll(m0) + __popcll(m1 & below)) : -1;
            if (l == 0 && info_local != 0 && sm.info == 0) sm.info = info_local;
        }
        // the rows just written by other lanes are read back below
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        {
            // L11^{-1} by forward substitution on the unit columns: lane
            // 4j + q owns column j, rows q, q+4, q+8, q+12; the solved entry
            // of step k is broadcast within the quad.
            const int j = l >> 2, q = l & 3;
            double xi[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) xi[s4] = (q + 4 * s4 == j) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < 15; ++k) {
                double xk;
                switch (k & 3) {
                    case 0: xk = quad_bcast_d<0>(xi[k >> 2]); break;
                    case 1: xk = quad_bcast_d<1>(xi[k >> 2]); break;
                    case 2: xk = quad_bcast_d<2>(xi[k >> 2]); break;
                    default: xk = quad_bcast_d<3>(xi[k >> 2]); break;
                }
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) {
                    const int i = q + 4 * s4;
                    const double coef = (i > k) ? sm.L[i * LD16 + k] : 0.0;
                    xi[s4] = fma(-coef, xk, xi[s4]);
                }
            }
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) sm.Li[(q + 4 * s4) * LD16 + j] = xi[s4];
        }
        __syncthreads();                                          // S2

        if constexpr (WITH_B) {
            // S3: stage B's moved rows beside A's (staging columns NP..NP+15);
            // lane l copies its rows l and l + 64, a row that stays goes to
            // the trash row
            constexpr int SLD = TileSmem<NT, WITH_B>::SLD;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = l + 64 * h;
                if (h * 64 >= NP) continue;
                int s = 32;
                if (i < mp) {
                    const int sl = sm.slotof[i];
                    s = sl >= 0 ? sl : 32;
                }
                const double* src = &sm.B[(k0 + (i < mp ? i : 0)) * LD16];
                double* dst = &sm.U[s * SLD + NP];
#pragma unroll
                for (int c = 0; c < 16; ++c) dst[c] = src[c];
            }
            __syncthreads();                                      // S3
            // S4a: B's rows into their new positions
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = l + 64 * h;
                if (h * 64 >= NP || i >= mp) continue;
                const int src = sm.rowmap[i];
                if (src != i) {
                    const double* from = &sm.U[sm.slotof[src] * SLD + NP];
                    double* to = &sm.B[(k0 + i) * LD16];
#pragma unroll
                    for (int c = 0; c < 16; ++c) to[c] = from[c];
                }
            }
            __syncthreads();                                      // S4a
            // S4b: B's panel rows beside U12
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = 4 * q + lq;
                sm.U[i * ULD + NP + lr] = sm.B[(k0 + i) * LD16 + lr];
            }
            __syncthreads();                                      // S4b
            // S5: Y_P = L11^{-1} B_P
            {
                double4_t uacc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const double a12 = sm.U[(4 * kk + lq) * ULD + NP + lr];
                    const double lb = sm.Li[lr * LD16 + 4 * kk + lq];
                    uacc = __builtin_amdgcn_mfma_f64_16x16x4f64(a12, lb, uacc, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    sm.U[lr * ULD + NP + 4 * r + lq] = uacc[r];
                    sm.B[(k0 + lr) * LD16 + 4 * r + lq] = uacc[r];
                }
            }
            __syncthreads();                                      // S5
            // S6: B block rows below the panel
            {
                double uf[4];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) uf[kk] = sm.U[(4 * kk + lq) * ULD + NP + lr];
                for (int I = P + 1; I < NT; ++I) {
                    double4_t cb;
#pragma unroll
                    for (int r = 0; r < 4; ++r) cb[r] = sm.B[(16 * I + lr) * LD16 + 4 * r + lq];
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) {
                        const double lfr = -sm.L[(16 * (I - P) + lr) * LD16 + 4 * kk + lq];
                        cb = __builtin_amdgcn_mfma_f64_16x16x4f64(uf[kk], lfr, cb, 0, 0, 0);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) sm.B[(16 * I + lr) * LD16 + 4 * r + lq] = cb[r];
                }
            }
            __syncthreads();                                      // S6
        } else {
            __syncthreads();                                      // S3
            __syncthreads();                                      // S4a
            __syncthreads();                                      // S4b
            __syncthreads();                                      // S5
            __syncthreads();                                      // S6
        }
    }

    if constexpr (WITH_B) {
        double* T = sm.U;
        const int nb_real = (n + 15) / 16;
        for (int I = nb_real - 1; I >= 0; --I) {
            __syncthreads();                                      // partials ready
            // lane 4c + q: right-hand side c, rows q, q+4, q+8, q+12 of the
            // block; the solved entry of row k is broadcast within the quad
            const int c = l >> 2, q = l & 3;
            double y[4], rd[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int i = q + 4 * s4;
                double v = sm.B[(16 * I + i) * LD16 + c];
#pragma unroll
                for (int qq = 0; qq < kTileOwners; ++qq) v -= T[(qq * 16 + i) * LD16 + c];
                y[s4] = v;
                rd[s4] = 1.0 / T[(48 + i) * LD16 + i];
            }
#pragma unroll
            for (int k = 15; k >= 0; --k) {
                const int sk = k >> 2;
                y[sk] = (q == (k & 3)) ? y[sk] * rd[sk] : y[sk];
                double xk;
                switch (k & 3) {
                    case 0: xk = quad_bcast_d<0>(y[sk]); break;
                    case 1: xk = quad_bcast_d<1>(y[sk]); break;
                    case 2: xk = quad_bcast_d<2>(y[sk]); break;
                    default: xk = quad_bcast_d<3>(y[sk]); break;
                }
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) {
                    const int i = q + 4 * s4;
                    const double coef = (i < k) ? T[(48 + i) * LD16 + k] : 0.0;
                    y[s4] = fma(-coef, xk, y[s4]);
                }
            }
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) sm.B[(16 * I + q + 4 * s4) * LD16 + c] = y[s4];
            __syncthreads();                                      // block row solved
        }
    }
}

template <int NT, bool WITH_B>
__global__ void __launch_bounds__(kTileThreads, 2)
getrf_tile_kernel(TileArgs ta, double** __restrict__ dA_array, int** __restrict__ ipiv_array,
                  int* __restrict__ info_array, double** __restrict__ dB_array)
{
    constexpr int LD16 = TileSmem<NT, WITH_B>::LD16;
    __shared__ TileSmem<NT, WITH_B> sm;
    const int t = threadIdx.x;
    const int w = t / kWave;
    const int l = t % kWave;
    const int mat = ta.map ? ta.map[blockIdx.x] : blockIdx.x;
    TileCtx<NT, WITH_B> cx{sm,
                           ta.n_arr ? ta.n_arr[mat] : ta.n,
                           ta.nrhs_arr ? ta.nrhs_arr[mat] : ta.nrhs,
                           ta.ldda_arr ? ta.ldda_arr[mat] : ta.ldda,
                           ta.lddb_arr ? ta.lddb_arr[mat] : ta.lddb,
                           gptr(dA_array[mat]), gptr(ipiv_array[mat]), gptr(WITH_B ? dB_array[mat] : nullptr),
                           l, l % 16, l / 16};
    if (cx.n <= 0) return;                // empty system: nothing to factor
    switch (w) {
        case 0: tile_wave_main<NT, WITH_B, 0>(cx); break;
        case 1: tile_wave_main<NT, WITH_B, 1>(cx); break;
        case 2: tile_wave_main<NT, WITH_B, 2>(cx); break;
        default: panel_wave_main<NT, WITH_B>(cx); break;
    }
    __syncthreads();                      // sm.info final
    if (t == 0) info_array[mat] = sm.info;
    if constexpr (WITH_B) {
        auto B = cx.B;
        for (int idx = t; idx < cx.n * cx.nrhs; idx += kTileThreads) {
            const int row = idx % cx.n, c = idx / cx.n;
            B[row + (size_t)c * cx.lddb] = sm.B[row * LD16 + c];
        }
    }
}

template <bool WITH_B>
static magma_int_t launch_nt(int nt, const TileArgs& a, double** dA_array, int** ipiv_array, int* info_array,
                             double** dB_array, int batch, hipStream_t stream)
{
    dim3 grid(batch), block(kTileThreads);
#define MGI_TILE_CASE(NTV)                                                                               \
    case NTV:                                                                                            \
        getrf_tile_kernel<NTV, WITH_B><<<grid, block, 0, stream>>>(a, dA_array, ipiv_array, info_array, \
                                                                   dB_array);                           \
        break;
    switch (nt) {
        MGI_TILE_CASE(1)
        MGI_TILE_CASE(2)
        MGI_TILE_CASE(3)
        MGI_TILE_CASE(4)
        MGI_TILE_CASE(5)
        MGI_TILE_CASE(6)
        MGI_TILE_CASE(7)
        MGI_TILE_CASE(8)
        default: return MAGMA_ERR_NOT_SUPPORTED;
    }
#undef MGI_TILE_CASE
    MGI_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

bool getrf_tile_supported(int n) { return n >= 1 && n <= 128; }

magma_int_t getrf_tile_launch_args(int nt, const TileArgs& a, double** dA_array, int** ipiv_array, int* info_array,
                                   int batch, hipStream_t stream)
{
    return launch_nt<false>(nt, a, dA_array, ipiv_array, info_array, nullptr, batch, stream);
}

magma_int_t gesv_tile_launch_args(int nt, const TileArgs& a, double** dA_array, int** ipiv_array, int* info_array,
                                  double** dB_array, int batch, hipStream_t stream)
{
    return launch_nt<true>(nt, a, dA_array, ipiv_array, info_array, dB_array, batch, stream);
}

magma_int_t getrf_tile_launch(int n, double** dA_array, int ldda, int** ipiv_array, int* info_array,
                              int batch, hipStream_t stream)
{
    if (!getrf_tile_supported(n)) return MAGMA_ERR_NOT_SUPPORTED;
    TileArgs a{n, 0, ldda, 0, nullptr, nullptr, nullptr, nullptr, nullptr};
    return launch_nt<false>((n + 15) / 16, a, dA_array, ipiv_array, info_array, nullptr, batch, stream);
}

bool gesv_tile_supported(int n, int nrhs) { return getrf_tile_supported(n) && nrhs >= 1 && nrhs <= 16; }

magma_int_t gesv_tile_launch(int n, int nrhs, double** dA_array, int ldda, int** ipiv_array, int* info_array,
                             double** dB_array, int lddb, int batch, hipStream_t stream)
{
    if (!gesv_tile_supported(n, nrhs)) return MAGMA_ERR_NOT_SUPPORTED;
    TileArgs a{n, nrhs, ldda, lddb, nullptr, nullptr, nullptr, nullptr, nullptr};
    return launch_nt<true>((n + 15) / 16, a, dA_array, ipiv_array, info_array, dB_array, batch, stream);
}

}  // namespace mgi
