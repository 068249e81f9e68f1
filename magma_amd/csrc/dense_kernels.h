// Launchers of the single-matrix (dense) LU and solve kernels, internal to
// libmagma_mi355x.so. All of them enqueue on `stream`; the scratch buffers
// are passed in by the caller (DenseWork / SolveWork own no memory).
#pragma once

#include "magma_internal.h"
#include "getrf_kernels.h"

namespace mgi {

// Panel width of the cooperative panel kernel and the default block width of
// the blocked LU (a multiple of the panel width): FP64 64 / 256, FP32 128 / 512.
template <typename T> struct DenseShape;
template <> struct DenseShape<double> { static constexpr int PW = 64;  static constexpr int NB = 256; };
template <> struct DenseShape<float>  { static constexpr int PW = 128; static constexpr int NB = 512; };
// Largest block width (row pairs of one block), and the strip height of the
// triangular solves (order of the diagonal-block inverses).
constexpr int kDenseMaxNB = 512;
constexpr int kDenseStrip = 128;

// Rows held by one workgroup of the panel kernel (one row per thread).
constexpr int kDensePanelRows = 256;
// Row-block size of the blocked triangular solves.
constexpr int kSolveBS = 128;

// Device scratch of the dense LU, carved from one buffer by dense_work_bytes.
struct DenseWork {
    void* crow;      // [2][gmax][1 + PW * sizeof(T)/4] tagged 64-bit words of the candidates
    int nwords;      // number of those words
    int* err;        // [0] a wait in the panel kernel exceeded its time limit
    int* pairs;      // [2][2*kDenseMaxNB] (source row, destination row) of a block's row moves
    int* npairs;     // [0] number of pairs
    void* X;         // kDenseStrip x kDenseStrip inverse of a diagonal block (T)
    void* Ah;        // FP32 storage: 16-bit A21 (rows, k contiguous), hi then lo halves of h_elems each
    void* Bh;        // FP32 storage: 16-bit U12 (columns, k contiguous), likewise
    size_t h_elems;  // capacity of each half
    int gmax;        // panel workgroups the buffers hold
    // the same row-move, inverse and 16-bit operand buffers for the second
    // stream of the look-ahead (the trailing update of the columns beyond
    // the next block runs there while the next block is factored)
    int* pairs2;
    int* npairs2;
    void* X2;
    void* Ah2;
    void* Bh2;
    // this scratch as the second stream uses it
    DenseWork side() const
    {
        DenseWork s = *this;
        s.pairs = pairs2; s.npairs = npairs2; s.X = X2; s.Ah = Ah2; s.Bh = Bh2;
        return s;
    }
};

size_t dense_work_bytes(int m, int n, int elem, int gmax);
void dense_work_carve(void* base, int m, int n, int elem, int gmax, DenseWork& w);
int dense_panel_gmax(int m);
int device_cu_count();

// Right-looking blocked LU with partial pivoting of one m x n column-major
// matrix in place. d_ipiv (device, min(m,n) entries) receives LAPACK's
// 1-based interchanges and d_info the first zero pivot (0 if none).
// mode: kUpdF64 for T = double; for T = float the trailing update format
// (kUpdF32, kUpdBF16, kUpdFP16, kUpdBF16x3). The triangular solves for U12
// run in the storage precision. Returns MAGMA_ERR_NOT_SUPPORTED when the
// matrix has more rows than the panel kernel can hold on the device.
// nb: block width (a multiple of the panel width, at most kDenseMaxNB), 0 for
// the default. side: a second stream for the look-ahead (null: everything on
// stream in order); the work on it is joined back into stream before return.
magma_int_t dense_getrf(int mode, int m, int n, double* dA, int ldda, int* d_ipiv, int* d_info,
                        const DenseWork& w, hipStream_t stream, int nb = 0, hipStream_t side = nullptr);
magma_int_t dense_sgetrf(int mode, int m, int n, float* dA, int ldda, int* d_ipiv, int* d_info,
                         const DenseWork& w, hipStream_t stream, int nb = 0, hipStream_t side = nullptr);
// The look-ahead's second stream, or null when the order is too small to
// gain. Owned by the caller (hipStreamDestroy).
hipStream_t dense_side_stream(int m, int elem);
// Reads the panel kernel's error word (synchronises the stream).
magma_int_t dense_check_err(const DenseWork& w, hipStream_t stream);

// ---------------------------------------------------------------------------
// Solves with the dense factors (getrs_dense.hip).
// ---------------------------------------------------------------------------
struct SolveWork {
    void* invL;      // nblk x BS x BS inverses of the unit-lower diagonal blocks (T)
    void* invU;      // nblk x BS x BS inverses of the upper diagonal blocks (T)
    int* perm;       // n: row map of the interchanges (row i of P*B is B[perm[i]])
    unsigned long long* xw;   // tagged words of the solved values (2 per value, 8-column chunks)
    int* ticket;     // [0] start order of the solve workgroups
    int* err;        // [0] a wait exceeded its time limit
    void* Y;         // n x nrhs_max staging (T)
    int nrhs_max;
};
size_t solve_work_bytes(int n, int nrhs_max, int elem);
void solve_work_carve(void* base, int n, int nrhs_max, int elem, SolveWork& s);

// Inverses of the diagonal blocks of the factors (after the LU, before the
// solves).
template <typename T>
magma_int_t dense_diag_inverses(int n, const T* dLU, int ldlu, const SolveWork& s, hipStream_t stream);
// perm from host ipiv (LAPACK 1-based): composed on the host, copied to s.perm.
magma_int_t dense_set_perm(int n, const int* h_ipiv, const SolveWork& s, hipStream_t stream);
// X := op(A)^{-1} X for the factors of A, X column-major n x nrhs (ldx),
// trans 0 = NoTrans, 1 = Trans. The row map and the diagonal inverses must
// be set. nrhs <= s.nrhs_max.
template <typename T>
magma_int_t dense_getrs(int trans, int n, int nrhs, const T* dLU, int ldlu, T* dX, int ldx, const SolveWork& s,
                        hipStream_t stream);

template <typename T>
magma_int_t dense_trsv_both(int trans, int n, int nrhs, const T* LU, int ld, T* Y, int ldy, const SolveWork& s,
                            hipStream_t st);
magma_int_t dense_solve_check_err(const SolveWork& s, hipStream_t stream);

// Refinement solver bits: solver = magma_refinement_t - Magma_REFINE_IRSTRS
// gives IRSTRS 0, IRDTRS 1, IRGMSTRS 2, IRGMDTRS 3, GMSTRS 4, GMDTRS 5,
// GMGMSTRS 6, GMGMDTRS 7 (include/magma_types.h order).
constexpr int kRefineDTRS = 1;      // correction solves on the FP64 copy of the factors
constexpr int kRefineInnerGM = 2;   // preconditioner: inner FGMRES over the triangular solves
constexpr int kRefineOuterGM = 4;   // outer FGMRES (else classical refinement)

// Refinement solve of one system (dense_drivers.hip): mode is the update
// format of the FP32 LU, solver the bits above. ipiv (host, n) receives the
// pivots used, dipiv (device, nullable) their row map. dworkd (nullable,
// n*nrhs doubles) holds the residual and dworks (nullable, n*(n+nrhs)
// floats) the FP32 factors. *iter follows MAGMA's convention; on fallback
// dA holds its FP64 LU.
magma_int_t dense_iteref(int trans, int n, int nrhs, double* dA, int ldda, int* ipiv, int* dipiv, double* dB,
                         int lddb, double* dX, int lddx, double* dworkd, float* dworks, int mode, int solver,
                         int* iter, int* info);

}  // namespace mgi
