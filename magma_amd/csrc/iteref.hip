// Mixed-precision LU solve with FP64 iterative refinement, batched.
//
// Per system A x = b (src/dxgesv_gmres_gpu.cpp:160-832 for the driver logic,
// src/dgmres_plu_gpu.cpp:79-316 for FGMRES):
//   1. A is rounded to FP32 once (with an overflow check, magmablas/zlag2c.cu);
//      an overflow sends the system to the FP64 fallback (iter = -2);
//   2. the FP32 copy is factored with partial pivoting; its trailing updates
//      take BF16, FP16, split-BF16 or FP32 inputs with FP32 accumulation
//      (MAGMA's SHT: FP16 inputs, src/xshgetrf_gpu.cpp:270-276). A failed or
//      non-finite factorisation sends the system to the fallback (iter = -3);
//   3. refinement, with FP64 residuals r = b - A x:
//        IRSTRS: x += LU^{-1} r until ||r||_inf <= ||x||_inf * cte for every
//                column, cte = ||A||_inf * eps * sqrt(n);
//        GMSTRS: flexible GMRES right-preconditioned by the LU (modified
//                Gram-Schmidt, Givens rotations kept on the device), stopping
//                when the residual estimate drops below cte * ||r0||_2;
//   4. a system that does not converge within itermax steps is re-solved in
//      FP64 (iter = -steps), which MAGMA's GMRES path does not do.
// Everything stays on the GPU; the host only loops over refinement steps and
// reads one counter per step.
#include "magma_internal.h"
#include "lu_device.h"
#include "getrf_kernels.h"
#include "dense_kernels.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

namespace mgi {

constexpr int kIrThreads = 256;
// single systems above this order go to the dense LU on the whole device
constexpr int kDenseIrMinN = 512;
// Restart of the FGMRES: one system restarts after ITERMAX = min(n, 500)
// iterations as src/dxgesv_gmres_gpu.cpp:574-575 does; the batched solver
// restarts after 50 (inner_itermax, :212) to bound the Krylov basis of a
// whole batch (n x 101 doubles per system instead of n x 1001).
constexpr int kGmresMaxRestart = 500;
constexpr int kGmresBatchedRestart = 50;
constexpr int kGmresCheck = 4;       // inner steps between host checks of the open cycles

struct IrState {
    int* done;          // 1 converged, -1 diverged or not converged, 0 active
    int* fail;          // 1 overflow on down-cast, 2 low-precision LU failed
    int* iter;          // refinement steps taken
    int* nrhs_eff;      // nrhs while active, 0 once finished (drives the solves)
    int* active_count;  // [0] systems still refining
    double* anrm;       // ||A||_inf
};

// ---------------------------------------------------------------------------
// A (FP64) -> As (FP32, ld = n) with ||A||_inf and an overflow flag. A system
// that overflows gets lu_n = 0: the low-precision LU skips it (it goes to the
// FP64 fallback), so no arithmetic runs on its infinite copy.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kIrThreads)
ir_lag2s_kernel(int n, double** __restrict__ dA_array, int ldda, float* __restrict__ As, int* __restrict__ fail,
                double* __restrict__ anrm, int* __restrict__ lu_n)
{
    __shared__ double red[kIrThreads / kWave];
    __shared__ int ovf;
    const int s = blockIdx.x;
    const double* A = dA_array[s];
    float* S = As + (size_t)s * n * n;
    const int t = threadIdx.x;
    if (t == 0) ovf = 0;
    __syncthreads();
    double rmax = 0.0;
    for (int i = t; i < n; i += kIrThreads) {
        double rs = 0.0;
        for (int j = 0; j < n; ++j) {
            const double a = A[i + (size_t)j * ldda];
            rs += fabs(a);
            if (fabs(a) > (double)FLT_MAX) ovf = 1;
            S[i + (size_t)j * n] = (float)a;
        }
        rmax = fmax(rmax, rs);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) rmax = fmax(rmax, shfl_xor_d(rmax, off));
    if ((t & 63) == 0) red[t / 64] = rmax;
    __syncthreads();
    if (t == 0) {
        double m = red[0];
        for (int q = 1; q < kIrThreads / kWave; ++q) m = fmax(m, red[q]);
        anrm[s] = m;
        fail[s] = ovf ? 1 : 0;
        lu_n[s] = ovf ? 0 : n;
    }
}

// Low-precision LU must be nonsingular and finite.
__global__ void __launch_bounds__(kIrThreads)
ir_check_lu_kernel(int n, const float* __restrict__ As, const int* __restrict__ info, int* __restrict__ fail)
{
    __shared__ int bad;
    const int s = blockIdx.x;
    if (fail[s] != 0) return;             // already headed for the fallback
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    const float* S = As + (size_t)s * n * n;
    for (size_t idx = threadIdx.x; idx < (size_t)n * n; idx += kIrThreads)
        if (!isfinite(S[idx])) bad = 1;
    __syncthreads();
    if (threadIdx.x == 0 && (bad || info[s] != 0)) fail[s] = 2;
}

__global__ void ir_ptrs_kernel(float** out, float* base, long stride, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = base + (size_t)i * stride;
}

__global__ void ir_ptrs_i_kernel(int** out, int* base, long stride, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = base + (size_t)i * stride;
}

// Start: R = B, X = 0 for the systems that refine; failed ones are finished
// at once (they go to the FP64 fallback).
__global__ void ir_init_kernel(int n, int nrhs, double** __restrict__ dB_array, int lddb,
                               double** __restrict__ dX_array, int lddx, double* __restrict__ R, IrState st)
{
    const int s = blockIdx.y;
    const bool ok = st.fail[s] == 0;
    const double* B = dB_array[s];
    double* X = dX_array[s];
    double* Rs = R + (size_t)s * n * nrhs;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n * nrhs; idx += gridDim.x * blockDim.x) {
        const int i = idx % n, c = idx / n;
        Rs[idx] = B[i + (size_t)c * lddb];
        X[i + (size_t)c * lddx] = 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st.done[s] = ok ? 0 : 0;
        st.iter[s] = 0;
        st.nrhs_eff[s] = ok ? nrhs : 0;
        if (ok) atomicAdd(st.active_count, 1);
    }
}

// Rs (FP32) = R; after the FP32 solve, X += Rs.
__global__ void ir_r2s_kernel(int n, int nrhs, const double* __restrict__ R, float* __restrict__ Rs,
                              const int* __restrict__ nrhs_eff)
{
    const int s = blockIdx.y;
    if (nrhs_eff[s] == 0) return;
    const size_t base = (size_t)s * n * nrhs;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n * nrhs; idx += gridDim.x * blockDim.x)
        Rs[base + idx] = (float)R[base + idx];
}

__global__ void ir_axpy_kernel(int n, int nrhs, const float* __restrict__ Rs, double** __restrict__ dX_array,
                               int lddx, const int* __restrict__ nrhs_eff)
{
    const int s = blockIdx.y;
    if (nrhs_eff[s] == 0) return;
    double* X = dX_array[s];
    const size_t base = (size_t)s * n * nrhs;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n * nrhs; idx += gridDim.x * blockDim.x) {
        const int i = idx % n, c = idx / n;
        X[i + (size_t)c * lddx] += (double)Rs[base + idx];
    }
}

// R = B - op(A) X in FP64 and the stopping test of src/dxgesv_gmres_gpu.cpp:
// converged when ||r_c||_inf <= ||x_c||_inf * cte for every column c, with
// both norms finite. A non-finite residual or solution (the refinement
// diverged) ends the system at once: it goes to the FP64 fallback.
// One workgroup per system; X staged in LDS column by column.
__global__ void __launch_bounds__(kIrThreads)
ir_residual_kernel(int trans, int n, int nrhs, double** __restrict__ dA_array, int ldda,
                   double** __restrict__ dB_array, int lddb, double** __restrict__ dX_array, int lddx,
                   double* __restrict__ R, IrState st, double eps, int step, int check)
{
    extern __shared__ __align__(16) unsigned char ir_raw[];
    double* xs = reinterpret_cast<double*>(ir_raw);          // n
    __shared__ double red[2][kIrThreads / kWave];
    __shared__ int all_ok;
    const int s = blockIdx.x;
    if (st.nrhs_eff[s] == 0) return;
    const double* A = dA_array[s];
    const double* B = dB_array[s];
    const double* X = dX_array[s];
    double* Rs = R + (size_t)s * n * nrhs;
    const int t = threadIdx.x;
    const double cte = st.anrm[s] * eps * sqrt((double)n);
    if (t == 0) all_ok = 1;
    for (int c = 0; c < nrhs; ++c) {
        __syncthreads();
        for (int i = t; i < n; i += kIrThreads) xs[i] = X[i + (size_t)c * lddx];
        __syncthreads();
        double rmax = 0.0, xmax = 0.0;
        for (int i = t; i < n; i += kIrThreads) {
            double acc = B[i + (size_t)c * lddb];
            if (trans == 0) {
                for (int j = 0; j < n; ++j) acc = fma(-A[i + (size_t)j * ldda], xs[j], acc);
            } else {
                for (int j = 0; j < n; ++j) acc = fma(-A[j + (size_t)i * ldda], xs[j], acc);
            }
            Rs[i + (size_t)c * n] = acc;
            // fmax drops NaN; a NaN entry counts as +inf instead
            const double ra = fabs(acc), xa = fabs(xs[i]);
            rmax = (ra == ra) ? fmax(rmax, ra) : INFINITY;
            xmax = (xa == xa) ? fmax(xmax, xa) : INFINITY;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            rmax = fmax(rmax, shfl_xor_d(rmax, off));
            xmax = fmax(xmax, shfl_xor_d(xmax, off));
        }
        if ((t & 63) == 0) { red[0][t / 64] = rmax; red[1][t / 64] = xmax; }
        __syncthreads();
        if (t == 0) {
            double rm = red[0][0], xm = red[1][0];
            for (int q = 1; q < kIrThreads / kWave; ++q) { rm = fmax(rm, red[0][q]); xm = fmax(xm, red[1][q]); }
            if (!isfinite(rm) || !isfinite(xm)) all_ok = -1;
            else if (all_ok == 1 && !(rm <= xm * cte)) all_ok = 0;
        }
    }
    __syncthreads();
    if (t == 0 && check) {
        st.iter[s] = step;
        if (all_ok != 0) {
            st.done[s] = all_ok;          // 1 converged, -1 diverged
            st.nrhs_eff[s] = 0;
            atomicSub(st.active_count, 1);
        }
    }
}

// ---------------------------------------------------------------------------
// FGMRES (one right-hand side column at a time), per-system state on device.
// ---------------------------------------------------------------------------
struct GmState {
    double* V;       // n x (m+1) per system
    double* Z;       // n x m
    double* H;       // (m+1) x m, column-major
    double* cs;      // m
    double* sn;      // m
    double* g;       // m+1
    double* eps1;    // stopping threshold (set on the first restart)
    double* ro;      // current residual estimate
    int* inner;      // columns of the current cycle
    int* cycle_done; // 1 once the current cycle has stopped
    int* cycle_open; // [0] systems whose current cycle is still open
    int* total;      // total inner iterations
    int m;           // restart length
};

// r = b - op(A) x (FP64) -> V_0 = r / ||r||_2; g = [beta, 0...]; eps1 on the
// first cycle. A zero residual finishes the system.
__global__ void __launch_bounds__(kIrThreads)
gm_start_kernel(int trans, int n, int col, double** __restrict__ dA_array, int ldda, double** __restrict__ dB_array,
                int lddb, double** __restrict__ dX_array, int lddx, IrState st, GmState gs, double eps, int first)
{
    extern __shared__ __align__(16) unsigned char gm_raw[];
    double* xs = reinterpret_cast<double*>(gm_raw);
    __shared__ double red[kIrThreads / kWave];
    __shared__ double beta_s;
    const int s = blockIdx.x;
    if (st.nrhs_eff[s] == 0) return;
    const double* A = dA_array[s];
    const double* b = dB_array[s] + (size_t)col * lddb;
    const double* x = dX_array[s] + (size_t)col * lddx;
    const int m = gs.m;
    double* V = gs.V + (size_t)s * n * (m + 1);
    const int t = threadIdx.x;
    for (int i = t; i < n; i += kIrThreads) xs[i] = x[i];
    __syncthreads();
    double ss = 0.0;
    for (int i = t; i < n; i += kIrThreads) {
        double acc = b[i];
        if (trans == 0) {
            for (int j = 0; j < n; ++j) acc = fma(-A[i + (size_t)j * ldda], xs[j], acc);
        } else {
            for (int j = 0; j < n; ++j) acc = fma(-A[j + (size_t)i * ldda], xs[j], acc);
        }
        V[i] = acc;
        ss = fma(acc, acc, ss);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) ss += shfl_xor_d(ss, off);
    if ((t & 63) == 0) red[t / 64] = ss;
    __syncthreads();
    if (t == 0) {
        double tot = 0.0;
        for (int q = 0; q < kIrThreads / kWave; ++q) tot += red[q];
        beta_s = sqrt(tot);
        double* g = gs.g + (size_t)s * (m + 1);
        for (int q = 0; q <= m; ++q) g[q] = 0.0;
        g[0] = beta_s;
        if (first) gs.eps1[s] = st.anrm[s] * eps * sqrt((double)n) * beta_s;
        gs.ro[s] = beta_s;
        gs.inner[s] = 0;
        // a zero residual needs no cycle; a non-finite one ends in the fallback
        gs.cycle_done[s] = (beta_s == 0.0 || !isfinite(beta_s)) ? 1 : 0;
        if (gs.cycle_done[s] == 0) atomicAdd(gs.cycle_open, 1);
    }
    __syncthreads();
    const double beta = beta_s;
    if (beta != 0.0 && isfinite(beta))
        for (int i = t; i < n; i += kIrThreads) V[i] = V[i] / beta;
}

// Rs = (float) V_i for the preconditioner solve (systems in an open cycle).
__global__ void gm_v2s_kernel(int n, const double* __restrict__ Vall, int m, int i, float* __restrict__ Rs,
                              const int* __restrict__ nrhs_eff, const int* __restrict__ cycle_done,
                              int* __restrict__ solve_flag)
{
    const int s = blockIdx.y;
    const bool on = nrhs_eff[s] != 0 && cycle_done[s] == 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) solve_flag[s] = on ? 1 : 0;
    if (!on) return;
    const double* V = Vall + (size_t)s * n * (m + 1) + (size_t)i * n;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x)
        Rs[(size_t)s * n + idx] = (float)V[idx];
}

// Z_i = (double) Rs; V_{i+1} = op(A) Z_i; modified Gram-Schmidt against
// V_0..V_i; H column i; Givens rotations; residual estimate; stop test.
__global__ void __launch_bounds__(kIrThreads)
gm_arnoldi_kernel(int trans, int n, int i, double** __restrict__ dA_array, int ldda, const float* __restrict__ Rs,
                  IrState st, GmState gs, int maxiter)
{
    extern __shared__ __align__(16) unsigned char ga_raw[];
    double* zs = reinterpret_cast<double*>(ga_raw);              // n
    double* ws = zs + n;                                          // n
    __shared__ double red[kIrThreads / kWave];
    __shared__ double bc;
    const int s = blockIdx.x;
    if (st.nrhs_eff[s] == 0 || gs.cycle_done[s] != 0) return;
    const int m = gs.m;
    const double* A = dA_array[s];
    double* V = gs.V + (size_t)s * n * (m + 1);
    double* Z = gs.Z + (size_t)s * n * m;
    double* H = gs.H + (size_t)s * (m + 1) * m;
    const int t = threadIdx.x;
    for (int r = t; r < n; r += kIrThreads) {
        const double z = (double)Rs[(size_t)s * n + r];
        zs[r] = z;
        Z[(size_t)i * n + r] = z;
    }
    __syncthreads();
    for (int r = t; r < n; r += kIrThreads) {
        double acc = 0.0;
        if (trans == 0) {
            for (int j = 0; j < n; ++j) acc = fma(A[r + (size_t)j * ldda], zs[j], acc);
        } else {
            for (int j = 0; j < n; ++j) acc = fma(A[j + (size_t)r * ldda], zs[j], acc);
        }
        ws[r] = acc;
    }
    __syncthreads();
    auto block_sum = [&](double v) -> double {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += shfl_xor_d(v, off);
        __syncthreads();
        if ((t & 63) == 0) red[t / 64] = v;
        __syncthreads();
        double tot = 0.0;
        for (int q = 0; q < kIrThreads / kWave; ++q) tot += red[q];
        return tot;
    };
    for (int j = 0; j <= i; ++j) {
        const double* Vj = V + (size_t)j * n;
        double d = 0.0;
        for (int r = t; r < n; r += kIrThreads) d = fma(Vj[r], ws[r], d);
        const double h = block_sum(d);
        if (t == 0) H[(size_t)i * (m + 1) + j] = h;
        for (int r = t; r < n; r += kIrThreads) ws[r] = fma(-h, Vj[r], ws[r]);
        __syncthreads();
    }
    double d = 0.0;
    for (int r = t; r < n; r += kIrThreads) d = fma(ws[r], ws[r], d);
    const double hn = sqrt(block_sum(d));
    double* Vn = V + (size_t)(i + 1) * n;
    for (int r = t; r < n; r += kIrThreads) Vn[r] = hn != 0.0 ? ws[r] / hn : ws[r];
    if (t == 0) {
        double* col = H + (size_t)i * (m + 1);
        col[i + 1] = hn;
        double* cs = gs.cs + (size_t)s * m;
        double* sn = gs.sn + (size_t)s * m;
        double* g = gs.g + (size_t)s * (m + 1);
        for (int k = 0; k < i; ++k) {                      // previous rotations
            const double a = col[k], b2 = col[k + 1];
            col[k] = cs[k] * a + sn[k] * b2;
            col[k + 1] = -sn[k] * a + cs[k] * b2;
        }
        double gam = sqrt(col[i] * col[i] + col[i + 1] * col[i + 1]);
        if (gam == 0.0) gam = DBL_MIN;
        cs[i] = col[i] / gam;
        sn[i] = col[i + 1] / gam;
        g[i + 1] = -sn[i] * g[i];
        g[i] = cs[i] * g[i];
        col[i] = cs[i] * col[i] + sn[i] * col[i + 1];
        col[i + 1] = 0.0;
        const double ro = fabs(g[i + 1]);
        gs.ro[s] = ro;
        gs.inner[s] = i + 1;
        gs.total[s] += 1;
        if (ro <= gs.eps1[s] || i + 1 >= m || gs.total[s] >= maxiter) {
            gs.cycle_done[s] = 1;
            atomicSub(gs.cycle_open, 1);
        }
    }
}

// End of a cycle: y = H^{-1} g, x += Z y; converged when ro <= eps1.
__global__ void __launch_bounds__(kIrThreads)
gm_update_kernel(int n, int col, double** __restrict__ dX_array, int lddx, IrState st, GmState gs, int maxiter)
{
    __shared__ double y[kGmresMaxRestart + 1];
    const int s = blockIdx.x;
    if (st.nrhs_eff[s] == 0) return;
    const int m = gs.m;
    const int k = gs.inner[s];
    const double* H = gs.H + (size_t)s * (m + 1) * m;
    const double* g = gs.g + (size_t)s * (m + 1);
    const double* Z = gs.Z + (size_t)s * n * m;
    const int t = threadIdx.x;
    if (t == 0) {
        for (int r = k - 1; r >= 0; --r) {
            double v = g[r];
            for (int c = r + 1; c < k; ++c) v -= H[(size_t)c * (m + 1) + r] * y[c];
            y[r] = v / H[(size_t)r * (m + 1) + r];
        }
    }
    __syncthreads();
    double* x = dX_array[s] + (size_t)col * lddx;
    for (int r = t; r < n; r += kIrThreads) {
        double acc = x[r];
        for (int c = 0; c < k; ++c) acc = fma(y[c], Z[(size_t)c * n + r], acc);
        x[r] = acc;
    }
    if (t == 0) {
        st.iter[s] = max(st.iter[s], gs.total[s]);
        if (gs.ro[s] <= gs.eps1[s]) {
            st.nrhs_eff[s] = 0;          // this column converged
            atomicSub(st.active_count, 1);
        } else if (gs.total[s] >= maxiter || !isfinite(gs.ro[s])) {
            st.nrhs_eff[s] = 0;
            st.done[s] = -1;             // not converged or diverged: FP64 fallback
            atomicSub(st.active_count, 1);
        }
    }
}

__global__ void gm_reset_kernel(int batch, int nrhs, IrState st, GmState gs)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= batch) return;
    gs.total[s] = 0;
    const bool on = st.fail[s] == 0 && st.done[s] >= 0;
    st.nrhs_eff[s] = on ? 1 : 0;
    if (on) atomicAdd(st.active_count, 1);
    (void)nrhs;
}

__global__ void ir_copy_kernel(int n, int nrhs, const int* __restrict__ map, double** __restrict__ dB_array,
                               int lddb, double** __restrict__ dX_array, int lddx)
{
    const int s = map[blockIdx.y];
    const double* B = dB_array[s];
    double* X = dX_array[s];
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n * nrhs; idx += gridDim.x * blockDim.x) {
        const int i = idx % n, c = idx / n;
        X[i + (size_t)c * lddx] = B[i + (size_t)c * lddb];
    }
}

__global__ void gm_finish_kernel(int batch, IrState st)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= batch) return;
    if (st.fail[s] == 0 && st.done[s] == 0) st.done[s] = 1;
}

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------
static int mode_of(magma_refinement_t f)
{
    switch (f) {
        case Magma_PREC_SS: case Magma_PREC_XSH_S: return kUpdF32;
        case Magma_PREC_SHT: case Magma_PREC_XSH_163232TC: return kUpdFP16;
        case Magma_PREC_SBT: return kUpdBF16;
        case Magma_PREC_SB3T: return kUpdBF16x3;
        case Magma_PREC_SF8T: return kUpdFP8;
        default: return -1;
    }
}

template <typename T>
static T* carve(char*& p, size_t count)
{
    T* r = reinterpret_cast<T*>(p);
    p += (count * sizeof(T) + 255) & ~size_t(255);
    return r;
}

// Batched refinement core. Inputs are left unchanged unless a system falls
// back to FP64 (then its A holds the FP64 LU and ipiv_fp64 its pivots).
// iter_out[s]: >= 0 refinement steps, -2 overflow, -3 low-precision LU failed,
// -k not converged after k steps. info_out[s]: FP64 fallback info.
magma_int_t iteref_batched(int trans, int n, int nrhs, double** dA_array, int ldda, double** dB_array, int lddb,
                           double** dX_array, int lddx, int mode, bool gmres, int itermax, int restart,
                           int* iter_out, int* info_out, int** lp_ipiv_out, int batch, magma_queue_t queue,
                           float* ws_float /* nullable: n*n*batch floats */)
{
    hipStream_t stream = queue->stream;
    const int m = std::max(1, std::min({kGmresMaxRestart, restart, itermax, n}));
    // workspace layout
    size_t bytes = 0;
    auto add = [&](size_t b) { bytes += (b + 255) & ~size_t(255); };
    add(ws_float ? 0 : sizeof(float) * (size_t)n * n * batch);        // As
    add(sizeof(float) * (size_t)n * std::max(nrhs, 1) * batch);       // Rs
    add(sizeof(double) * (size_t)n * std::max(nrhs, 1) * batch);      // R
    add(sizeof(int) * (size_t)n * batch);                             // ipiv
    add(sizeof(float*) * batch);
    add(sizeof(int*) * batch);
    add(sizeof(float*) * batch);                                      // Rs pointers
    add(sizeof(int) * batch * 7 + 256);                               // states
    add(sizeof(double) * batch);                                      // anrm
    if (gmres) {
        add(sizeof(double) * (size_t)n * (m + 1) * batch);
        add(sizeof(double) * (size_t)n * m * batch);
        add(sizeof(double) * (size_t)(m + 1) * m * batch);
        add(sizeof(double) * (size_t)(3 * m + 1 + 2) * batch);
        add(sizeof(int) * batch * 4);
    }
    char* base = (char*)queue_scratch(queue, bytes);
    if (base == nullptr) return MAGMA_ERR_DEVICE_ALLOC;
    char* p = base;
    float* As = ws_float ? ws_float : carve<float>(p, (size_t)n * n * batch);
    float* Rs = carve<float>(p, (size_t)n * std::max(nrhs, 1) * batch);
    double* R = carve<double>(p, (size_t)n * std::max(nrhs, 1) * batch);
    int* ipiv = carve<int>(p, (size_t)n * batch);
    float** As_ptr = carve<float*>(p, batch);
    int** ipiv_ptr = carve<int*>(p, batch);
    float** Rs_ptr = carve<float*>(p, batch);
    int* ints = carve<int>(p, (size_t)batch * 7 + 64);
    double* anrm = carve<double>(p, batch);
    IrState st{ints, ints + batch, ints + 2 * batch, ints + 3 * batch, ints + 7 * batch, anrm};
    int* lp_info = ints + 4 * batch;
    int* solve_flag = ints + 5 * batch;
    int* lu_n = ints + 6 * batch;       // per-system order of the low-precision LU (0 = skipped)
    GmState gs{};
    if (gmres) {
        gs.V = carve<double>(p, (size_t)n * (m + 1) * batch);
        gs.Z = carve<double>(p, (size_t)n * m * batch);
        gs.H = carve<double>(p, (size_t)(m + 1) * m * batch);
        double* dd = carve<double>(p, (size_t)(3 * m + 1 + 2) * batch);
        gs.cs = dd;
        gs.sn = dd + (size_t)m * batch;
        gs.g = dd + (size_t)2 * m * batch;
        gs.eps1 = dd + (size_t)(3 * m + 1) * batch;
        gs.ro = gs.eps1 + batch;
        int* gi = carve<int>(p, (size_t)batch * 4);
        gs.inner = gi;
        gs.cycle_done = gi + batch;
        gs.total = gi + 2 * batch;
        gs.cycle_open = ints + 7 * batch + 1;
        gs.m = m;
    }
    const double eps = DBL_EPSILON / 2;   // dlamch('Epsilon'): relative machine precision
    const int g256 = ceildiv(batch, 256);

    ir_ptrs_kernel<<<g256, 256, 0, stream>>>(As_ptr, As, (long)n * n, batch);
    ir_ptrs_i_kernel<<<g256, 256, 0, stream>>>(ipiv_ptr, ipiv, (long)n, batch);
    ir_ptrs_kernel<<<g256, 256, 0, stream>>>(Rs_ptr, Rs, (long)n * std::max(nrhs, 1), batch);
    MGI_LAUNCH_CHECK();
    // 1. down-cast with overflow check and ||A||_inf
    ir_lag2s_kernel<<<batch, kIrThreads, 0, stream>>>(n, dA_array, ldda, As, st.fail, anrm, lu_n);
    MGI_LAUNCH_CHECK();
    // 2. low-precision LU of the systems whose down-cast stayed finite
    magma_int_t rc = sgetrf_blocked_launch(mode, n, n, n, lu_n, lu_n, nullptr, n, n, As_ptr, ipiv_ptr, lp_info,
                                           batch, stream);
    if (rc != MAGMA_SUCCESS) return rc;
    ir_check_lu_kernel<<<batch, kIrThreads, 0, stream>>>(n, As, lp_info, st.fail);
    MGI_LAUNCH_CHECK();

    const size_t xs_lds = sizeof(double) * n;
    int h_active = 0;
    if (!gmres) {
        // 3a. classical refinement, all columns together
        MGI_HIP_CHECK(hipMemsetAsync(st.active_count, 0, sizeof(int), stream));
        dim3 g2(std::max(1, std::min(64, ceildiv(n * nrhs, 256))), batch);
        ir_init_kernel<<<g2, 256, 0, stream>>>(n, nrhs, dB_array, lddb, dX_array, lddx, R, st);
        MGI_LAUNCH_CHECK();
        RsSizes rs{n, nrhs, n, n, n, nullptr, st.nrhs_eff, nullptr, nullptr, nullptr};
        for (int step = 0; step <= itermax; ++step) {
            ir_r2s_kernel<<<g2, 256, 0, stream>>>(n, nrhs, R, Rs, st.nrhs_eff);
            rc = getrs_generic_launch_t<float>(trans, rs, nrhs, As_ptr, ipiv_ptr, Rs_ptr, batch, stream);
            if (rc != MAGMA_SUCCESS) return rc;
            ir_axpy_kernel<<<g2, 256, 0, stream>>>(n, nrhs, Rs, dX_array, lddx, st.nrhs_eff);
            ir_residual_kernel<<<batch, kIrThreads, xs_lds, stream>>>(trans, n, nrhs, dA_array, ldda, dB_array, lddb,
                                                                       dX_array, lddx, R, st, eps, step, 1);
            MGI_LAUNCH_CHECK();
            MGI_HIP_CHECK(hipMemcpyAsync(&h_active, st.active_count, sizeof(int), hipMemcpyDeviceToHost, stream));
            MGI_HIP_CHECK(hipStreamSynchronize(stream));
            if (h_active == 0) break;
        }
    } else {
        // 3b. FGMRES per column (src/dgmres_plu_gpu.cpp:79-316). One vector
        // per system is preconditioned at a time: Rs holds it at stride n.
        ir_ptrs_kernel<<<g256, 256, 0, stream>>>(Rs_ptr, Rs, (long)n, batch);
        MGI_LAUNCH_CHECK();
        MGI_HIP_CHECK(hipMemsetAsync(st.done, 0, sizeof(int) * batch, stream));
        MGI_HIP_CHECK(hipMemsetAsync(st.iter, 0, sizeof(int) * batch, stream));
        dim3 g2(std::max(1, std::min(64, ceildiv(n * nrhs, 256))), batch);
        // X = 0 (ir_init_kernel also primes R, unused here)
        MGI_HIP_CHECK(hipMemsetAsync(st.active_count, 0, sizeof(int), stream));
        ir_init_kernel<<<g2, 256, 0, stream>>>(n, nrhs, dB_array, lddb, dX_array, lddx, R, st);
        MGI_LAUNCH_CHECK();
        RsSizes rs1{n, 1, n, n, n, nullptr, solve_flag, nullptr, nullptr, nullptr};
        for (int col = 0; col < nrhs; ++col) {
            MGI_HIP_CHECK(hipMemsetAsync(st.active_count, 0, sizeof(int), stream));
            gm_reset_kernel<<<g256, 256, 0, stream>>>(batch, nrhs, st, gs);
            int first = 1;
            for (int cycle = 0; cycle <= itermax; ++cycle) {
                MGI_HIP_CHECK(hipMemsetAsync(gs.cycle_open, 0, sizeof(int), stream));
                gm_start_kernel<<<batch, kIrThreads, xs_lds, stream>>>(trans, n, col, dA_array, ldda, dB_array, lddb,
                                                                       dX_array, lddx, st, gs, eps, first);
                MGI_LAUNCH_CHECK();
                first = 0;
                for (int i = 0; i < m; ++i) {
                    // every few inner steps: stop the cycle once every system's is closed
                    if (i > 0 && i % kGmresCheck == 0) {
                        int h_open = 0;
                        MGI_HIP_CHECK(hipMemcpyAsync(&h_open, gs.cycle_open, sizeof(int), hipMemcpyDeviceToHost,
                                                     stream));
                        MGI_HIP_CHECK(hipStreamSynchronize(stream));
                        if (h_open == 0) break;
                    }
                    gm_v2s_kernel<<<dim3(std::max(1, ceildiv(n, 256)), batch), 256, 0, stream>>>(
                        n, gs.V, m, i, Rs, st.nrhs_eff, gs.cycle_done, solve_flag);
                    rc = getrs_generic_launch_t<float>(trans, rs1, 1, As_ptr, ipiv_ptr, Rs_ptr, batch, stream);
                    if (rc != MAGMA_SUCCESS) return rc;
                    gm_arnoldi_kernel<<<batch, kIrThreads, 2 * xs_lds, stream>>>(trans, n, i, dA_array, ldda, Rs, st,
                                                                                 gs, itermax);
                    MGI_LAUNCH_CHECK();
                }
                gm_update_kernel<<<batch, kIrThreads, 0, stream>>>(n, col, dX_array, lddx, st, gs, itermax);
                MGI_LAUNCH_CHECK();
                MGI_HIP_CHECK(hipMemcpyAsync(&h_active, st.active_count, sizeof(int), hipMemcpyDeviceToHost, stream));
                MGI_HIP_CHECK(hipStreamSynchronize(stream));
                if (h_active == 0) break;
            }
        }
        gm_finish_kernel<<<g256, 256, 0, stream>>>(batch, st);
        MGI_LAUNCH_CHECK();
    }

    // 4. outcomes and FP64 fallback for the rest
    std::vector<int> h_done(batch), h_fail(batch), h_iter(batch);
    MGI_HIP_CHECK(hipMemcpyAsync(h_done.data(), st.done, sizeof(int) * batch, hipMemcpyDeviceToHost, stream));
    MGI_HIP_CHECK(hipMemcpyAsync(h_fail.data(), st.fail, sizeof(int) * batch, hipMemcpyDeviceToHost, stream));
    MGI_HIP_CHECK(hipMemcpyAsync(h_iter.data(), st.iter, sizeof(int) * batch, hipMemcpyDeviceToHost, stream));
    MGI_HIP_CHECK(hipStreamSynchronize(stream));
    std::vector<int> fb;
    std::vector<int> h_it(batch);
    for (int s = 0; s < batch; ++s) {
        if (h_fail[s] == 1) { h_it[s] = -2; fb.push_back(s); }
        else if (h_fail[s] == 2) { h_it[s] = -3; fb.push_back(s); }
        else if (h_done[s] != 1) { h_it[s] = -std::max(1, h_iter[s]); fb.push_back(s); }
        else h_it[s] = h_iter[s];
    }
    MGI_HIP_CHECK(hipMemcpyAsync(iter_out, h_it.data(), sizeof(int) * batch, hipMemcpyHostToDevice, stream));
    MGI_HIP_CHECK(hipMemsetAsync(info_out, 0, sizeof(int) * batch, stream));
    if (lp_ipiv_out) *lp_ipiv_out = ipiv;
    if (!fb.empty()) {
        // FP64 LU + solve of the fallback systems in place (A -> LU, X = A \ B)
        int* d_map = nullptr;
        int* d_ipiv64 = nullptr;
        int** d_ipiv64_ptr = nullptr;
        const int nf = (int)fb.size();
        MGI_HIP_CHECK(hipMalloc(&d_map, sizeof(int) * nf));
        MGI_HIP_CHECK(hipMalloc(&d_ipiv64, sizeof(int) * (size_t)n * batch));
        MGI_HIP_CHECK(hipMalloc(&d_ipiv64_ptr, sizeof(int*) * batch));
        MGI_HIP_CHECK(hipMemcpyAsync(d_map, fb.data(), sizeof(int) * nf, hipMemcpyHostToDevice, stream));
        ir_ptrs_i_kernel<<<g256, 256, 0, stream>>>(d_ipiv64_ptr, d_ipiv64, (long)n, batch);
        magma_int_t frc;
        if (getrf_tile_supported(n)) {
            TileArgs a{n, nrhs, ldda, lddx, d_map, nullptr, nullptr, nullptr, nullptr};
            frc = getrf_tile_launch_args((n + 15) / 16, a, dA_array, d_ipiv64_ptr, info_out, nf, stream);
        } else {
            SizeArgs sz{n, n, ldda, nullptr, nullptr, nullptr, d_map};
            frc = getrf_blocked_launch_sz(sz, n, n, dA_array, d_ipiv64_ptr, info_out, nf, stream);
        }
        if (frc == MAGMA_SUCCESS) {
            // X = B for the fallback systems, then solve in place
            ir_copy_kernel<<<dim3(std::max(1, std::min(64, ceildiv(n * nrhs, 256))), nf), 256, 0, stream>>>(
                n, nrhs, d_map, dB_array, lddb, dX_array, lddx);
            MGI_LAUNCH_CHECK();
            RsSizes rs{n, nrhs, ldda, lddx, n, nullptr, nullptr, nullptr, nullptr, d_map};
            frc = getrs_generic_launch_t<double>(trans, rs, nrhs, dA_array, d_ipiv64_ptr, dX_array, nf, stream);
        }
        MGI_HIP_CHECK(hipStreamSynchronize(stream));
        if (lp_ipiv_out) {
            // the caller wants the pivots actually used: FP64 ones for fallbacks
            for (int s : fb)
                MGI_HIP_CHECK(hipMemcpy(ipiv + (size_t)s * n, d_ipiv64 + (size_t)s * n, sizeof(int) * n,
                                        hipMemcpyDeviceToDevice));
        }
        hipFree(d_map);
        hipFree(d_ipiv64);
        hipFree(d_ipiv64_ptr);
        if (frc != MAGMA_SUCCESS) return frc;
    }
    // h_it is pageable host memory read by the queued copy
    MGI_HIP_CHECK(hipStreamSynchronize(stream));
    return MAGMA_SUCCESS;
}

}  // namespace mgi

using namespace mgi;

extern "C" {

magma_int_t magma_dxgesv_iteref_batched(magma_int_t n, magma_int_t nrhs, double** dA_array, magma_int_t ldda,
                                        double** dB_array, magma_int_t lddb, double** dX_array, magma_int_t lddx,
                                        magma_refinement_t facto_type, magma_refinement_t solver_type,
                                        magma_int_t itermax, magma_int_t* iter_array, magma_int_t* info_array,
                                        magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    const int mode = mode_of(facto_type);
    if (n < 0) info = -1;
    else if (nrhs < 0) info = -2;
    else if (ldda < std::max(1, n)) info = -4;
    else if (lddb < std::max(1, n)) info = -6;
    else if (lddx < std::max(1, n)) info = -8;
    else if (mode < 0 || mode == kUpdFP8) info = -9;     // the FP8 experiment runs on the dense LU only
    else if (solver_type != Magma_REFINE_IRSTRS && solver_type != Magma_REFINE_GMSTRS) info = -10;
    else if (itermax < 1) info = -11;
    else if (batchCount < 0) info = -14;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    // the refinement kernels put the batch in grid y
    if (batchCount > kMaxGridYZ) return MAGMA_ERR_NOT_SUPPORTED;
    if (!getrf_blocked_supported(n)) return MAGMA_ERR_NOT_SUPPORTED;
    hipSetDevice(queue->device);
    return iteref_batched(0, n, nrhs, dA_array, ldda, dB_array, lddb, dX_array, lddx, mode,
                          solver_type == Magma_REFINE_GMSTRS, itermax, kGmresBatchedRestart, iter_array, info_array,
                          nullptr, batchCount, queue, nullptr);
}

// src/dxgesv_gmres_gpu.cpp:160 (expert driver) on one system.
magma_int_t magma_dxgesv_gmres_gpu(magma_trans_t trans, magma_int_t n, magma_int_t nrhs, magmaDouble_ptr dA,
                                   magma_int_t ldda, magma_int_t* ipiv, magmaInt_ptr dipiv, magmaDouble_ptr dB,
                                   magma_int_t lddb, magmaDouble_ptr dX, magma_int_t lddx, magmaDouble_ptr dworkd,
                                   magmaFloat_ptr dworks, magma_refinement_t facto_type,
                                   magma_refinement_t solver_type, magma_int_t* iter, magma_int_t* info,
                                   real_Double_t* facto_time)
{
    (void)dworkd;
    *iter = 0;
    *info = 0;
    const int mode = mode_of(facto_type);
    if (trans != MagmaNoTrans && trans != MagmaTrans) *info = -1;
    else if (n < 0) *info = -2;
    else if (nrhs < 0) *info = -3;
    else if (ldda < std::max(1, n)) *info = -5;
    else if (lddb < std::max(1, n)) *info = -9;
    else if (lddx < std::max(1, n)) *info = -11;
    else if (mode < 0) *info = -14;
    else if (solver_type < Magma_REFINE_IRSTRS || solver_type > Magma_REFINE_GMGMDTRS) *info = -15;
    if (*info != 0) {
        magma_xerbla(__func__, -(*info));
        return *info;
    }
    if (n == 0 || nrhs == 0) return 0;
    const bool s_plain = solver_type == Magma_REFINE_IRSTRS || solver_type == Magma_REFINE_GMSTRS;
    if (!s_plain || mode == kUpdFP8 || n > kDenseIrMinN || !getrf_blocked_supported(n)) {
        // one large system, or a mode the batched engine does not run: the
        // dense LU on the whole device
        const double t0 = magma_wtime();
        int it = 0, inf = 0;
        const magma_int_t rc = dense_iteref(trans == MagmaNoTrans ? 0 : 1, n, nrhs, dA, ldda, ipiv, dipiv, dB, lddb,
                                            dX, lddx, dworkd, dworks, mode, (int)(solver_type - Magma_REFINE_IRSTRS),
                                            &it, &inf);
        if (facto_time) *facto_time = magma_wtime() - t0;
        if (rc != MAGMA_SUCCESS) {
            *info = rc;
            return rc;
        }
        *iter = it;
        *info = inf;
        return *info;
    }
    magma_queue_t q = nullptr;
    int dev = 0;
    hipGetDevice(&dev);
    magma_queue_create_internal(dev, &q, __func__, __FILE__, __LINE__);
    if (q == nullptr) {
        *info = MAGMA_ERR_UNKNOWN;
        return MAGMA_ERR_UNKNOWN;
    }
    const double t0 = magma_wtime();
    double** ptrs = nullptr;
    int* small = nullptr;
    magma_int_t rc = MAGMA_SUCCESS;
    if (hipMalloc(&ptrs, sizeof(double*) * 3) != hipSuccess || hipMalloc(&small, sizeof(int) * 2) != hipSuccess) {
        rc = MAGMA_ERR_DEVICE_ALLOC;
    } else {
        double* hp[3] = {dA, dB, dX};
        hipMemcpy(ptrs, hp, sizeof(hp), hipMemcpyHostToDevice);
        int* lp_ipiv = nullptr;
        // The caller's FP32 workspace (n*(n+nrhs) floats) holds the low-precision LU.
        rc = iteref_batched(trans == MagmaNoTrans ? 0 : 1, n, nrhs, ptrs, ldda, ptrs + 1, lddb, ptrs + 2, lddx, mode,
                            solver_type == Magma_REFINE_GMSTRS, std::min(n, 500), std::min(n, 500), small,
                            small + 1, &lp_ipiv, 1, q, dworks);
        if (rc == MAGMA_SUCCESS) {
            int h[2];
            hipMemcpy(h, small, sizeof(h), hipMemcpyDeviceToHost);
            *iter = h[0];
            *info = h[1];
            std::vector<int> hp_ipiv(n), newp(n);
            hipMemcpy(hp_ipiv.data(), lp_ipiv, sizeof(int) * n, hipMemcpyDeviceToHost);
            for (int i = 0; i < n; ++i) ipiv[i] = hp_ipiv[i];
            // dipiv: the row map of the interchanges (control/auxiliary.cpp:81)
            std::vector<int> perm(n);
            for (int i = 0; i < n; ++i) perm[i] = i;
            for (int k = 0; k < n; ++k) std::swap(perm[k], perm[ipiv[k] - 1]);
            if (trans == MagmaNoTrans) newp = perm;
            else for (int i = 0; i < n; ++i) newp[perm[i]] = i;
            if (dipiv) hipMemcpy(dipiv, newp.data(), sizeof(int) * n, hipMemcpyHostToDevice);
        }
    }
    if (facto_time) *facto_time = magma_wtime() - t0;
    hipFree(ptrs);
    hipFree(small);
    magma_queue_destroy_internal(q, __func__, __FILE__, __LINE__);
    if (rc != MAGMA_SUCCESS) {
        *info = rc;
        return rc;
    }
    return *info;
}

// include/magma_htc.h:27: FP16-input LU + FGMRES (src/dxgesv_gmres_gpu.cpp:1143-1152).
magma_int_t magma_dhgesv_iteref_gpu(magma_trans_t trans, magma_int_t n, magma_int_t nrhs, magmaDouble_ptr dA,
                                    magma_int_t ldda, magma_int_t* ipiv, magmaInt_ptr dipiv, magmaDouble_ptr dB,
                                    magma_int_t lddb, magmaDouble_ptr dX, magma_int_t lddx, magmaDouble_ptr dworkd,
                                    magmaFloat_ptr dworks, magma_int_t* iter, magma_int_t* info)
{
    real_Double_t ft;
    return magma_dxgesv_gmres_gpu(trans, n, nrhs, dA, ldda, ipiv, dipiv, dB, lddb, dX, lddx, dworkd, dworks,
                                  Magma_PREC_SHT, Magma_REFINE_GMSTRS, iter, info, &ft);
}

// src/dxgesv_gmres_gpu.cpp:966: FP32 LU + FGMRES.
magma_int_t magma_dsgesv_iteref_gpu(magma_trans_t trans, magma_int_t n, magma_int_t nrhs, magmaDouble_ptr dA,
                                    magma_int_t ldda, magma_int_t* ipiv, magmaInt_ptr dipiv, magmaDouble_ptr dB,
                                    magma_int_t lddb, magmaDouble_ptr dX, magma_int_t lddx, magmaDouble_ptr dworkd,
                                    magmaFloat_ptr dworks, magma_int_t* iter, magma_int_t* info)
{
    real_Double_t ft;
    return magma_dxgesv_gmres_gpu(trans, n, nrhs, dA, ldda, ipiv, dipiv, dB, lddb, dX, lddx, dworkd, dworks,
                                  Magma_PREC_SS, Magma_REFINE_GMSTRS, iter, info, &ft);
}

magma_int_t magma_dbgesv_iteref_gpu(magma_trans_t trans, magma_int_t n, magma_int_t nrhs, magmaDouble_ptr dA,
                                    magma_int_t ldda, magma_int_t* ipiv, magmaInt_ptr dipiv, magmaDouble_ptr dB,
                                    magma_int_t lddb, magmaDouble_ptr dX, magma_int_t lddx, magmaDouble_ptr dworkd,
                                    magmaFloat_ptr dworks, magma_int_t* iter, magma_int_t* info)
{
    real_Double_t ft;
    return magma_dxgesv_gmres_gpu(trans, n, nrhs, dA, ldda, ipiv, dipiv, dB, lddb, dX, lddx, dworkd, dworks,
                                  Magma_PREC_SBT, Magma_REFINE_GMSTRS, iter, info, &ft);
}

}  // extern "C"
