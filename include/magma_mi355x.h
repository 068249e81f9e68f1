/*
 * magma_mi355x.h -- C ABI of the MI355X-native batched LU factor/solve library
 * (libmagma_mi355x.so).
 *
 * Every declaration below keeps the exact signature, argument meaning, return
 * convention and enum values of the MAGMA 2.10 routine it replaces, so that a
 * caller compiled against MAGMA's headers links against this library instead.
 * Each entry cites the MAGMA prototype (file:line, relative to the MAGMA source
 * tree) whose double-precision ("d") codegen form it replaces.
 *
 * magma_int_t is 32-bit (MAGMA's default LP64 build, include/magma_types.h:68-72).
 * Device pointers are plain HIP device addresses; pointer arrays (double **,
 * magma_int_t **) live in device memory, as in MAGMA.
 */
#ifndef MAGMA_MI355X_H
#define MAGMA_MI355X_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Types (include/magma_types.h:68-72, 100, 227)                              */
/* ------------------------------------------------------------------------- */
typedef int            magma_int_t;
typedef magma_int_t    magma_device_t;
typedef double         real_Double_t;
typedef void*          magma_ptr;
typedef double*        magmaDouble_ptr;
typedef float*         magmaFloat_ptr;
typedef magma_int_t*   magmaInt_ptr;

struct magma_queue;                       /* opaque: wraps one hipStream_t */
typedef struct magma_queue* magma_queue_t;

/* Return codes (include/magma_types.h:549-568) */
#define MAGMA_SUCCESS               0
#define MAGMA_ERR                  -100
#define MAGMA_ERR_NOT_INITIALIZED  -101
#define MAGMA_ERR_NOT_SUPPORTED    -103
#define MAGMA_ERR_ILLEGAL_VALUE    -104
#define MAGMA_ERR_HOST_ALLOC       -112
#define MAGMA_ERR_DEVICE_ALLOC     -113
#define MAGMA_ERR_INVALID_PTR      -115
#define MAGMA_ERR_UNKNOWN          -116
#define MAGMA_ERR_NOT_IMPLEMENTED  -117

/* Parameter enums (include/magma_types.h:613-640), same numeric values. */
typedef enum { MagmaFalse = 0, MagmaTrue = 1 } magma_bool_t;
typedef enum { MagmaNoTrans = 111, MagmaTrans = 112, MagmaConjTrans = 113 } magma_trans_t;
typedef enum { MagmaUpper = 121, MagmaLower = 122, MagmaFull = 123 } magma_uplo_t;
typedef enum { MagmaNonUnit = 131, MagmaUnit = 132 } magma_diag_t;
typedef enum { MagmaLeft = 141, MagmaRight = 142 } magma_side_t;

/* Mixed-precision refinement selectors (include/magma_types.h:861-898).
 * The reference values are kept; the two BF16 factorisation kinds are new
 * (MAGMA has no BF16 LU) and use values the reference leaves unassigned. */
typedef enum {
    Magma_PREC_SS         = 900,   /* FP32 LU, FP32 GEMM                          */
    Magma_PREC_SHT        = 905,   /* FP32 LU, FP16-input GEMM, FP32 accumulate    */
    Magma_PREC_XSH_S      = 915,
    Magma_PREC_XSH_163232TC = 917,
    Magma_PREC_SBT        = 940,   /* new: FP32 LU, BF16-input GEMM, FP32 accum.   */
    Magma_PREC_SB3T       = 941,   /* new: FP32 LU, 3-term BF16 split GEMM         */
    Magma_PREC_SF8T       = 942,   /* new: FP32 LU, FP8 (e4m3) GEMM, FP32 accum.;
                                      an experiment, one-system drivers only       */
    Magma_REFINE_IRSTRS   = 920,   /* classical IR, low-precision triangular solve */
    Magma_REFINE_IRDTRS   = 921,
    Magma_REFINE_IRGMSTRS = 922,
    Magma_REFINE_IRGMDTRS = 923,
    Magma_REFINE_GMSTRS   = 924,   /* FGMRES preconditioned by the low-prec. LU    */
    Magma_REFINE_GMDTRS   = 925,
    Magma_REFINE_GMGMSTRS = 926,
    Magma_REFINE_GMGMDTRS = 927
} magma_refinement_t;

/* Mixed-precision GEMM selectors of magma_xshgetrf_gpu
 * (include/magma_types.h:885-908), same values. */
typedef enum {
    Magma_MP_BASE_SS              = 950,
    Magma_MP_BASE_DD              = 951,
    Magma_MP_BASE_XHS             = 952,
    Magma_MP_BASE_XSH             = 953,
    Magma_MP_BASE_XHD             = 954,
    Magma_MP_BASE_XDH             = 955,
    Magma_MP_ENABLE_DFLT_MATH     = 960,
    Magma_MP_ENABLE_TC_MATH       = 961,
    Magma_MP_SGEMM                = 962,
    Magma_MP_HGEMM                = 963,
    Magma_MP_GEMEX_I32_O32_C32    = 964,
    Magma_MP_GEMEX_I16_O32_C32    = 965,
    Magma_MP_GEMEX_I16_O16_C32    = 966,
    Magma_MP_GEMEX_I16_O16_C16    = 967,
    Magma_MP_TC_SGEMM             = 968,
    Magma_MP_TC_HGEMM             = 969,
    Magma_MP_TC_GEMEX_I32_O32_C32 = 970,
    Magma_MP_TC_GEMEX_I16_O32_C32 = 971,
    Magma_MP_TC_GEMEX_I16_O16_C32 = 972,
    Magma_MP_TC_GEMEX_I16_O16_C16 = 973
} magma_mp_type_t;

/* ------------------------------------------------------------------------- */
/* Runtime (include/magma_auxiliary.h, include/magma_copy.h)                  */
/* ------------------------------------------------------------------------- */
/* include/magma_auxiliary.h:24-25 */
magma_int_t magma_init(void);
magma_int_t magma_finalize(void);

/* include/magma_auxiliary.h:48-49 */
real_Double_t magma_wtime(void);
real_Double_t magma_sync_wtime(magma_queue_t queue);

/* include/magma_auxiliary.h:82-103 (magma_free is a macro over _internal) */
magma_int_t magma_malloc(magma_ptr *ptr_ptr, size_t bytes);
magma_int_t magma_malloc_cpu(void **ptr_ptr, size_t bytes);
magma_int_t magma_free_cpu(void *ptr);
magma_int_t magma_free_internal(magma_ptr ptr, const char *func, const char *file, int line);
#define magma_free(ptr) magma_free_internal(ptr, __func__, __FILE__, __LINE__)

/* include/magma_auxiliary.h:230 */
void magma_getdevice(magma_device_t *dev);
void magma_setdevice(magma_device_t dev);

/* include/magma_auxiliary.h:259-300 (the public names are macros) */
void magma_queue_create_internal(magma_device_t device, magma_queue_t *queue_ptr,
                                 const char *func, const char *file, int line);
/* MAGMA's HIP form also takes hipBLAS/hipSPARSE handles; this library uses no
 * vendor BLAS, so those two arguments are accepted and ignored. */
void magma_queue_create_from_hip_internal(magma_device_t device, void *hip_stream,
                                          void *hipblas_handle, void *hipsparse_handle,
                                          magma_queue_t *queue_ptr,
                                          const char *func, const char *file, int line);
void magma_queue_destroy_internal(magma_queue_t queue, const char *func, const char *file, int line);
void magma_queue_sync_internal(magma_queue_t queue, const char *func, const char *file, int line);
/* include/magma_types.h:227 */
void *magma_queue_get_hip_stream(magma_queue_t queue);
magma_device_t magma_queue_get_device(magma_queue_t queue);
#define magma_queue_create(dev, qp)  magma_queue_create_internal(dev, qp, __func__, __FILE__, __LINE__)
#define magma_queue_destroy(q)       magma_queue_destroy_internal(q, __func__, __FILE__, __LINE__)
#define magma_queue_sync(q)          magma_queue_sync_internal(q, __func__, __FILE__, __LINE__)

/* include/magma_copy.h:26-30 and the matrix forms */
void magma_setvector_internal(magma_int_t n, magma_int_t elemSize, const void *hx_src, magma_int_t incx,
                              magma_ptr dy_dst, magma_int_t incy, magma_queue_t queue,
                              const char *func, const char *file, int line);
void magma_getvector_internal(magma_int_t n, magma_int_t elemSize, const void *dx_src, magma_int_t incx,
                              void *hy_dst, magma_int_t incy, magma_queue_t queue,
                              const char *func, const char *file, int line);
void magma_setmatrix_internal(magma_int_t m, magma_int_t n, magma_int_t elemSize,
                              const void *hA_src, magma_int_t lda, magma_ptr dB_dst, magma_int_t lddb,
                              magma_queue_t queue, const char *func, const char *file, int line);
void magma_getmatrix_internal(magma_int_t m, magma_int_t n, magma_int_t elemSize,
                              const void *dA_src, magma_int_t ldda, void *hB_dst, magma_int_t ldb,
                              magma_queue_t queue, const char *func, const char *file, int line);

/* control/xerbla.cpp:51, include/magma_auxiliary.h:357-359 */
void magma_xerbla(const char *srname, magma_int_t neg_info);
const char *magma_strerror(magma_int_t error);

/* include/magma_zbatched.h:28 and include/magma_batched.h:63: build a device
 * pointer array out[i] = input + i*batch_offset + row + column*lda. */
void magma_dset_pointer(double **output_array, double *input, magma_int_t lda,
                        magma_int_t row, magma_int_t column, magma_int_t batch_offset,
                        magma_int_t batchCount, magma_queue_t queue);
void magma_iset_pointer(magma_int_t **output_array, magma_int_t *input, magma_int_t lda,
                        magma_int_t row, magma_int_t column, magma_int_t batchSize,
                        magma_int_t batchCount, magma_queue_t queue);

/* Block-size queries (control/get_batched_crossover.cpp:300,336). The values
 * describe this library's kernels, not MAGMA's. */
void magma_get_dgetrf_batched_nbparam(magma_int_t n, magma_int_t *nb, magma_int_t *recnb);
void magma_get_dgetrf_vbatched_nbparam(magma_int_t max_m, magma_int_t max_n, magma_int_t *nb, magma_int_t *recnb);

/* ------------------------------------------------------------------------- */
/* Batched LU (the hot path)                                                  */
/* ------------------------------------------------------------------------- */
/* include/magma_zbatched.h:848 -- P*A = L*U for batchCount m-by-n matrices.
 * ipiv_array[i] holds min(m,n) 1-based LAPACK interchanges; info_array[i] is
 * 0 or the 1-based index of the first exactly-zero pivot. Host-synchronous at
 * return, as in MAGMA. */
magma_int_t magma_dgetrf_batched(magma_int_t m, magma_int_t n, double **dA_array, magma_int_t ldda,
                                 magma_int_t **ipiv_array, magma_int_t *info_array,
                                 magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:464 -- solve op(A) X = B with the factors of
 * magma_dgetrf_batched. trans = MagmaTrans follows LAPACK dgetrs('T')
 * (src/zgetrs_gpu.cpp:203-227), not the batched reference's transposed branch.
 * Asynchronous on queue. */
magma_int_t magma_dgetrs_batched(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
                                 double **dA_array, magma_int_t ldda, magma_int_t **dipiv_array,
                                 double **dB_array, magma_int_t lddb,
                                 magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:1009 -- A X = B; A is overwritten by its LU factors,
 * B by X. Asynchronous on queue. */
magma_int_t magma_dgesv_batched(magma_int_t n, magma_int_t nrhs, double **dA_array, magma_int_t ldda,
                                magma_int_t **dipiv_array, double **dB_array, magma_int_t lddb,
                                magma_int_t *dinfo_array, magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zvbatched.h:57 -- variable-size batched LU. m, n, ldda are
 * device arrays of batchCount entries. Host-synchronous at return. */
magma_int_t magma_dgetrf_vbatched(magma_int_t *m, magma_int_t *n, double **dA_array, magma_int_t *ldda,
                                  magma_int_t **ipiv_array, magma_int_t *info_array,
                                  magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zvbatched.h:73 -- asynchronous form with caller workspace;
 * *lwork < 0 is a workspace query (bytes). The max_* arguments bound the
 * batch (max_m >= every m[i], and so on). */
magma_int_t magma_dgetrf_vbatched_max_nocheck_work(magma_int_t *m, magma_int_t *n,
        magma_int_t max_m, magma_int_t max_n, magma_int_t max_minmn, magma_int_t max_mxn,
        double **dA_array, magma_int_t *ldda, magma_int_t **dipiv_array, magma_int_t *info_array,
        void *work, magma_int_t *lwork, magma_int_t batchCount, magma_queue_t queue);

/* No MAGMA counterpart (MAGMA has no getrs/gesv for variable sizes). Solve
 * A_i X_i = B_i with the factors of magma_dgetrf_vbatched; n, nrhs, ldda, lddb
 * are device arrays; max_n / max_nrhs bound them. Asynchronous. */
magma_int_t magma_dgetrs_vbatched(magma_trans_t trans, magma_int_t *n, magma_int_t *nrhs,
        magma_int_t max_n, magma_int_t max_nrhs,
        double **dA_array, magma_int_t *ldda, magma_int_t **dipiv_array,
        double **dB_array, magma_int_t *lddb, magma_int_t batchCount, magma_queue_t queue);

/* No MAGMA counterpart: factor and solve variable-size systems; A_i is
 * overwritten by its LU factors and B_i by X_i. Asynchronous. */
magma_int_t magma_dgesv_vbatched(magma_int_t *n, magma_int_t *nrhs,
        magma_int_t max_n, magma_int_t max_nrhs,
        double **dA_array, magma_int_t *ldda, magma_int_t **dipiv_array,
        double **dB_array, magma_int_t *lddb, magma_int_t *dinfo_array,
        magma_int_t batchCount, magma_queue_t queue);

/* Variable-size plans (no MAGMA counterpart). A plan groups a batch into
 * n-buckets once (sizes given on the host) and keeps the device index maps,
 * so executing it allocates nothing and never synchronises; it can be
 * recorded once into a hipGraph and replayed. h_variants (nullable) holds a
 * kernel choice per bucket id (0 auto, 1 register-tile, 2/3/4 blocked with
 * panel width 32/16/8, 5 blocked with two 32-wide panels per 64-wide block,
 * orders <= 512), e.g. from magma_amd.autotune. h_nrhs == NULL makes
 * a factor-only plan. */
struct magma_vbatched_plan;
typedef struct magma_vbatched_plan* magma_vbatched_plan_t;
magma_int_t magma_dgesv_vbatched_plan_create(const magma_int_t *h_n, const magma_int_t *h_nrhs,
        magma_int_t batchCount, const magma_int_t *h_variants, magma_queue_t queue,
        magma_vbatched_plan_t *plan);
magma_int_t magma_vbatched_plan_num_buckets(magma_vbatched_plan_t plan);
/* info5 = {bucket id (-1: rectangular group), variant, count, max_n, max_nrhs} */
magma_int_t magma_vbatched_plan_bucket(magma_vbatched_plan_t plan, magma_int_t i, magma_int_t *info5);
magma_int_t magma_dgesv_vbatched_plan_exec(magma_vbatched_plan_t plan, magma_int_t *n, magma_int_t *nrhs,
        double **dA_array, magma_int_t *ldda, magma_int_t **dipiv_array,
        double **dB_array, magma_int_t *lddb, magma_int_t *dinfo_array, magma_queue_t queue);
magma_int_t magma_dgesv_vbatched_plan_capture(magma_vbatched_plan_t plan, magma_int_t *n, magma_int_t *nrhs,
        double **dA_array, magma_int_t *ldda, magma_int_t **dipiv_array,
        double **dB_array, magma_int_t *lddb, magma_int_t *dinfo_array, magma_queue_t queue);
magma_int_t magma_vbatched_plan_replay(magma_vbatched_plan_t plan, magma_queue_t queue);
void magma_vbatched_plan_destroy(magma_vbatched_plan_t plan);

/* ------------------------------------------------------------------------- */
/* Single precision batched LU: the "s" codegen forms of the prototypes at    */
/* include/magma_zbatched.h:848, :464 and :1009 (include/magma_sbatched.h).   */
/* ------------------------------------------------------------------------- */
magma_int_t magma_sgetrf_batched(magma_int_t m, magma_int_t n, float **dA_array, magma_int_t ldda,
                                 magma_int_t **ipiv_array, magma_int_t *info_array,
                                 magma_int_t batchCount, magma_queue_t queue);
magma_int_t magma_sgetrs_batched(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
                                 float **dA_array, magma_int_t ldda, magma_int_t **dipiv_array,
                                 float **dB_array, magma_int_t lddb,
                                 magma_int_t batchCount, magma_queue_t queue);
magma_int_t magma_sgesv_batched(magma_int_t n, magma_int_t nrhs, float **dA_array, magma_int_t ldda,
                                magma_int_t **dipiv_array, float **dB_array, magma_int_t lddb,
                                magma_int_t *dinfo_array, magma_int_t batchCount, magma_queue_t queue);
/* include/magma_zbatched.h:28, s form */
void magma_sset_pointer(float **output_array, float *input, magma_int_t lda,
                        magma_int_t row, magma_int_t column, magma_int_t batch_offset,
                        magma_int_t batchCount, magma_queue_t queue);

/* ------------------------------------------------------------------------- */
/* Batched LU without pivoting, random butterflies, batched inverse            */
/* ------------------------------------------------------------------------- */
/* include/magma_zbatched.h:977 (src/zgetrf_nopiv_batched.cpp), d form: LU with
 * the diagonal as pivots; info_array[i] = first exactly-zero pivot (1-based).
 * m <= 2048. Host-synchronous. */
magma_int_t magma_dgetrf_nopiv_batched(magma_int_t m, magma_int_t n, double **dA_array, magma_int_t ldda,
                                       magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue);
/* magma_zbatched.h (magma_zgetrs_nopiv_batched): solve with those factors. */
magma_int_t magma_dgetrs_nopiv_batched(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
                                       double **dA_array, magma_int_t ldda, double **dB_array, magma_int_t lddb,
                                       magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue);
/* magma_zbatched.h (magma_zgesv_nopiv_batched): factor and solve. */
magma_int_t magma_dgesv_nopiv_batched(magma_int_t n, magma_int_t nrhs, double **dA_array, magma_int_t ldda,
                                      double **dB_array, magma_int_t lddb, magma_int_t *info_array,
                                      magma_int_t batchCount, magma_queue_t queue);
/* include/magma_zvbatched.h:218 (src/zgetrf_nopiv_vbatched.cpp): m, n, ldda are
 * device arrays; max(m) <= 2048. Host-synchronous. */
magma_int_t magma_dgetrf_nopiv_vbatched(magma_int_t *m, magma_int_t *n, double **dA_array, magma_int_t *ldda,
                                        magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue);
/* include/magma_zbatched.h:1001 (src/zgesv_rbt_batched.cpp): A := U^T A V and
 * B := U^T B with two-level random butterflies (entries exp((r-1/2)/10),
 * r ~ U[0,1] from a fixed seed), LU without pivoting, solve, B := V B.
 * A is overwritten by the factors of U^T A V. Host-synchronous. */
magma_int_t magma_dgesv_rbt_batched(magma_int_t n, magma_int_t nrhs, double **dA_array, magma_int_t ldda,
                                    double **dB_array, magma_int_t lddb, magma_int_t *dinfo_array,
                                    magma_int_t batchCount, magma_queue_t queue);
/* include/magma_zbatched.h:872 (src/zgetri_outofplace_batched.cpp): inverse from
 * the factors and pivots of magma_dgetrf_batched, into dinvA_array (lddia).
 * Host-synchronous. */
magma_int_t magma_dgetri_outofplace_batched(magma_int_t n, double **dA_array, magma_int_t ldda,
                                            magma_int_t **dipiv_array, double **dinvA_array, magma_int_t lddia,
                                            magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue);

/* ------------------------------------------------------------------------- */
/* One large matrix on the whole device (dense LU)                            */
/* ------------------------------------------------------------------------- */
/* include/magma_z.h:612 (src/zgetrf_gpu.cpp:542), d form: LU with partial
 * pivoting of the m x n matrix dA in place; ipiv is host memory (min(m,n)
 * LAPACK interchanges, 1-based); info > 0: U(info,info) is exactly zero.
 * Host-synchronous. */
magma_int_t magma_dgetrf_gpu(magma_int_t m, magma_int_t n, magmaDouble_ptr dA, magma_int_t ldda,
                             magma_int_t *ipiv, magma_int_t *info);
/* include/magma_z.h:638 (src/zgetrf_gpu.cpp:607): the all-GPU variant; same
 * arguments and results as magma_dgetrf_gpu. */
magma_int_t magma_dgetrf_native(magma_int_t m, magma_int_t n, magmaDouble_ptr dA, magma_int_t ldda,
                                magma_int_t *ipiv, magma_int_t *info);
/* include/magma_z.h:713 (src/zgetrs_gpu.cpp:292), d form: solve op(A) X = B
 * with the factors of magma_dgetrf_gpu (trans as LAPACK dgetrs). */
magma_int_t magma_dgetrs_gpu(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
                             magmaDouble_ptr dA, magma_int_t ldda, magma_int_t *ipiv,
                             magmaDouble_ptr dB, magma_int_t lddb, magma_int_t *info);
/* s forms of include/magma_z.h:612 and :713 (include/magma_s.h). */
magma_int_t magma_sgetrf_gpu(magma_int_t m, magma_int_t n, magmaFloat_ptr dA, magma_int_t ldda,
                             magma_int_t *ipiv, magma_int_t *info);
magma_int_t magma_sgetrs_gpu(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
                             magmaFloat_ptr dA, magma_int_t ldda, magma_int_t *ipiv,
                             magmaFloat_ptr dB, magma_int_t lddb, magma_int_t *info);

/* include/magma_htc.h:119 (src/xshgetrf_gpu.cpp:86): FP32 LU with FP16-input,
 * FP32-accumulate trailing updates on the matrix cores when enable_tc is
 * Magma_MP_ENABLE_TC_MATH (FP32 updates for Magma_MP_ENABLE_DFLT_MATH or an
 * FP32 mp_algo_type). ipiv is host memory. Host-synchronous. */
magma_int_t magma_xshgetrf_gpu(magma_int_t m, magma_int_t n, magmaFloat_ptr dA, magma_int_t ldda,
                               magma_int_t *ipiv, magma_int_t *info,
                               magma_mp_type_t enable_tc, magma_mp_type_t mp_algo_type);
/* No MAGMA counterpart: the FP32 LU with the update format of a refinement
 * factorisation type (Magma_PREC_SS, _SHT, _SBT, _SB3T); -7 for another. */
magma_int_t magma_sgetrf_prec_gpu(magma_int_t m, magma_int_t n, magmaFloat_ptr dA, magma_int_t ldda,
                                  magma_int_t *ipiv, magma_int_t *info, magma_refinement_t facto_type);

/* ------------------------------------------------------------------------- */
/* Mixed-precision iterative refinement                                        */
/* ------------------------------------------------------------------------- */
/* include/magma_htc.h:27 -- FP16-input/FP32-accumulate LU, then FGMRES in
 * FP64 preconditioned by it (src/dxgesv_gmres_gpu.cpp:1122-1157). ipiv is host
 * memory, dipiv device memory; dworkd >= n*nrhs doubles, dworks >= n*(n+nrhs)
 * floats. iter > 0: refinement converged in iter steps; iter < 0: FP64
 * fallback was used (-2 overflow on down-cast, -3 low-precision LU failed,
 * -k no convergence after k steps). Host-synchronous. */
magma_int_t magma_dhgesv_iteref_gpu(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
        magmaDouble_ptr dA, magma_int_t ldda, magma_int_t *ipiv, magmaInt_ptr dipiv,
        magmaDouble_ptr dB, magma_int_t lddb, magmaDouble_ptr dX, magma_int_t lddx,
        magmaDouble_ptr dworkd, magmaFloat_ptr dworks, magma_int_t *iter, magma_int_t *info);

/* src/dxgesv_gmres_gpu.cpp:966 -- FP32 LU (FP32 updates) + FGMRES in FP64
 * preconditioned by it (Magma_PREC_SS, Magma_REFINE_GMSTRS); same contract as
 * magma_dhgesv_iteref_gpu. */
magma_int_t magma_dsgesv_iteref_gpu(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
        magmaDouble_ptr dA, magma_int_t ldda, magma_int_t *ipiv, magmaInt_ptr dipiv,
        magmaDouble_ptr dB, magma_int_t lddb, magmaDouble_ptr dX, magma_int_t lddx,
        magmaDouble_ptr dworkd, magmaFloat_ptr dworks, magma_int_t *iter, magma_int_t *info);

/* src/dxgesv_gmres_gpu.cpp:160 -- expert driver: facto_type in {SS, SHT,
 * XSH_163232TC (same as SHT), SBT, SB3T}, solver_type any of the eight
 * Magma_REFINE_* values (:328-359): IR or FGMRES outer level, optionally an
 * inner FGMRES (50 iterations to 1e-4) as preconditioner, with the
 * triangular solves on the FP32 factors (..STRS) or on their FP64 copy
 * (..DTRS). The outer FGMRES restarts after min(n, 500) iterations. Every
 * mode that does not reach ||r||_inf <= ||x||_inf ||A||_inf eps sqrt(n)
 * falls back to FP64 (iter < 0). facto_time (seconds) may be NULL. */
magma_int_t magma_dxgesv_gmres_gpu(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
        magmaDouble_ptr dA, magma_int_t ldda, magma_int_t *ipiv, magmaInt_ptr dipiv,
        magmaDouble_ptr dB, magma_int_t lddb, magmaDouble_ptr dX, magma_int_t lddx,
        magmaDouble_ptr dworkd, magmaFloat_ptr dworks,
        magma_refinement_t facto_type, magma_refinement_t solver_type,
        magma_int_t *iter, magma_int_t *info, real_Double_t *facto_time);

/* No MAGMA counterpart: BF16-input/FP32-accumulate LU + FP64 refinement with
 * the same contract as magma_dhgesv_iteref_gpu. */
magma_int_t magma_dbgesv_iteref_gpu(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
        magmaDouble_ptr dA, magma_int_t ldda, magma_int_t *ipiv, magmaInt_ptr dipiv,
        magmaDouble_ptr dB, magma_int_t lddb, magmaDouble_ptr dX, magma_int_t lddx,
        magmaDouble_ptr dworkd, magmaFloat_ptr dworks, magma_int_t *iter, magma_int_t *info);

/* No MAGMA counterpart: batched mixed-precision solve of A_i X_i = B_i
 * (n <= 2048, any nrhs; the low-precision LU, the refinement steps and the
 * FP64 fallback are each one batched launch sequence on the queue's stream,
 * and the call returns after the refinement has finished). A_i and B_i are
 * read only; X_i is written. facto_type selects the low-precision LU (SHT, SBT, SB3T, SS),
 * solver_type IRSTRS or GMSTRS (FGMRES restarted every 50 iterations, which
 * bounds the Krylov basis of the whole batch); itermax bounds refinement steps.
 * iter_array[i] follows the iter convention of magma_dhgesv_iteref_gpu;
 * a system that does not converge is re-solved in FP64 inside the launch.
 * info_array[i] > 0 if the FP64 fallback found an exactly-zero pivot. */
magma_int_t magma_dxgesv_iteref_batched(magma_int_t n, magma_int_t nrhs,
        double **dA_array, magma_int_t ldda, double **dB_array, magma_int_t lddb,
        double **dX_array, magma_int_t lddx,
        magma_refinement_t facto_type, magma_refinement_t solver_type, magma_int_t itermax,
        magma_int_t *iter_array, magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue);

/* ------------------------------------------------------------------------- */
/* Adjoint of the batched solve (differentiable solve)                        */
/* ------------------------------------------------------------------------- */
/* No MAGMA counterpart: the backward pass of X = A^{-1} B needs
 * grad_A = -grad_B X^T. For contiguous batches of column-major n x nrhs
 * grad_B and X (leading dimensions ldb, ldx; strideB, strideX elements
 * between systems) it writes the row-major n x n grad_A (row stride lda,
 * system stride strideA). elem: 8 for FP64, 4 for FP32. Asynchronous on the
 * queue. */
magma_int_t magma_adjoint_outer_batched(magma_int_t elem, magma_int_t n, magma_int_t nrhs, const void *dgB,
                                        magma_int_t ldb, long long strideB, const void *dX, magma_int_t ldx,
                                        long long strideX, void *dgA, magma_int_t lda, long long strideA,
                                        magma_int_t batchCount, magma_queue_t queue);

/* No MAGMA counterpart. For callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the LU of magma_dgetrf_native (same kernels, same results),
 * with its scratch in a fixed 1 MB device array of the library instead of
 * allocations. d_ipiv (min(m,n) entries) and d_info are device memory; the
 * call is asynchronous on queue's stream. m <= 65,536 rows (and at most 256
 * rows per compute unit): larger m returns MAGMA_ERR_NOT_SUPPORTED.
 * Calls that overlap on one device share that array and race: serialise them
 * (one queue, or events between queues). A wait that exceeds its time limit
 * sets *d_info = MAGMA_ERR_UNKNOWN. */
magma_int_t magma_dgetrf_native_noalloc(magma_int_t m, magma_int_t n, magmaDouble_ptr dA, magma_int_t ldda,
                                        magmaInt_ptr d_ipiv, magmaInt_ptr d_info, magma_queue_t queue);

#ifdef __cplusplus
}
#endif

#endif /* MAGMA_MI355X_H */
