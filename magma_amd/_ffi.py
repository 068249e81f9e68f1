"""ctypes binding of libmagma_mi355x.so (the C ABI in include/magma_mi355x.h).

This module only declares prototypes; it loads the shared library built by
``magma_amd.build`` from the package folder and raises if it is missing, so
nothing in the package can silently run without the HIP kernels.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libmagma_mi355x.so"
HEADER = PKG.parent / "include" / "magma_mi355x.h"

c_int = ctypes.c_int
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
c_size_t = ctypes.c_size_t
i32p = ctypes.POINTER(ctypes.c_int)

MagmaNoTrans, MagmaTrans, MagmaConjTrans = 111, 112, 113
Magma_PREC_SS, Magma_PREC_SHT, Magma_PREC_XSH_163232TC = 900, 905, 917
Magma_PREC_SBT, Magma_PREC_SB3T, Magma_PREC_SF8T = 940, 941, 942
(Magma_REFINE_IRSTRS, Magma_REFINE_IRDTRS, Magma_REFINE_IRGMSTRS, Magma_REFINE_IRGMDTRS, Magma_REFINE_GMSTRS,
 Magma_REFINE_GMDTRS, Magma_REFINE_GMGMSTRS, Magma_REFINE_GMGMDTRS) = range(920, 928)
Magma_MP_ENABLE_DFLT_MATH, Magma_MP_ENABLE_TC_MATH = 960, 961
Magma_MP_SGEMM, Magma_MP_GEMEX_I16_O32_C32 = 962, 965
MAGMA_ERR_NOT_SUPPORTED = -103

# name -> (restype, argtypes); pointers to device memory are passed as c_void_p.
_PROTOS = {
    "magma_init": (c_int, []),
    "magma_finalize": (c_int, []),
    "magma_wtime": (c_double, []),
    "magma_sync_wtime": (c_double, [c_void_p]),
    "magma_malloc": (c_int, [ctypes.POINTER(c_void_p), c_size_t]),
    "magma_malloc_cpu": (c_int, [ctypes.POINTER(c_void_p), c_size_t]),
    "magma_free_cpu": (c_int, [c_void_p]),
    "magma_free_internal": (c_int, [c_void_p, c_char_p, c_char_p, c_int]),
    "magma_getdevice": (None, [i32p]),
    "magma_setdevice": (None, [c_int]),
    "magma_queue_create_internal": (None, [c_int, ctypes.POINTER(c_void_p), c_char_p, c_char_p, c_int]),
    "magma_queue_create_from_hip_internal": (None, [c_int, c_void_p, c_void_p, c_void_p,
                                                    ctypes.POINTER(c_void_p), c_char_p, c_char_p, c_int]),
    "magma_queue_destroy_internal": (None, [c_void_p, c_char_p, c_char_p, c_int]),
    "magma_queue_sync_internal": (None, [c_void_p, c_char_p, c_char_p, c_int]),
    "magma_queue_get_hip_stream": (c_void_p, [c_void_p]),
    "magma_queue_get_device": (c_int, [c_void_p]),
    "magma_setvector_internal": (None, [c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p,
                                        c_char_p, c_char_p, c_int]),
    "magma_getvector_internal": (None, [c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p,
                                        c_char_p, c_char_p, c_int]),
    "magma_setmatrix_internal": (None, [c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p,
                                        c_char_p, c_char_p, c_int]),
    "magma_getmatrix_internal": (None, [c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p,
                                        c_char_p, c_char_p, c_int]),
    "magma_xerbla": (None, [c_char_p, c_int]),
    "magma_strerror": (c_char_p, [c_int]),
    "magma_dset_pointer": (None, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p]),
    "magma_iset_pointer": (None, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p]),
    "magma_get_dgetrf_batched_nbparam": (None, [c_int, i32p, i32p]),
    "magma_get_dgetrf_vbatched_nbparam": (None, [c_int, c_int, i32p, i32p]),
    "magma_dgetrf_batched": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p]),
    "magma_dgetrs_batched": (c_int, [c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int,
                                     c_int, c_void_p]),
    "magma_dgesv_batched": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                    c_int, c_void_p]),
    "magma_dgetrf_vbatched": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                      c_void_p]),
    "magma_dgetrf_vbatched_max_nocheck_work": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int,
                                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                                       i32p, c_int, c_void_p]),
    "magma_dgetrs_vbatched": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "magma_dgesv_vbatched": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "magma_sset_pointer": (None, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p]),
    "magma_sgetrf_batched": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p]),
    "magma_sgetrs_batched": (c_int, [c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int,
                                     c_int, c_void_p]),
    "magma_sgesv_batched": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                    c_int, c_void_p]),
    "magma_dgesv_vbatched_plan_create": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                                 ctypes.POINTER(c_void_p)]),
    "magma_vbatched_plan_num_buckets": (c_int, [c_void_p]),
    "magma_vbatched_plan_bucket": (c_int, [c_void_p, c_int, c_void_p]),
    "magma_dgesv_vbatched_plan_exec": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                               c_void_p, c_void_p, c_void_p, c_void_p]),
    "magma_dgesv_vbatched_plan_capture": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                                  c_void_p, c_void_p, c_void_p, c_void_p]),
    "magma_vbatched_plan_replay": (c_int, [c_void_p, c_void_p]),
    "magma_vbatched_plan_destroy": (None, [c_void_p]),
    "magma_dhgesv_iteref_gpu": (c_int, [c_int, c_int, c_int, c_void_p, c_int, i32p, c_void_p, c_void_p,
                                        c_int, c_void_p, c_int, c_void_p, c_void_p, i32p, i32p]),
    "magma_dbgesv_iteref_gpu": (c_int, [c_int, c_int, c_int, c_void_p, c_int, i32p, c_void_p, c_void_p,
                                        c_int, c_void_p, c_int, c_void_p, c_void_p, i32p, i32p]),
    "magma_adjoint_outer_batched": (c_int, [c_int, c_int, c_int, c_void_p, c_int, ctypes.c_longlong, c_void_p, c_int,
                                            ctypes.c_longlong, c_void_p, c_int, ctypes.c_longlong, c_int, c_void_p]),
    "magma_dsgesv_iteref_gpu": (c_int, [c_int, c_int, c_int, c_void_p, c_int, i32p, c_void_p, c_void_p,
                                        c_int, c_void_p, c_int, c_void_p, c_void_p, i32p, i32p]),
    "magma_dxgesv_gmres_gpu": (c_int, [c_int, c_int, c_int, c_void_p, c_int, i32p, c_void_p, c_void_p,
                                       c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, i32p, i32p,
                                       ctypes.POINTER(c_double)]),
    "magma_dgetrf_nopiv_batched": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p]),
    "magma_dgetrs_nopiv_batched": (c_int, [c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p,
                                           c_int, c_void_p]),
    "magma_dgesv_nopiv_batched": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int,
                                          c_void_p]),
    "magma_dgetrf_nopiv_vbatched": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "magma_dgesv_rbt_batched": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int,
                                        c_void_p]),
    "magma_dgetri_outofplace_batched": (c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                                c_int, c_void_p]),
    "magma_dgetrf_gpu": (c_int, [c_int, c_int, c_void_p, c_int, i32p, i32p]),
    "magma_dgetrf_native": (c_int, [c_int, c_int, c_void_p, c_int, i32p, i32p]),
    "magma_sgetrf_gpu": (c_int, [c_int, c_int, c_void_p, c_int, i32p, i32p]),
    "magma_dgetrs_gpu": (c_int, [c_int, c_int, c_int, c_void_p, c_int, i32p, c_void_p, c_int, i32p]),
    "magma_xshgetrf_gpu": (c_int, [c_int, c_int, c_void_p, c_int, i32p, i32p, c_int, c_int]),
    "magma_sgetrf_prec_gpu": (c_int, [c_int, c_int, c_void_p, c_int, i32p, i32p, c_int]),
    "magma_sgetrs_gpu": (c_int, [c_int, c_int, c_int, c_void_p, c_int, i32p, c_void_p, c_int, i32p]),
    "magma_dgetrf_native_noalloc": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "magma_dxgesv_iteref_batched": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int,
                                            c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p]),
}


def header_functions(header: Path = HEADER) -> list[str]:
    """Names of the functions declared in the public header."""
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"#define[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(magma_\w+)\s*\(", text)))


_lib: ctypes.CDLL | None = None


def load(path: Path | None = None) -> ctypes.CDLL:
    """Load libmagma_mi355x.so and attach prototypes. Raises if it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP library first (python -m magma_amd.build); "
            "there is no non-GPU fallback for the magma_amd solvers")
    L = ctypes.CDLL(str(p))
    for name, (res, args) in _PROTOS.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = L
    return L
