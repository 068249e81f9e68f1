"""C-ABI checks that need no GPU: the library loads, exports every function
the public header declares, and rejects bad arguments with MAGMA's codes
before touching the device."""
import ctypes

import pytest

import magma_amd._ffi as ffi


@pytest.fixture(scope="module")
def lib():
    return ffi.load()


def test_exports_every_declared_symbol(lib):
    names = ffi.header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_abi_enum_values():
    text = ffi.HEADER.read_text()
    for token in ("MagmaNoTrans = 111", "MagmaTrans = 112", "MagmaConjTrans = 113",
                  "Magma_PREC_SHT        = 905", "Magma_REFINE_GMSTRS   = 924",
                  "Magma_REFINE_IRSTRS   = 920", "typedef int            magma_int_t"):
        assert token in text


def test_getrf_batched_argument_codes(lib):
    # src/zgetrf_batched.cpp:94-105: m<0 -> -1, n<0 -> -2, ldda<max(1,m) -> -4
    assert lib.magma_dgetrf_batched(-1, 4, None, 4, None, None, 1, None) == -1
    assert lib.magma_dgetrf_batched(4, -1, None, 4, None, None, 1, None) == -2
    assert lib.magma_dgetrf_batched(4, 4, None, 3, None, None, 1, None) == -4
    # quick return without touching memory
    assert lib.magma_dgetrf_batched(0, 4, None, 1, None, None, 1, None) == 0
    assert lib.magma_dgetrf_batched(4, 0, None, 4, None, None, 1, None) == 0


def test_getrs_batched_argument_codes(lib):
    # src/zgetrs_batched.cpp:93-112
    assert lib.magma_dgetrs_batched(999, 4, 1, None, 4, None, None, 4, 1, None) == -1
    assert lib.magma_dgetrs_batched(111, -1, 1, None, 4, None, None, 4, 1, None) == -2
    assert lib.magma_dgetrs_batched(111, 4, -1, None, 4, None, None, 4, 1, None) == -3
    assert lib.magma_dgetrs_batched(111, 4, 1, None, 3, None, None, 4, 1, None) == -5
    assert lib.magma_dgetrs_batched(111, 4, 1, None, 4, None, None, 3, 1, None) == -8
    assert lib.magma_dgetrs_batched(111, 4, 0, None, 4, None, None, 4, 1, None) == 0


def test_gesv_batched_argument_codes(lib):
    # src/zgesv_batched.cpp:104-116
    assert lib.magma_dgesv_batched(-1, 1, None, 4, None, None, 4, None, 1, None) == -1
    assert lib.magma_dgesv_batched(4, -1, None, 4, None, None, 4, None, 1, None) == -2
    assert lib.magma_dgesv_batched(4, 1, None, 3, None, None, 4, None, 1, None) == -4
    assert lib.magma_dgesv_batched(4, 1, None, 4, None, None, 3, None, 1, None) == -6
    assert lib.magma_dgesv_batched(0, 1, None, 1, None, None, 1, None, 1, None) == 0


def test_strerror_and_nbparam(lib):
    assert lib.magma_strerror(0) == b"success"
    assert b"device memory" in lib.magma_strerror(-113)
    assert lib.magma_strerror(-3) == b"invalid argument"
    nb, rec = ctypes.c_int(), ctypes.c_int()
    lib.magma_get_dgetrf_batched_nbparam(128, ctypes.byref(nb), ctypes.byref(rec))
    assert (nb.value, rec.value) == (16, 16)            # register-tile LU: 16-column panels
    lib.magma_get_dgetrf_batched_nbparam(400, ctypes.byref(nb), ctypes.byref(rec))
    assert (nb.value, rec.value) == (64, 32)            # blocked LU: 64-wide blocks of 32-wide panels
    lib.magma_get_dgetrf_batched_nbparam(1000, ctypes.byref(nb), ctypes.byref(rec))
    assert (nb.value, rec.value) == (16, 16)            # 4 rows per thread: 16-wide panels
    lib.magma_get_dgetrf_vbatched_nbparam(100, 100, ctypes.byref(nb), ctypes.byref(rec))
    assert (nb.value, rec.value) == (16, 16)
    lib.magma_get_dgetrf_vbatched_nbparam(200, 200, ctypes.byref(nb), ctypes.byref(rec))
    assert (nb.value, rec.value) == (32, 32)


def test_xerbla_message(lib, capfd):
    lib.magma_xerbla(b"magma_dgetrf_batched", 4)
    err = capfd.readouterr().err
    assert "parameter 4 had an illegal value (info = -4)" in err


def test_wtime_monotone(lib):
    a = lib.magma_wtime()
    b = lib.magma_wtime()
    assert b >= a


def test_refinement_argument_codes(lib):
    """magma_dxgesv_gmres_gpu (src/dxgesv_gmres_gpu.cpp:219-359): every
    Magma_REFINE_* solver type is accepted, anything else is -15; an unknown
    factorisation type is -14; n = 0 returns before any device work."""
    it, info = ctypes.c_int(7), ctypes.c_int(7)
    ft = ctypes.c_double(0)
    ip = (ctypes.c_int * 4)()

    def call(trans, n, facto, solver, ldda=4):
        return lib.magma_dxgesv_gmres_gpu(trans, n, 1, None, ldda, ip, None, None, max(n, 1), None, max(n, 1),
                                          None, None, facto, solver, ctypes.byref(it), ctypes.byref(info),
                                          ctypes.byref(ft))
    assert call(999, 4, ffi.Magma_PREC_SS, ffi.Magma_REFINE_IRSTRS) == -1
    assert call(111, -1, ffi.Magma_PREC_SS, ffi.Magma_REFINE_IRSTRS) == -2
    assert call(111, 4, ffi.Magma_PREC_SS, ffi.Magma_REFINE_IRSTRS, ldda=3) == -5
    assert call(111, 4, 12345, ffi.Magma_REFINE_IRSTRS) == -14
    assert call(111, 4, ffi.Magma_PREC_SS, 919) == -15
    assert call(111, 4, ffi.Magma_PREC_SS, 928) == -15
    for solver in range(ffi.Magma_REFINE_IRSTRS, ffi.Magma_REFINE_GMGMDTRS + 1):
        for facto in (ffi.Magma_PREC_SS, ffi.Magma_PREC_SHT, ffi.Magma_PREC_SBT, ffi.Magma_PREC_SB3T,
                      ffi.Magma_PREC_SF8T):
            assert call(111, 0, facto, solver) == 0
            assert (it.value, info.value) == (0, 0)
    # the batched solver runs IRSTRS / GMSTRS with 16-bit or FP32 updates;
    # the FP8 experiment is for the one-system drivers only
    q = ctypes.c_void_p(0)
    assert lib.magma_dxgesv_iteref_batched(4, 1, None, 4, None, 4, None, 4, ffi.Magma_PREC_SF8T,
                                           ffi.Magma_REFINE_IRSTRS, 10, None, None, 1, q) == -9
    assert lib.magma_dxgesv_iteref_batched(4, 1, None, 4, None, 4, None, 4, ffi.Magma_PREC_SBT,
                                           ffi.Magma_REFINE_IRDTRS, 10, None, None, 1, q) == -10


def test_prec_lu_and_adjoint_argument_codes(lib):
    info = ctypes.c_int(0)
    ip = (ctypes.c_int * 4)()
    assert lib.magma_sgetrf_prec_gpu(4, 4, None, 4, ip, ctypes.byref(info), 12345) == -7
    assert lib.magma_sgetrf_prec_gpu(-1, 4, None, 4, ip, ctypes.byref(info), ffi.Magma_PREC_SF8T) == -1
    assert lib.magma_adjoint_outer_batched(2, 4, 1, None, 4, 4, None, 4, 4, None, 4, 16, 1, None) == -1
    assert lib.magma_adjoint_outer_batched(8, 4, 1, None, 3, 4, None, 4, 4, None, 4, 16, 1, None) == -5
    assert lib.magma_adjoint_outer_batched(8, 0, 1, None, 1, 4, None, 1, 4, None, 1, 16, 1, None) == 0
