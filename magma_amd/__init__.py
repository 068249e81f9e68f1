"""magma_amd -- MI355X-native batched LU factor/solve behind MAGMA's C ABI.

The compute lives in libmagma_mi355x.so (HIP kernels for gfx950, built by
``python -m magma_amd.build``); this package is a thin binding over it.
"""
from ._ffi import (MagmaNoTrans, MagmaTrans, MagmaConjTrans, Magma_PREC_SS, Magma_PREC_SHT,  # noqa: F401
                   Magma_PREC_SBT, Magma_PREC_SB3T, Magma_REFINE_IRSTRS, Magma_REFINE_GMSTRS)
from . import _ffi  # noqa: F401

__all__ = ["load", "MagmaNoTrans", "MagmaTrans"]


def load():
    """Load the HIP library (raises if it has not been built)."""
    return _ffi.load()


def __getattr__(name):
    # Lazy import of the torch-based mirror so that ``import magma_amd`` works
    # on machines used only for building.
    if name in ("batched",):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
