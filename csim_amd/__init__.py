"""csim_amd: MI355X-native csim E-field propagation path.

The product is ``libcsx.so`` (HIP kernels for gfx950 + C++ host, C ABI in
``include/csx.h``).  This package is its Python host mirror: ``Context`` and
the operator wrappers (``csim_amd.ops``) and the element-chain API that
follows csim's input-script elements (``csim_amd.chain``).
"""
from ._lib import CsxError, LIB_PATH, PRECISIONS, lib, last_error  # noqa: F401

__all__ = ["CsxError", "LIB_PATH", "PRECISIONS", "lib", "last_error"]
