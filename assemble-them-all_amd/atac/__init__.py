"""atac: MI355X-native point-vs-SDF contact (BodySDFObj queries feeding
ForceGeneralSDFContact penalty/friction wrenches, their Jacobians and adjoint).

The arithmetic runs in gfx950 HIP kernels behind the C ABI of
libatac_contact.so (include/atac.h); this package is the host-side front end.
Importing it does not load the library; the first call that needs the GPU does,
and raises AtacError when the library has not been built.
"""
# PyTorch-ROCm carries its own HIP runtime. It is loaded before any of this package's
# native libraries (libatac_contact.so, the _redmax module) so that the process holds
# one HIP runtime: loaded the other way round, the library's device queries find no
# device once torch has initialised its own.
try:
    import torch  # noqa: F401
except ImportError:
    pass

from ._lib import AtacError, EVAL_CHECKPOINT, EVAL_DERIV, EVAL_FORCE, EVAL_P6, EVAL_PRECISE, LIB_PATH  # noqa: F401

__all__ = ["AtacError", "LIB_PATH", "EVAL_FORCE", "EVAL_DERIV", "EVAL_P6", "EVAL_CHECKPOINT", "EVAL_PRECISE"]
