"""opensubdiv_amd -- MI355X-native (gfx950) Osd evaluator.

The product is the C ABI library ``lib/libosdhip.so`` (HIP kernels for
EvalStencils / EvalPatches) with C++ drop-in headers under ``include/``, and
the host table compiler ``lib/libosdfar.so``.  This Python package is a thin
ctypes mirror used by tests, the benchmark and Python callers.
"""
from .osd import (BufferDescriptor, HipEvaluator, HipPatchTable, HipStencilTable,  # noqa: F401
                  HipVertexBuffer, PATCH_ARRAY_DTYPE, PATCH_COORD_DTYPE, PATCH_PARAM_DTYPE)

__version__ = "0.1.0"
