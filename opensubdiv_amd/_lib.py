"""ctypes binding of the C ABI in include/osdhip.h (libosdhip.so) and of the
host table compiler in include/osdfar.h (libosdfar.so).

The libraries are built in-tree by ``opensubdiv_amd.build`` into
``opensubdiv_amd/lib``.  Loading fails loudly when a library is missing:
there is no CPU fallback for the evaluator.
"""
from __future__ import annotations

import ctypes as C
import os

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
HIP_LIB = os.path.join(LIBDIR, "libosdhip.so")
FAR_LIB = os.path.join(LIBDIR, "libosdfar.so")

OK = 0
ERR_ARGS = -1
ERR_UNSUPPORTED = -2


class Desc(C.Structure):
    """osdhip_buffer_desc == Osd::BufferDescriptor (offset, length, stride)."""
    _fields_ = [("offset", C.c_int), ("length", C.c_int), ("stride", C.c_int)]


def desc(d) -> Desc:
    if d is None:
        return Desc(0, 0, 0)
    if isinstance(d, Desc):
        return d
    return Desc(int(d.offset), int(d.length), int(d.stride))


class TableOptions(C.Structure):
    """osdhip_table_options (layout, keepCsr, group, kernel, order)."""
    _fields_ = [("layout", C.c_int), ("keepCsr", C.c_int), ("group", C.c_int),
                ("kernel", C.c_int), ("order", C.c_int)]


class BatchStencils(C.Structure):
    """osdhip_batch_stencils."""
    _fields_ = [("table", C.c_void_p), ("src", C.c_void_p),
                ("srcVertexStride", C.c_longlong), ("srcFrameStride", C.c_longlong),
                ("dst", C.c_void_p),
                ("dstVertexStride", C.c_longlong), ("dstFrameStride", C.c_longlong),
                ("frames", C.c_int), ("stage", C.c_int)]


class BatchPatches(C.Structure):
    """osdhip_batch_patches."""
    _fields_ = [("numCoords", C.c_int), ("coords", C.c_void_p), ("arrays", C.c_void_p),
                ("patchIndices", C.c_void_p), ("params", C.c_void_p), ("src", C.c_void_p),
                ("srcVertexStride", C.c_longlong), ("srcFrameStride", C.c_longlong),
                ("P", C.c_void_p), ("dPdu", C.c_void_p), ("dPdv", C.c_void_p),
                ("outVertexStride", C.c_longlong), ("outFrameStride", C.c_longlong),
                ("frames", C.c_int), ("stage", C.c_int)]


_vp = C.c_void_p
_i = C.c_int
_D = Desc

_HIP_SIGS = {
    "osdhip_eval_stencils": [_vp, _D, _vp, _D, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "osdhip_eval_stencils_d1": [_vp, _D, _vp, _D, _vp, _D, _vp, _D,
                                _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "osdhip_eval_stencils_d2": [_vp, _D, _vp, _D, _vp, _D, _vp, _D, _vp, _D, _vp, _D, _vp, _D,
                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "osdhip_eval_patches": [_vp, _D, _vp, _D, _vp, _D, _vp, _D, _vp, _D, _vp, _D, _vp, _D,
                            _i, _vp, _vp, _vp, _vp, _vp],
    "osdhip_eval_patches_sorted": [_vp, _D, _vp, _D, _vp, _D, _vp, _D, _vp, _D, _vp, _D, _vp,
                                   _D, _i, _vp, _vp, _vp, _vp, _vp],
    "osdhip_eval_patches_sorted_noalloc": [_vp, _D, _vp, _D, _vp, _D, _vp, _D, _vp, _D, _vp, _D,
                                           _vp, _D, _i, _vp, _vp, _vp, _vp, _vp],
    "osdhip_stencil_table_create": [_i, _vp, _vp, _vp, _i, _vp, _vp],
    "osdhip_stencil_table_create_ex": [_i, _vp, _vp, _vp, _i, _vp, _vp, _vp],
    "osdhip_table_options_init": [_vp],
    "osdhip_stencil_table_set": [_vp, C.c_char_p, _i],
    "osdhip_stencil_table_options": [_vp, _vp],
    "osdhip_stencil_table_destroy": [_vp],
    "osdhip_stencil_table_info": [_vp, _vp, _vp, _vp, _vp],
    "osdhip_stencil_table_csr": [_vp, _vp, _vp, _vp, _vp],
    "osdhip_stencil_table_layout": [_vp, _vp, _vp, _vp, _vp],
    "osdhip_stencil_table_slices": [_vp, _vp, _vp],
    "osdhip_limit_normals": [_vp, _D, _vp, _D, _vp, _D, _i, _i, _vp],
    "osdhip_patch_map_create": [_vp, _i, _vp, _i, _vp],
    "osdhip_patch_map_find": [_vp, _i, _vp, _vp, _vp, _vp, _vp],
    "osdhip_patch_map_destroy": [_vp],
    "osdhip_patch_map_find_faces": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp],
    "osdhip_tess_uniform_quads": [_i, _vp, _i, _vp, _vp, _vp, _vp],
    "osdhip_tess_count": [_i, _i, _i, _vp],
    "osdhip_tess_coords": [_i, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp],
    "osdhip_stencil_table_eval": [_vp, _vp, _D, _i, _vp, _vp, _i, _i, _vp],
    "osdhip_malloc": [C.POINTER(C.c_void_p), C.c_size_t],
    "osdhip_free": [_vp],
    "osdhip_memcpy_htod": [_vp, _vp, C.c_size_t, _vp],
    "osdhip_memcpy_dtoh": [_vp, _vp, C.c_size_t, _vp],
    "osdhip_memcpy_dtod": [_vp, _vp, C.c_size_t, _vp],
    "osdhip_stream_synchronize": [_vp],
    "osdhip_device_synchronize": [],
    "osdhip_device_info": [C.c_char_p, _i],
    "osdhip_version": [],
    "osdhip_batch_create": [_i, _i, _vp, _i, _vp, _vp],
    "osdhip_batch_destroy": [_vp],
    "osdhip_batch_launch": [_vp, _vp],
    "osdhip_batch_run": [_vp, _i, _vp],
    "osdhip_batch_run_plain": [_vp, _vp],
    "osdhip_batch_run_stage": [_vp, _i, _i, _vp],
    "osdhip_batch_run_plain_stage": [_vp, _i, _vp],
    "osdhip_batch_num_stages": [_vp],
    "osdhip_batch_set": [_vp, C.c_char_p, _i],
    "osdhip_batch_get": [_vp, C.c_char_p],
    "osdhip_batch_info": [_vp, _vp, _vp, _vp, _vp, _vp],
    "osdhip_batch_tune": [_vp, _i, _vp, _vp, _i],
    "osdhip_batch_patch_staging": [_vp, _i, _vp, _vp, _vp, _vp, _vp],
    "osdhip_batch_epoch_bytes": [_i, _i, _vp, _i, _vp, _i],
    "osdhip_batch_epoch": [_i, _i, _vp, _i, _vp, _i, _vp, C.c_longlong, _vp, C.c_longlong, _vp],
}

_hip = None


def hip_symbols():
    return list(_HIP_SIGS)


def load_hip():
    """Load libosdhip.so (raises if it has not been built)."""
    global _hip
    if _hip is not None:
        return _hip
    if not os.path.exists(HIP_LIB):
        raise RuntimeError(f"libosdhip.so not built ({HIP_LIB}); run opensubdiv_amd.build")
    # One HIP runtime per process.  torch ships its own libamdhip64 (same
    # soname as the system one libosdhip.so links against): imported first,
    # its runtime is the one libosdhip then binds to; loaded the other way
    # round, the process would hold both, and streams or device pointers of
    # one are invalid in the other.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(HIP_LIB)
    for name, args in _HIP_SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        if name == "osdhip_version":
            fn.restype = C.c_char_p
        elif name == "osdhip_batch_epoch_bytes":
            fn.restype = C.c_longlong
        elif name in ("osdhip_stencil_table_create", "osdhip_stencil_table_create_ex",
                      "osdhip_batch_create", "osdhip_patch_map_create"):
            fn.restype = C.c_void_p
        elif name in ("osdhip_stencil_table_destroy", "osdhip_batch_destroy",
                      "osdhip_patch_map_destroy", "osdhip_table_options_init"):
            fn.restype = None
        else:
            fn.restype = C.c_int
    _hip = lib
    return lib


def check(rc: int, what: str = "osdhip call"):
    if rc != OK:
        raise RuntimeError(f"{what} failed with code {rc}")
    return rc
