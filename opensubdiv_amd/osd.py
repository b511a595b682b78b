"""Python mirror of the Osd HIP evaluator interface (tests, benchmark, and
Python callers).  Mirrors the reference's static evaluator API
(osd/cudaEvaluator.h:94-1265, osd/cpuEvaluator.h) with the same names,
argument meaning and bool error results; device memory is held in torch
tensors on ``cuda`` (plumbing only -- every computation runs in the HIP
kernels of libosdhip.so, there is no CPU fallback).

    vb = HipVertexBuffer.Create(3, nCoarse + nRefined)
    vb.UpdateData(coarse_xyz, 0, nCoarse)
    table = HipStencilTable.Create(farLikeTable)
    HipEvaluator.EvalStencils(vb, BufferDescriptor(0, 3, 3),
                              vb, BufferDescriptor(nCoarse * 3, 3, 3), table)
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib

PATCH_COORD_DTYPE = np.dtype([("arrayIndex", "<i4"), ("patchIndex", "<i4"),
                              ("vertIndex", "<i4"), ("s", "<f4"), ("t", "<f4")])
PATCH_ARRAY_DTYPE = np.dtype([("regDesc", "<i4"), ("desc", "<i4"), ("numPatches", "<i4"),
                              ("indexBase", "<i4"), ("stride", "<i4"),
                              ("primitiveIdBase", "<i4")])
PATCH_PARAM_DTYPE = np.dtype([("field0", "<u4"), ("field1", "<u4"), ("sharpness", "<f4")])

QUADS, TRIANGLES, LOOP, REGULAR, GREGORY_BASIS, GREGORY_TRIANGLE = 3, 4, 5, 6, 9, 10


@dataclass
class BufferDescriptor:
    """Osd::BufferDescriptor (osd/bufferDescriptor.h:61-103), in floats."""
    offset: int = 0
    length: int = 0
    stride: int = 0

    def GetLocalOffset(self) -> int:
        return self.offset % self.stride if self.stride > 0 else 0

    def IsValid(self) -> bool:
        return self.length > 0 and self.length <= self.stride - self.GetLocalOffset()


def _torch():
    import torch
    return torch


def _ptr(x) -> int | None:
    """Device pointer of a buffer-like object (torch tensor, HipVertexBuffer)."""
    if x is None:
        return None
    if hasattr(x, "BindHipBuffer"):
        return x.BindHipBuffer()
    if hasattr(x, "data_ptr"):
        if not x.is_cuda:
            raise ValueError("device buffer expected")
        return x.data_ptr()
    return int(x)


def _stream(ctx=None) -> int | None:
    if ctx is not None:
        return int(ctx)
    torch = _torch()
    return torch.cuda.current_stream().cuda_stream


def _device_tensor(a: np.ndarray):
    torch = _torch()
    a = np.ascontiguousarray(a)
    if a.dtype.fields is not None:
        a = a.view(np.int32 if a.dtype.itemsize % 4 == 0 else np.uint8)
    return torch.from_numpy(a.copy()).to("cuda")


class HipVertexBuffer:
    """Osd::CudaVertexBuffer counterpart (osd/cudaVertexBuffer.h:37-68)."""

    def __init__(self, numElements: int, numVertices: int):
        torch = _torch()
        self._n = int(numElements)
        self._v = int(numVertices)
        self.data = torch.zeros(self._n * self._v, dtype=torch.float32, device="cuda")

    @classmethod
    def Create(cls, numElements, numVertices, deviceContext=None):
        return cls(numElements, numVertices)

    def UpdateData(self, src, startVertex, numVertices, deviceContext=None):
        torch = _torch()
        a = np.ascontiguousarray(np.asarray(src, np.float32).reshape(-1))
        n = int(numVertices) * self._n
        self.data[startVertex * self._n: startVertex * self._n + n].copy_(
            torch.from_numpy(a[:n]), non_blocking=False)

    def GetNumElements(self):
        return self._n

    def GetNumVertices(self):
        return self._v

    def BindHipBuffer(self):
        return self.data.data_ptr()

    def numpy(self):
        return self.data.cpu().numpy().reshape(self._v, self._n)


class HipStencilTable:
    """Device mirror of a (Limit)StencilTable (osd/cudaEvaluator.h:52-91),
    built by osdhip_stencil_table_create: the Far CSR arrays plus the 64-row
    sliced layout the table kernel streams.

    Accepts any object with GetSizes/GetOffsets/GetControlIndices/GetWeights
    (and optionally GetDuWeights ... GetDvvWeights) returning array-likes, or
    a dict with keys sizes/offsets/indices/weights[/du/dv/duu/duv/dvv].

    Device options (osdhip_table_options, fixed per table): ``layout`` 0
    (record slices), 1 (split streams) or 2 (locality-ordered slices staged
    in LDS); ``keep_csr`` keeps the Far CSR arrays on the device; ``group``
    (4, 8, 16) and ``kernel`` (0 slices, 1 CSR) select the evaluation
    kernel and can be changed later with ``set``; ``order`` (layout 2) 0
    groups rows by dominant control vertex and then by size within row
    windows when that streams fewer padded entries, 2 by dominant control
    vertex only, 1 keeps Far's row order."""

    _KEYS = ["weights", "du", "dv", "duu", "duv", "dvv"]
    _GETTERS = ["GetWeights", "GetDuWeights", "GetDvWeights", "GetDuuWeights",
                "GetDuvWeights", "GetDvvWeights"]

    def __init__(self, table, stream=None, layout=0, keep_csr=True, group=8, kernel=0, order=0):
        import ctypes
        opts = _lib.TableOptions(int(layout), int(bool(keep_csr)), int(group), int(kernel),
                                 int(order))
        if hasattr(table, "weights_ptr"):            # far.RawStencilTable
            lib = _lib.load_hip()
            wptrs = (ctypes.c_void_p * 1)(table.weights_ptr)
            self._n, self._nw = int(table.num_stencils), 1
            self.h = lib.osdhip_stencil_table_create_ex(self._n, table.sizes_ptr, None,
                                                        table.indices_ptr, 1,
                                                        ctypes.cast(wptrs, ctypes.c_void_p),
                                                        ctypes.byref(opts), _stream(stream))
            if not self.h:
                raise MemoryError("osdhip_stencil_table_create failed")
            self._csr_pointers(lib)
            return
        get = (lambda k, g: table.get(k)) if isinstance(table, dict) else \
              (lambda k, g: getattr(table, g)() if hasattr(table, g) else None)
        sizes = np.ascontiguousarray(get("sizes", "GetSizes"), np.int32)
        offsets = get("offsets", "GetOffsets")
        offsets = None if offsets is None or len(offsets) == 0 else \
            np.ascontiguousarray(offsets, np.int32)
        indices = np.ascontiguousarray(get("indices", "GetControlIndices"), np.int32)
        ws = []
        for i, (k, g) in enumerate(zip(self._KEYS, self._GETTERS)):
            v = get(k, g)
            # derivative arrays are optional (absent or empty); the position
            # weights are always passed, also when every stencil is empty
            if v is None or (i > 0 and len(v) == 0):
                break
            ws.append(np.ascontiguousarray(v, np.float32))
        nw = 6 if len(ws) >= 6 else (3 if len(ws) >= 3 else 1)
        ws = ws[:nw]
        lib = _lib.load_hip()
        wptrs = (ctypes.c_void_p * nw)(*[w.ctypes.data for w in ws])
        self._n = int(sizes.shape[0])
        self._nw = nw
        self.h = lib.osdhip_stencil_table_create_ex(
            self._n, sizes.ctypes.data, None if offsets is None else offsets.ctypes.data,
            indices.ctypes.data, nw, ctypes.cast(wptrs, ctypes.c_void_p), ctypes.byref(opts),
            _stream(stream))
        if not self.h:
            raise MemoryError("osdhip_stencil_table_create failed")
        self._csr_pointers(lib)

    def _csr_pointers(self, lib):
        import ctypes
        p = [ctypes.c_void_p() for _ in range(3)]
        w6 = (ctypes.c_void_p * 6)()
        lib.osdhip_stencil_table_csr(self.h, *[ctypes.byref(x) for x in p],
                                     ctypes.cast(w6, ctypes.c_void_p))
        self._csr = [x.value for x in p]
        self._w = list(w6)

    @classmethod
    def Create(cls, table, deviceContext=None, **options):
        return cls(table, deviceContext, **options)

    def set(self, key, value):
        """Changes the evaluation option "group" or "kernel" of this table;
        returns the previous value."""
        rc = _lib.load_hip().osdhip_stencil_table_set(self.h, key.encode(), int(value))
        if rc < 0:
            raise ValueError(f"osdhip_stencil_table_set({key!r}, {value}) failed: {rc}")
        return rc

    def options(self):
        import ctypes
        o = _lib.TableOptions()
        _lib.check(_lib.load_hip().osdhip_stencil_table_options(self.h, ctypes.byref(o)),
                   "osdhip_stencil_table_options")
        return {"layout": o.layout, "keep_csr": bool(o.keepCsr), "group": o.group,
                "kernel": o.kernel, "order": o.order}

    def GetSizesBuffer(self): return self._csr[0]
    def GetOffsetsBuffer(self): return self._csr[1]
    def GetIndicesBuffer(self): return self._csr[2]
    def GetWeightsBuffer(self): return self._w[0]
    def GetDuWeightsBuffer(self): return self._w[1]
    def GetDvWeightsBuffer(self): return self._w[2]
    def GetDuuWeightsBuffer(self): return self._w[3]
    def GetDuvWeightsBuffer(self): return self._w[4]
    def GetDvvWeightsBuffer(self): return self._w[5]
    def GetNumStencils(self): return self._n
    def GetNumWeightArrays(self): return self._nw

    def device_bytes(self):
        import ctypes
        b = ctypes.c_longlong()
        _lib.load_hip().osdhip_stencil_table_info(self.h, None, None, None, ctypes.byref(b))
        return b.value

    def layout(self):
        """Device layout: dict(layout, slices, listed_slices, list_entries,
        record_entries, size_window)."""
        import ctypes
        v = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()]
        lib = _lib.load_hip()
        _lib.check(lib.osdhip_stencil_table_layout(self.h, *[ctypes.byref(x) for x in v]),
                   "osdhip_stencil_table_layout")
        e, wdw = ctypes.c_longlong(), ctypes.c_int()
        _lib.check(lib.osdhip_stencil_table_slices(self.h, ctypes.byref(e), ctypes.byref(wdw)),
                   "osdhip_stencil_table_slices")
        out = dict(zip(("layout", "slices", "listed_slices", "list_entries"), (x.value for x in v)))
        out.update(record_entries=e.value, size_window=wdw.value)
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib._hip is not None:
            _lib._hip.osdhip_stencil_table_destroy(self.h)
            self.h = None


class HipPatchTable:
    """Device mirror of a flattened patch table (osd/cudaPatchTable.h:45-100):
    PatchArray[], patch index buffer, PatchParam[] as CpuPatchTable lays them
    out (osd/cpuPatchTable.cpp:35-158)."""

    def __init__(self, arrays, indices, params, varying_arrays=None, varying_indices=None,
                 fvar_channels=()):
        self.arrays = _device_tensor(np.asarray(arrays).astype(PATCH_ARRAY_DTYPE, copy=False))
        self.indices = _device_tensor(np.asarray(indices, np.int32))
        self.params = _device_tensor(np.asarray(params).astype(PATCH_PARAM_DTYPE, copy=False))
        self.varying_arrays = None if varying_arrays is None else _device_tensor(
            np.asarray(varying_arrays).astype(PATCH_ARRAY_DTYPE, copy=False))
        self.varying_indices = None if varying_indices is None else _device_tensor(
            np.asarray(varying_indices, np.int32))
        # face-varying channels: (arrays, indices, params) per channel
        self.fvar = [(_device_tensor(np.asarray(a).astype(PATCH_ARRAY_DTYPE, copy=False)),
                      _device_tensor(np.asarray(i, np.int32)),
                      _device_tensor(np.asarray(p).astype(PATCH_PARAM_DTYPE, copy=False)))
                     for a, i, p in fvar_channels]

    @classmethod
    def Create(cls, flat, deviceContext=None):
        return cls(flat["arrays"], flat["indices"], flat["params"],
                   flat.get("varying_arrays"), flat.get("varying_indices"),
                   flat.get("fvar_channels", ()))

    def GetPatchArrayBuffer(self): return self.arrays.data_ptr()
    def GetPatchIndexBuffer(self): return self.indices.data_ptr()
    def GetPatchParamBuffer(self): return self.params.data_ptr()
    def GetVaryingPatchArrayBuffer(self):
        return None if self.varying_arrays is None else self.varying_arrays.data_ptr()
    def GetVaryingPatchIndexBuffer(self):
        return None if self.varying_indices is None else self.varying_indices.data_ptr()
    def GetNumFVarChannels(self): return len(self.fvar)
    def GetFVarPatchArrayBuffer(self, fvarChannel=0): return self.fvar[fvarChannel][0].data_ptr()
    def GetFVarPatchIndexBuffer(self, fvarChannel=0): return self.fvar[fvarChannel][1].data_ptr()
    def GetFVarPatchParamBuffer(self, fvarChannel=0): return self.fvar[fvarChannel][2].data_ptr()


def limit_normals(du, duDesc, dv, dvDesc, normals, normalsDesc, count, normalize=True,
                  deviceContext=None):
    """Limit normals after EvalPatches: cross(dPdu, dPdv) over xyz, scaled to
    unit length when `normalize` -- the reference tutorial's steps
    (tutorials/far/tutorial_2_3/far_tutorial_2_3.cpp:73-87) on the device
    (osdhip_limit_normals).  Returns bool."""
    lib = _lib.load_hip()
    rc = lib.osdhip_limit_normals(_ptr(du), _lib.desc(duDesc), _ptr(dv), _lib.desc(dvDesc),
                                  _ptr(normals), _lib.desc(normalsDesc), int(count),
                                  int(bool(normalize)), _stream(deviceContext))
    return rc == _lib.OK


class HipPatchMap:
    """Far::PatchMap on the device (osdhip_patch_map_*): built from a patch
    table's PatchArrays and PatchParams (numpy, host), FindPatches maps
    device (ptex face, u, v) arrays to device PatchCoords in one launch."""

    def __init__(self, arrays, params, deviceContext=None):
        a = np.ascontiguousarray(arrays, PATCH_ARRAY_DTYPE)
        p = np.ascontiguousarray(params, PATCH_PARAM_DTYPE)
        self.h = _lib.load_hip().osdhip_patch_map_create(a.ctypes.data, len(a), p.ctypes.data,
                                                         len(p), _stream(deviceContext))
        if not self.h:
            raise ValueError("osdhip_patch_map_create: empty or inconsistent patch table")

    def FindPatches(self, n, ptexFaces, u, v, patchCoords, deviceContext=None):
        """Device int32 faces, float64 u / v, PatchCoord output; returns bool."""
        return _lib.load_hip().osdhip_patch_map_find(self.h, int(n), _ptr(ptexFaces), _ptr(u),
                                                     _ptr(v), _ptr(patchCoords),
                                                     _stream(deviceContext)) == _lib.OK

    def FindFaceCoords(self, n, faces, u, v, faceSizes, ptexBase, patchCoords,
                       deviceContext=None):
        """Bfr-style queries: device int32 base faces, float64 (u, v) in
        Bfr's parameterization of each face, int32 face sizes and first
        ptex index per base face; PatchCoord output; returns bool."""
        return _lib.load_hip().osdhip_patch_map_find_faces(
            self.h, int(n), _ptr(faces), _ptr(u), _ptr(v), _ptr(faceSizes), _ptr(ptexBase),
            int(len(faceSizes)), _ptr(patchCoords), _stream(deviceContext)) == _lib.OK

    def close(self):
        if getattr(self, "h", None) and _lib._hip is not None:
            _lib._hip.osdhip_patch_map_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class HipEvaluator:
    """Static evaluator (osd/cudaEvaluator.h:94-1265); all methods return
    bool like the reference.  ``deviceContext`` may be a hipStream_t (int);
    None = torch's current stream."""

    @staticmethod
    def EvalStencils(srcBuffer, srcDesc, dstBuffer, dstDesc, *rest):
        """EvalStencils(src, srcDesc, dst, dstDesc, table[, instance, ctx])
        EvalStencils(src, srcDesc, dst, dstDesc, du, duDesc, dv, dvDesc, table[, ...])
        EvalStencils(src, srcDesc, dst, dstDesc, du, duDesc, dv, dvDesc,
                     duu, duuDesc, duv, duvDesc, dvv, dvvDesc, table[, ...])"""
        lib = _lib.load_hip()
        outs = [(dstBuffer, dstDesc)]
        rest = list(rest)
        while rest and not isinstance(rest[0], HipStencilTable):
            outs.append((rest[0], rest[1]))
            rest = rest[2:]
        table = rest[0]
        ctx = rest[2] if len(rest) > 2 else None
        if table.GetNumStencils() == 0:
            return False                   # osd/cpuEvaluator.h:82-83
        if len(outs) not in (1, 3, 6):
            raise TypeError("EvalStencils takes 1, 3 or 6 outputs")
        import ctypes
        ptrs = (ctypes.c_void_p * len(outs))(*[_ptr(b) for b, _ in outs])
        descs = (_lib.Desc * len(outs))(*[_lib.desc(d) for _, d in outs])
        rc = lib.osdhip_stencil_table_eval(table.h, _ptr(srcBuffer), _lib.desc(srcDesc), len(outs),
                                           ctypes.cast(ptrs, ctypes.c_void_p),
                                           ctypes.cast(descs, ctypes.c_void_p), 0,
                                           table.GetNumStencils(), _stream(ctx))
        return rc == _lib.OK

    @staticmethod
    def _patches(srcBuffer, srcDesc, outs, numPatchCoords, patchCoords, arrays, indices, params,
                 ctx):
        lib = _lib.load_hip()
        args = []
        for k in range(6):
            if k < len(outs) and outs[k][0] is not None:
                args += [_ptr(outs[k][0]), _lib.desc(outs[k][1])]
            else:
                args += [None, _lib.Desc(0, 0, 0)]
        rc = lib.osdhip_eval_patches(_ptr(srcBuffer), _lib.desc(srcDesc), *args,
                                     int(numPatchCoords), _ptr(patchCoords), arrays, indices,
                                     params, _stream(ctx))
        return rc == _lib.OK

    @staticmethod
    def EvalPatches(srcBuffer, srcDesc, dstBuffer, dstDesc, *rest):
        """EvalPatches(src, srcDesc, dst, dstDesc, [du, duDesc, dv, dvDesc,
        [duu, duuDesc, duv, duvDesc, dvv, dvvDesc,]] numPatchCoords,
        patchCoords, patchTable[, instance, ctx])"""
        outs = [(dstBuffer, dstDesc)]
        rest = list(rest)
        while rest and not isinstance(rest[0], int):
            outs.append((rest[0], rest[1]))
            rest = rest[2:]
        n, coords, table = rest[0], rest[1], rest[2]
        ctx = rest[4] if len(rest) > 4 else None
        return HipEvaluator._patches(srcBuffer, srcDesc, outs, n, coords,
                                     table.GetPatchArrayBuffer(), table.GetPatchIndexBuffer(),
                                     table.GetPatchParamBuffer(), ctx)

    @staticmethod
    def EvalPatchesVarying(srcBuffer, srcDesc, dstBuffer, dstDesc, *rest):
        """Varying evaluation through the table's varying arrays and indices
        (osd/cudaEvaluator.h:857-1040)."""
        outs = [(dstBuffer, dstDesc)]
        rest = list(rest)
        while rest and not isinstance(rest[0], int):
            outs.append((rest[0], rest[1]))
            rest = rest[2:]
        n, coords, table = rest[0], rest[1], rest[2]
        ctx = rest[4] if len(rest) > 4 else None
        if table.GetVaryingPatchArrayBuffer() is None:
            return False
        return HipEvaluator._patches(srcBuffer, srcDesc, outs, n, coords,
                                     table.GetVaryingPatchArrayBuffer(),
                                     table.GetVaryingPatchIndexBuffer(),
                                     table.GetPatchParamBuffer(), ctx)

    @staticmethod
    def EvalPatchesFaceVarying(srcBuffer, srcDesc, dstBuffer, dstDesc, *rest):
        """Face-varying evaluation through channel `fvarChannel`'s arrays,
        indices and params (osd/cudaEvaluator.h:1050-1253):
        EvalPatchesFaceVarying(src, srcDesc, dst, dstDesc, [du, duDesc, dv,
        dvDesc, [...]] numPatchCoords, patchCoords, patchTable, fvarChannel
        [, instance, ctx])"""
        outs = [(dstBuffer, dstDesc)]
        rest = list(rest)
        while rest and not isinstance(rest[0], int):
            outs.append((rest[0], rest[1]))
            rest = rest[2:]
        n, coords, table, ch = rest[0], rest[1], rest[2], rest[3]
        ctx = rest[5] if len(rest) > 5 else None
        if ch < 0 or ch >= table.GetNumFVarChannels():
            return False
        return HipEvaluator._patches(srcBuffer, srcDesc, outs, n, coords,
                                     table.GetFVarPatchArrayBuffer(ch),
                                     table.GetFVarPatchIndexBuffer(ch),
                                     table.GetFVarPatchParamBuffer(ch), ctx)

    @staticmethod
    def Synchronize(deviceContext=None):
        lib = _lib.load_hip()
        return lib.osdhip_stream_synchronize(_stream(deviceContext)) == _lib.OK
