"""Python binding of the host table compiler (libosdfar.so, include/osdfar.h).

    topo = Topology(mesh)                 # far/topologyDescriptor.h
    topo.refine_uniform(3)                # TopologyRefiner::RefineUniform
    table = topo.stencil_table()          # StencilTableFactory::Create
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib

_far = None
vp, i, ll = C.c_void_p, C.c_int, C.c_longlong


def load_far():
    global _far
    if _far is not None:
        return _far
    if not os.path.exists(_lib.FAR_LIB):
        raise RuntimeError(f"libosdfar.so not built ({_lib.FAR_LIB}); run opensubdiv_amd.build")
    L = C.CDLL(_lib.FAR_LIB)
    sig = {
        "osdfar_topology_create": ([i, i, vp, vp, i, vp, vp, i, vp, vp, i], vp),
        "osdfar_topology_destroy": ([vp], None),
        "osdfar_refine_uniform": ([vp, i], i),
        "osdfar_level_info": ([vp, i, vp, vp, vp], i),
        "osdfar_level_face_vertices": ([vp, i, vp], i),
        "osdfar_stencil_table_create": ([vp, i, i, i, i], vp),
        "osdfar_stencil_table_info": ([vp, vp, vp, vp, vp], i),
        "osdfar_stencil_table_copy": ([vp, vp, vp, vp, vp, i, vp, vp], i),
        "osdfar_stencil_table_destroy": ([vp], None),
    }
    for k, (a, r) in sig.items():
        f = getattr(L, k)
        f.argtypes = a
        f.restype = r
    _far = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


class StencilTable(dict):
    """sizes / offsets / indices / weights (+ du, dv, ...) numpy arrays;
    also answers the Far getters used by HipStencilTable.Create."""

    def GetSizes(self): return self["sizes"]
    def GetOffsets(self): return self["offsets"]
    def GetControlIndices(self): return self["indices"]
    def GetWeights(self): return self["weights"]
    def GetDuWeights(self): return self.get("du")
    def GetDvWeights(self): return self.get("dv")
    def GetDuuWeights(self): return self.get("duu")
    def GetDuvWeights(self): return self.get("duv")
    def GetDvvWeights(self): return self.get("dvv")
    def GetNumStencils(self): return len(self["sizes"])
    def GetNumControlVertices(self): return self["num_control"]


def _read_table(h, double=False):
    L = load_far()
    ns, nnz, nc, nw = C.c_int(), C.c_longlong(), C.c_int(), C.c_int()
    L.osdfar_stencil_table_info(h, C.byref(ns), C.byref(nnz), C.byref(nc), C.byref(nw))
    out = StencilTable(sizes=np.zeros(ns.value, np.int32), indices=np.zeros(nnz.value, np.int32),
                       num_control=nc.value)
    off64 = np.zeros(ns.value, np.int64)
    keys = ["weights", "du", "dv", "duu", "duv", "dvv"][:nw.value]
    for k, key in enumerate(keys):
        w = np.zeros(nnz.value, np.float64 if double else np.float32)
        rc = L.osdfar_stencil_table_copy(h, _p(out["sizes"]) if k == 0 else None, None,
                                         _p(off64) if k == 0 else None,
                                         _p(out["indices"]) if k == 0 else None, k,
                                         None if double else _p(w), _p(w) if double else None)
        assert rc == 0, rc
        out[key] = w
    out["offsets"] = off64.astype(np.int32) if nnz.value < 2 ** 31 else off64
    return out


class Topology:
    """Control mesh + uniform refinement levels (TopologyRefiner)."""

    def __init__(self, mesh, boundary=1):
        L = load_far()
        fv = np.ascontiguousarray(mesh.face_verts, np.int32).reshape(-1)
        nvpf = np.full(mesh.nfaces, mesh.face_verts.shape[1], np.int32)
        cp = np.ascontiguousarray(mesh.crease_pairs, np.int32).reshape(-1)
        cw = np.ascontiguousarray(mesh.crease_weights, np.float32)
        cv = np.ascontiguousarray(mesh.corner_verts, np.int32)
        cwv = np.ascontiguousarray(mesh.corner_weights, np.float32)
        self.h = L.osdfar_topology_create(mesh.nverts, mesh.nfaces, _p(nvpf), _p(fv), len(cw),
                                          _p(cp) if len(cw) else None,
                                          _p(cw) if len(cw) else None, len(cv),
                                          _p(cv) if len(cv) else None,
                                          _p(cwv) if len(cv) else None, boundary)
        if not self.h:
            raise ValueError("topology rejected by the table compiler")
        self.num_control = mesh.nverts

    def refine_uniform(self, level):
        assert load_far().osdfar_refine_uniform(self.h, level) == 0

    def level_info(self, level):
        nv, nf, ne = C.c_int(), C.c_int(), C.c_int()
        assert load_far().osdfar_level_info(self.h, level, C.byref(nv), C.byref(nf),
                                            C.byref(ne)) == 0
        return nv.value, nf.value, ne.value

    def level_face_vertices(self, level):
        nv, nf, ne = self.level_info(level)
        out = np.zeros(nf * 4, np.int32)
        assert load_far().osdfar_level_face_vertices(self.h, level, _p(out)) == 0
        return out.reshape(nf, 4)

    def stencil_table(self, mode=0, intermediate=False, control=False, double=False):
        h = load_far().osdfar_stencil_table_create(self.h, mode, int(intermediate), int(control),
                                                   int(double))
        try:
            return _read_table(h, double)
        finally:
            load_far().osdfar_stencil_table_destroy(h)

    def __del__(self):
        if getattr(self, "h", None) and _far is not None:
            _far.osdfar_topology_destroy(self.h)
            self.h = None
