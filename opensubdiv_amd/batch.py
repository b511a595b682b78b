"""Batched evaluation: many meshes x many frames per graph launch.

Thin Python handle over the ``osdhip_batch_*`` C ABI (include/osdhip.h).
Segments reference device buffers (torch tensors) and HipStencilTables; the
Batch keeps them alive for its own lifetime.

Stencil segment (dict):
    table   HipStencilTable with one weight array (record layout)
    src     device tensor; element (v, f, k) at src + v*src_v + f*src_f + k
    dst     device tensor; element (row, f, k) at dst + row*dst_v + f*dst_f + k
    frames, src_v, src_f, dst_v, dst_f
    src_off / dst_off (optional): float offsets into src / dst
    stage (optional, default 0): stages run in order within a launch

Patch segment (dict):
    coords, arrays, indices, params   device tensors (Osd layouts)
    src, src_v, src_f                 patch points [control | refined | end cap]
    P, dPdu, dPdv (each optional), out_v, out_f, frames
"""
from __future__ import annotations

import ctypes
import json

from . import _lib


def _p(t, off=0):
    if t is None:
        return None
    return t.data_ptr() + 4 * int(off)


class Batch:
    KNOBS = ("a_frames", "a_group", "a_item_slices", "a_persistent", "b_frames", "b_persistent",
             "b_stage")

    def __init__(self, length, stencils=(), patches=(), stream=None):
        self._lib = _lib.load_hip()
        self._keep = [list(stencils), list(patches)]
        S = (_lib.BatchStencils * max(1, len(stencils)))()
        for i, s in enumerate(stencils):
            S[i] = _lib.BatchStencils(s["table"].h, _p(s["src"], s.get("src_off", 0)),
                                      int(s["src_v"]), int(s["src_f"]),
                                      _p(s["dst"], s.get("dst_off", 0)),
                                      int(s["dst_v"]), int(s["dst_f"]), int(s["frames"]),
                                      int(s.get("stage", 0)))
        P = (_lib.BatchPatches * max(1, len(patches)))()
        for i, p in enumerate(patches):
            P[i] = _lib.BatchPatches(int(p["n"]), _p(p["coords"]), _p(p["arrays"]),
                                     _p(p["indices"]), _p(p["params"]), _p(p["src"]),
                                     int(p["src_v"]), int(p["src_f"]), _p(p.get("P")),
                                     _p(p.get("dPdu")), _p(p.get("dPdv")), int(p["out_v"]),
                                     int(p["out_f"]), int(p["frames"]), int(p.get("stage", 0)))
        self.h = self._lib.osdhip_batch_create(int(length), len(stencils),
                                               ctypes.cast(S, ctypes.c_void_p), len(patches),
                                               ctypes.cast(P, ctypes.c_void_p), stream)
        if not self.h:
            raise ValueError("osdhip_batch_create rejected the segments")

    def launch(self, stream=None):
        _lib.check(self._lib.osdhip_batch_launch(self.h, stream), "osdhip_batch_launch")

    def run(self, phases=3, stream=None):
        _lib.check(self._lib.osdhip_batch_run(self.h, int(phases), stream), "osdhip_batch_run")

    def run_plain(self, stream=None):
        _lib.check(self._lib.osdhip_batch_run_plain(self.h, stream), "osdhip_batch_run_plain")

    def run_stage(self, stage, phases=3, stream=None):
        _lib.check(self._lib.osdhip_batch_run_stage(self.h, int(stage), int(phases), stream),
                   "osdhip_batch_run_stage")

    def run_plain_stage(self, stage, stream=None):
        _lib.check(self._lib.osdhip_batch_run_plain_stage(self.h, int(stage), stream),
                   "osdhip_batch_run_plain_stage")

    def num_stages(self):
        return self._lib.osdhip_batch_num_stages(self.h)

    def set(self, key, value):
        return self._lib.osdhip_batch_set(self.h, key.encode(), int(value))

    def get(self, key):
        return self._lib.osdhip_batch_get(self.h, key.encode())

    def config(self):
        return {k: self.get(k) for k in self.KNOBS}

    def tune(self, reps=3, stream=None):
        buf = ctypes.create_string_buffer(65536)
        _lib.check(self._lib.osdhip_batch_tune(self.h, int(reps), stream,
                                               ctypes.cast(buf, ctypes.c_void_p), len(buf)),
                   "osdhip_batch_tune")
        return json.loads(buf.value.decode())

    def info(self):
        v = [ctypes.c_longlong() for _ in range(5)]
        _lib.check(self._lib.osdhip_batch_info(self.h, *[ctypes.byref(x) for x in v]),
                   "osdhip_batch_info")
        keys = ("stencil_rows", "frame_rows", "patch_coords", "patch_frames", "table_bytes")
        return dict(zip(keys, (x.value for x in v)))

    def patch_staging(self, segment):
        """Patch segment `segment`'s point staging lists (numpy arrays):
        nuniq per 64-coord wave (-1: not staged), ptsOff, pts (ascending
        per wave) and 4 slot words per coord; None without lists."""
        import numpy as np
        n = ctypes.c_longlong()
        _lib.check(self._lib.osdhip_batch_patch_staging(self.h, int(segment), None, None, None,
                                                        None, ctypes.byref(n)),
                   "osdhip_batch_patch_staging")
        if n.value < 0:
            return None
        ncoords = int(self._keep[1][segment]["n"])
        nw = (ncoords + 63) // 64
        nuniq = np.zeros(nw, np.int32)
        off = np.zeros(nw, np.int64)
        pts = np.zeros(max(1, n.value), np.int32)
        slots = np.zeros(4 * ncoords, np.uint32)
        _lib.check(self._lib.osdhip_batch_patch_staging(
            self.h, int(segment), nuniq.ctypes.data, off.ctypes.data, pts.ctypes.data,
            slots.ctypes.data, ctypes.byref(n)), "osdhip_batch_patch_staging")
        return {"nuniq": nuniq, "ptsOff": off, "pts": pts[:n.value], "slots": slots}

    def close(self):
        if getattr(self, "h", None):
            self._lib.osdhip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
