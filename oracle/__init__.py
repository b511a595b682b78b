"""CPU oracle for the Osd evaluator hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's CPU-baseline leg import
this package, and only as the checker / baseline.  It wraps:

* ``liboracle.so`` -- the C restatement of CpuEvaluator::EvalStencils /
  EvalPatches and the Osd patch basis (osd_oracle.c; float and double);
* ``_ref/libosdref.so`` -- the real OpenSubdiv Far + Osd CPU/OMP evaluators
  compiled from the reference sources (built where the reference tree is
  present; travels as a prebuilt file otherwise).  See ref.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libosdref.so")
REF_TREE = "/root/reference"


def build(verbose=False):
    """make -C oracle (the reference part only where its tree exists)."""
    r = subprocess.run(["make", "-C", HERE, "-j8", "all"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout[-4000:])
    if verbose:
        print(r.stdout[-2000:])


class Desc(C.Structure):
    _fields_ = [("offset", C.c_int), ("length", C.c_int), ("stride", C.c_int)]


def D(d):
    if d is None:
        return Desc(0, 0, 0)
    if isinstance(d, (tuple, list)):
        return Desc(*[int(x) for x in d])
    return Desc(int(d.offset), int(d.length), int(d.stride))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        L = C.CDLL(ORACLE_LIB)
        vp, i, dd = C.c_void_p, C.c_int, Desc
        sigs = {
            "oracle_eval_stencils_f32": [vp, dd, vp, dd, vp, vp, vp, vp, i, i],
            "oracle_eval_stencils_d1_f32": [vp, dd, vp, dd, vp, dd, vp, dd, vp, vp, vp, vp, vp, vp, i, i],
            "oracle_eval_stencils_d2_f32": [vp, dd, vp, dd, vp, dd, vp, dd, vp, dd, vp, dd, vp, dd,
                                            vp, vp, vp, vp, vp, vp, vp, vp, vp, i, i],
            "oracle_eval_stencils_f64": [vp, dd, vp, dd, vp, vp, vp, vp, i, i],
            "oracle_eval_stencils_omp_f32": [vp, dd, vp, dd, vp, vp, vp, vp, i, i, i],
            "oracle_patch_basis_f32": [i, C.c_uint, C.c_uint, C.c_float, C.c_float,
                                       vp, vp, vp, vp, vp, vp],
            "oracle_patch_basis_f64": [i, C.c_uint, C.c_uint, C.c_double, C.c_double,
                                       vp, vp, vp, vp, vp, vp],
            "oracle_eval_patches_f32": [vp, dd, vp, dd, vp, dd, vp, dd, vp, dd, vp, dd, vp, dd,
                                        i, vp, vp, vp, vp],
            "oracle_eval_patches_f64": [vp, dd, vp, dd, vp, dd, vp, dd, i, vp, vp, vp, vp],
            "oracle_eval_patches_omp_f32": [vp, dd, vp, dd, vp, dd, vp, dd, i, vp, vp, vp, vp, i],
            "oracle_limit_normals_f32": [vp, dd, vp, dd, vp, dd, i, i],
            "oracle_tess_uniform_quad_f32": [i, vp],
            "oracle_tess_coords_f32": [i, i, i, vp, vp],
        }
        for k, a in sigs.items():
            f = getattr(L, k)
            f.argtypes = a
            f.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------- stencils

def eval_stencils(src, src_desc, dst, dst_desc, sizes, offsets, indices, weights,
                  start=0, end=None, threads=0):
    """float32 CpuEvalStencils (threads>0: OpenMP variant). Returns bool."""
    end = len(sizes) if end is None else end
    if threads:
        return bool(lib().oracle_eval_stencils_omp_f32(_p(src), D(src_desc), _p(dst), D(dst_desc),
                                                       _p(sizes), _p(offsets), _p(indices),
                                                       _p(weights), start, end, threads))
    return bool(lib().oracle_eval_stencils_f32(_p(src), D(src_desc), _p(dst), D(dst_desc),
                                               _p(sizes), _p(offsets), _p(indices), _p(weights),
                                               start, end))


def eval_stencils_d1(src, sd, outs, descs, sizes, offsets, indices, ws, start=0, end=None):
    end = len(sizes) if end is None else end
    a = []
    for o, d in zip(outs, descs):
        a += [_p(o), D(d)]
    return bool(lib().oracle_eval_stencils_d1_f32(_p(src), D(sd), *a, _p(sizes), _p(offsets),
                                                  _p(indices), *[_p(w) for w in ws], start, end))


def eval_stencils_d2(src, sd, outs, descs, sizes, offsets, indices, ws, start=0, end=None):
    end = len(sizes) if end is None else end
    a = []
    for o, d in zip(outs, descs):
        a += [_p(o), D(d)]
    return bool(lib().oracle_eval_stencils_d2_f32(_p(src), D(sd), *a, _p(sizes), _p(offsets),
                                                  _p(indices), *[_p(w) for w in ws], start, end))


def eval_stencils_f64(src, src_desc, dst, dst_desc, sizes, offsets, indices, weights,
                      start=0, end=None):
    end = len(sizes) if end is None else end
    return bool(lib().oracle_eval_stencils_f64(_p(src), D(src_desc), _p(dst), D(dst_desc),
                                               _p(sizes), _p(offsets), _p(indices), _p(weights),
                                               start, end))


def apply_stencils(src_rows, sizes, offsets, indices, weights, dtype=np.float32):
    """Convenience: rows of src (V, L) -> (N, L) with the oracle."""
    src_rows = np.ascontiguousarray(src_rows, dtype)
    V, L = src_rows.shape
    n = len(sizes)
    out = np.zeros((n, L), dtype)
    sizes = np.ascontiguousarray(sizes, np.int32)
    offsets = np.ascontiguousarray(offsets, np.int32)
    indices = np.ascontiguousarray(indices, np.int32)
    weights = np.ascontiguousarray(weights, dtype)
    f = eval_stencils if dtype == np.float32 else eval_stencils_f64
    ok = f(src_rows, (0, L, L), out, (0, L, L), sizes, offsets, indices, weights)
    assert ok
    return out


# ----------------------------------------------------------------- patches

def patch_basis(ptype, field0, field1, s, t, deriv=1, dtype=np.float32):
    """(n, wP, wDs, wDt, wDss, wDst, wDtt) from the restated basis."""
    w = [np.zeros(20, dtype) for _ in range(6)]
    want = [True, deriv >= 1, deriv >= 1, deriv >= 2, deriv >= 2, deriv >= 2]
    ptrs = [_p(w[k]) if want[k] else None for k in range(6)]
    f = lib().oracle_patch_basis_f32 if dtype == np.float32 else lib().oracle_patch_basis_f64
    n = f(int(ptype), int(field0) & 0xffffffff, int(field1) & 0xffffffff, float(s), float(t), *ptrs)
    return (n, *w)


def eval_patches(src, sd, outs, descs, coords, arrays, indices, params, threads=0):
    """float32 CpuEvaluator::EvalPatches; outs/descs: up to 6 (None allowed)."""
    outs = list(outs) + [None] * (6 - len(outs))
    descs = list(descs) + [None] * (6 - len(descs))
    n = len(coords)
    if threads:
        a = []
        for o, d in zip(outs[:3], descs[:3]):
            a += [_p(o), D(d)]
        return bool(lib().oracle_eval_patches_omp_f32(_p(src), D(sd), *a, n, _p(coords),
                                                      _p(arrays), _p(indices), _p(params),
                                                      threads))
    a = []
    for o, d in zip(outs, descs):
        a += [_p(o), D(d)]
    return bool(lib().oracle_eval_patches_f32(_p(src), D(sd), *a, n, _p(coords), _p(arrays),
                                              _p(indices), _p(params)))


def eval_patches_f64(src, sd, outs, descs, coords, arrays, indices, params):
    outs = list(outs) + [None] * (3 - len(outs))
    descs = list(descs) + [None] * (3 - len(descs))
    a = []
    for o, d in zip(outs, descs):
        a += [_p(o), D(d)]
    return bool(lib().oracle_eval_patches_f64(_p(src), D(sd), *a, len(coords), _p(coords),
                                              _p(arrays), _p(indices), _p(params)))


# ---------------------------------------------------------- limit normals

def limit_normals(du, du_desc, dv, dv_desc, n, n_desc, count, normalize=True):
    """float32 cross(dPdu, dPdv) [normalized], the reference tutorial's
    steps (far_tutorial_2_3.cpp:73-87). Returns bool."""
    return bool(lib().oracle_limit_normals_f32(_p(du), D(du_desc), _p(dv), D(dv_desc), _p(n),
                                               D(n_desc), int(count), int(bool(normalize))))


def tess_coords(scheme, face_size, rates):
    """Bfr::Tessellation::GetCoords<float> of a face of `face_size` corners
    under `scheme` (0 bilinear, 1 Catmark, 2 Loop) at `rates` (1 uniform,
    2 inner of a quad, N outer, N + 1 / N + 2 outer and inner): (n, 2)
    float32, or None for invalid arguments."""
    r = np.ascontiguousarray(rates, np.int32)
    n = lib().oracle_tess_coords_f32(int(scheme), int(face_size), len(r), _p(r), None)
    if n < 0:
        return None
    uv = np.zeros((n, 2), np.float32)
    assert lib().oracle_tess_coords_f32(int(scheme), int(face_size), len(r), _p(r), _p(uv)) == n
    return uv


def tess_uniform_quad(rate):
    """(rate + 1)^2 x 2 float32 (u, v) of Bfr's uniform quad pattern
    (bfr/tessellation.cpp:158-174, 929-990)."""
    uv = np.zeros(((rate + 1) ** 2, 2), np.float32)
    n = lib().oracle_tess_uniform_quad_f32(int(rate), _p(uv))
    assert n == len(uv), n
    return uv
