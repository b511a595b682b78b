"""Synthetic meshes and control-point frames for the evaluator workloads.

All generators are deterministic functions of their seed (numpy PCG64).

* ``unit_cube`` -- the 8-vertex / 6-quad cube of
  ``tutorials/osd/tutorial_0/osd_tutorial_0.cpp:42-60`` (every corner has
  valence 3), optionally perturbed by seeded uniform noise.
* ``cube_sphere`` -- a closed quad mesh: each cube face is an n x n grid,
  vertices are shared along cube edges (6 n^2 + 2 vertices, 6 n^2 faces,
  12 n^2 edges, 8 valence-3 corners), projected to the unit sphere with
  seeded radial noise; a seeded fraction of edges is creased with sharpness
  uniform in [0, smax).
* ``cube_sphere_uvs`` -- a face-varying UV channel for ``cube_sphere``: one
  chart per cube face, laid out in a 3 x 2 atlas, so values split along the
  twelve cube edges (seams) and the eight corners carry three values each.
* ``deform_frames`` -- seeded sinusoidal displacement along the vertex
  normal, one frame per call index, plus extra uniform primvars.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class QuadMesh:
    """Closed quad mesh topology + positions + creases (Far TopologyDescriptor
    fields, far/topologyDescriptor.h:40-80)."""

    nverts: int
    face_verts: np.ndarray            # (F, 4) int32, counter-clockwise from outside
    positions: np.ndarray             # (V, 3) float64
    crease_pairs: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    crease_weights: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.float32))
    corner_verts: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.int32))
    corner_weights: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.float32))

    @property
    def nfaces(self) -> int:
        return int(self.face_verts.shape[0])

    def edges(self) -> np.ndarray:
        """Unique undirected edges (E, 2), lower index first, sorted."""
        fv = self.face_verts
        e = np.stack([fv, np.roll(fv, -1, axis=1)], axis=2).reshape(-1, 2)
        e = np.sort(e, axis=1)
        return np.unique(e, axis=0)


_CUBE_VERTS = np.array([[-0.5, -0.5, 0.5], [0.5, -0.5, 0.5], [-0.5, 0.5, 0.5],
                        [0.5, 0.5, 0.5], [-0.5, 0.5, -0.5], [0.5, 0.5, -0.5],
                        [-0.5, -0.5, -0.5], [0.5, -0.5, -0.5]], dtype=np.float64)
_CUBE_FACES = np.array([[0, 1, 3, 2], [2, 3, 5, 4], [4, 5, 7, 6],
                        [6, 7, 1, 0], [1, 7, 5, 3], [6, 0, 2, 4]], dtype=np.int32)


def unit_cube(seed: int | None = 0, noise: float = 0.05) -> QuadMesh:
    """The tutorial cube; xyz += U(-noise, noise) when seed is not None."""
    p = _CUBE_VERTS.copy()
    if seed is not None:
        rng = np.random.default_rng(seed)
        p = p + rng.uniform(-noise, noise, size=p.shape)
    return QuadMesh(8, _CUBE_FACES.copy(), p)


# Cube face frames: (fixed axis, fixed value, u axis, v axis) with u x v
# pointing outward, so faces are counter-clockwise seen from outside.
_FACES = [(0, 1, 1, 2), (0, 0, 2, 1), (1, 1, 2, 0), (1, 0, 0, 2), (2, 1, 0, 1), (2, 0, 1, 0)]


def cube_sphere(n: int, seed: int = 0, radial_noise: float = 0.01,
                crease_fraction: float = 0.05, sharpness_max: float = 4.0) -> QuadMesh:
    """Cube-sphere with n x n quads per cube face (see module docstring)."""
    n = int(n)
    a = np.arange(n + 1)
    A, B = np.meshgrid(a, a, indexing="ij")          # A along u, B along v
    keys, quads = [], []
    for fixed, val, ua, va in _FACES:
        lat = np.zeros((n + 1, n + 1, 3), np.int64)
        lat[..., fixed] = val * n
        lat[..., ua] = A
        lat[..., va] = B
        keys.append(lat.reshape(-1, 3))
    allkeys = np.concatenate(keys)                   # (6 (n+1)^2, 3)
    code = (allkeys[:, 0] * (n + 1) + allkeys[:, 1]) * (n + 1) + allkeys[:, 2]
    uniq, inv = np.unique(code, return_inverse=True)
    nv = uniq.shape[0]
    per = (n + 1) * (n + 1)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    for f in range(6):
        ids = inv[f * per:(f + 1) * per].reshape(n + 1, n + 1)
        q = np.stack([ids[ii, jj], ids[ii + 1, jj], ids[ii + 1, jj + 1], ids[ii, jj + 1]], axis=-1)
        quads.append(q.reshape(-1, 4))
    fv = np.concatenate(quads).astype(np.int32)
    lat = np.zeros((nv, 3), np.float64)
    lat[inv] = allkeys
    p = lat / n * 2.0 - 1.0
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    rng = np.random.default_rng(seed)
    p *= (1.0 + radial_noise * rng.uniform(-1.0, 1.0, size=(nv, 1)))
    mesh = QuadMesh(nv, fv, p)
    if crease_fraction > 0:
        e = mesh.edges()
        k = int(round(crease_fraction * e.shape[0]))
        pick = np.sort(rng.choice(e.shape[0], size=k, replace=False))
        mesh.crease_pairs = e[pick].astype(np.int32)
        mesh.crease_weights = rng.uniform(0.0, sharpness_max, size=k).astype(np.float32)
    return mesh


def cube_sphere_uvs(n: int):
    """Face-varying UVs of ``cube_sphere(n)``: (numValues, value index per
    face-vertex (6 n^2, 4), (numValues, 2) float64 values).  Chart c holds the
    (n + 1)^2 lattice values of cube face c at u = (c % 3 + a / n) / 3,
    v = (c // 3 + b / n) / 2; faces index them in the order cube_sphere emits
    its quads."""
    n = int(n)
    per = (n + 1) * (n + 1)
    a, b = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    idx, uv = [], []
    for c in range(6):
        lid = c * per + a * (n + 1) + b
        idx.append(np.stack([lid[ii, jj], lid[ii + 1, jj], lid[ii + 1, jj + 1], lid[ii, jj + 1]],
                            axis=-1).reshape(-1, 4))
        uv.append(np.stack([((c % 3) + a / n) / 3.0, ((c // 3) + b / n) / 2.0],
                           axis=-1).reshape(-1, 2))
    return 6 * per, np.concatenate(idx).astype(np.int32), np.concatenate(uv)


def vertex_normals(mesh: QuadMesh) -> np.ndarray:
    """Area-weighted vertex normals (unit length)."""
    p = mesh.positions
    fv = mesh.face_verts
    d1 = p[fv[:, 2]] - p[fv[:, 0]]
    d2 = p[fv[:, 3]] - p[fv[:, 1]]
    fn = np.cross(d1, d2)
    vn = np.zeros_like(p)
    for k in range(4):
        np.add.at(vn, fv[:, k], fn)
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
    return vn


def deform_frames(mesh: QuadMesh, nframes: int, seed: int, amplitude: float = 0.1,
                  extra_primvars: int = 4) -> np.ndarray:
    """(nframes, V, 3 + extra) float32 control-point frames: frame k is
    P + a sin(2 pi f_k <d_k, P> + phi_k) n(P), f_k ~ U[1, 8], d_k a random
    unit direction, phi_k ~ U[0, 2 pi); extra primvars ~ U[0, 1)."""
    rng = np.random.default_rng(seed)
    p = mesh.positions
    nrm = vertex_normals(mesh)
    out = np.empty((nframes, p.shape[0], 3 + extra_primvars), np.float32)
    for k in range(nframes):
        f = rng.uniform(1.0, 8.0)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        phi = rng.uniform(0.0, 2 * np.pi)
        disp = amplitude * np.sin(2 * np.pi * f * (p @ d) + phi)
        out[k, :, :3] = p + disp[:, None] * nrm
        if extra_primvars:
            out[k, :, 3:] = rng.uniform(0.0, 1.0, size=(p.shape[0], extra_primvars))
    return out


def random_patch_locations(nfaces: int, n: int, seed: int):
    """n seeded (ptex face, s, t) triples: face ~ U{0..nfaces-1}, (s,t) ~ U[0,1)^2."""
    rng = np.random.default_rng(seed)
    f = rng.integers(0, nfaces, size=n, dtype=np.int64).astype(np.int32)
    s = rng.uniform(0.0, 1.0, size=n)
    t = rng.uniform(0.0, 1.0, size=n)
    return f, s, t
