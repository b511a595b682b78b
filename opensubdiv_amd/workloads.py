"""Multi-mesh workloads: many cube-sphere meshes, each with many frames.

``sample_meshes`` draws the batched configuration: per mesh a cube-sphere
with n x n quads per cube face, n = floor(8 * 16^U) (log-uniform in
[8, 128]), 5 % of edges creased with sharpness U[0, 4), uniform refinement
level U{2, 3, 4}, U{1..16} frames of control points (xyz displaced along the
normal by 0.1 sin(2 pi f <d, P> + phi), plus 4 primvars U[0, 1)), and
log-uniform 64K..1M PatchCoords (random ptex face, (s, t) ~ U[0, 1)^2) on
the feature-adaptive patch table of the same isolation level.

``compile_mesh`` builds one mesh's host data with the table compiler
(libosdfar): the level-l refinement table, the adaptive patch table and its
patch-point stencils, and the coords (listed in patch order -- the same
random set, ordered so neighbouring coords share control points).
``DeviceMesh`` uploads it: one vertex buffer [control | patch points] with
frames interleaved per vertex and each frame's record padded to a multiple
of 4 floats (element (v, f, k) at (v * F + f) * Lp + k, Lp = 8 for L = 7:
aligned 32-byte records, four frames of a vertex in one 128-byte line for
the gathers), frame-major
refinement and P / dPdu / dPdv outputs (element (r, f, k) at
(f * rows + r) * L + k: rows of one frame adjacent, as in per-frame vertex
buffers; measured faster than interleaving a row's frames), and the
segments a ``Batch`` (or the one-mesh-per-launch path) evaluates.
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor, as_completed
from dataclasses import dataclass

import numpy as np

from . import far, meshes


@dataclass
class MeshSpec:
    index: int
    n: int
    level: int
    frames: int
    ncoords: int
    seed: int


def sample_meshes(count=256, seed=0, n_range=(8, 128), levels=(2, 4), frames=(1, 16),
                  coords=(65536, 1 << 20)):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(math.floor(n_range[0] * (n_range[1] / n_range[0]) ** rng.uniform()))
        lvl = int(rng.integers(levels[0], levels[1] + 1))
        nf = int(rng.integers(frames[0], frames[1] + 1))
        nc = int(math.exp(rng.uniform(math.log(coords[0]), math.log(coords[1]))))
        out.append(MeshSpec(i, n, lvl, nf, nc, int(seed) * 100003 + 7919 * i + 1))
    return out


def spec_rows(s: MeshSpec):
    """Closed forms: control vertices and level-l vertices of the cube-sphere
    (V = F + 2 for a closed genus-0 quad mesh)."""
    return 6 * s.n * s.n + 2, 6 * s.n * s.n * 4 ** s.level + 2


# Measured on creased cube-spheres (5 % creased edges) by level: table
# entries per refined vertex, patch points per level-l vertex, entries per
# patch point, patches per patch point.
_LEVEL_STATS = {1: (6.3, 0.9, 7.0, 0.5), 2: (10.6, 0.52, 8.8, 0.51), 3: (13.2, 0.22, 10.3, 0.45),
                4: (14.6, 0.075, 11.0, 0.43)}


def estimate_device_bytes(s: MeshSpec, length=7):
    """Estimate of one mesh's device footprint (tables in the slice layout
    with 15 % padding, vertex / output buffers, patch tables and coords),
    from the measured per-level statistics plus 5 %; used to fit a sample
    into an HBM budget."""
    nc, nr = spec_rows(s)
    epr, ppr, eppt, patches_pp = _LEVEL_STATS.get(s.level, _LEVEL_STATS[4])
    npts = int(ppr * nr) + 480
    F, L = s.frames, length
    tables = 1.15 * 8 * (epr * nr + eppt * npts) + 12 * (nr + npts)
    patches = int(patches_pp * npts) * (16 * 4 + 12) + 20 * s.ncoords
    buffers = 4 * F * L * (nc + npts + nr + 3 * s.ncoords)
    return int(1.05 * (tables + patches + buffers))


def sample_fitting(count, seed, budget_bytes, length=7, max_tries=64):
    """First seed >= `seed` whose sample fits the budget; returns (seed, specs)."""
    for k in range(max_tries):
        specs = sample_meshes(count, seed + k)
        if sum(estimate_device_bytes(s, length) for s in specs) <= budget_bytes:
            return seed + k, specs
    raise ValueError("no sample fits the budget")


@dataclass
class HostMesh:
    spec: MeshSpec
    length: int
    nctrl: int
    refine: object          # far.RawStencilTable
    points: dict            # patch-point stencils (numpy)
    arrays: np.ndarray
    indices: np.ndarray
    params: np.ndarray
    coords: np.ndarray      # PATCH_COORD_DTYPE, in patch order
    frames: np.ndarray      # (F, nctrl, L) float32
    source: object = None   # table_cache TableFile the arrays above view, if loaded

    def release(self):
        if self.refine is not None:
            self.refine.close()
            self.refine = None
        if self.source is not None:
            self.source.close()
            self.source = None


def mesh_recipe(spec: MeshSpec) -> str:
    """What compile_mesh builds from a spec's topology (the table cache's
    recipe string): everything but the frames, which are not cached."""
    return (f"cube-sphere tables v3: refine vertex L{spec.level}; patches iso {spec.level} "
            f"gregory-basis end caps; coords {spec.ncoords} random @ seed {spec.seed + 1} "
            f"in patch order")


def compile_mesh(spec: MeshSpec, length=7, threads=1, cache=None) -> HostMesh:
    """One mesh's host data; with a table_cache.TableCache, the tables and
    coords come from its file for this topology and recipe when there is a
    valid one, and are compiled and stored otherwise."""
    far.set_threads(threads)
    mesh = meshes.cube_sphere(spec.n, seed=spec.seed, radial_noise=0.01, crease_fraction=0.05,
                              sharpness_max=4.0)
    frames = meshes.deform_frames(mesh, spec.frames, seed=spec.seed + 2,
                                  extra_primvars=length - 3)
    key = cache.key(mesh, mesh_recipe(spec)) if cache is not None else None
    if cache is not None:
        tf = cache.load(key)
        if tf is not None:
            from .osd import PATCH_ARRAY_DTYPE, PATCH_COORD_DTYPE, PATCH_PARAM_DTYPE
            from .table_cache import MappedStencilTable
            try:
                points = far.StencilTable(
                    {k: tf.array("points." + k) for k in ("sizes", "indices", "weights")},
                    num_control=mesh.nverts)
                off = np.zeros(len(points["sizes"]), np.int64)
                np.cumsum(points["sizes"][:-1], out=off[1:])
                points["offsets"] = off.astype(np.int32) if len(points["indices"]) < 2 ** 31 \
                    else off
                h = HostMesh(spec, length, mesh.nverts, MappedStencilTable(tf, "refine",
                                                                           mesh.nverts),
                             points, tf.array("patch.arrays", PATCH_ARRAY_DTYPE),
                             tf.array("patch.indices"),
                             tf.array("patch.params", PATCH_PARAM_DTYPE),
                             tf.array("coords", PATCH_COORD_DTYPE), frames, tf)
            except (KeyError, ValueError):
                tf.close()                  # a file of another layout: a miss
                cache.count(False, undo_hit=True)
            else:
                return h
    topo = far.Topology(mesh)
    topo.cache_stencils(spec.level)
    refine = topo.stencil_table_raw()
    pt = topo.patch_table(spec.level)
    topo.drop_stencils()
    points = pt.point_stencils()
    f, s, t = meshes.random_patch_locations(mesh.nfaces, spec.ncoords, seed=spec.seed + 1)
    coords = pt.find(f, s, t)
    coords = coords[np.argsort(coords["patchIndex"], kind="stable")]
    if cache is not None:
        from .table_cache import raw_table_arrays
        arrays = raw_table_arrays(refine, "refine")
        arrays.update({"points." + k: points[k] for k in ("sizes", "indices", "weights")})
        arrays.update({"patch.arrays": pt.arrays, "patch.indices": pt.indices,
                       "patch.params": pt.params, "coords": coords})
        cache.store(key, arrays)
    return HostMesh(spec, length, mesh.nverts, refine, points, pt.arrays, pt.indices, pt.params,
                    coords, frames)


class DeviceMesh:
    """One mesh's device data and its batch segments (see module docstring)."""

    def __init__(self, h: HostMesh, stream=None, own_outputs=True, layout=2, keep_csr=False,
                 order=2):
        import torch
        from .osd import HipStencilTable, _device_tensor
        s = h.spec
        F, L = s.frames, h.length
        self.spec, self.length = s, L
        self.lp = Lp = (L + 3) // 4 * 4
        self.nctrl = h.nctrl
        self.npts = int(len(h.points["sizes"]))
        self.nrefine = int(h.refine.num_stencils)
        self.nnz_refine = int(h.refine.nnz)
        self.nnz_points = int(len(h.points["indices"]))
        self.ncoords = int(len(h.coords))
        self.coord_points = coord_point_count(h.arrays, h.params, h.coords)
        self.npatches = int(len(h.params))
        self.npatch_indices = int(len(h.indices))
        dev = "cuda"
        self.vbuf = torch.zeros((self.nctrl + self.npts) * F * Lp, dtype=torch.float32, device=dev)
        ctrl = np.zeros((self.nctrl, F, Lp), np.float32)
        ctrl[:, :, :L] = h.frames.transpose(1, 0, 2)
        self.vbuf[:ctrl.size].copy_(torch.from_numpy(ctrl.reshape(-1)))
        self.stage = 0
        self.refined = None
        self.out = None
        if own_outputs:
            self.bind_outputs(torch.empty(self.output_floats(), dtype=torch.float32, device=dev))
        # stencil tables in slice layout `layout` (0 record slices, 2
        # locality-ordered, rows by dominant vertex (order 2: a slice's
        # control vertices are staged once per frame chunk, so fewer of them
        # beats the fewer padded entries of order 0's size windows), or in
        # Far's order (order 1)); the batch kernels need no CSR arrays
        self.refine_table = HipStencilTable(h.refine, stream, layout=layout, keep_csr=keep_csr,
                                            order=order)
        self.point_table = HipStencilTable(h.points, stream, layout=layout, keep_csr=keep_csr,
                                           order=order)
        self.coords = _device_tensor(h.coords)
        self.arrays = _device_tensor(h.arrays)
        self.indices = _device_tensor(h.indices)
        self.params = _device_tensor(h.params)

    def output_floats(self):
        """Refined vertices + P, dPdu, dPdv for every frame."""
        F, L = self.spec.frames, self.length
        return (self.nrefine + 3 * self.ncoords) * F * L

    def bind_outputs(self, region, stage=0):
        """Places the outputs in `region` (a float tensor of output_floats()
        elements, possibly a view of a shared arena) and tags the segments
        with `stage`."""
        F, L = self.spec.frames, self.length
        nr, nc = self.nrefine * F * L, self.ncoords * F * L
        self.refined = region[:nr]
        self.out = [region[nr + k * nc:nr + (k + 1) * nc] for k in range(3)]
        self.stage = int(stage)

    def stencil_segments(self):
        F, L, Lp = self.spec.frames, self.length, self.lp
        v = F * Lp
        return [dict(table=self.refine_table, src=self.vbuf, src_v=v, src_f=Lp,
                     dst=self.refined, dst_v=L, dst_f=self.nrefine * L, frames=F,
                     stage=self.stage),
                dict(table=self.point_table, src=self.vbuf, src_v=v, src_f=Lp,
                     dst=self.vbuf, dst_off=self.nctrl * v, dst_v=v, dst_f=Lp, frames=F,
                     stage=self.stage)]

    def patch_segment(self):
        F, L, Lp = self.spec.frames, self.length, self.lp
        v = F * Lp
        return dict(n=self.ncoords, coords=self.coords, arrays=self.arrays,
                    indices=self.indices, params=self.params, src=self.vbuf, src_v=v,
                    src_f=Lp, P=self.out[0], dPdu=self.out[1], dPdv=self.out[2], out_v=L,
                    out_f=self.ncoords * L, frames=F, stage=self.stage)

    def device_bytes(self):
        t = self.refine_table.device_bytes() + self.point_table.device_bytes()
        b = 4 * (self.vbuf.numel() + self.output_floats())
        return t + b + 4 * (self.coords.numel() + self.arrays.numel() + self.indices.numel() +
                            self.params.numel())

    def points_per_step(self):
        """Stencil outputs (refined vertices) and PatchCoords, per frame."""
        return self.spec.frames * (self.nrefine + self.ncoords)

    def valu_ops(self):
        """Floating-point operations one step cannot avoid when every sum is
        formed as the reference forms it -- a rounded product, then a
        rounded add, per term, so no fused multiply-add: per frame and
        float, a multiply and an add per stencil entry (refinement and
        patch-point tables) and three of each per patch point of every
        PatchCoord (P, dP/du, dP/dv)."""
        F, L = self.spec.frames, self.length
        return F * L * (2 * (self.nnz_refine + self.nnz_points) + 6 * self.coord_points)

    def algorithmic_bytes(self):
        """Minimum HBM traffic of one step: tables, coords and patch tables
        once (frames are extra columns); per frame, control points read and
        refined / patch points / P, dPdu, dPdv written once, and the patch
        points read once by the patch phase."""
        F, L = self.spec.frames, self.length
        e = 4 * L
        once = (4 * self.nrefine + 8 * self.nnz_refine + 4 * self.npts + 8 * self.nnz_points +
                20 * self.ncoords + 4 * self.npatch_indices + 12 * self.npatches)
        per_frame = (e * (self.nctrl + self.nrefine) + e * (self.nctrl + self.npts) +
                     e * (self.nctrl + self.npts) + 3 * e * self.ncoords)
        return once + F * per_frame


# control points of each patch type (Far::PatchDescriptor::GetNumControlVertices)
_TYPE_POINTS = {3: 4, 4: 3, 5: 12, 6: 16, 9: 20, 10: 18}


def coord_point_count(arrays, params, coords):
    """Control points summed over PatchCoords: each coord's patch type is
    its array's descriptor, or the array's regular descriptor when the
    patch's param says regular (bit 5 of field1)."""
    if len(coords) == 0:
        return 0
    a = arrays[coords["arrayIndex"]]
    f1 = params["field1"][coords["patchIndex"]].astype(np.int64)
    ptype = np.where((f1 >> 5) & 1, a["regDesc"], a["desc"])
    table = np.zeros(16, np.int64)
    for t, n in _TYPE_POINTS.items():
        table[t] = n
    return int(table[ptype].sum())


def stage_outputs(dms, arena_floats):
    """Assigns meshes, in order, to stages whose outputs fit one arena of
    `arena_floats` floats (a mesh that fits nowhere alone raises), allocates
    the arena and binds every mesh's outputs into it; meshes of different
    stages share the arena.  Returns (arena, number of stages)."""
    import torch
    need = [d.output_floats() for d in dms]
    if max(need) > arena_floats:
        raise MemoryError("one mesh's outputs exceed the output arena")
    stage, used, place = 0, 0, []
    for n in need:
        n64 = (n + 63) // 64 * 64
        if used + n64 > arena_floats:
            stage, used = stage + 1, 0
        place.append((stage, used))
        used += n64
    peak = max(o + (n + 63) // 64 * 64 for (st, o), n in zip(place, need))
    arena = torch.empty(peak, dtype=torch.float32, device="cuda")
    for d, (st, o), n in zip(dms, place, need):
        d.bind_outputs(arena[o:o + n], st)
    return arena, stage + 1


def build_device_meshes(specs, length=7, stream=None, workers=8, log=None, own_outputs=True,
                        cache=None, layout=2, order=2):
    """Compiles the meshes on `workers` host threads (largest first), or
    loads them from a table_cache.TableCache, and uploads each as it
    completes (tables in slice layout `layout`, row order `order`); returns
    DeviceMeshes in spec order (with their own output buffers, or none for
    stage_outputs)."""
    largest_first = sorted(specs, key=lambda s: -spec_rows(s)[1])
    done = {}
    with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
        futs = {ex.submit(compile_mesh, s, length, 1, cache): s for s in largest_first}
        for k, fut in enumerate(as_completed(futs)):
            h = fut.result()
            done[h.spec.index] = DeviceMesh(h, stream, own_outputs, layout=layout, order=order)
            h.release()
            del h
            if log and (k + 1) % 16 == 0:
                log(f"[workload] {k + 1}/{len(specs)} meshes on the device")
    return [done[s.index] for s in specs]
