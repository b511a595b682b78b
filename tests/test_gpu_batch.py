"""GPU parity of the batched engine (osdhip_batch_*): many meshes x many
frames in one graph launch.

Every kernel variant (frames per chunk, gather group, slices per item,
static / persistent scheduling), the captured graph (replayed several times
with other work in between) and the autotuned configuration must produce
the same bits as the one-mesh-per-launch path (osdhip_stencil_table_eval /
osdhip_eval_patches per mesh and frame), which in turn equals the oracle
(CpuEvaluator order) on the same inputs."""
import numpy as np
import pytest

import oracle
from opensubdiv_amd import _lib, far, meshes
from opensubdiv_amd.batch import Batch
from opensubdiv_amd.workloads import DeviceMesh, MeshSpec, compile_mesh

pytestmark = pytest.mark.gpu

# small meshes: n x n quads per cube face, refinement / isolation level,
# frames, coords -- Gregory end caps at the valence-3 corners; mesh 5 has
# fewer coords than patches, so its waves of coords reference too many
# points to be staged in LDS; mesh 6 is refined to level 4, whose rows are
# the longest of the batched workload (16 entries, the most the tiled kernel
# keeps in registers) and whose slices list the most control vertices
SPECS = [MeshSpec(0, 3, 2, 3, 700, 11), MeshSpec(1, 2, 1, 1, 300, 12),
         MeshSpec(2, 5, 2, 5, 2500, 13), MeshSpec(3, 4, 3, 2, 1800, 14),
         MeshSpec(4, 2, 2, 4, 64, 15), MeshSpec(5, 12, 2, 2, 700, 16),
         MeshSpec(6, 6, 4, 3, 2000, 17)]


def build(length, layout=0, order=0):
    import torch
    hosts = [compile_mesh(s, length) for s in SPECS]
    dms = [DeviceMesh(h, layout=layout, order=order) for h in hosts]
    for h in hosts:
        h.release()
    torch.cuda.synchronize()
    return hosts, dms


def outputs(dms):
    import torch
    torch.cuda.synchronize()
    res = []
    for d in dms:
        res.append([d.refined.cpu().numpy().copy(),
                    d.vbuf[d.nctrl * d.spec.frames * d.lp:].cpu().numpy().copy()] +
                   [o.cpu().numpy().copy() for o in d.out])
    return res


def poison(dms):
    import torch
    for d in dms:
        d.refined.fill_(float("nan"))
        d.vbuf[d.nctrl * d.spec.frames * d.lp:].fill_(float("nan"))
        for o in d.out:
            o.fill_(float("nan"))
    torch.cuda.synchronize()


def same(a, b, what):
    for i, (x, y) in enumerate(zip(a, b)):
        for k, (u, v) in enumerate(zip(x, y)):
            assert np.array_equal(u.view(np.uint32), v.view(np.uint32)), f"{what}: mesh {i} buffer {k}"


# (floats per vertex, table layout, row order): record slices and
# locality-ordered slices (LDS-staged control points), rows by dominant
# vertex and size windows (0), in Far's order (1) or by dominant vertex
# only (2)
@pytest.fixture(scope="module",
                params=[(7, 0, 0), (3, 0, 0), (7, 2, 0), (3, 2, 0), (7, 2, 1), (7, 2, 2)],
                ids=["L7-rec", "L3-rec", "L7-tiled", "L3-tiled", "L7-tiledfar", "L7-tileddom"])
def workload(request):
    length, layout, order = request.param
    hosts, dms = build(length, layout, order)
    if layout == 2:
        for d in dms:
            info = d.refine_table.layout()
            assert info["layout"] == 2 and info["listed_slices"] > 0
            assert order == 1 or info["listed_slices"] >= 0.9 * info["slices"]
    batch = Batch(length, [s for d in dms for s in d.stencil_segments()],
                  [d.patch_segment() for d in dms])
    batch.layout = layout
    # some waves of coords are staged, some are not
    assert 0 < batch.get("staged_coords") < sum(d.ncoords for d in dms)
    yield length, hosts, dms, batch
    batch.close()


def test_plain_path_matches_oracle(workload):
    """One mesh and frame per launch, checked against the oracle's
    CpuEvaluator-order sums on the same inputs (refinement, patch points,
    P / dPdu / dPdv)."""
    length, hosts, dms, batch = workload
    poison(dms)
    batch.run_plain()
    got = outputs(dms)
    for h, d, g in zip(hosts, dms, got):
        s = h.spec
        F, L = s.frames, length
        mesh = meshes.cube_sphere(s.n, seed=s.seed, radial_noise=0.01, crease_fraction=0.05,
                                  sharpness_max=4.0)
        topo = far.Topology(mesh)
        topo.refine_uniform(s.level)
        ref = topo.stencil_table()
        pts = h.points
        refined = g[0].reshape(F, -1, L).transpose(1, 0, 2)      # frame-major outputs
        ptsout = g[1].reshape(-1, F, d.lp)[:, :, :L]             # frames interleaved, padded
        for f in range(F):
            ctrl = np.ascontiguousarray(h.frames[f])
            want = oracle.apply_stencils(ctrl, ref["sizes"], ref["offsets"], ref["indices"],
                                         ref["weights"])
            np.testing.assert_array_equal(refined[:, f], want, err_msg=f"mesh {s.index} refine f{f}")
            wp = oracle.apply_stencils(ctrl, pts["sizes"], pts["offsets"], pts["indices"],
                                       pts["weights"])
            np.testing.assert_array_equal(ptsout[:, f], wp, err_msg=f"mesh {s.index} points f{f}")
            src = np.ascontiguousarray(np.concatenate([ctrl, wp]))
            n = len(h.coords)
            dst = [np.zeros((n, L), np.float32) for _ in range(3)]
            assert oracle.eval_patches(src, (0, L, L), dst, [(0, L, L)] * 3, h.coords, h.arrays,
                                       h.indices, h.params)
            for k in range(3):
                np.testing.assert_array_equal(g[2 + k].reshape(F, -1, L)[f], dst[k],
                                              err_msg=f"mesh {s.index} patch output {k} f{f}")


def test_patch_staging_lists_stay_in_bounds(workload):
    """The point staging lists k_patch_slots builds at batch creation, read
    back and checked on the host before any kernel relies on them: per wave
    a count in [-1, 256] and contiguous offsets; listed points strictly
    ascending and inside the segment's point buffer; every regular-interior
    coord's 16 slot bytes below its wave's count, naming exactly the
    patch's 16 points.  (The staged patch kernel reads LDS slots and the
    point buffer through these lists.)"""
    from opensubdiv_amd.osd import PATCH_ARRAY_DTYPE, PATCH_COORD_DTYPE, PATCH_PARAM_DTYPE
    length, hosts, dms, batch = workload
    staged = 0
    for seg, d in enumerate(dms):
        lists = batch.patch_staging(seg)
        assert lists is not None
        nu, off, pts, slots = lists["nuniq"], lists["ptsOff"], lists["pts"], lists["slots"]
        nw = (d.ncoords + 63) // 64
        assert len(nu) == nw and ((nu >= -1) & (nu <= 256)).all()
        np.testing.assert_array_equal(off, np.concatenate([[0], np.cumsum(np.maximum(nu, 0))[:-1]]))
        assert len(pts) == int(np.maximum(nu, 0).sum())
        npoints = d.nctrl + d.npts                       # rows of the segment's point buffer
        coords = d.coords.cpu().numpy().view(PATCH_COORD_DTYPE).reshape(-1)
        arrays = d.arrays.cpu().numpy().view(PATCH_ARRAY_DTYPE).reshape(-1)
        params = d.params.cpu().numpy().view(PATCH_PARAM_DTYPE).reshape(-1)
        indices = d.indices.cpu().numpy()
        sb = slots.view(np.uint8).reshape(-1, 16)
        for w in range(nw):
            if nu[w] < 0:
                continue
            lst = pts[off[w]:off[w] + nu[w]]
            assert ((lst >= 0) & (lst < npoints)).all(), (seg, w)
            assert (np.diff(lst) > 0).all(), (seg, w)
            for i in range(64 * w, min(64 * w + 64, d.ncoords)):
                c = coords[i]
                a = arrays[c["arrayIndex"]]
                f1 = int(params[c["patchIndex"]]["field1"])
                ptype = a["regDesc"] if (f1 >> 5) & 1 else a["desc"]
                if ptype != 6 or (f1 >> 7) & 0x1F:
                    continue                              # not regular-interior: no slots
                b = a["indexBase"] + a["stride"] * (c["patchIndex"] - a["primitiveIdBase"])
                assert (sb[i] < nu[w]).all(), (seg, i)
                np.testing.assert_array_equal(lst[sb[i]], indices[b:b + 16], err_msg=f"{seg} {i}")
                staged += 1
    assert staged > 0


def test_graph_launch_matches_plain_and_replays(workload):
    length, hosts, dms, batch = workload
    poison(dms)
    batch.run_plain()
    plain = outputs(dms)
    for rep in range(3):
        poison(dms)
        batch.launch()
        same(plain, outputs(dms), f"graph replay {rep}")
        batch.run_plain()          # other work on the stream between replays


VARIANTS = {0: [dict(a_frames=1, a_group=4), dict(a_frames=1, a_group=8),
                dict(a_frames=2, a_group=4), dict(a_frames=4, a_group=2),
                dict(a_frames=0, a_group=4)],
            2: [dict(a_frames=1, a_group=4), dict(a_frames=1, a_group=8),
                dict(a_frames=2, a_group=4), dict(a_frames=2, a_group=8),
                dict(a_frames=4, a_group=4), dict(a_frames=4, a_group=8)]}
TUNED_A = {0: 30, 2: 36}


@pytest.mark.parametrize("persistent", [0, 1])
def test_every_variant_matches_plain(workload, persistent):
    length, hosts, dms, batch = workload
    poison(dms)
    batch.run_plain()
    plain = outputs(dms)
    for variant in VARIANTS[batch.layout]:
        check_variant(batch, dms, plain, variant, persistent)


def check_variant(batch, dms, plain, variant, persistent):
    """One phase-A variant (dict of knobs) with several item sizes and
    phase-B settings, eager and graph-launched, against the plain path."""
    for k, v in variant.items():
        assert batch.set(k, v) == 0
    for slices, bfr, stg in ((1, 1, 1), (3, 2, 0), (16, 4, 1), (2, 2, 1)):
        assert batch.set("a_item_slices", slices) == 0
        assert batch.set("a_persistent", persistent) == 0
        assert batch.set("b_frames", bfr) == 0
        assert batch.set("b_persistent", persistent) == 0
        assert batch.set("b_stage", stg) == 0
        poison(dms)
        batch.run(3)
        same(plain, outputs(dms), f"eager {variant} slices {slices} b_frames {bfr} stage {stg}")
        poison(dms)
        batch.launch()
        same(plain, outputs(dms), f"graph {variant} slices {slices} b_frames {bfr} stage {stg}")


def test_vertex_major_refined_outputs(workload):
    """Refined outputs with a row's frames adjacent (element (r, f, k) at
    (r * F + f) * L + k; the layout-2 kernel stores them through LDS as row
    runs) equal the plain path and the frame-major results, for every
    frames-per-chunk variant."""
    import torch
    length, hosts, dms, batch = workload
    poison(dms)
    batch.run_plain()
    ref = outputs(dms)
    segs, outs = [], []
    for d in dms:
        s0, s1 = d.stencil_segments()
        F = d.spec.frames
        o = torch.full((d.nrefine * F * length,), float("nan"), device="cuda")
        segs += [dict(s0, dst=o, dst_v=F * length, dst_f=length), s1]
        outs.append(o)
    b2 = Batch(length, segs, [])
    try:
        for fc in (1, 2, 4):
            assert b2.set("a_frames", fc) == 0
            assert b2.set("a_group", 2 if fc == 4 and batch.layout == 0 else 4) == 0
            for o in outs:
                o.fill_(float("nan"))
            b2.run(1)
            torch.cuda.synchronize()
            for d, o, r in zip(dms, outs, ref):
                F = d.spec.frames
                got = o.cpu().numpy().reshape(-1, F, length)
                want = r[0].reshape(F, -1, length).transpose(1, 0, 2)
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
                    f"mesh {d.spec.index} frames per chunk {fc}"
        for o in outs:
            o.fill_(float("nan"))
        b2.run_plain()
        torch.cuda.synchronize()
        for d, o, r in zip(dms, outs, ref):
            F = d.spec.frames
            want = r[0].reshape(F, -1, length).transpose(1, 0, 2)
            assert np.array_equal(o.cpu().numpy().reshape(-1, F, length).view(np.uint32),
                                  want.view(np.uint32))
    finally:
        b2.close()


def test_tuned_configuration_matches_plain(workload):
    length, hosts, dms, batch = workload
    poison(dms)
    batch.run_plain()
    plain = outputs(dms)
    report = batch.tune(reps=1)
    assert len(report["phase_a"]) == TUNED_A[batch.layout] and len(report["phase_b"]) == 12
    poison(dms)
    batch.launch()
    same(plain, outputs(dms), f"tuned {batch.config()}")


def test_batch_knob_and_segment_errors(workload):
    length, hosts, dms, batch = workload
    assert batch.set("a_frames", 3) == _lib.ERR_ARGS
    assert batch.set("a_group", 16) == _lib.ERR_ARGS
    assert batch.set("b_frames", 8) == _lib.ERR_ARGS
    # a pair that has no kernel instance is refused at launch
    cfg = batch.config()
    bad = (4, 8) if batch.layout == 0 else (0, 4)
    assert batch.set("a_frames", bad[0]) == 0 and batch.set("a_group", bad[1]) == 0
    with pytest.raises(RuntimeError):
        batch.run(1)
    assert batch.set("a_frames", cfg["a_frames"]) == 0 and batch.set("a_group", cfg["a_group"]) == 0
    assert batch.set("a_persistent", 2) == _lib.ERR_ARGS
    assert batch.set("b_stage", 2) == _lib.ERR_ARGS
    assert batch.set("nope", 1) == _lib.ERR_UNSUPPORTED
    info = batch.info()
    assert info["stencil_rows"] == sum(d.nrefine + d.npts for d in dms)
    assert info["patch_coords"] == sum(d.ncoords for d in dms)
    # a derivative (3-weight) table is not a batch stencil segment
    from opensubdiv_amd import HipStencilTable
    rng = np.random.default_rng(3)
    t = {"sizes": np.full(10, 2, np.int32), "offsets": np.arange(0, 20, 2, dtype=np.int32),
         "indices": rng.integers(0, 4, 20).astype(np.int32)}
    for k in ("weights", "du", "dv"):
        t[k] = rng.uniform(size=20).astype(np.float32)
    t3 = HipStencilTable(t)
    seg = dict(dms[0].stencil_segments()[0])
    seg["table"] = t3
    with pytest.raises(ValueError):
        Batch(length, [seg], [])
    # tables of both slice layouts in one batch
    t1 = {"sizes": t["sizes"], "offsets": t["offsets"], "indices": t["indices"],
          "weights": t["weights"]}
    other = HipStencilTable(t1, layout=2 - batch.layout)
    seg2 = dict(seg)
    seg2["table"] = other
    with pytest.raises(ValueError):
        Batch(length, [dms[0].stencil_segments()[0], seg2], [])


def test_staged_outputs_share_an_arena():
    """Meshes in several stages whose outputs share one arena (each output
    still written once per launch): every stage equals the plain path bit
    for bit right after it runs, and the arena after a graph launch equals
    the arena after the plain path (the last writes of each region)."""
    import torch
    from opensubdiv_amd.workloads import stage_outputs
    length = 7
    # the level-4 mesh (the largest outputs) between smaller ones, so that
    # an arena of its size holds three stages: meshes before it, it alone,
    # meshes after it
    hosts = [compile_mesh(s, length) for s in SPECS]
    dms = [DeviceMesh(h, own_outputs=False, layout=0) for h in hosts]
    need = [d.output_floats() for d in dms]
    big = dms[need.index(max(need))]
    dms = [d for d in dms if d is not big]
    dms.insert(len(dms) // 2, big)
    need = [d.output_floats() for d in dms]
    arena, nstages = stage_outputs(dms, max(need) + 64)
    assert nstages >= 3 and arena.numel() < sum(need)
    batch = Batch(length, [s for d in dms for s in d.stencil_segments()],
                  [d.patch_segment() for d in dms])
    assert batch.num_stages() == nstages
    for st in range(nstages):
        members = [d for d in dms if d.stage == st]
        assert members
        arena.fill_(float("nan"))
        batch.run_plain_stage(st)
        want = outputs(members)
        arena.fill_(float("nan"))
        batch.run_stage(st)
        same(want, outputs(members), f"stage {st} batched")
    arena.fill_(float("nan"))
    batch.run_plain()
    torch.cuda.synchronize()
    final_plain = arena.cpu().numpy().copy()
    arena.fill_(float("nan"))
    batch.launch()
    torch.cuda.synchronize()
    assert np.array_equal(arena.cpu().numpy().view(np.uint32), final_plain.view(np.uint32))
    batch.close()
    for h in hosts:
        h.release()


def test_meshes_loaded_from_the_table_cache_match_compiled(tmp_path):
    """Meshes whose tables and coords come from table files (stored by a
    first compile, then mapped) evaluate to the same bits as freshly
    compiled ones, on the plain path and in a graph launch."""
    from opensubdiv_amd.table_cache import TableCache
    length = 7
    cache = TableCache(tmp_path)
    for s in SPECS:
        compile_mesh(s, length, cache=cache).release()
    assert (cache.hits, cache.misses) == (0, len(SPECS))
    cached = [compile_mesh(s, length, cache=cache) for s in SPECS]
    assert cache.hits == len(SPECS)
    dms = [DeviceMesh(h, layout=2) for h in cached]
    for h in cached:
        h.release()
    _, fresh = build(length, 2)
    results = []
    for group in (dms, fresh):
        batch = Batch(length, [s for d in group for s in d.stencil_segments()],
                      [d.patch_segment() for d in group])
        poison(group)
        batch.run_plain()
        plain = outputs(group)
        poison(group)
        batch.launch()
        same(plain, outputs(group), "graph launch")
        results.append(plain)
        batch.close()
    same(results[0], results[1], "cached vs compiled")
