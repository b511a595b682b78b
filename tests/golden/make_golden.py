"""Writes the golden fixtures of tests/golden/*.npz from the REAL OpenSubdiv
reference (Far factories + Osd::CpuEvaluator), built from the reference
sources into oracle/_ref/libosdref.so by oracle/Makefile.

Run in the build container (where the reference tree exists):
    make -C oracle && python tests/golden/make_golden.py [fixture function ...]

Each fixture holds inputs (tables, control points, coords) and the
reference's outputs; nothing of the reference's source goes into them.
  cube_l3.npz           cfg0: tutorial cube, seeded noise, uniform L3 stencils
                        (generateIntermediateLevels=false), CpuEvaluator
                        output, double table + double output (FP64 oracle)
  cubesphere_l2.npz     creased cube-sphere n=4, uniform L2 stencils with
                        intermediate levels, xyz + 4 primvars (stride 7)
  cubesphere_limit.npz  LimitStencilTable (du/dv + 2nd) at seeded locations,
                        CpuEvaluator outputs (1st and 2nd derivative arities)
  cubesphere_patches.npz iso-3 feature-adaptive PatchTable (B-spline +
                        GregoryBasis end caps), appended local-point
                        stencils, 4096 PatchMap coords, CpuEvaluator
                        EvalPatches P/du/dv and 2nd derivatives, and Far
                        EvaluateBasis<double> weights for the first 256
  loop_patches.npz      triangular patch types: Loop-scheme meshes (a
                        creased triangulated cube-sphere, closed, iso 2; an
                        open triangle grid with EDGE_AND_CORNER boundaries,
                        iso 3; a one-quad-wide strip, whose patches touch
                        both boundaries, iso 2) whose adaptive tables hold
                        LOOP box-spline patches (edge and vertex boundary
                        masks) and
                        GREGORY_TRIANGLE end caps, and a uniform level-2
                        Loop table of TRIANGLES; per mesh: tables, refined
                        control points, PatchMap coords, CpuEvaluator
                        EvalPatches outputs (P/du/dv and the 2nd-derivative
                        arity) and Far EvaluateBasis<double> weights
  bfr_surfaces.npz      Bfr::Surface<double>::Evaluate (approximation levels
                        2 smooth / 6 sharp) of three meshes -- a creased
                        cube-sphere, a pentagonal prism (n-gon faces, so
                        Bfr's quad sub-face parameterization), an open
                        creased grid -- at seeded (face, u, v) queries in
                        Bfr's parameterization: meshes, queries, P/Du/Dv
  tess_coords.npz       Bfr::Tessellation::GetCoords<float> for 300 seeded
                        (scheme, face size, rates) patterns: quads, Loop
                        triangles and Catmark n-gon sub-faces (3-12 sides),
                        uniform, inner-only, outer and outer + inner rates
  cubesphere_fvar.npz   a face-varying UV channel (one chart per cube face,
                        seams along the cube edges) on the creased
                        cube-sphere n=4, iso 3, two ways: Far's defaults
                        (FVAR_LINEAR_ALL, legacy linear QUADS patches) and
                        smooth UVs (FVAR_LINEAR_CORNERS_PLUS1, B-spline and
                        Gregory face-varying patches with face-varying local
                        points); per variant: the face-varying stencil table,
                        the refined UV buffer, the flattened face-varying
                        patch arrays / indices / params, PatchMap coords and
                        CpuEvaluator EvalPatchesFaceVarying outputs (P/du/dv
                        and the 2nd-derivative arity)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from opensubdiv_amd import meshes  # noqa: E402
from opensubdiv_amd.osd import PATCH_COORD_DTYPE  # noqa: E402


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


def cube_l3():
    m = meshes.unit_cube(seed=0)
    topo = ref.Topology(m)
    t = ref.stencils_uniform(topo, 3)
    td = ref.stencils_uniform(topo, 3, double=True)
    src = m.positions.astype(np.float32)
    nC, nS = 8, len(t["sizes"])
    buf = np.zeros((nC + nS, 3), np.float32)
    buf[:nC] = src
    assert ref.cpu_eval_stencils(buf, (0, 3, 3), buf, (nC * 3, 3, 3), t)
    out64 = np.zeros((nS, 3))
    from oracle import eval_stencils_f64
    srcd = m.positions.astype(np.float64)
    assert eval_stencils_f64(srcd, (0, 3, 3), out64, (0, 3, 3), td["sizes"], td["offsets"],
                             td["indices"], td["weights"])
    save("cube_l3.npz", src=src, src64=srcd, sizes=t["sizes"], offsets=t["offsets"],
         indices=t["indices"], weights=t["weights"], weights64=td["weights"],
         cpu_out=buf[nC:].copy(), out64=out64)


def cubesphere_l2():
    m = meshes.cube_sphere(4, seed=11)
    topo = ref.Topology(m)
    t = ref.stencils_uniform(topo, 2, intermediate=True)
    frames = meshes.deform_frames(m, 1, seed=5)[0]           # (V, 7)
    nC, nS = m.nverts, len(t["sizes"])
    buf = np.zeros((nC + nS, 7), np.float32)
    buf[:nC] = frames
    assert ref.cpu_eval_stencils(buf, (0, 7, 7), buf, (nC * 7, 7, 7), t)
    # the same table on xyz only, sharing the stride-7 buffer (offset 0, length 3)
    buf3 = np.zeros((nC + nS, 7), np.float32)
    buf3[:nC] = frames
    assert ref.cpu_eval_stencils(buf3, (0, 3, 7), buf3, (nC * 7, 3, 7), t)
    save("cubesphere_l2.npz", face_verts=m.face_verts, crease_pairs=m.crease_pairs,
         crease_weights=m.crease_weights, src=frames, sizes=t["sizes"], offsets=t["offsets"],
         indices=t["indices"], weights=t["weights"], cpu_out7=buf[nC:].copy(),
         cpu_out3=buf3[nC:, :3].copy())


def cubesphere_limit():
    m = meshes.cube_sphere(4, seed=12)
    topo = ref.Topology(m)
    f, s, t = meshes.random_patch_locations(m.nfaces, 512, seed=13)
    order = np.argsort(f, kind="stable")
    f, s, t = f[order], s[order], t[order]
    lt = ref.stencils_limit(topo, 3, f, s, t, d2=True)
    src = m.positions.astype(np.float32)
    n = len(lt["sizes"])
    outs = [np.zeros((n, 3), np.float32) for _ in range(6)]
    assert ref.cpu_eval_stencils_d2(src, (0, 3, 3), outs, [(0, 3, 3)] * 6, lt)
    outs1 = [np.zeros((n, 3), np.float32) for _ in range(3)]
    assert ref.cpu_eval_stencils_d1(src, (0, 3, 3), outs1, [(0, 3, 3)] * 3, lt)
    save("cubesphere_limit.npz", face_verts=m.face_verts, crease_pairs=m.crease_pairs,
         crease_weights=m.crease_weights, src=src, ptex=f, s=s, t=t, sizes=lt["sizes"],
         offsets=lt["offsets"], indices=lt["indices"], weights=lt["weights"], du=lt["du"],
         dv=lt["dv"], duu=lt["duu"], duv=lt["duv"], dvv=lt["dvv"],
         cpu_p=outs[0], cpu_du=outs[1], cpu_dv=outs[2], cpu_duu=outs[3], cpu_duv=outs[4],
         cpu_dvv=outs[5], cpu1_p=outs1[0], cpu1_du=outs1[1], cpu1_dv=outs1[2])


def cubesphere_patches():
    m = meshes.cube_sphere(4, seed=14)
    topo = ref.Topology(m)
    P = ref.Patches(topo, 3)
    st = P.stencils
    nC, nT = P.num_coarse, P.num_vertices_total
    src = np.zeros((nT, 3), np.float32)
    src[:nC] = m.positions.astype(np.float32)
    assert ref.cpu_eval_stencils(src, (0, 3, 3), src, (nC * 3, 3, 3), st)
    f, s, t = meshes.random_patch_locations(m.nfaces, 4096, seed=15)
    h = P.find(f, s, t)
    assert (h[:, 0] >= 0).all()
    coords = np.zeros(len(f), PATCH_COORD_DTYPE)
    coords["arrayIndex"], coords["patchIndex"], coords["vertIndex"] = h[:, 0], h[:, 1], h[:, 2]
    coords["s"], coords["t"] = s.astype(np.float32), t.astype(np.float32)
    n = len(coords)
    o = [np.zeros((n, 3), np.float32) for _ in range(6)]
    assert ref.cpu_eval_patches(src, (0, 3, 3), o[:3], [(0, 3, 3)] * 3, coords, P.arrays,
                                P.indices, P.params)
    pos_only = np.zeros((n, 3), np.float32)
    assert ref.cpu_eval_patches(src, (0, 3, 3), [pos_only], [(0, 3, 3)], coords, P.arrays,
                                P.indices, P.params)
    o2 = [np.zeros((n, 3), np.float32) for _ in range(6)]
    assert ref.cpu_eval_patches_d2(src, (0, 3, 3), o2, [(0, 3, 3)] * 6, coords, P.arrays,
                                   P.indices, P.params)
    k = 256
    npts, wP, wDs, wDt = P.eval_basis_f64(h[:k], coords["s"][:k].astype(np.float64),
                                          coords["t"][:k].astype(np.float64))
    save("cubesphere_patches.npz", face_verts=m.face_verts, crease_pairs=m.crease_pairs,
         crease_weights=m.crease_weights, coarse=m.positions.astype(np.float32),
         num_coarse=np.int32(nC), st_sizes=st["sizes"], st_offsets=st["offsets"],
         st_indices=st["indices"], st_weights=st["weights"], src=src,
         arrays=P.arrays.view(np.int32).reshape(-1, 6), indices=P.indices,
         params=P.params.view(np.uint32).reshape(-1, 3), coords=coords.view(np.int32).reshape(-1, 5),
         ptex=f, u=s, v=t,
         cpu_p=o[0], cpu_du=o[1], cpu_dv=o[2], cpu_p_only=pos_only,
         cpu2_p=o2[0], cpu2_du=o2[1], cpu2_dv=o2[2], cpu2_duu=o2[3], cpu2_duv=o2[4],
         cpu2_dvv=o2[5], basis_n=npts, basis_wP=wP, basis_wDs=wDs, basis_wDt=wDt)


def cubesphere_fvar():
    m = meshes.cube_sphere(4, seed=14)
    nval, fidx, uv = meshes.cube_sphere_uvs(4)
    out = {"face_verts": m.face_verts, "crease_pairs": m.crease_pairs,
           "crease_weights": m.crease_weights, "fvar_indices": fidx,
           "uv": uv.astype(np.float32)}
    f, s, t = meshes.random_patch_locations(m.nfaces, 2048, seed=16)
    for name, interp, legacy in (("linear", ref.FVAR_LINEAR_ALL, True),
                                 ("smooth", ref.FVAR_LINEAR_CORNERS_PLUS1, False)):
        topo = ref.Topology(m)
        topo.set_fvar(nval, fidx, interp)
        P = ref.Patches(topo, 3, fvar=True, fvar_legacy_linear=legacy)
        fv = P.fvar
        st = fv["stencils"]
        nb, ns = fv["num_base"], len(st["sizes"])
        assert nb == nval
        buf = np.zeros((nb + ns, 2), np.float32)
        buf[:nb] = out["uv"]
        assert ref.cpu_eval_stencils(buf, (0, 2, 2), buf, (nb * 2, 2, 2), st)
        h = P.find(f, s, t)
        assert (h[:, 0] >= 0).all()
        coords = np.zeros(len(f), PATCH_COORD_DTYPE)
        coords["arrayIndex"], coords["patchIndex"], coords["vertIndex"] = h[:, 0], h[:, 1], h[:, 2]
        coords["s"], coords["t"] = s.astype(np.float32), t.astype(np.float32)
        n = len(coords)
        o = [np.zeros((n, 2), np.float32) for _ in range(3)]
        assert ref.cpu_eval_patches(buf, (0, 2, 2), o, [(0, 2, 2)] * 3, coords, fv["arrays"],
                                    fv["indices"], fv["params"])
        o2 = [np.zeros((n, 2), np.float32) for _ in range(6)]
        assert ref.cpu_eval_patches_d2(buf, (0, 2, 2), o2, [(0, 2, 2)] * 6, coords,
                                       fv["arrays"], fv["indices"], fv["params"])
        out.update({
            f"{name}_st_sizes": st["sizes"], f"{name}_st_offsets": st["offsets"],
            f"{name}_st_indices": st["indices"], f"{name}_st_weights": st["weights"],
            f"{name}_buffer": buf,
            f"{name}_arrays": fv["arrays"].view(np.int32).reshape(-1, 6),
            f"{name}_indices": fv["indices"],
            f"{name}_params": fv["params"].view(np.uint32).reshape(-1, 3),
            f"{name}_vertex_arrays": P.arrays.view(np.int32).reshape(-1, 6),
            f"{name}_vertex_indices": P.indices,
            f"{name}_vertex_params": P.params.view(np.uint32).reshape(-1, 3),
            f"{name}_coords": coords.view(np.int32).reshape(-1, 5),
            f"{name}_cpu_p": o[0], f"{name}_cpu_du": o[1], f"{name}_cpu_dv": o[2],
            f"{name}_cpu2_p": o2[0], f"{name}_cpu2_du": o2[1], f"{name}_cpu2_dv": o2[2],
            f"{name}_cpu2_duu": o2[3], f"{name}_cpu2_duv": o2[4], f"{name}_cpu2_dvv": o2[5]})
    save("cubesphere_fvar.npz", **out)


def bfr_surfaces():
    import math
    sys.path.insert(0, os.path.dirname(HERE))
    import test_far_compiler as T
    out = {}
    cases = (("sphere", meshes.cube_sphere(4, seed=5, crease_fraction=0.1), 1),
             ("prism", T.POLY_MESHES["pentagonal-prism"](), 1),
             ("grid", T.POLY_MESHES["open-grid-creased"](), 2))
    for k, (name, m, boundary) in enumerate(cases):
        sizes = getattr(m, "face_sizes", None)
        sizes = np.full(m.nfaces, 4, np.int32) if sizes is None else np.asarray(sizes, np.int32)
        rng = np.random.default_rng(40 + k)
        n = 1500
        faces = rng.integers(0, m.nfaces, n).astype(np.int32)
        s, t = rng.random(n), rng.random(n)
        u, v = s.copy(), t.copy()
        for i in range(n):                                 # n-gon: a random sub-face's tile
            fs = int(sizes[faces[i]])
            if fs != 4:
                sub = int(rng.integers(fs))
                udim = (2 + (fs > 4)) if fs < 10 else 1 + int(math.sqrt(fs - 1))
                u[i] = sub % udim + 0.5 * s[i] * 0.999
                v[i] = sub // udim + 0.5 * t[i] * 0.999
        P, Du, Dv, done = ref.bfr_evaluate(ref.Topology(m, boundary=boundary), m.positions,
                                           faces, u, v)
        assert done == n
        out.update({f"{name}_face_sizes": sizes,
                    f"{name}_face_verts": np.asarray(m.face_verts, np.int32).reshape(-1),
                    f"{name}_crease_pairs": np.asarray(m.crease_pairs, np.int32).reshape(-1, 2),
                    f"{name}_crease_weights": np.asarray(m.crease_weights, np.float32),
                    f"{name}_positions": np.asarray(m.positions, np.float64),
                    f"{name}_boundary": np.int32(boundary),
                    f"{name}_faces": faces, f"{name}_u": u, f"{name}_v": v,
                    f"{name}_P": P, f"{name}_Du": Du, f"{name}_Dv": Dv})
    save("bfr_surfaces.npz", **out)


def tess_coords():
    rng = np.random.default_rng(77)
    kinds = [(1, 4), (0, 4), (2, 3), (1, 3), (1, 5), (1, 6), (1, 7), (1, 9), (1, 10), (1, 12)]
    cfg, rates_flat, coords = [], [], []
    for k in range(300):
        scheme, N = kinds[k % len(kinds)]
        if k < 60:
            rates = [int(rng.integers(1, 17))]
        else:
            m = int(rng.choice([N, N + 1, N + 2 if N == 4 else N + 1, 2 if N == 4 else 1]))
            rates = [int(x) for x in rng.integers(1, 13, m)]
        uv = ref.tess_coords(scheme, N, rates)
        assert uv is not None
        cfg.append((scheme, N, len(rates), len(rates_flat), len(uv),
                    sum(len(c) for c in coords)))
        rates_flat.extend(rates)
        coords.append(uv)
    save("tess_coords.npz", configs=np.array(cfg, np.int32),
         rates=np.array(rates_flat, np.int32), coords=np.concatenate(coords))


def tri_sphere(n, seed):
    """The cube-sphere's quads split along one diagonal (a closed Loop
    mesh, valences 4 to 6), 5 % of its edges creased."""
    m = meshes.cube_sphere(n, seed=seed, crease_fraction=0.0)
    q = m.face_verts
    t = np.concatenate([q[:, [0, 1, 2]], q[:, [0, 2, 3]]]).astype(np.int32)
    tm = meshes.QuadMesh(m.nverts, t, m.positions)
    rng = np.random.default_rng(seed + 1)
    e = tm.edges()
    k = int(0.05 * len(e))
    pick = np.sort(rng.choice(len(e), k, replace=False))
    tm.crease_pairs = e[pick].astype(np.int32)
    tm.crease_weights = rng.uniform(0, 4, k).astype(np.float32)
    return tm


def tri_grid(nx, ny, seed):
    """An open nx x ny grid of quads split into triangles, heights noisy."""
    X, Y = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    ids = X * (ny + 1) + Y
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    q = np.stack([ids[i, j], ids[i + 1, j], ids[i + 1, j + 1], ids[i, j + 1]], -1).reshape(-1, 4)
    t = np.concatenate([q[:, [0, 1, 2]], q[:, [0, 2, 3]]]).astype(np.int32)
    rng = np.random.default_rng(seed)
    p = np.stack([X.ravel() / nx, Y.ravel() / nx, 0.1 * rng.uniform(-1, 1, X.size)], -1)
    return meshes.QuadMesh(X.size, t, p)


def loop_patches():
    out = {}
    cases = [("sphere", tri_sphere(3, 21), dict(boundary=1), dict(iso_level=2)),
             ("grid", tri_grid(5, 5, 22), dict(boundary=2), dict(iso_level=3)),
             ("strip", tri_grid(5, 1, 24), dict(boundary=1), dict(iso_level=2)),
             ("uniform", tri_sphere(2, 23), dict(boundary=1), dict(iso_level=2, uniform=True))]
    for k, (name, m, topt, popt) in enumerate(cases):
        topo = ref.Topology(m, scheme=ref.SCHEME_LOOP, **topt)
        P = ref.Patches(topo, **popt)
        st = P.stencils
        nC, nT = P.num_coarse, P.num_vertices_total
        src = np.zeros((nT, 3), np.float32)
        src[:nC] = m.positions.astype(np.float32)
        assert ref.cpu_eval_stencils(src, (0, 3, 3), src, (nC * 3, 3, 3), st)
        rng = np.random.default_rng(30 + k)
        nq = 1024
        f = rng.integers(0, m.nfaces, nq).astype(np.int32)
        u, v = rng.uniform(size=nq), rng.uniform(size=nq)
        flip = u + v > 1
        u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
        h = P.find(f, u, v)
        assert (h[:, 0] >= 0).all()
        coords = np.zeros(nq, PATCH_COORD_DTYPE)
        coords["arrayIndex"], coords["patchIndex"], coords["vertIndex"] = h[:, 0], h[:, 1], h[:, 2]
        coords["s"], coords["t"] = u.astype(np.float32), v.astype(np.float32)
        o = [np.zeros((nq, 3), np.float32) for _ in range(3)]
        assert ref.cpu_eval_patches(src, (0, 3, 3), o, [(0, 3, 3)] * 3, coords, P.arrays,
                                    P.indices, P.params)
        o2 = [np.zeros((nq, 3), np.float32) for _ in range(6)]
        assert ref.cpu_eval_patches_d2(src, (0, 3, 3), o2, [(0, 3, 3)] * 6, coords, P.arrays,
                                       P.indices, P.params)
        nb = 128
        npts, wP, wDs, wDt = P.eval_basis_f64(h[:nb], coords["s"][:nb].astype(np.float64),
                                              coords["t"][:nb].astype(np.float64))
        types = sorted(set(P.arrays["regDesc"].tolist()) | set(P.arrays["desc"].tolist()))
        print(name, "types", types, "patches", len(P.params), "coords", nq)
        out.update({f"{name}_{key}": val for key, val in dict(
            src=src, arrays=P.arrays.view(np.int32).reshape(-1, 6), indices=P.indices,
            params=P.params.view(np.uint32).reshape(-1, 3),
            coords=coords.view(np.int32).reshape(-1, 5),
            cpu_p=o[0], cpu_du=o[1], cpu_dv=o[2],
            cpu2_p=o2[0], cpu2_du=o2[1], cpu2_dv=o2[2], cpu2_duu=o2[3], cpu2_duv=o2[4],
            cpu2_dvv=o2[5], basis_n=npts, basis_wP=wP, basis_wDs=wDs, basis_wDt=wDt).items()})
    save("loop_patches.npz", **out)


if __name__ == "__main__":
    if not ref.available():
        sys.exit("oracle/_ref/libosdref.so not built (make -C oracle)")
    # names on the command line write only those fixtures
    makers = [cube_l3, cubesphere_l2, cubesphere_limit, cubesphere_patches, loop_patches,
              cubesphere_fvar, bfr_surfaces, tess_coords]
    wanted = set(sys.argv[1:])
    for fn in makers:
        if not wanted or fn.__name__ in wanted:
            fn()
