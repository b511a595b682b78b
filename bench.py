"""Benchmark of the MI355X Osd evaluator (one GPU).

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--only cfg1]

Workloads (synthetic, seeded; tables built by the project's own table
compiler, which reproduces Far's tables):

* cfg1 (headline): cube-sphere, 6 x 128 x 128 quads (98,306 vertices),
  float32 xyz on the unit sphere with 1% seeded radial noise, 5% of edges
  creased with sharpness U[0,4).  Limit stencil table with du/dv weights at
  the 6,291,458 vertex locations of uniform level 3 (Far
  LimitStencilTableFactory over the feature-adaptive iso-3 patch table).
  One step = one fused EvalStencils launch producing P, dP/du, dP/dv for
  every location (one frame, inputs resident in HBM).
  value = limit points per second.
* cfg1_refine: the uniform level-3 refinement table of the same mesh, P.
* cfg2: EvalPatches (P, dP/du, dP/dv) at 4,194,304 PatchCoords (ptex face
  uniform, (u,v) uniform [0,1)^2, PatchMap lookup) on the iso-3
  feature-adaptive patch table (B-spline + Gregory end caps) of the mesh.

The JSON line carries the roofline of the headline kernel (algorithmic HBM
bytes per launch / average launch time from HIP events on the launch
stream) and the CPU baseline: Osd::CpuEvaluator built from the reference
sources when that library is present (else the C restatement), timed on
this machine's host cores over the same table; OmpEvaluator beside it.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E, GB/s
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")
METRIC = "limit points/s + HBM GB/s vs roofline; served p50/p99 latency at fixed QPS"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# --------------------------------------------------------------- inputs

def build_inputs(need_refine=True, need_patches=True):
    from opensubdiv_amd import far, meshes
    t0 = time.time()
    mesh = meshes.cube_sphere(128, seed=1, radial_noise=0.01, crease_fraction=0.05,
                              sharpness_max=4.0)
    out = {"mesh": mesh}
    topo = far.Topology(mesh)
    topo.refine_uniform(3)
    if need_refine:
        out["refine"] = topo.stencil_table()
        log(f"[bench] refinement table built ({time.time() - t0:.1f}s)")
    ptex, u, v = topo.vertex_locations(3)
    pt = far.Topology(mesh).patch_table(3)
    out["patches"] = pt
    log(f"[bench] patch table built ({time.time() - t0:.1f}s)")
    out["limit"] = pt.limit_stencils(ptex, u, v)
    log(f"[bench] limit table built ({time.time() - t0:.1f}s)")
    if need_patches:
        out["points"] = pt.point_stencils()
        f, s, t = meshes.random_patch_locations(mesh.nfaces, 4194304, seed=2)
        out["queries"] = (f, s, t)
        out["coords"] = pt.find(f, s, t)
        assert (out["coords"]["arrayIndex"] >= 0).all()
    log(f"[bench] inputs built in {time.time() - t0:.1f}s: limit table "
        f"{len(out['limit']['sizes'])} x {len(out['limit']['indices'])} entries, "
        f"{len(pt.params)} patches")
    return out


# ------------------------------------------------------- algorithmic bytes

def stencil_bytes(n_st, nnz, n_weights, L, n_src, n_out):
    """One EvalStencils launch: sizes (4 B / stencil), index + n_weights
    weights (4 B each) per entry, n_out outputs of L floats, control points
    read once."""
    return 4 * n_st + (4 + 4 * n_weights) * nnz + 4 * L * n_out * n_st + 4 * L * n_src


def patch_bytes(n_pc, L, n_out, n_indices, n_patches, n_arrays, n_cv):
    """One EvalPatches launch: PatchCoords (20 B) and outputs per coord, the
    patch tables (indices, 12 B params, 24 B arrays) and control points
    counted once."""
    return (20 + 4 * L * n_out) * n_pc + 4 * n_indices + 12 * n_patches + 24 * n_arrays + \
        4 * L * n_cv


# ---------------------------------------------------------------- timing

def time_launches(step, steps, warmup, stream):
    """warmup + steps launches; returns (wall seconds of the timed region,
    per-launch milliseconds from HIP events recorded on `stream`)."""
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, [a.elapsed_time(b) for a, b in ev]


def cpu_time(run, budget_s, max_runs=9):
    run()                                   # warm-up
    times = []
    t_end = time.time() + budget_s
    while time.time() < t_end and len(times) < max_runs:
        t0 = time.perf_counter()
        assert run()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), len(times)


def traffic_for(key, kernel):
    """HBM bytes per launch from the committed counter profile, when it was
    measured on the same kernel (template arguments included)."""
    if os.path.exists(PROFILE_TRAFFIC):
        tr = json.load(open(PROFILE_TRAFFIC)).get(key)
        if tr and tr.get("kernel_short") == kernel:
            return tr.get("bytes_per_launch")
    return None


# ------------------------------------------------------------- workloads

def table_eval_fn(lib, table, src, srcdesc, outs, descs, stream):
    D = type(srcdesc)
    n = table.GetNumStencils()
    optr = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
    dptr = (D * len(outs))(*descs)

    def step():
        rc = lib.osdhip_stencil_table_eval(table.h, src.data_ptr(), srcdesc, len(outs),
                                           ctypes.cast(optr, ctypes.c_void_p),
                                           ctypes.cast(dptr, ctypes.c_void_p), 0, n, stream)
        if rc != 0:
            raise RuntimeError(f"osdhip_stencil_table_eval failed: {rc}")
    return step


def kernel_labels(nw, lc):
    return {"rec_g8": f"k_stencils_rec<{nw},{lc},8>", "rec_g4": f"k_stencils_rec<{nw},{lc},4>",
            "rec_g16": f"k_stencils_rec<{nw},{lc},16>",
            "split": f"k_stencils_sliced<{nw},{lc}>", "csr_lds": f"k_stencils_csr<{nw},{lc}>"}


def run_stencils(args, lib, name, table_np, src_np, nw, kernel_label, with_cpu):
    """EvalStencils of a (limit) stencil table through HipStencilTable."""
    import torch
    from opensubdiv_amd import HipStencilTable, _lib
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    D = _lib.Desc
    L = src_np.shape[1]
    n = len(table_np["sizes"])
    nnz = len(table_np["indices"])
    t0 = time.time()
    lib.osdhip_config(b"table_layout", 0)
    htable = HipStencilTable.Create(table_np, s)
    log(f"[bench] {name}: device table {htable.device_bytes() / 1e6:.0f} MB in "
        f"{time.time() - t0:.1f}s")
    d_src = torch.from_numpy(np.ascontiguousarray(src_np)).to("cuda")
    outs = [torch.empty(n * L, dtype=torch.float32, device="cuda") for _ in range(nw)]
    alg = stencil_bytes(n, nnz, nw, L, src_np.shape[0], nw)
    # autotune: a few untimed launches of each kernel variant (record slices
    # with gather groups of 8 / 4, split-stream slices, CSR with LDS
    # staging), then the timed region runs exactly `steps` launches of the
    # fastest
    VARIANTS = {"rec_g8": (0, 8, 0), "rec_g4": (0, 4, 0), "rec_g16": (0, 16, 0),
                "split": (1, 8, 0), "csr_lds": (0, 8, 1)}
    variants = {}
    for vname, (layout, group, kernel) in VARIANTS.items():
        lib.osdhip_config(b"slice_group", group)
        lib.osdhip_config(b"stencil_kernel", kernel)
        tab = htable
        if layout != 0:
            lib.osdhip_config(b"table_layout", layout)
            tab = HipStencilTable.Create(table_np, s)
            lib.osdhip_config(b"table_layout", 0)
        fn = table_eval_fn(lib, tab, d_src, D(0, L, L), outs, [D(0, L, L)] * nw, s)
        _, ms = time_launches(fn, 5, 2, stream)
        avg = float(np.median(ms))
        variants[vname] = {"kernel_ms": avg, "GBps": alg / (avg * 1e-3) / 1e9}
        if tab is not htable:
            del tab, fn
            torch.cuda.synchronize()
    best = min(variants, key=lambda k: variants[k]["kernel_ms"])
    layout, group, kernel = VARIANTS[best]
    if layout != 0:
        del htable
        lib.osdhip_config(b"table_layout", layout)
        htable = HipStencilTable.Create(table_np, s)
        lib.osdhip_config(b"table_layout", 0)
    lib.osdhip_config(b"slice_group", group)
    lib.osdhip_config(b"stencil_kernel", kernel)
    step = table_eval_fn(lib, htable, d_src, D(0, L, L), outs, [D(0, L, L)] * nw, s)
    wall, ms = time_launches(step, args.steps, args.warmup, stream)
    lib.osdhip_config(b"slice_group", 8)
    lib.osdhip_config(b"stencil_kernel", 0)
    got = [o.cpu().numpy().reshape(n, L) for o in outs]
    avg = float(np.mean(ms))
    res = {"points_per_s": n * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps,
           "kernel_ms": avg, "stencils": n, "table_entries": nnz, "floats_per_point": L,
           "outputs": nw, "algorithmic_bytes": alg, "GBps": alg / (avg * 1e-3) / 1e9,
           "kernel": f"{kernel_label[best]} ({best})",
           "autotune": variants}
    if with_cpu:
        from oracle import ref
        import oracle
        dst = [np.zeros((n, L), np.float32) for _ in range(nw)]
        if ref.available():
            kind = "reference"
            if nw == 1:
                run = lambda: ref.cpu_eval_stencils(src_np, (0, L, L), dst[0], (0, L, L),  # noqa
                                                    table_np)
            else:
                run = lambda: ref.cpu_eval_stencils_d1(src_np, (0, L, L), dst, [(0, L, L)] * 3,  # noqa
                                                       table_np)
        else:
            kind = "port"
            if nw == 1:
                run = lambda: oracle.eval_stencils(src_np, (0, L, L), dst[0], (0, L, L),  # noqa
                                                   table_np["sizes"], table_np["offsets"],
                                                   table_np["indices"], table_np["weights"])
            else:
                run = lambda: oracle.eval_stencils_d1(  # noqa
                    src_np, (0, L, L), dst, [(0, L, L)] * 3, table_np["sizes"],
                    table_np["offsets"], table_np["indices"],
                    [table_np["weights"], table_np["du"], table_np["dv"]])
        med, runs = cpu_time(run, args.cpu_budget)
        res["cpu_baseline"] = {"value": n / med, "unit": "points/s", "cores": 1, "kind": kind,
                               "sample": f"full table ({n} stencils, {nw} outputs) x {runs} "
                                         f"runs, median {med * 1e3:.1f} ms"}
        if nw == 1 and ref.available():
            threads = min(os.cpu_count() or 1, 16)
            omp = lambda: ref.cpu_eval_stencils(src_np, (0, L, L), dst[0], (0, L, L),  # noqa
                                                table_np, threads=threads)
            m2, r2 = cpu_time(omp, args.cpu_budget / 2)
            res["cpu_baseline"]["omp"] = {"value": n / m2, "cores": threads,
                                          "kind": "reference OmpEvaluator",
                                          "sample": f"{r2} runs, median {m2 * 1e3:.1f} ms"}
            want = [np.zeros((n, L), np.float32)]
            ref.cpu_eval_stencils(src_np, (0, L, L), want[0], (0, L, L), table_np)
        else:
            want = dst
        res["parity"] = {"bit_exact_vs_cpu": all(np.array_equal(a, b) for a, b in zip(got, want))}
    del htable
    return res


def run_cube(args, lib, with_cpu):
    """cfg0: the tutorial cube (osd_tutorial_0.cpp:43-62, xyz + U(-0.05, 0.05)
    noise, seed 0), uniform level-3 stencils (386 rows), evaluated in place as
    the tutorial does (src {0,3,3}, dst {8*3,3,3} in one buffer).  A CPU
    configuration: the point of comparison is CpuEvaluator's time per call;
    the HIP launch is reported beside it."""
    import torch
    from opensubdiv_amd import HipStencilTable, _lib, far, meshes
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    D = _lib.Desc
    m = meshes.unit_cube(seed=0)
    topo = far.Topology(m)
    topo.refine_uniform(3)
    t = topo.stencil_table()
    n = len(t["sizes"])
    buf = np.zeros((8 + n, 3), np.float32)
    buf[:8] = m.positions.astype(np.float32)
    d_buf = torch.from_numpy(buf).to("cuda")
    table = HipStencilTable.Create(t, s)
    step = table_eval_fn(lib, table, d_buf, D(0, 3, 3), [d_buf.view(-1)[8 * 3:]], [D(0, 3, 3)], s)
    reps = 200
    wall, ms = time_launches(step, reps, 20, stream)
    res = {"stencils": n, "table_entries": int(len(t["indices"])),
           "hip_us_per_call": float(np.median(ms)) * 1e3,
           "hip_points_per_s": n / (float(np.median(ms)) * 1e-3)}
    if with_cpu:
        from oracle import ref
        got = d_buf.cpu().numpy()
        host = buf.copy()
        if ref.available():
            ref.cpu_eval_stencils(host, (0, 3, 3), host, (8 * 3, 3, 3), t)
            calls = 2000
            t0 = time.perf_counter()
            for _ in range(calls):
                ref.cpu_eval_stencils(host, (0, 3, 3), host, (8 * 3, 3, 3), t)
            dt = (time.perf_counter() - t0) / calls
            res["cpu_baseline"] = {"value": n / dt, "unit": "points/s", "us_per_call": dt * 1e6,
                                   "cores": 1, "kind": "reference",
                                   "sample": f"{calls} calls (includes the ctypes call overhead)"}
            res["parity"] = {"bit_exact_vs_cpu": bool(np.array_equal(got, host))}
    return res


def device_patch_map(args, inputs, stream):
    """The same PatchCoords made on the device (osdhip_patch_map_find, one
    launch for every query) against the host map that made them: identical
    records; lookups/s beside the host map's (OpenMP, all host threads)."""
    import torch
    from opensubdiv_amd import far
    from opensubdiv_amd.osd import PATCH_COORD_DTYPE, HipPatchMap
    pt, coords = inputs["patches"], inputs["coords"]
    f, s, t = inputs["queries"]
    n = len(f)
    pm = HipPatchMap(pt.arrays, pt.params, stream.cuda_stream)
    d_f = torch.from_numpy(np.ascontiguousarray(f, np.int32)).to("cuda")
    d_u = torch.from_numpy(np.ascontiguousarray(s, np.float64)).to("cuda")
    d_v = torch.from_numpy(np.ascontiguousarray(t, np.float64)).to("cuda")
    d_out = torch.empty(n * 5, dtype=torch.int32, device="cuda")

    def step():
        if not pm.FindPatches(n, d_f, d_u, d_v, d_out, stream.cuda_stream):
            raise RuntimeError("osdhip_patch_map_find failed")
    wall, ms = time_launches(step, 10, 2, stream)
    got = d_out.cpu().numpy().view(PATCH_COORD_DTYPE)
    same = bool(got.tobytes() == np.ascontiguousarray(coords).tobytes())
    far.set_threads(0)
    t0 = time.perf_counter()
    pt.find(f, s, t)
    host_s = time.perf_counter() - t0
    pm.close()
    return {"lookups_per_s": n / (float(np.mean(ms)) * 1e-3), "kernel_ms": float(np.mean(ms)),
            "identical_to_host_map": same, "host_lookups_per_s": n / host_s,
            "host_threads": os.cpu_count()}


def run_patches(args, lib, inputs, with_cpu):
    import torch
    from opensubdiv_amd import _lib
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    D = _lib.Desc
    mesh, pt, pts, coords = inputs["mesh"], inputs["patches"], inputs["points"], inputs["coords"]
    from opensubdiv_amd import HipStencilTable
    ctrl = mesh.positions.astype(np.float32)
    # vertex buffer [control | refined | end-cap points], filled on the device
    # by the patch-point stencils (the Osd::Mesh::Refine step)
    nctrl, npts = len(ctrl), len(pts["sizes"])
    d_buf = torch.zeros((nctrl + npts) * 3, dtype=torch.float32, device="cuda")
    d_buf[:nctrl * 3] = torch.from_numpy(np.ascontiguousarray(ctrl).reshape(-1)).to("cuda")
    ptab = HipStencilTable.Create(pts, s)
    outp = (ctypes.c_void_p * 1)(d_buf.data_ptr())
    descp = (D * 1)(D(nctrl * 3, 3, 3))
    rc = lib.osdhip_stencil_table_eval(ptab.h, d_buf.data_ptr(), D(0, 3, 3), 1,
                                       ctypes.cast(outp, ctypes.c_void_p),
                                       ctypes.cast(descp, ctypes.c_void_p), 0, npts, s)
    assert rc == 0, rc
    torch.cuda.synchronize()
    del ptab
    d_src = d_buf
    L, n = 3, len(coords)

    def dev(a):
        a = np.ascontiguousarray(a)
        if a.dtype.fields is not None:
            a = a.view(np.int32)
        return torch.from_numpy(a.copy()).to("cuda")
    d_coords = dev(coords)
    d_arrays, d_indices, d_params = dev(pt.arrays), dev(pt.indices), dev(pt.params)
    outs = [torch.empty(n * L, dtype=torch.float32, device="cuda") for _ in range(3)]
    d3 = D(0, L, L)
    z = D(0, 0, 0)

    def step():
        rc = lib.osdhip_eval_patches(d_src.data_ptr(), d3, outs[0].data_ptr(), d3,
                                     outs[1].data_ptr(), d3, outs[2].data_ptr(), d3,
                                     None, z, None, z, None, z, n, d_coords.data_ptr(),
                                     d_arrays.data_ptr(), d_indices.data_ptr(),
                                     d_params.data_ptr(), s)
        if rc != 0:
            raise RuntimeError(f"osdhip_eval_patches failed: {rc}")
    alg = patch_bytes(n, L, 3, len(pt.indices), len(pt.params), len(pt.arrays), nctrl + npts)
    # autotune: caller order vs patch-sorted order (radix sort inside each
    # launch, included in the timed step)
    variants = {}
    for vname, smin in (("caller_order", 0), ("patch_sorted", 16384)):
        lib.osdhip_config(b"patch_sort_min", smin)
        _, ms = time_launches(step, 5, 2, stream)
        avg = float(np.median(ms))
        variants[vname] = {"step_ms": avg, "GBps": alg / (avg * 1e-3) / 1e9}
    best = min(variants, key=lambda k: variants[k]["step_ms"])
    lib.osdhip_config(b"patch_sort_min", 0 if best == "caller_order" else 16384)
    wall, ms = time_launches(step, args.steps, args.warmup, stream)
    avg = float(np.mean(ms))
    lib.osdhip_config(b"patch_sort_min", 0)
    res = {"points_per_s": n * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps,
           "kernel_ms": avg, "patch_coords": n, "patches": len(pt.params),
           "patch_map": device_patch_map(args, inputs, stream),
           "gregory_patches": int(sum(a["numPatches"] for a in pt.arrays if a["desc"] == 9)),
           "algorithmic_bytes": alg, "GBps": alg / (avg * 1e-3) / 1e9,
           "kernel": f"k_patches<true,false,3> ({best})", "autotune": variants}
    if with_cpu:
        src_np = d_src.cpu().numpy().reshape(-1, 3)
        import oracle
        from oracle import ref
        got = [o.cpu().numpy().reshape(n, L) for o in outs]
        dst = [np.zeros((n, L), np.float32) for _ in range(3)]
        if ref.available():
            kind = "reference"
            run = lambda: ref.cpu_eval_patches(src_np, (0, L, L), dst, [(0, L, L)] * 3,  # noqa
                                               coords, pt.arrays, pt.indices, pt.params)
        else:
            kind = "port"
            run = lambda: oracle.eval_patches(src_np, (0, L, L), dst, [(0, L, L)] * 3,  # noqa
                                              coords, pt.arrays, pt.indices, pt.params)
        med, runs = cpu_time(run, args.cpu_budget, max_runs=3)
        res["cpu_baseline"] = {"value": n / med, "unit": "points/s", "cores": 1, "kind": kind,
                               "sample": f"all {n} coords x {runs} runs, median "
                                         f"{med * 1e3:.1f} ms"}
        res["parity"] = {"bit_exact_vs_cpu": all(np.array_equal(a, b) for a, b in zip(got, dst))}
    return res


def buffer_checksums(dms):
    """Exact order-independent checksum of every output buffer: the sum of
    its float bit patterns as int64 (equal bits -> equal sums)."""
    import torch
    chunk = 1 << 26
    out = []
    for d in dms:
        bufs = [d.refined, d.vbuf[d.nctrl * d.spec.frames * d.lp:]] + d.out
        sums = []
        for b in bufs:
            v = b.view(torch.int32)
            sums.append(sum(int(v[i:i + chunk].to(torch.int64).sum().item())
                            for i in range(0, v.numel(), chunk)))
        out.append(sums)
    return out


def table_cache(args):
    """The --table-cache directory as a TableCache (None without one)."""
    if not args.table_cache:
        return None
    from opensubdiv_amd.table_cache import TableCache
    return TableCache(args.table_cache)


def cache_summary(cache):
    return None if cache is None else {"hits": cache.hits, "misses": cache.misses}


def run_batched(args, lib, device):
    """cfg3: 256 cube-sphere meshes (n log-uniform in [8,128], 5% creased
    edges, level 2-4, 1-16 frames of xyz + 4 primvars, 64K-1M patch-ordered
    PatchCoords) in one graph launch per step, against the same work one
    mesh and one frame per launch."""
    import torch
    from opensubdiv_amd import workloads
    from opensubdiv_amd.batch import Batch
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    L = 7
    seed = args.seed
    specs = workloads.sample_meshes(args.meshes, seed)
    t0 = time.time()
    lib.osdhip_config(b"table_csr_mirror", 0)
    old_layout = lib.osdhip_config(b"table_layout", args.table_layout)
    workers = min(16, os.cpu_count() or 1)
    cache = table_cache(args)
    dms = workloads.build_device_meshes(specs, L, s, workers=workers, log=log, own_outputs=False,
                                        cache=cache)
    lib.osdhip_config(b"table_csr_mirror", 1)
    lib.osdhip_config(b"table_layout", old_layout)
    torch.cuda.synchronize()
    build_s = time.time() - t0
    # outputs (refined vertices, P / dPdu / dPdv of every frame) stream
    # through one arena of the memory left after the resident tables and
    # vertex buffers; meshes fill stages in order
    free, total = torch.cuda.mem_get_info()
    reserve = 8 << 30
    arena_floats = max(0, int(free - reserve) // 4)
    arena, nstages = workloads.stage_outputs(dms, arena_floats)
    dev_bytes = sum(d.device_bytes() for d in dms)
    out_bytes = 4 * sum(d.output_floats() for d in dms)
    log(f"[cfg3] {len(dms)} meshes built and uploaded in {build_s:.1f}s; resident "
        f"{(dev_bytes - out_bytes) / 1e9:.1f} GB, outputs {out_bytes / 1e9:.1f} GB per step "
        f"through a {4 * arena.numel() / 1e9:.1f} GB arena in {nstages} stage(s)")
    batch = Batch(L, [x for d in dms for x in d.stencil_segments()],
                  [d.patch_segment() for d in dms], s)
    # parity at full size: every stage, batched kernels vs one mesh and
    # frame per launch, exact checksums of every output buffer
    parity = True
    for st in range(nstages):
        members = [d for d in dms if d.stage == st]
        batch.run_stage(st, 3, s)
        a = buffer_checksums(members)
        batch.run_plain_stage(st, s)
        b = buffer_checksums(members)
        parity &= a == b
    log(f"[cfg3] parity batched == plain: {parity}")
    t0 = time.time()
    report = batch.tune(reps=2, stream=s)
    log(f"[cfg3] tuned in {time.time() - t0:.1f}s: {batch.config()}")
    points = sum(d.points_per_step() for d in dms)
    alg = sum(d.algorithmic_bytes() for d in dms)
    wall, ms = time_launches(lambda: batch.launch(s), args.steps, args.warmup, stream)
    kms = float(np.median(ms))
    plain_steps = max(2, min(args.steps, 3))
    pwall, pms = time_launches(lambda: batch.run_plain(s), plain_steps, 1, stream)
    pkms = float(np.median(pms))
    res = {
        "metric": METRIC, "value": points * args.steps / wall, "unit": "points/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"cfg3: {len(dms)} cube-sphere meshes (seed {seed}), n log-uniform "
                               "[8,128], 5% creased edges, refinement level 2-4, 1-16 frames of "
                               "xyz + 4 primvars, 64K-1M PatchCoords each; one graph launch per "
                               "step (stencils then patches)",
                   "meshes": len(dms), "seed": seed, "floats_per_point": L, "device": device,
                   "stages": nstages, "table_layout": args.table_layout,
                   "output_bytes_per_step": out_bytes,
                   "refined_rows_per_step": sum(d.spec.frames * d.nrefine for d in dms),
                   "patch_points_per_step": sum(d.spec.frames * d.ncoords for d in dms),
                   "table_entries": sum(d.nnz_refine + d.nnz_points for d in dms),
                   "device_bytes": dev_bytes, "build_s": build_s,
                   "table_cache": cache_summary(cache)},
        "roofline": {"bound": "hbm", "achieved": alg / (kms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "k_batch_stencils + k_batch_patches (graph)",
                     "algorithmic_bytes": alg, "kernel_ms": kms},
        "plain": {"points_per_s": points * plain_steps / pwall, "step_ms": pkms,
                  "launches_per_step": sum(2 * d.spec.frames + d.spec.frames for d in dms),
                  "speedup_batched": pkms / kms},
        "autotune": report, "tuned": batch.config(),
        "parity": {"batched_equals_plain_checksums": parity,
                   "method": "per stage: int64 sums of every output buffer's bit patterns"},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = batched_cpu_baseline(dms, specs, args.cpu_budget)
    batch.close()
    return res


def run_served(args, lib, device):
    """cfg4: a pool of 32 resident topologies drawn like cfg3; requests (one
    new frame of xyz + 4 primvars, then the topology's refinement or its
    limit points) arrive as a Poisson process and are served with
    continuous batching under fixed device / pinned-host budgets."""
    import torch
    from concurrent.futures import ThreadPoolExecutor
    from opensubdiv_amd import serve, workloads
    L = 7
    specs = workloads.sample_meshes(32, seed=args.seed + 1)
    t0 = time.time()
    workers = min(16, os.cpu_count() or 1)
    cache = table_cache(args)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        hosts = list(ex.map(lambda sp: workloads.compile_mesh(sp, L, 1, cache), specs))
    lib.osdhip_config(b"table_csr_mirror", 0)
    old_layout = lib.osdhip_config(b"table_layout", args.table_layout)
    residents = serve.residents_from_hosts(hosts)
    lib.osdhip_config(b"table_csr_mirror", 1)
    lib.osdhip_config(b"table_layout", old_layout)
    for h in hosts:
        h.release()
    torch.cuda.synchronize()
    build_s = time.time() - t0
    dev_budget, host_budget = 32 << 30, 24 << 30
    log(f"[cfg4] {len(residents)} resident topologies in {build_s:.1f}s; budgets "
        f"{dev_budget >> 30} GiB device, {host_budget >> 30} GiB pinned host")
    server = serve.Server(residents, L, dev_budget, host_budget, slots=args.slots,
                          max_epoch_points=int(args.epoch_points))

    def points(r):
        t = residents[r.topo]
        return t.nrefine if r.kind == "refine" else t.ncoords

    # capacity: a saturating burst (all arrive at t = 0), after a warm-up burst
    for rep in range(2):
        burst = serve.poisson_requests(96, 1e9, len(residents), seed=100 + rep)
        for r in burst:
            r.arrival = 0.0
        tb = time.perf_counter()
        server.run(burst, max_seconds=300)
        burst_s = time.perf_counter() - tb
    capacity = len(burst) / burst_s
    rate = args.load * capacity
    reqs = serve.poisson_requests(args.requests, rate, len(residents), seed=7)
    keep = {r.rid for r in reqs[:16]}
    server.epochs, server.epoch_points = [], []
    tr = time.perf_counter()
    done, results = server.run(reqs, keep=keep, max_seconds=600)
    wall = time.perf_counter() - tr
    lat = np.array([r.done - r.arrival for r in done]) * 1e3
    pts = sum(points(r) for r in done)
    makespan = max(r.done for r in done) - min(r.arrival for r in done)
    # the same requests alone (one per epoch, nothing else in flight)
    same = True
    for r in [x for x in done if x.rid in keep]:
        r2 = serve.Request(r.rid, r.topo, r.kind, 0.0, r.frame)
        _, alone = server.run([r2], keep={r.rid}, max_seconds=60)
        same &= bool(np.array_equal(alone[r.rid].view(np.uint32), results[r.rid].view(np.uint32)))
    return {
        "metric": METRIC, "value": pts / makespan, "unit": "points/s", "n_gpus": 1,
        "steps": len(done), "warmup": 2, "ms_per_step": wall * 1e3 / len(done),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cfg4: 32 resident cube-sphere topologies drawn like cfg3 "
                               "(n log-uniform [8,128], level 2-4, 64K-1M PatchCoords); each "
                               "request = one new frame (xyz + 4 primvars) + refinement or limit "
                               "points (50/50); Poisson arrivals; continuous batching",
                   "topologies": len(residents), "floats_per_point": L, "device": device,
                   "device_budget_bytes": dev_budget, "host_budget_bytes": host_budget,
                   "slots": args.slots, "max_epoch_points": int(args.epoch_points),
                   "table_layout": args.table_layout,
                   "build_s": build_s, "table_cache": cache_summary(cache)},
        "served": {"requests": len(done), "rate_per_s": rate, "capacity_per_s": capacity,
                   "load": args.load, "p50_ms": float(np.percentile(lat, 50)),
                   "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                   "max_ms": float(lat.max()), "epochs": len(server.epochs),
                   "mean_requests_per_epoch": float(np.mean(server.epochs)),
                   "max_requests_per_epoch": int(np.max(server.epochs)),
                   "latency": "arrival -> results copied to pinned host memory"},
        "parity": {"alone_equals_mix_bits": same, "checked_requests": len(keep)},
    }


def batched_cpu_baseline(dms, specs, budget_s):
    """Reference CpuEvaluator on a bounded sample of the batched workload:
    the refinement stencils and EvalPatches of the smallest meshes, one
    frame each, until about `budget_s` seconds of CPU work."""
    from oracle import ref
    import oracle
    from opensubdiv_amd import far, meshes
    t_end = time.time() + budget_s
    pts, secs, used = 0, 0.0, 0
    for d in sorted(dms, key=lambda x: x.nrefine):
        if time.time() > t_end:
            break
        sp, L, F = d.spec, d.length, d.spec.frames
        mesh = meshes.cube_sphere(sp.n, seed=sp.seed, radial_noise=0.01, crease_fraction=0.05,
                                  sharpness_max=4.0)
        topo = far.Topology(mesh)
        topo.refine_uniform(sp.level)
        table = topo.stencil_table()
        vb = np.ascontiguousarray(d.vbuf.cpu().numpy().reshape(-1, F, d.lp)[:, 0, :L])
        ctrl = np.ascontiguousarray(vb[:d.nctrl])
        dst = np.zeros((d.nrefine, L), np.float32)
        n = d.ncoords
        outs = [np.zeros((n, L), np.float32) for _ in range(3)]
        coords = d.coords.cpu().numpy().view(oracle_coord_dtype())
        arrays = d.arrays.cpu().numpy().view(oracle_array_dtype())
        params = d.params.cpu().numpy().view(oracle_param_dtype())
        indices = d.indices.cpu().numpy()
        t0 = time.perf_counter()
        if ref.available():
            ref.cpu_eval_stencils(ctrl, (0, L, L), dst, (0, L, L), table)
            ref.cpu_eval_patches(vb, (0, L, L), outs, [(0, L, L)] * 3, coords, arrays, indices,
                                 params)
            kind = "reference"
        else:
            oracle.eval_stencils(ctrl, (0, L, L), dst, (0, L, L), table["sizes"],
                                 table["offsets"], table["indices"], table["weights"])
            oracle.eval_patches(vb, (0, L, L), outs, [(0, L, L)] * 3, coords, arrays, indices,
                                params)
            kind = "port"
        secs += time.perf_counter() - t0
        pts += d.nrefine + n
        used += 1
    return {"value": pts / secs if secs else None, "unit": "points/s", "cores": 1, "kind": kind,
            "sample": f"{used} smallest meshes, frame 0: refinement stencils + EvalPatches "
                      f"(P, dPdu, dPdv) of {pts} points in {secs:.1f}s"}


def oracle_coord_dtype():
    from opensubdiv_amd.osd import PATCH_COORD_DTYPE
    return PATCH_COORD_DTYPE


def oracle_array_dtype():
    from opensubdiv_amd.osd import PATCH_ARRAY_DTYPE
    return PATCH_ARRAY_DTYPE


def oracle_param_dtype():
    from opensubdiv_amd.osd import PATCH_PARAM_DTYPE
    return PATCH_PARAM_DTYPE


def copy_probe(args):
    """1 GiB device-to-device copy: reads and writes exactly 2^30 bytes per
    launch, so its FETCH_SIZE / WRITE_SIZE calibrate the counters and its
    rate is the measured copy bandwidth (read + write bytes / time)."""
    import torch
    n = 1 << 28
    a = torch.rand(n, device="cuda")
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream()
    wall, ms = time_launches(lambda: b.copy_(a), args.steps, args.warmup, stream)
    kms = float(np.median(ms))
    _, rms = time_launches(lambda: a.sum(), args.steps, args.warmup, stream)
    rkms = float(np.median(rms))
    print(json.dumps({"probe": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "kernel_ms": kms, "GBps": 8 * n / (kms * 1e-3) / 1e9,
                      "read_probe": {"bytes_read": 4 * n, "kernel_ms": rkms,
                                     "GBps": 4 * n / (rkms * 1e-3) / 1e9, "op": "torch.sum"},
                      "steps": args.steps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds per CPU baseline")
    ap.add_argument("--only", default=None, choices=[None, "cfg1", "cfg1_refine", "cfg2", "cfg3", "cfg4"])
    ap.add_argument("--requests", type=int, default=1500, help="cfg4: requests served")
    ap.add_argument("--load", type=float, default=0.7,
                    help="cfg4: arrival rate as a fraction of the measured capacity")
    ap.add_argument("--slots", type=int, default=4, help="cfg4: epochs in flight")
    ap.add_argument("--epoch-points", type=float, default=32e6,
                    help="cfg4: cap on the points batched into one epoch")
    ap.add_argument("--meshes", type=int, default=256, help="cfg3: meshes in the batch")
    ap.add_argument("--seed", type=int, default=0, help="cfg3: first sample seed")
    ap.add_argument("--table-cache", default=None,
                    help="cfg3/cfg4: directory of compiled-table files; tables found there "
                         "(same topology and recipe) are mapped instead of compiled, the "
                         "rest are compiled and stored")
    ap.add_argument("--table-layout", type=int, default=2, choices=[0, 2],
                    help="cfg3/cfg4: stencil table layout (0 record slices, 2 locality-ordered)")

    ap.add_argument("--copy-probe", action="store_true",
                    help="time a plain device-to-device copy of 1 GiB and a 1 GiB read "
                         "(torch.sum) instead: measured HBM bandwidths, and the reference "
                         "kernel for FETCH_SIZE/WRITE_SIZE scaling")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the evaluator shards by mesh across GPUs without collectives)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe:
        copy_probe(args)
        return
    from opensubdiv_amd import _lib
    lib = _lib.load_hip()
    name = ctypes.create_string_buffer(256)
    lib.osdhip_device_info(name, 256)
    if args.only == "cfg3":
        print(json.dumps(run_batched(args, lib, name.value.decode())), flush=True)
        return
    if args.only == "cfg4":
        print(json.dumps(run_served(args, lib, name.value.decode())), flush=True)
        return
    want = {"cfg1", "cfg1_refine", "cfg2"} if args.only is None else {args.only, "cfg1"}
    inputs = build_inputs(need_refine="cfg1_refine" in want, need_patches="cfg2" in want)
    with_cpu = not args.no_cpu
    src = inputs["mesh"].positions.astype(np.float32)
    workloads = {}
    workloads["cfg1"] = run_stencils(args, lib, "cfg1 limit stencils", inputs["limit"], src, 3,
                                     kernel_labels(3, 3), with_cpu)
    if "cfg1_refine" in want:
        workloads["cfg1_refine"] = run_stencils(args, lib, "cfg1 refinement", inputs["refine"],
                                                src, 1, kernel_labels(1, 3), with_cpu)
    if "cfg2" in want:
        workloads["cfg2"] = run_patches(args, lib, inputs, with_cpu)
    if args.only is None:
        workloads["cfg0"] = run_cube(args, lib, with_cpu)
    head = workloads["cfg1"]
    out = {
        "metric": METRIC,
        "value": head["points_per_s"],
        "unit": "points/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "cfg1: cube-sphere 6x128x128 quads (98,306 verts), 5% creased "
                               "edges, limit stencils (P, du, dv) at the 6,291,458 uniform "
                               "level-3 vertex locations, one fused launch per frame",
                   "stencils": head["stencils"], "table_entries": head["table_entries"],
                   "floats_per_point": 3, "device": name.value.decode()},
        "roofline": {"bound": "hbm", "achieved": head["GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": head["GBps"] / HBM_PEAK_GBS,
                     "traffic": traffic_for("cfg1_limit", head["kernel"].split(" ")[0]),
                     "kernel": head["kernel"],
                     "algorithmic_bytes": head["algorithmic_bytes"],
                     "kernel_ms": head["kernel_ms"]},
        "cpu_baseline": head.get("cpu_baseline"),
        "parity": head.get("parity"),
        "workloads": {k: v for k, v in workloads.items() if k != "cfg1"},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
