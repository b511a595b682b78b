"""Benchmark of the MI355X Osd evaluator (one GPU).

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--only cfg3|cfg4|cfg1|...]

Workloads (synthetic, seeded; tables built by the project's own table
compiler, which reproduces Far's tables):

* cfg3 (headline): 256 cube-sphere meshes (seed 0), n log-uniform in
  [8, 128], 5% creased edges, uniform refinement level 2-4, 1-16 frames of
  float32 xyz + 4 primvars, 64K-1M PatchCoords each.  One step = every
  mesh's refinement of every frame, its patch points, and P / dPdu / dPdv at
  its coords for every frame, in one graph launch per output stage.
  value = refined rows + patch coords per second, over all frames.
* cfg4: the serving loop over 32 resident topologies drawn like cfg3,
  Poisson arrivals at a fixed fraction of the measured capacity; p50 / p99
  latency beside points/s.
* cfg1: cube-sphere, 6 x 128 x 128 quads (98,306 vertices), float32 xyz on
  the unit sphere with 1% seeded radial noise, 5% of edges creased with
  sharpness U[0,4).  Limit stencil table with du/dv weights at the 6,291,458
  vertex locations of uniform level 3; one fused EvalStencils launch.
* cfg1_refine: the uniform level-3 refinement table of the same mesh, P.
* cfg2: EvalPatches (P, dP/du, dP/dv) at 4,194,304 PatchCoords (ptex face
  uniform, (u,v) uniform [0,1)^2, PatchMap lookup) on the iso-3
  feature-adaptive patch table (B-spline + Gregory end caps) of the mesh.
* cfg0: the tutorial cube, uniform level 3 (a CPU configuration).

The JSON line carries the roofline of the headline step (algorithmic HBM
bytes per step / the step's time from HIP events on the launch stream, as a
fraction of 8 TB/s and of the copy rate measured in the same run), the CPU
baseline (Osd::CpuEvaluator and OmpEvaluator built from the reference
sources when that library is present, else the C restatement, on a bounded
sample of the same workload), parity (bit-exact checks against the
reference CPU evaluator and position-sensitive fingerprints of every output
buffer) and each workload's error against an FP64 evaluation of FP64
tables.
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
# MI355X FP32 VALU without packed math: 256 CUs x 4 SIMDs x 16 lanes at
# 2.4 GHz, one multiply or add per lane per clock (the published 157.3
# TFLOP/s vector figure counts packed FMAs, 4 flops per lane per clock)
VALU_PEAK_OPS = 256 * 4 * 16 * 2.4e9
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")
METRIC = "limit points/s + HBM GB/s vs roofline; served p50/p99 latency at fixed QPS"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# --------------------------------------------------------------- inputs

def build_inputs(need_refine=True, need_patches=True, fp64=True):
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
    if need_refine and fp64:
        out["refine64"] = topo.stencil_table(double=True)
    ptex, u, v = topo.vertex_locations(3)
    pt = far.Topology(mesh).patch_table(3)
    out["patches"] = pt
    log(f"[bench] patch table built ({time.time() - t0:.1f}s)")
    out["limit"] = pt.limit_stencils(ptex, u, v)
    if fp64:
        out["limit64"] = pt.limit_stencils(ptex, u, v, double=True)
    log(f"[bench] limit table built ({time.time() - t0:.1f}s)")
    if need_patches:
        out["points"] = pt.point_stencils()
        if fp64:
            out["points64"] = pt.point_stencils(double=True)
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


def traffic_entry(key):
    """The committed counter profile's entry for a workload, if any."""
    if os.path.exists(PROFILE_TRAFFIC):
        return json.load(open(PROFILE_TRAFFIC)).get(key)
    return None


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
            "split": f"k_stencils_sliced<{nw},{lc}>", "csr_lds": f"k_stencils_csr<{nw},{lc}>",
            "rec_pair": f"k_stencils_rec_paired<{lc},16>",
            "tiled8": f"k_stencils_tiled<{nw},{lc},8>", "tiled4": f"k_stencils_tiled<{nw},{lc},4>",
            "tiled4_dom": f"k_stencils_tiled<{nw},{lc},4> (rows by dominant vertex only)"}


def bbox_rel_err(x, ref):
    """max |x - ref| over the bounding-box diagonal of ref (per column set)."""
    x = np.asarray(x, np.float64).reshape(len(ref), -1)
    ref = np.asarray(ref, np.float64).reshape(len(ref), -1)
    if ref.size == 0:
        return 0.0
    diag = float(np.linalg.norm(ref.max(axis=0) - ref.min(axis=0))) or 1.0
    return float(np.abs(x - ref).max() / diag)


def omp_threads():
    return min(os.cpu_count() or 1, 16)


OMP_NOTE = ("reference OmpEvaluator: default OpenMP schedule, and its per-thread accumulators "
            "share cache lines (osd/ompKernel.cpp:90-111), so it can run slower than one core")


def run_stencils(args, lib, name, table_np, src_np, nw, kernel_label, with_cpu, table64=None):
    """EvalStencils of a (limit) stencil table through HipStencilTable.
    table64: the same table in double (Far StencilTableReal<double>) for
    the error against an FP64 evaluation."""
    import torch
    from opensubdiv_amd import HipStencilTable, _lib
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    D = _lib.Desc
    L = src_np.shape[1]
    n = len(table_np["sizes"])
    nnz = len(table_np["indices"])
    t0 = time.time()
    htable = HipStencilTable.Create(table_np, s)
    log(f"[bench] {name}: device table {htable.device_bytes() / 1e6:.0f} MB in "
        f"{time.time() - t0:.1f}s")
    d_src = torch.from_numpy(np.ascontiguousarray(src_np)).to("cuda")
    outs = [torch.empty(n * L, dtype=torch.float32, device="cuda") for _ in range(nw)]
    alg = stencil_bytes(n, nnz, nw, L, src_np.shape[0], nw)
    # autotune: a few untimed launches of each kernel variant (record slices
    # with gather groups of 16 / 8 / 4, split-stream slices, CSR with LDS
    # staging, locality-ordered slices with control points staged in LDS,
    # and for one weight array paired records; per-table options),
    # then the timed region runs exactly `steps` launches of the fastest
    # (layout, group, kernel, layout-2 row order)
    VARIANTS = {"rec_g8": (0, 8, 0, 0), "rec_g4": (0, 4, 0, 0), "rec_g16": (0, 16, 0, 0),
                "split": (1, 8, 0, 0), "csr_lds": (0, 8, 1, 0), "tiled8": (2, 8, 0, 0),
                "tiled4": (2, 4, 0, 0), "tiled4_dom": (2, 4, 0, 2)}
    if nw == 1:
        VARIANTS["rec_pair"] = (3, 16, 0, 0)
    if args.variant:                      # one variant only (profiling runs)
        if args.variant not in VARIANTS:
            raise SystemExit(f"--variant {args.variant}: not one of {sorted(VARIANTS)}")
        VARIANTS = {args.variant: VARIANTS[args.variant]}
    variants = {}
    for vname, (layout, group, kernel, order) in VARIANTS.items():
        tab = htable
        if layout != 0:
            tab = HipStencilTable.Create(table_np, s, layout=layout, group=group, order=order)
        else:
            htable.set("group", group)
            htable.set("kernel", kernel)
        fn = table_eval_fn(lib, tab, d_src, D(0, L, L), outs, [D(0, L, L)] * nw, s)
        _, ms = time_launches(fn, 5, 2, stream)
        avg = float(np.median(ms))
        variants[vname] = {"kernel_ms": avg, "GBps": alg / (avg * 1e-3) / 1e9}
        if layout == 2:
            info = tab.layout()
            variants[vname].update(record_entries=info["record_entries"],
                                   size_window=info["size_window"],
                                   listed_slices=info["listed_slices"], slices=info["slices"])
        if tab is not htable:
            del tab, fn
            torch.cuda.synchronize()
    best = min(variants, key=lambda k: variants[k]["kernel_ms"])
    layout, group, kernel, order = VARIANTS[best]
    if layout != 0:
        del htable
        htable = HipStencilTable.Create(table_np, s, layout=layout, group=group, order=order)
    else:
        htable.set("group", group)
        htable.set("kernel", kernel)
    step = table_eval_fn(lib, htable, d_src, D(0, L, L), outs, [D(0, L, L)] * nw, s)
    wall, ms = time_launches(step, args.steps, args.warmup, stream)
    got = [o.cpu().numpy().reshape(n, L) for o in outs]
    avg = float(np.mean(ms))
    res = {"points_per_s": n * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps,
           "kernel_ms": avg, "stencils": n, "table_entries": nnz, "floats_per_point": L,
           "outputs": nw, "algorithmic_bytes": alg, "GBps": alg / (avg * 1e-3) / 1e9,
           "frac_of_peak": alg / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "kernel": f"{kernel_label[best]} ({best})",
           "autotune": variants}
    if table64 is not None:
        import oracle
        src64 = np.ascontiguousarray(src_np, np.float64)
        errs = {}
        for k, key in enumerate(("weights", "du", "dv")[:nw]):
            want = np.zeros((n, L), np.float64)
            assert oracle.eval_stencils_f64(src64, (0, L, L), want, (0, L, L), table64["sizes"],
                                            table64["offsets"], table64["indices"],
                                            table64[key])
            errs[("P", "dPdu", "dPdv")[k]] = bbox_rel_err(got[k], want)
        res["fp64_error"] = {"bbox_rel_max": max(errs.values()), "per_output": errs,
                             "reference": "FP64 evaluation of the table compiled in double "
                                          "(Far StencilTableReal<double> semantics)"}
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
        res["parity"] = {"bit_exact_vs_cpu": all(np.array_equal(a, b) for a, b in zip(got, dst))}
        threads = omp_threads()
        odst = [np.zeros((n, L), np.float32) for _ in range(nw)]
        if ref.available():
            okind = "reference OmpEvaluator"
            if nw == 1:
                omp = lambda: ref.cpu_eval_stencils(src_np, (0, L, L), odst[0], (0, L, L),  # noqa
                                                    table_np, threads=threads)
            else:
                omp = lambda: ref.cpu_eval_stencils_d1(src_np, (0, L, L), odst,  # noqa
                                                       [(0, L, L)] * 3, table_np,
                                                       threads=threads)
        else:
            okind = "port (OpenMP restatement)"
            omp = (lambda: oracle.eval_stencils(src_np, (0, L, L), odst[0], (0, L, L),  # noqa
                                                table_np["sizes"], table_np["offsets"],
                                                table_np["indices"], table_np["weights"],
                                                threads=threads)) if nw == 1 else None
        if omp is not None:
            m2, r2 = cpu_time(omp, args.cpu_budget / 2)
            res["cpu_baseline"]["omp"] = {
                "value": n / m2, "cores": threads, "kind": okind, "note": OMP_NOTE,
                "sample": f"full table x {r2} runs, median {m2 * 1e3:.1f} ms",
                "bit_exact_vs_cpu": all(np.array_equal(a, b) for a, b in zip(odst, dst))}
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

    def make_step(entry):
        fn = getattr(lib, entry)

        def step():
            rc = fn(d_src.data_ptr(), d3, outs[0].data_ptr(), d3, outs[1].data_ptr(), d3,
                    outs[2].data_ptr(), d3, None, z, None, z, None, z, n, d_coords.data_ptr(),
                    d_arrays.data_ptr(), d_indices.data_ptr(), d_params.data_ptr(), s)
            if rc != 0:
                raise RuntimeError(f"{entry} failed: {rc}")
        return step
    alg = patch_bytes(n, L, 3, len(pt.indices), len(pt.params), len(pt.arrays), nctrl + npts)
    # autotune: caller order vs patch-sorted order (binning by patch
    # inside each call, included in the timed step)
    entries = {"caller_order": "osdhip_eval_patches", "patch_sorted": "osdhip_eval_patches_sorted"}
    variants = {}
    for vname, entry in entries.items():
        _, ms = time_launches(make_step(entry), 5, 2, stream)
        avg = float(np.median(ms))
        variants[vname] = {"step_ms": avg, "GBps": alg / (avg * 1e-3) / 1e9}
    best = min(variants, key=lambda k: variants[k]["step_ms"])
    wall, ms = time_launches(make_step(entries[best]), args.steps, args.warmup, stream)
    avg = float(np.mean(ms))
    res = {"points_per_s": n * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps,
           "kernel_ms": avg, "patch_coords": n, "patches": len(pt.params),
           "patch_map": device_patch_map(args, inputs, stream),
           "gregory_patches": int(sum(a["numPatches"] for a in pt.arrays if a["desc"] == 9)),
           "algorithmic_bytes": alg, "GBps": alg / (avg * 1e-3) / 1e9,
           "frac_of_peak": alg / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "kernel": f"k_patches<true,false,3> ({best})", "autotune": variants}
    got = [o.cpu().numpy().reshape(n, L) for o in outs]
    if inputs.get("points64") is not None:
        import oracle
        p64 = inputs["points64"]
        ctrl64 = np.ascontiguousarray(mesh.positions, np.float64)
        pts64 = oracle.apply_stencils(ctrl64, p64["sizes"], p64["offsets"], p64["indices"],
                                      p64["weights"], np.float64)
        src64 = np.ascontiguousarray(np.concatenate([ctrl64, pts64]))
        want = [np.zeros((n, L), np.float64) for _ in range(3)]
        assert oracle.eval_patches_f64(src64, (0, L, L), want, [(0, L, L)] * 3, coords,
                                       pt.arrays, pt.indices, pt.params)
        errs = {k: bbox_rel_err(g, w) for k, g, w in zip(("P", "dPdu", "dPdv"), got, want)}
        res["fp64_error"] = {"bbox_rel_max": max(errs.values()), "per_output": errs,
                             "reference": "FP64 patch points (double point stencils) and "
                                          "FP64 basis (Far EvaluateBasis<double> semantics)"}
    if with_cpu:
        src_np = d_src.cpu().numpy().reshape(-1, 3)
        import oracle
        from oracle import ref
        dst = [np.zeros((n, L), np.float32) for _ in range(3)]
        odst = [np.zeros((n, L), np.float32) for _ in range(3)]
        threads = omp_threads()
        if ref.available():
            kind, okind = "reference", "reference OmpEvaluator"
            run = lambda: ref.cpu_eval_patches(src_np, (0, L, L), dst, [(0, L, L)] * 3,  # noqa
                                               coords, pt.arrays, pt.indices, pt.params)
            omp = lambda: ref.cpu_eval_patches(src_np, (0, L, L), odst, [(0, L, L)] * 3,  # noqa
                                               coords, pt.arrays, pt.indices, pt.params,
                                               threads=threads)
        else:
            kind, okind = "port", "port (OpenMP restatement)"
            run = lambda: oracle.eval_patches(src_np, (0, L, L), dst, [(0, L, L)] * 3,  # noqa
                                              coords, pt.arrays, pt.indices, pt.params)
            omp = lambda: oracle.eval_patches(src_np, (0, L, L), odst, [(0, L, L)] * 3,  # noqa
                                              coords, pt.arrays, pt.indices, pt.params,
                                              threads=threads)
        med, runs = cpu_time(run, args.cpu_budget, max_runs=3)
        res["cpu_baseline"] = {"value": n / med, "unit": "points/s", "cores": 1, "kind": kind,
                               "sample": f"all {n} coords x {runs} runs, median "
                                         f"{med * 1e3:.1f} ms"}
        m2, r2 = cpu_time(omp, args.cpu_budget / 2, max_runs=5)
        res["cpu_baseline"]["omp"] = {
            "value": n / m2, "cores": threads, "kind": okind, "note": OMP_NOTE,
            "sample": f"all coords x {r2} runs, median {m2 * 1e3:.1f} ms",
            "bit_exact_vs_cpu": all(np.array_equal(a, b) for a, b in zip(odst, dst))}
        res["parity"] = {"bit_exact_vs_cpu": all(np.array_equal(a, b) for a, b in zip(got, dst))}
    return res


def fingerprint(buf, chunk=1 << 26):
    """Position-sensitive fingerprint of a float32 device buffer: the sum
    over i of bits[i] * w(i) modulo 2^64, with w(i) an odd 31-bit hash of
    the position.  Equal buffers give equal fingerprints; a permutation of
    rows, a shifted range or any changed value changes it (with
    overwhelming probability).  Computed on the device in chunks."""
    import torch
    v = buf.view(torch.int32).reshape(-1)
    total = 0
    for i in range(0, v.numel(), chunk):
        x = v[i:i + chunk].to(torch.int64) & 0xFFFFFFFF
        w = torch.arange(i, i + x.numel(), device=x.device, dtype=torch.int64)
        w = ((w * 0x9E3779B1 + 0x7F4A7C15) & 0x7FFFFFFF) | 1
        total = (total + int((x * w).sum().item())) % (1 << 64)
        del x, w
    return total


def mesh_fingerprints(dms):
    """Fingerprints of every output buffer of every mesh: refined rows,
    patch points (the vertex buffer past the control points) and P, dPdu,
    dPdv."""
    out = []
    for d in dms:
        bufs = [d.refined, d.vbuf[d.nctrl * d.spec.frames * d.lp:]] + d.out
        out.append([fingerprint(b) for b in bufs])
    return out


def check_mesh_vs_reference(d, frames, fp64=True):
    """One batched mesh against the reference CPU evaluator at full size:
    its tables recompiled on the host (the compiler is deterministic), then
    for each frame in `frames` the refinement rows, the patch points and
    P / dPdu / dPdv at every coord are compared bit for bit with
    Osd::CpuEvaluator (the C restatement where the reference library is
    absent) on the same inputs; and the error of the device results against
    an FP64 evaluation of FP64 tables (frame frames[0])."""
    import oracle
    from oracle import ref
    from opensubdiv_amd import far, meshes
    from opensubdiv_amd.workloads import compile_mesh
    sp, L, F, Lp = d.spec, d.length, d.spec.frames, d.lp
    h = compile_mesh(sp, L, threads=0)
    try:
        mesh = meshes.cube_sphere(sp.n, seed=sp.seed, radial_noise=0.01, crease_fraction=0.05,
                                  sharpness_max=4.0)
        topo = far.Topology(mesh)
        topo.refine_uniform(sp.level)
        table = topo.stencil_table()
        refined = d.refined.view(F, d.nrefine, L)
        vpts = d.vbuf[d.nctrl * F * Lp:].view(d.npts, F, Lp)
        outs = [o.view(F, d.ncoords, L) for o in d.out]
        use_ref = ref.available()
        same = True
        for f in frames:
            ctrl = np.ascontiguousarray(h.frames[f])
            dst = np.zeros((d.nrefine, L), np.float32)
            pts = np.zeros((d.npts, L), np.float32)
            po = [np.zeros((d.ncoords, L), np.float32) for _ in range(3)]
            if use_ref:
                ok = ref.cpu_eval_stencils(ctrl, (0, L, L), dst, (0, L, L), table) and \
                    ref.cpu_eval_stencils(ctrl, (0, L, L), pts, (0, L, L), h.points)
                src = np.ascontiguousarray(np.concatenate([ctrl, pts]))
                ok = ok and ref.cpu_eval_patches(src, (0, L, L), po, [(0, L, L)] * 3, h.coords,
                                                 h.arrays, h.indices, h.params)
            else:
                dst = oracle.apply_stencils(ctrl, table["sizes"], table["offsets"],
                                            table["indices"], table["weights"])
                pts = oracle.apply_stencils(ctrl, h.points["sizes"], h.points["offsets"],
                                            h.points["indices"], h.points["weights"])
                src = np.ascontiguousarray(np.concatenate([ctrl, pts]))
                ok = oracle.eval_patches(src, (0, L, L), po, [(0, L, L)] * 3, h.coords, h.arrays,
                                         h.indices, h.params)
            same &= bool(ok)
            same &= np.array_equal(refined[f].cpu().numpy().view(np.uint32), dst.view(np.uint32))
            same &= np.array_equal(np.ascontiguousarray(vpts[:, f, :L].cpu().numpy()).view(
                np.uint32), pts.view(np.uint32))
            for o, w in zip(outs, po):
                same &= np.array_equal(o[f].cpu().numpy().view(np.uint32), w.view(np.uint32))
        res = {"mesh": sp.index, "n": sp.n, "level": sp.level, "frames_checked": list(frames),
               "refined_rows": d.nrefine, "patch_points": d.npts, "coords": d.ncoords,
               "bit_exact_vs_cpu": bool(same),
               "checker": "reference CpuEvaluator" if use_ref else "oracle restatement"}
        if fp64:
            f = frames[0]
            c64 = np.ascontiguousarray(h.frames[f], np.float64)
            t64 = topo.stencil_table(double=True)
            r64 = oracle.apply_stencils(c64, t64["sizes"], t64["offsets"], t64["indices"],
                                        t64["weights"], np.float64)
            pt = far.Topology(mesh).patch_table(sp.level)
            p64t = pt.point_stencils(double=True)
            p64 = oracle.apply_stencils(c64, p64t["sizes"], p64t["offsets"], p64t["indices"],
                                        p64t["weights"], np.float64)
            s64 = np.ascontiguousarray(np.concatenate([c64, p64]))
            w64 = [np.zeros((d.ncoords, L), np.float64) for _ in range(3)]
            assert oracle.eval_patches_f64(s64, (0, L, L), w64, [(0, L, L)] * 3, h.coords,
                                           h.arrays, h.indices, h.params)
            errs = {"refined": bbox_rel_err(refined[f].cpu().numpy(), r64)}
            for k, o, w in zip(("P", "dPdu", "dPdv"), outs, w64):
                errs[k] = bbox_rel_err(o[f].cpu().numpy(), w)
            res["fp64_error"] = {"bbox_rel_max": max(errs.values()), "per_output": errs}
        return res
    finally:
        h.release()


def table_cache(args):
    """The --table-cache directory as a TableCache (None without one)."""
    if not args.table_cache:
        return None
    from opensubdiv_amd.table_cache import TableCache
    return TableCache(args.table_cache)


def cache_summary(cache):
    return None if cache is None else {"hits": cache.hits, "misses": cache.misses}


def sample_for_check(members, max_rows=3_000_000):
    """The mesh of a stage checked against the reference CPU evaluator: the
    largest whose refinement has at most `max_rows` rows (bounded host
    time), else the smallest."""
    small = [d for d in members if d.nrefine <= max_rows]
    return max(small, key=lambda d: d.nrefine) if small else min(members, key=lambda d: d.nrefine)


def run_batched(args, lib, device, copy_gbps=None):
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
    workers = min(16, os.cpu_count() or 1)
    cache = table_cache(args)
    dms = workloads.build_device_meshes(specs, L, s, workers=workers, log=log, own_outputs=False,
                                        cache=cache, layout=args.table_layout,
                                        order=args.table_order)
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
    if args.config:
        for kv in args.config.split(","):
            k, v = kv.split("=")
            batch.set(k.strip(), int(v))
    parity = None
    if not args.no_parity:
        # parity at full size, every stage: (1) one sampled mesh's batched
        # outputs against the reference CPU evaluator, bit for bit, plus its
        # error against FP64; (2) position-sensitive fingerprints of every
        # output buffer, batched kernels vs one mesh and frame per launch
        parity = {"stages": []}
        for st in range(nstages):
            members = [d for d in dms if d.stage == st]
            batch.run_stage(st, 3, s)
            torch.cuda.synchronize()
            a = mesh_fingerprints(members)
            chk = sample_for_check(members)
            frames = sorted({0, chk.spec.frames - 1})
            sample = check_mesh_vs_reference(chk, frames)
            batch.run_plain_stage(st, s)
            torch.cuda.synchronize()
            b = mesh_fingerprints(members)
            parity["stages"].append({"meshes": len(members), "buffers": 5 * len(members),
                                     "batched_equals_plain": a == b, "sampled_mesh": sample})
            log(f"[cfg3] stage {st}: batched == plain fingerprints {a == b}; mesh "
                f"{chk.spec.index} (n {chk.spec.n}, level {chk.spec.level}) bit-exact vs CPU "
                f"{sample['bit_exact_vs_cpu']}, fp64 error {sample['fp64_error']['bbox_rel_max']:.2e}")
        parity["batched_equals_plain_all"] = all(x["batched_equals_plain"]
                                                 for x in parity["stages"])
        parity["sampled_bit_exact_vs_cpu_all"] = all(x["sampled_mesh"]["bit_exact_vs_cpu"]
                                                     for x in parity["stages"])
        parity["method"] = ("per stage: position-sensitive fingerprints (sum of bits x position "
                            "hash mod 2^64) of all 5 output buffers of every mesh, batched vs one "
                            "mesh and frame per launch; and one sampled mesh per stage compared "
                            "bit for bit with the reference CpuEvaluator on its first and last "
                            "frames")
    report = None
    if not args.no_tune:
        t0 = time.time()
        report = batch.tune(reps=2, stream=s)
        log(f"[cfg3] tuned in {time.time() - t0:.1f}s: {batch.config()}")
    points = sum(d.points_per_step() for d in dms)
    alg = sum(d.algorithmic_bytes() for d in dms)
    ops_a = sum(d.spec.frames * d.length * 2 * (d.nnz_refine + d.nnz_points) for d in dms)
    ops = sum(d.valu_ops() for d in dms)
    wall, ms = time_launches(lambda: batch.launch(s), args.steps, args.warmup, stream)
    kms = float(np.median(ms))
    achieved = alg / (kms * 1e-3) / 1e9
    plain = None
    if not args.no_plain:
        plain_steps = max(2, min(args.steps, 3))
        pwall, pms = time_launches(lambda: batch.run_plain(s), plain_steps, 1, stream)
        pkms = float(np.median(pms))
        plain = {"points_per_s": points * plain_steps / pwall, "step_ms": pkms,
                 "launches_per_step": sum(2 * d.spec.frames + d.spec.frames for d in dms),
                 "speedup_batched": pkms / kms}
    tr = traffic_entry("cfg3")
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
                   "table_order": args.table_order,
                   "output_bytes_per_step": out_bytes,
                   "refined_rows_per_step": sum(d.spec.frames * d.nrefine for d in dms),
                   "patch_points_per_step": sum(d.spec.frames * d.ncoords for d in dms),
                   "table_entries": sum(d.nnz_refine + d.nnz_points for d in dms),
                   "device_bytes": dev_bytes, "build_s": build_s,
                   "table_cache": cache_summary(cache)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tr.get("bytes_per_step") if tr else None,
                     "copy_rate_GBps": copy_gbps,
                     "frac_of_copy_rate": achieved / copy_gbps if copy_gbps else None,
                     "kernel": "one step: k_batch_stencils_tiled (phase A) + k_batch_patches* "
                               "(phase B), graph launches of every stage",
                     "algorithmic_bytes": alg, "kernel_ms": kms,
                     "output_bytes": out_bytes,
                     "valu": valu_roofline(ops, ops_a, kms, report)},
        "plain": plain,
        "autotune": report, "tuned": batch.config(),
        "parity": parity,
    }
    if tr:
        res["roofline"]["traffic_source"] = tr.get("source")
    if parity:
        errs = [x["sampled_mesh"]["fp64_error"]["bbox_rel_max"] for x in parity["stages"]]
        res["fp64_error"] = {"bbox_rel_max": max(errs), "per_stage": errs,
                             "reference": "sampled mesh per stage: FP64 tables, FP64 control "
                                          "points (promoted), FP64 basis"}
    if not args.no_cpu:
        res["cpu_baseline"] = batched_cpu_baseline(dms, args.cpu_budget)
    batch.close()
    del arena, dms
    return res


def valu_roofline(ops, ops_a, kms, report):
    """The step against the FP32 VALU rate: bit-exact sums take a multiply
    and an add per term (no FMA), and packed FP32 math is not used, so
    `ops` / VALU_PEAK_OPS is the step's arithmetic floor.  Per phase from
    the autotuner's best phase times when it ran."""
    out = {"ops": ops, "unit": "unfused FP32 multiplies and adds (lane ops)",
           "peak_ops_per_s": VALU_PEAK_OPS, "floor_ms": ops / VALU_PEAK_OPS * 1e3,
           "frac": ops / (kms * 1e-3) / VALU_PEAK_OPS}
    if report:
        for ph, n in (("phase_a", ops_a), ("phase_b", ops - ops_a)):
            runs = report.get(ph) or []
            if runs:
                best = min(r["ms"] for r in runs)
                out[ph] = {"ops": n, "best_ms": best, "floor_ms": n / VALU_PEAK_OPS * 1e3,
                           "frac": n / (best * 1e-3) / VALU_PEAK_OPS}
    return out


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
    residents = serve.residents_from_hosts(hosts, layout=args.table_layout)
    for h in hosts:
        h.release()
    torch.cuda.synchronize()
    build_s = time.time() - t0
    dev_budget, host_budget = 32 << 30, 24 << 30
    log(f"[cfg4] {len(residents)} resident topologies in {build_s:.1f}s; budgets "
        f"{dev_budget >> 30} GiB device, {host_budget >> 30} GiB pinned host")
    server = serve.Server(residents, L, dev_budget, host_budget, slots=args.slots,
                          max_epoch_points=int(args.epoch_points), copy_order=args.copy_order)

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
                   "copy_order": args.copy_order,
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


def batched_cpu_baseline(dms, budget_s):
    """Reference CpuEvaluator on a bounded sample of the batched workload:
    the refinement stencils, patch points and EvalPatches of the smallest
    meshes, one frame each, until about `budget_s` seconds of CPU work; then
    OmpEvaluator on the same sample (all host threads, at most 16)."""
    from oracle import ref
    import oracle
    from opensubdiv_amd.workloads import compile_mesh
    use_ref = ref.available()
    threads = omp_threads()
    t_end = time.time() + budget_s
    pts, secs, osecs, used = 0, 0.0, 0.0, 0
    for d in sorted(dms, key=lambda x: x.nrefine):
        if time.time() > t_end:
            break
        sp, L = d.spec, d.length
        h = compile_mesh(sp, L, threads=0)
        try:
            ctrl = np.ascontiguousarray(h.frames[0])
            from opensubdiv_amd import far, meshes
            mesh = meshes.cube_sphere(sp.n, seed=sp.seed, radial_noise=0.01,
                                      crease_fraction=0.05, sharpness_max=4.0)
            topo = far.Topology(mesh)
            topo.refine_uniform(sp.level)
            table = topo.stencil_table()
            n = len(h.coords)
            for omp in (0, threads):
                dst = np.zeros((len(table["sizes"]), L), np.float32)
                pbuf = np.zeros((len(h.points["sizes"]), L), np.float32)
                outs = [np.zeros((n, L), np.float32) for _ in range(3)]
                t0 = time.perf_counter()
                if use_ref:
                    ref.cpu_eval_stencils(ctrl, (0, L, L), dst, (0, L, L), table, threads=omp)
                    ref.cpu_eval_stencils(ctrl, (0, L, L), pbuf, (0, L, L), h.points,
                                          threads=omp)
                    src = np.concatenate([ctrl, pbuf])
                    ref.cpu_eval_patches(src, (0, L, L), outs, [(0, L, L)] * 3, h.coords,
                                         h.arrays, h.indices, h.params, threads=omp)
                else:
                    oracle.eval_stencils(ctrl, (0, L, L), dst, (0, L, L), table["sizes"],
                                         table["offsets"], table["indices"], table["weights"],
                                         threads=omp)
                    oracle.eval_stencils(ctrl, (0, L, L), pbuf, (0, L, L), h.points["sizes"],
                                         h.points["offsets"], h.points["indices"],
                                         h.points["weights"], threads=omp)
                    src = np.concatenate([ctrl, pbuf])
                    oracle.eval_patches(src, (0, L, L), outs, [(0, L, L)] * 3, h.coords,
                                        h.arrays, h.indices, h.params, threads=omp)
                dt = time.perf_counter() - t0
                if omp:
                    osecs += dt
                else:
                    secs += dt
        finally:
            h.release()
        pts += d.nrefine + n
        used += 1
    kind = "reference" if use_ref else "port"
    return {"value": pts / secs if secs else None, "unit": "points/s", "cores": 1, "kind": kind,
            "sample": f"{used} smallest meshes, frame 0: refinement stencils, patch points and "
                      f"EvalPatches (P, dPdu, dPdv); {pts} points in {secs:.1f}s",
            "omp": {"value": pts / osecs if osecs else None, "cores": threads,
                    "kind": "reference OmpEvaluator" if use_ref else "port (OpenMP restatement)",
                    "note": OMP_NOTE, "sample": f"the same {pts} points in {osecs:.1f}s"}}


def oracle_coord_dtype():
    from opensubdiv_amd.osd import PATCH_COORD_DTYPE
    return PATCH_COORD_DTYPE


def oracle_array_dtype():
    from opensubdiv_amd.osd import PATCH_ARRAY_DTYPE
    return PATCH_ARRAY_DTYPE


def oracle_param_dtype():
    from opensubdiv_amd.osd import PATCH_PARAM_DTYPE
    return PATCH_PARAM_DTYPE


def measure_copy(steps=10, warmup=2):
    """1 GiB device-to-device copy: reads and writes exactly 2^30 bytes per
    launch, so its FETCH_SIZE / WRITE_SIZE calibrate the counters and its
    rate (read + write bytes / time) is the measured copy bandwidth."""
    import torch
    n = 1 << 28
    a = torch.rand(n, device="cuda")
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream()
    _, ms = time_launches(lambda: b.copy_(a), steps, warmup, stream)
    kms = float(np.median(ms))
    _, rms = time_launches(lambda: a.sum(), steps, warmup, stream)
    rkms = float(np.median(rms))
    _, wms = time_launches(lambda: b.fill_(1.0), steps, warmup, stream)
    wkms = float(np.median(wms))
    del a, b
    torch.cuda.empty_cache()
    return {"bytes_read": 4 * n, "bytes_written": 4 * n, "kernel_ms": kms,
            "GBps": 8 * n / (kms * 1e-3) / 1e9,
            "read_probe": {"bytes_read": 4 * n, "kernel_ms": rkms,
                           "GBps": 4 * n / (rkms * 1e-3) / 1e9, "op": "torch.sum"},
            "write_probe": {"bytes_written": 4 * n, "kernel_ms": wkms,
                            "GBps": 4 * n / (wkms * 1e-3) / 1e9, "op": "torch.fill_"}}


def copy_probe(args):
    print(json.dumps({"probe": "copy", **measure_copy(args.steps, args.warmup),
                      "steps": args.steps}), flush=True)


def run_single_mesh(args, lib, device, want, with_cpu):
    """cfg1 (limit stencils), cfg1_refine, cfg2 (EvalPatches) and cfg0 (the
    tutorial cube): the one-mesh workloads."""
    inputs = build_inputs(need_refine="cfg1_refine" in want, need_patches="cfg2" in want,
                          fp64=not args.no_fp64)
    src = inputs["mesh"].positions.astype(np.float32)
    out = {}
    if "cfg1" in want:
        out["cfg1"] = run_stencils(args, lib, "cfg1 limit stencils", inputs["limit"], src, 3,
                                   kernel_labels(3, 3), with_cpu, inputs.get("limit64"))
        out["cfg1"]["traffic"] = traffic_for("cfg1_limit", out["cfg1"]["kernel"].split(" ")[0])
    if "cfg1_refine" in want:
        out["cfg1_refine"] = run_stencils(args, lib, "cfg1 refinement", inputs["refine"], src, 1,
                                          kernel_labels(1, 3), with_cpu, inputs.get("refine64"))
        out["cfg1_refine"]["traffic"] = traffic_for(
            "cfg1_refine", out["cfg1_refine"]["kernel"].split(" ")[0])
    if "cfg2" in want:
        out["cfg2"] = run_patches(args, lib, inputs, with_cpu)
    if "cfg0" in want:
        out["cfg0"] = run_cube(args, lib, with_cpu)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 error checks")
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds per CPU baseline")
    ap.add_argument("--only", default=None,
                    choices=[None, "cfg0", "cfg1", "cfg1_refine", "cfg2", "cfg3", "cfg4"])
    ap.add_argument("--requests", type=int, default=1500, help="cfg4: requests served")
    ap.add_argument("--load", type=float, default=0.7,
                    help="cfg4: arrival rate as a fraction of the measured capacity")
    ap.add_argument("--slots", type=int, default=4, help="cfg4: epochs in flight")
    ap.add_argument("--copy-order", default="srpt", choices=["srpt", "edf"],
                    help="cfg4: which result download goes next (fewest bytes left, or "
                         "earliest ideal completion)")
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
    ap.add_argument("--table-order", type=int, default=2, choices=[0, 1, 2],
                    help="cfg3: layout-2 row order (2 by dominant control vertex, 0 then by "
                         "size within row windows, 1 Far's order)")
    ap.add_argument("--config", default=None,
                    help="cfg3: batch knobs key=value,... set before (or instead of) tuning")
    ap.add_argument("--no-tune", action="store_true", help="cfg3: skip the autotuner")
    ap.add_argument("--variant", default=None,
                    help="cfg1 / cfg1_refine: time this table variant only (e.g. tiled4)")
    ap.add_argument("--no-parity", action="store_true", help="cfg3: skip the parity checks")
    ap.add_argument("--no-plain", action="store_true",
                    help="cfg3: skip timing the one-mesh-per-launch path")
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
    device = name.value.decode()
    with_cpu = not args.no_cpu
    copy = measure_copy()
    log(f"[bench] 1 GiB device copy: {copy['GBps']:.0f} GB/s (read + write)")
    if args.only == "cfg4":
        print(json.dumps(run_served(args, lib, device)), flush=True)
        return
    if args.only in ("cfg0", "cfg1", "cfg1_refine", "cfg2"):
        print(json.dumps(run_single_mesh(args, lib, device, {args.only}, with_cpu)[args.only]),
              flush=True)
        return
    # the headline: cfg3, the batched workload
    out = run_batched(args, lib, device, copy["GBps"])
    out["copy_probe"] = copy
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if args.only == "cfg3":
        print(json.dumps(out), flush=True)
        return
    log("[bench] cfg4 (served)")
    served = run_served(args, lib, device)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    out["served"] = dict(served["served"], parity=served["parity"],
                         value_points_per_s=served["value"])
    out["workloads"] = {"cfg4": served}
    log("[bench] cfg1 / cfg1_refine / cfg2 / cfg0")
    out["workloads"].update(run_single_mesh(args, lib, device,
                                            {"cfg0", "cfg1", "cfg1_refine", "cfg2"}, with_cpu))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
