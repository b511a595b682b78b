"""Benchmark of the MI355X Osd evaluator (one GPU).

    python bench.py [--steps K] [--warmup W] [--workload cfg1] [--no-cpu]

Workload "cfg1" (default): cube-sphere, 6 x 128 x 128 quads (98,306
vertices), float32 xyz on the unit sphere with 1% seeded radial noise, 5%
of edges creased with sharpness U[0,4); uniform level-3 Catmull-Clark
stencil table (6,291,458 stencils) built by the project's own table
compiler.  One step = one EvalStencils pass over the whole table (one
frame), inputs resident in HBM.  value = refined points per second.

The JSON line carries the roofline of the dominant kernel (algorithmic HBM
bytes per launch / average launch time from HIP events on the launch
stream) and a CPU baseline (Osd::CpuEvaluator built from the reference
sources when that library is present, else the C restatement), timed on
this machine's host cores over the same table.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E, GB/s
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_cfg1():
    from opensubdiv_amd import far, meshes
    t0 = time.time()
    mesh = meshes.cube_sphere(128, seed=1, radial_noise=0.01, crease_fraction=0.05,
                              sharpness_max=4.0)
    topo = far.Topology(mesh)
    topo.refine_uniform(3)
    table = topo.stencil_table()
    log(f"[bench] cfg1 table: {len(table['sizes'])} stencils, {len(table['indices'])} entries, "
        f"built in {time.time() - t0:.1f}s")
    return mesh, table


def stencil_bytes(n_st, nnz, n_weights, L, n_src, n_out_arrays):
    """Algorithmic HBM bytes of one EvalStencils launch: sizes + offsets
    (4 B each per stencil), indices (4 B) and n_weights weight arrays per
    entry, the outputs, and the control points read once."""
    return 8 * n_st + (4 + 4 * n_weights) * nnz + 4 * L * n_out_arrays * n_st + 4 * L * n_src


def cpu_baseline_stencils(src, table, L, budget_s=12.0, threads=0):
    """Times CpuEvaluator::EvalStencils (threads=0: one thread) or
    OmpEvaluator::EvalStencils (threads>0) over the full table, repeated
    within a time budget; falls back to the C restatement when the reference
    library is absent."""
    import oracle
    from oracle import ref
    n = len(table["sizes"])
    dst = np.zeros((n, L), np.float32)
    if ref.available():
        kind = "reference"
        run = lambda: ref.cpu_eval_stencils(src, (0, L, L), dst, (0, L, L), table,  # noqa: E731
                                            threads=threads)
    else:
        kind = "port"
        run = lambda: oracle.eval_stencils(src, (0, L, L), dst, (0, L, L), table["sizes"],  # noqa
                                           table["offsets"], table["indices"], table["weights"],
                                           threads=threads)
    run()                                   # warm-up
    times = []
    t_end = time.time() + budget_s
    while time.time() < t_end and len(times) < 9:
        t0 = time.perf_counter()
        assert run()
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": n / med, "unit": "points/s", "cores": threads or 1, "kind": kind,
            "sample": f"full table ({n} stencils) x {len(times)} runs, median {med * 1e3:.1f} ms",
            "dst": dst}


def time_launches(step, steps, warmup, stream):
    """Runs warmup + steps launches; returns (wall seconds of the timed
    region, per-launch milliseconds from HIP events on `stream`)."""
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


def run_cfg1(args):
    import ctypes

    import torch
    from opensubdiv_amd import HipStencilTable, _lib
    lib = _lib.load_hip()
    mesh, table = build_cfg1()
    L = 3
    n_src = mesh.nverts
    src_np = mesh.positions.astype(np.float32)
    n = len(table["sizes"])
    nnz = len(table["indices"])

    d_src = torch.from_numpy(np.ascontiguousarray(src_np)).to("cuda")
    d_dst = torch.empty(n * L, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    D = _lib.Desc
    t0 = time.time()
    htable = HipStencilTable.Create(table, s)
    log(f"[bench] device table: {htable.device_bytes() / 1e6:.0f} MB, {time.time() - t0:.1f}s")
    outs = (ctypes.c_void_p * 1)(d_dst.data_ptr())
    descs = (D * 1)(D(0, L, L))

    def table_step():
        rc = lib.osdhip_stencil_table_eval(htable.h, d_src.data_ptr(), D(0, L, L), 1,
                                           ctypes.cast(outs, ctypes.c_void_p),
                                           ctypes.cast(descs, ctypes.c_void_p), 0, n, s)
        if rc != 0:
            raise RuntimeError(f"osdhip_stencil_table_eval failed: {rc}")

    def raw_step():
        rc = lib.osdhip_eval_stencils(d_src.data_ptr(), D(0, L, L), d_dst.data_ptr(), D(0, L, L),
                                      htable.GetSizesBuffer(), htable.GetOffsetsBuffer(),
                                      htable.GetIndicesBuffer(), htable.GetWeightsBuffer(), 0, n, s)
        if rc != 0:
            raise RuntimeError(f"osdhip_eval_stencils failed: {rc}")

    alg = stencil_bytes(n, nnz, 1, L, n_src, 1)
    variants = {}
    results = {}
    for name, kernel, fn in [("table_csr_lds", 1, table_step), ("raw_csr_lds", None, raw_step),
                             ("table_sliced", 0, table_step)]:
        if kernel is not None:
            lib.osdhip_config(b"stencil_kernel", kernel)
        wall, ms = time_launches(fn, args.steps, args.warmup, stream)
        avg = float(np.mean(ms))
        variants[name] = {"kernel_ms": avg, "GBps": alg / (avg * 1e-3) / 1e9,
                          "points_per_s": n * args.steps / wall}
        results[name] = (wall, ms, d_dst.cpu().numpy().reshape(n, L).copy())
    best = "table_sliced"                       # HipEvaluator::EvalStencils path
    wall, ms, got = results[best]
    avg_ms = float(np.mean(ms))
    achieved = alg / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": "limit points/s + HBM GB/s vs roofline; served p50/p99 latency at fixed QPS",
        "value": n * args.steps / wall,
        "unit": "points/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "cfg1: cube-sphere 6x128x128 quads, 5% creased edges, uniform "
                               "level-3 stencil table, xyz, one mesh per launch",
                   "stencils": n, "table_entries": nnz, "control_points": n_src,
                   "floats_per_point": L, "path": "HipStencilTable device layout"},
    }
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        tr = json.load(open(PROFILE_TRAFFIC)).get("cfg1_stencils")
        if tr:
            traffic = tr["bytes_per_launch"]
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                       "kernel": "k_stencils_sliced<1,3>", "algorithmic_bytes": alg,
                       "kernel_ms": avg_ms}
    out["variants"] = variants
    if not args.no_cpu:
        cb = cpu_baseline_stencils(src_np, table, L)
        ref_out = cb.pop("dst")
        nthreads = os.cpu_count() or 1
        omp = cpu_baseline_stencils(src_np, table, L, budget_s=6.0, threads=min(nthreads, 16))
        omp.pop("dst")
        cb["omp"] = {k: omp[k] for k in ("value", "cores", "kind", "sample")}
        out["cpu_baseline"] = cb
        out["parity"] = {k: bool(np.array_equal(r[2], ref_out)) for k, r in results.items()}
    else:
        out["cpu_baseline"] = None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cfg1", choices=["cfg1"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the evaluator shards by mesh across GPUs without collectives)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    out = run_cfg1(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
