"""GPU tests of the serving loop (continuous batching over resident
topologies): every request's result is the same bits alone and in a mix of
concurrent requests, and equals the plain per-mesh entry points."""
import numpy as np
import pytest

from opensubdiv_amd import _lib
from opensubdiv_amd.serve import Request, Server, residents_from_hosts
from opensubdiv_amd.workloads import MeshSpec, compile_mesh

pytestmark = pytest.mark.gpu

# the last topology is refined to level 4 (the widest stencil slices)
SPECS = [MeshSpec(0, 3, 2, 4, 900, 21), MeshSpec(1, 2, 3, 4, 400, 22),
         MeshSpec(2, 4, 2, 4, 1500, 23), MeshSpec(3, 4, 4, 2, 2000, 24)]
L = 7


@pytest.fixture(scope="module", params=[0, 2], ids=["rec", "tiled"])
def residents(request):
    hosts = [compile_mesh(s, L) for s in SPECS]
    res = residents_from_hosts(hosts, layout=request.param)
    yield res
    for h in hosts:
        h.release()


def make_requests():
    # several requests per topology, both kinds, different frames: the first
    # six arrive together (one multi-request epoch), the rest staggered
    out = []
    k = 0
    for rep in range(3):
        for t in range(len(SPECS)):
            for kind in ("refine", "limit"):
                out.append(Request(k, t, kind, 0.0 if k < 6 else 0.0002 * k, (k + rep) % 4))
                k += 1
    return out


def plain(res, r):
    """The same request through osdhip_stencil_table_eval / osdhip_eval_patches."""
    import torch
    from opensubdiv_amd.osd import _stream
    lib = _lib.load_hip()
    t = res[r.topo]
    D = _lib.Desc
    frame = torch.from_numpy(np.ascontiguousarray(t.frames[r.frame % t.frames.shape[0]])).cuda()
    st = _stream()
    if r.kind == "refine":
        out = torch.full((t.nrefine * L,), float("nan"), device="cuda")
        import ctypes
        o = (ctypes.c_void_p * 1)(out.data_ptr())
        d = (D * 1)(D(0, L, L))
        assert lib.osdhip_stencil_table_eval(t.refine_table.h, frame.data_ptr(), D(0, L, L), 1,
                                             ctypes.cast(o, ctypes.c_void_p),
                                             ctypes.cast(d, ctypes.c_void_p), 0, t.nrefine, st) == 0
        torch.cuda.synchronize()
        return out.cpu().numpy()
    import ctypes
    vb = torch.full(((t.nctrl + t.npts) * L,), float("nan"), device="cuda")
    vb[:t.nctrl * L] = frame.reshape(-1)
    o = (ctypes.c_void_p * 1)(vb.data_ptr())
    d = (D * 1)(D(t.nctrl * L, L, L))
    assert lib.osdhip_stencil_table_eval(t.point_table.h, vb.data_ptr(), D(0, L, L), 1,
                                         ctypes.cast(o, ctypes.c_void_p),
                                         ctypes.cast(d, ctypes.c_void_p), 0, t.npts, st) == 0
    outs = [torch.full((t.ncoords * L,), float("nan"), device="cuda") for _ in range(3)]
    dd = D(0, L, L)
    z = D(0, 0, 0)
    assert lib.osdhip_eval_patches(vb.data_ptr(), D(0, L, L), outs[0].data_ptr(), dd,
                                   outs[1].data_ptr(), dd, outs[2].data_ptr(), dd, None, z, None,
                                   z, None, z, t.ncoords, t.coords.data_ptr(), t.arrays.data_ptr(),
                                   t.indices.data_ptr(), t.params.data_ptr(), st) == 0
    torch.cuda.synchronize()
    return np.concatenate([o.cpu().numpy() for o in outs])


def test_results_alone_equal_results_in_the_mix(residents):
    reqs = make_requests()
    server = Server(residents, L, device_budget=64 << 20, host_budget=64 << 20, slots=3)
    done, mix = server.run(reqs, keep={r.rid for r in reqs}, max_seconds=60)
    assert len(done) == len(reqs)
    assert all(r.done >= r.arrival for r in done)
    assert server.dev_arena.used() == 0 and server.host_arena.used() == 0
    # some epochs really batched several requests together
    assert sum(server.epochs) == len(reqs) and max(server.epochs) > 1
    for r in make_requests():
        alone_server = Server(residents, L, device_budget=64 << 20, host_budget=64 << 20, slots=1)
        r.arrival = 0.0
        _, alone = alone_server.run([r], keep={r.rid}, max_seconds=30)
        a, m = alone[r.rid], mix[r.rid]
        assert np.array_equal(a.view(np.uint32), m.view(np.uint32)), f"request {r.rid} {r.kind}"
        p = plain(residents, r)
        assert np.array_equal(a.view(np.uint32), p.view(np.uint32)), f"plain {r.rid} {r.kind}"


def test_tight_budget_serialises_but_completes(residents):
    """A device budget that fits about one request at a time: admission
    waits for memory to free up, every request still completes with the
    same result."""
    reqs = make_requests()[:8]
    server = Server(residents, L, device_budget=64 << 20, host_budget=64 << 20, slots=2)
    need = max(sum(server.dev_floats(r)) for r in reqs) * 4
    tight = Server(residents, L, device_budget=need + (1 << 16), host_budget=64 << 20, slots=2)
    done, res_t = tight.run(reqs, keep={r.rid for r in reqs}, max_seconds=60)
    assert len(done) == len(reqs)
    for r in make_requests()[:8]:
        r.arrival = 0.0
        _, alone = server.run([r], keep={r.rid}, max_seconds=30)
        assert np.array_equal(alone[r.rid].view(np.uint32), res_t[r.rid].view(np.uint32))


def test_epoch_cap_splits_the_burst_with_same_results(residents):
    reqs = make_requests()
    base = Server(residents, L, device_budget=64 << 20, host_budget=64 << 20, slots=3)
    _, want = base.run(reqs, keep={r.rid for r in reqs}, max_seconds=60)
    capped = Server(residents, L, device_budget=64 << 20, host_budget=64 << 20, slots=3,
                    max_epoch_points=3000)
    reqs2 = make_requests()
    done, got = capped.run(reqs2, keep={r.rid for r in reqs2}, max_seconds=60)
    assert len(done) == len(reqs2)
    for n, pts in zip(capped.epochs, capped.epoch_points):
        assert n == 1 or pts <= 3000
    for r in reqs2:
        assert np.array_equal(want[r.rid].view(np.uint32), got[r.rid].view(np.uint32))


def test_chunked_downloads_keep_results_and_let_small_requests_finish_first(residents):
    """Results downloaded in many small chunks (shortest remaining transfer
    first) are the same bits as whole downloads; a small request that
    arrives with a large one is not held behind the large one's transfer."""
    reqs = make_requests()
    whole = Server(residents, L, device_budget=64 << 20, host_budget=64 << 20, slots=3,
                   chunk_bytes=1 << 30)
    _, want = whole.run(reqs, keep={r.rid for r in reqs}, max_seconds=60)
    chunked = Server(residents, L, device_budget=64 << 20, host_budget=64 << 20, slots=3,
                     chunk_bytes=1 << 12, copy_depth=2, copy_order="srpt")
    reqs2 = make_requests()
    done, got = chunked.run(reqs2, keep={r.rid for r in reqs2}, max_seconds=60)
    assert len(done) == len(reqs2)
    assert chunked.dev_arena.used() == 0 and chunked.host_arena.used() == 0
    for r in reqs2:
        assert np.array_equal(want[r.rid].view(np.uint32), got[r.rid].view(np.uint32))
    # one large and one small request arriving together, in one epoch
    big = max(range(len(SPECS)), key=lambda t: residents[t].nrefine)
    small = min(range(len(SPECS)), key=lambda t: residents[t].nrefine)
    pair = [Request(100, big, "refine", 0.0, 0), Request(101, small, "refine", 0.0, 1)]
    assert chunked.host_floats(pair[0]) > 8 * chunked.host_floats(pair[1])
    done, _ = chunked.run(pair, max_seconds=60)
    t = {r.rid: r.done for r in done}
    assert t[101] <= t[100]


def test_level4_requests_match_the_oracle(residents):
    """Both request kinds on the level-4 topology, served in a mix with the
    others, against the oracle's CpuEvaluator-order sums on the same frame:
    refinement rows, and P / dPdu / dPdv at the topology's coords (patch
    points refined from the same frame first)."""
    import oracle
    from opensubdiv_amd import far, meshes
    t = len(SPECS) - 1
    reqs = make_requests()
    server = Server(residents, L, device_budget=64 << 20, host_budget=64 << 20, slots=3)
    done, mix = server.run(reqs, keep={r.rid for r in reqs}, max_seconds=60)
    assert len(done) == len(reqs)
    h = compile_mesh(SPECS[t], L)
    try:
        mesh = meshes.cube_sphere(SPECS[t].n, seed=SPECS[t].seed, radial_noise=0.01,
                                  crease_fraction=0.05, sharpness_max=4.0)
        topo = far.Topology(mesh)
        topo.refine_uniform(SPECS[t].level)
        ref = topo.stencil_table()
        checked = set()
        for r in reqs:
            if r.topo != t:
                continue
            ctrl = np.ascontiguousarray(residents[t].frames[r.frame % residents[t].frames.shape[0]])
            if r.kind == "refine":
                want = oracle.apply_stencils(ctrl, ref["sizes"], ref["offsets"], ref["indices"],
                                             ref["weights"]).reshape(-1)
            else:
                pts = oracle.apply_stencils(ctrl, h.points["sizes"], h.points["offsets"],
                                            h.points["indices"], h.points["weights"])
                src = np.ascontiguousarray(np.concatenate([ctrl, pts]))
                n = len(h.coords)
                outs = [np.zeros((n, L), np.float32) for _ in range(3)]
                assert oracle.eval_patches(src, (0, L, L), outs, [(0, L, L)] * 3, h.coords,
                                           h.arrays, h.indices, h.params)
                want = np.concatenate([o.reshape(-1) for o in outs])
            assert np.array_equal(mix[r.rid].view(np.uint32), want.view(np.uint32)), \
                f"request {r.rid} {r.kind}"
            checked.add(r.kind)
        assert checked == {"refine", "limit"}
    finally:
        h.release()
