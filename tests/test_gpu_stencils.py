"""GPU parity of HIP EvalStencils (C ABI) against the reference's golden
outputs and the CPU oracle.  The HIP kernels accumulate w*x in table order
without FMA, so the expected result is bit-exact equality."""
import numpy as np
import pytest

import oracle
from conftest import golden
from gpu_util import dev, empty, host, ptr, stream
from opensubdiv_amd import _lib

pytestmark = pytest.mark.gpu
D = _lib.Desc


def eval1(lib, src_t, sd, dst_t, dd, tbl, start=0, end=None):
    end = int(len(tbl["sizes"])) if end is None else end
    return lib.osdhip_eval_stencils(ptr(src_t), D(*sd), ptr(dst_t), D(*dd), ptr(tbl["d_sizes"]),
                                    ptr(tbl["d_offsets"]), ptr(tbl["d_indices"]),
                                    ptr(tbl["d_w"][0]), start, end, stream())


def upload(t, keys=("weights",)):
    out = dict(t)
    out["d_sizes"] = dev(np.asarray(t["sizes"], np.int32))
    out["d_offsets"] = dev(np.asarray(t["offsets"], np.int32))
    out["d_indices"] = dev(np.asarray(t["indices"], np.int32))
    out["d_w"] = [dev(np.asarray(t[k], np.float32)) for k in keys]
    return out


def test_cube_l3_in_place_bit_exact(hip):
    """osd_tutorial_0 call shape: one buffer, src {0,3,3}, dst {nCoarse*3,3,3}."""
    g = golden("cube_l3.npz")
    tbl = upload(g)
    nS = len(g["sizes"])
    buf = np.zeros((8 + nS, 3), np.float32)
    buf[:8] = g["src"]
    b = dev(buf)
    assert eval1(hip, b, (0, 3, 3), b, (8 * 3, 3, 3), tbl) == 0
    out = host(b, shape=(8 + nS, 3))
    np.testing.assert_array_equal(out[8:], g["cpu_out"])
    np.testing.assert_array_equal(out[:8], g["src"])


def test_cubesphere_stride7_and_length3_in_stride7(hip):
    g = golden("cubesphere_l2.npz")
    tbl = upload(g)
    n = len(g["sizes"])
    src = dev(g["src"])
    d7 = empty(n * 7)
    assert eval1(hip, src, (0, 7, 7), d7, (0, 7, 7), tbl) == 0
    np.testing.assert_array_equal(host(d7, shape=(n, 7)), g["cpu_out7"])
    d3 = empty(n * 7, 0.0)
    assert eval1(hip, src, (0, 3, 7), d3, (0, 3, 7), tbl) == 0
    o = host(d3, shape=(n, 7))
    np.testing.assert_array_equal(o[:, :3], g["cpu_out3"])
    assert np.all(o[:, 3:] == 0)        # lanes outside the descriptor untouched


def test_limit_stencils_fused_derivatives(hip):
    g = golden("cubesphere_limit.npz")
    keys = ("weights", "du", "dv", "duu", "duv", "dvv")
    tbl = upload(g, keys)
    n = len(g["sizes"])
    src = dev(g["src"])
    # first derivatives, interleaved P/du/dv in one buffer {0,3,9}/{3,3,9}/{6,3,9}
    inter = empty(n * 9)
    rc = hip.osdhip_eval_stencils_d1(ptr(src), D(0, 3, 3), ptr(inter), D(0, 3, 9), ptr(inter),
                                     D(3, 3, 9), ptr(inter), D(6, 3, 9), ptr(tbl["d_sizes"]),
                                     ptr(tbl["d_offsets"]), ptr(tbl["d_indices"]),
                                     *[ptr(w) for w in tbl["d_w"][:3]], 0, n, stream())
    assert rc == 0
    o = host(inter, shape=(n, 9))
    np.testing.assert_array_equal(o[:, 0:3], g["cpu1_p"])
    np.testing.assert_array_equal(o[:, 3:6], g["cpu1_du"])
    np.testing.assert_array_equal(o[:, 6:9], g["cpu1_dv"])
    outs = [empty(n * 3) for _ in range(6)]
    args = []
    for t in outs:
        args += [ptr(t), D(0, 3, 3)]
    rc = hip.osdhip_eval_stencils_d2(ptr(src), D(0, 3, 3), *args, ptr(tbl["d_sizes"]),
                                     ptr(tbl["d_offsets"]), ptr(tbl["d_indices"]),
                                     *[ptr(w) for w in tbl["d_w"]], 0, n, stream())
    assert rc == 0
    for t, k in zip(outs, ("cpu_p", "cpu_du", "cpu_dv", "cpu_duu", "cpu_duv", "cpu_dvv")):
        np.testing.assert_array_equal(host(t, shape=(n, 3)), g[k], err_msg=k)


def random_table(rng, n, nsrc, max_size, nw=1, empty_frac=0.05):
    sizes = rng.integers(1, max_size + 1, size=n).astype(np.int32)
    sizes[rng.uniform(size=n) < empty_frac] = 0
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
    nnz = int(sizes.sum())
    t = {"sizes": sizes, "offsets": offsets,
         "indices": rng.integers(0, nsrc, size=nnz).astype(np.int32)}
    for k in ("weights", "du", "dv", "duu", "duv", "dvv")[:nw]:
        t[k] = rng.uniform(-1, 1, size=nnz).astype(np.float32)
    return t


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 7, 8, 9, 12, 17])
def test_random_tables_all_lengths_bit_exact(hip, L):
    rng = np.random.default_rng(100 + L)
    nsrc, n = 777, 5000
    t = random_table(rng, n, nsrc, 24)
    tbl = upload(t)
    # src with an element offset and a padded stride; dst with its own layout
    sstride, soff = L + 3, 2
    dstride, doff = L + 1, 5 * (L + 1)
    src = rng.normal(size=(nsrc, sstride)).astype(np.float32)
    s_t = dev(src)
    d_t = empty(doff + n * dstride)
    assert eval1(hip, s_t, (soff, L, sstride), d_t, (doff, L, dstride), tbl) == 0
    got = host(d_t)[doff:].reshape(n, dstride)[:, :L]
    want = np.zeros((n, dstride), np.float32)
    assert oracle.eval_stencils(src, (soff, L, sstride), want, (0, L, dstride), t["sizes"],
                                t["offsets"], t["indices"], t["weights"])
    np.testing.assert_array_equal(got, want[:, :L])


@pytest.mark.parametrize("nw", [3, 6])
@pytest.mark.parametrize("L", [3, 4, 7])
def test_random_derivative_tables_bit_exact(hip, nw, L):
    rng = np.random.default_rng(200 + nw * 10 + L)
    nsrc, n = 500, 3000
    t = random_table(rng, n, nsrc, 20, nw)
    keys = ("weights", "du", "dv", "duu", "duv", "dvv")[:nw]
    tbl = upload(t, keys)
    src = rng.normal(size=(nsrc, L)).astype(np.float32)
    s_t = dev(src)
    outs = [empty(n * L) for _ in range(nw)]
    args = []
    for o in outs:
        args += [ptr(o), D(0, L, L)]
    fn = hip.osdhip_eval_stencils_d1 if nw == 3 else hip.osdhip_eval_stencils_d2
    assert fn(ptr(s_t), D(0, L, L), *args, ptr(tbl["d_sizes"]), ptr(tbl["d_offsets"]),
              ptr(tbl["d_indices"]), *[ptr(w) for w in tbl["d_w"]], 0, n, stream()) == 0
    for k, o in enumerate(outs):
        want = oracle.apply_stencils(src, t["sizes"], t["offsets"], t["indices"], t[keys[k]])
        np.testing.assert_array_equal(host(o, shape=(n, L)), want, err_msg=keys[k])


def test_large_stencils_take_the_unstaged_path(hip):
    """Stencils of up to 200 entries: a wave's table slice exceeds the LDS
    staging capacity and the kernel reads the table directly."""
    rng = np.random.default_rng(7)
    t = random_table(rng, 1000, 300, 200, 3)
    keys = ("weights", "du", "dv")
    tbl = upload(t, keys)
    src = rng.normal(size=(300, 3)).astype(np.float32)
    s_t = dev(src)
    d = empty(1000 * 3)
    assert eval1(hip, s_t, (0, 3, 3), d, (0, 3, 3), tbl) == 0
    want = oracle.apply_stencils(src, t["sizes"], t["offsets"], t["indices"], t["weights"])
    np.testing.assert_array_equal(host(d, shape=(1000, 3)), want)


def test_start_end_subrange_writes_relative_rows(hip):
    rng = np.random.default_rng(9)
    t = random_table(rng, 4000, 100, 16)
    tbl = upload(t)
    src = rng.normal(size=(100, 3)).astype(np.float32)
    s_t = dev(src)
    full = oracle.apply_stencils(src, t["sizes"], t["offsets"], t["indices"], t["weights"])
    for start, end in [(0, 1), (1, 2), (63, 65), (1000, 3999), (130, 4000)]:
        d = empty(4000 * 3)
        assert eval1(hip, s_t, (0, 3, 3), d, (0, 3, 3), tbl, start, end) == 0
        o = host(d, shape=(4000, 3))
        np.testing.assert_array_equal(o[:end - start], full[start:end])
        assert np.all(np.isnan(o[end - start:]))


def test_argument_errors_and_noops(hip):
    rng = np.random.default_rng(11)
    t = random_table(rng, 100, 10, 8)
    tbl = upload(t)
    s_t = dev(rng.normal(size=(10, 3)).astype(np.float32))
    d = empty(300)
    assert eval1(hip, s_t, (0, 3, 3), d, (0, 3, 3), tbl, 5, 5) == 0        # empty range
    assert eval1(hip, s_t, (0, 3, 3), d, (0, 3, 3), tbl, 7, 2) == 0
    assert np.all(np.isnan(host(d)))
    assert eval1(hip, s_t, (0, 3, 3), d, (0, 2, 3), tbl) == _lib.ERR_ARGS  # length mismatch
    assert eval1(hip, None, (0, 3, 3), d, (0, 3, 3), tbl) == _lib.ERR_ARGS
    assert eval1(hip, s_t, (0, 3, 3), None, (0, 3, 3), tbl) == _lib.ERR_ARGS
    assert eval1(hip, s_t, (0, 4, 3), d, (0, 4, 3), tbl) == _lib.ERR_ARGS  # length > stride
    assert np.all(np.isnan(host(d)))


def test_python_mirror_eval_stencils_on_vertex_buffer(hip):
    """HipVertexBuffer / HipStencilTable / HipEvaluator.EvalStencils, as
    Osd::Mesh::Refine calls it (in-place buffer)."""
    from opensubdiv_amd import BufferDescriptor, HipEvaluator, HipStencilTable, HipVertexBuffer
    g = golden("cube_l3.npz")
    nS = len(g["sizes"])
    vb = HipVertexBuffer.Create(3, 8 + nS)
    vb.UpdateData(g["src"], 0, 8)
    table = HipStencilTable.Create({k: g[k] for k in ("sizes", "offsets", "indices", "weights")})
    assert HipEvaluator.EvalStencils(vb, BufferDescriptor(0, 3, 3), vb,
                                     BufferDescriptor(8 * 3, 3, 3), table)
    assert HipEvaluator.Synchronize()
    np.testing.assert_array_equal(vb.numpy()[8:], g["cpu_out"])
    assert not HipEvaluator.EvalStencils(vb, BufferDescriptor(0, 3, 3), vb,
                                         BufferDescriptor(8 * 3, 2, 3), table)


# ---------------------------------------------------------- device tables

def table_eval(hip, h, src, sd, outs, descs, start, end):
    import ctypes
    ptrs = (ctypes.c_void_p * len(outs))(*[ptr(o) for o in outs])
    ds = (D * len(outs))(*[D(*d) for d in descs])
    return hip.osdhip_stencil_table_eval(h, ptr(src), D(*sd), len(outs),
                                         ctypes.cast(ptrs, ctypes.c_void_p),
                                         ctypes.cast(ds, ctypes.c_void_p), start, end, stream())


# (table_layout, slice_group, stencil_kernel, row order): record slices with
# gather groups of 8 and 4, split-stream slices, the CSR kernel on the
# mirror, and locality-ordered slices, rows by dominant vertex or in Far's
# order (random tables: mostly gathers from memory)
VARIANTS = {"rec8": (0, 8, 0, 0), "rec4": (0, 4, 0, 0), "rec16": (0, 16, 0, 0),
            "split": (1, 8, 0, 0), "csr": (0, 8, 1, 0), "tiled8": (2, 8, 0, 0),
            "tiled4": (2, 4, 0, 0), "tiledfar": (2, 8, 0, 1), "pair16": (3, 16, 0, 0),
            "pair4": (3, 4, 0, 0)}


@pytest.mark.parametrize("variant", sorted(VARIANTS))
@pytest.mark.parametrize("nw,L", [(1, 3), (1, 7), (1, 12), (3, 3), (3, 7), (6, 4)])
def test_device_table_bit_exact(hip, variant, nw, L):
    """HipStencilTable device layouts (record / split slices, CSR-staged
    kernel) against the oracle, including sub-ranges that start and end
    inside a 64-row slice, empty rows, and P-only evaluation of a derivative
    table."""
    from opensubdiv_amd import HipStencilTable
    layout, group, kernel, order = VARIANTS[variant]
    rng = np.random.default_rng(300 + 10 * nw + L + 7 * layout + group + kernel)
    nsrc, n = 600, 4099
    t = random_table(rng, n, nsrc, 18, nw)
    keys = ("weights", "du", "dv", "duu", "duv", "dvv")[:nw]
    if layout == 3 and nw != 1:
        # paired records hold one weight array
        with pytest.raises(MemoryError):
            HipStencilTable.Create(t, layout=layout, group=group)
        return
    table = HipStencilTable.Create(t, layout=layout, group=group, kernel=kernel, order=order)
    assert table.options() == {"layout": layout, "keep_csr": True, "group": group,
                               "kernel": kernel, "order": order}
    assert table.GetNumWeightArrays() == nw
    src = rng.normal(size=(nsrc, L)).astype(np.float32)
    s_t = dev(src)
    cases = [(0, n, nw), (37, 2000, nw), (64, 128, nw), (4000, 4099, nw)]
    if nw > 1:
        cases.append((5, 3000, 1))          # P only from a derivative table
    for start, end, nout in cases:
        outs = [empty(n * L) for _ in range(nout)]
        assert table_eval(hip, table.h, s_t, (0, L, L), outs, [(0, L, L)] * nout, start,
                          end) == 0
        for k in range(nout):
            want = oracle.apply_stencils(src, t["sizes"], t["offsets"], t["indices"],
                                         t[keys[k]])[start:end]
            got = host(outs[k], shape=(n, L))
            np.testing.assert_array_equal(got[:end - start], want,
                                          err_msg=f"{keys[k]} [{start},{end})")
            assert np.all(np.isnan(got[end - start:]))


def test_device_table_option_errors(hip):
    """Options are checked at creation and per table: out-of-range layouts,
    groups or kernels are rejected, and a table's evaluation options change
    only for that table."""
    import ctypes
    from opensubdiv_amd import HipStencilTable
    t = {"sizes": np.full(70, 2, np.int32), "offsets": np.arange(0, 140, 2, dtype=np.int32),
         "indices": np.arange(140, dtype=np.int32) % 9,
         "weights": np.linspace(0, 1, 140).astype(np.float32)}
    for bad in (dict(layout=4), dict(layout=-1), dict(group=5), dict(group=32), dict(kernel=2),
                dict(kernel=1, keep_csr=False)):
        with pytest.raises(MemoryError):
            HipStencilTable.Create(t, **bad)
    a = HipStencilTable.Create(t, group=4)
    b = HipStencilTable.Create(t, layout=2, keep_csr=False)
    assert a.set("group", 16) == 4 and a.options()["group"] == 16
    assert b.options() == {"layout": 2, "keep_csr": False, "group": 8, "kernel": 0, "order": 0}
    assert hip.osdhip_stencil_table_set(a.h, b"group", 5) == _lib.ERR_ARGS
    assert hip.osdhip_stencil_table_set(b.h, b"kernel", 1) == _lib.ERR_ARGS   # no CSR kept
    assert hip.osdhip_stencil_table_set(a.h, b"layout", 2) == _lib.ERR_UNSUPPORTED
    assert hip.osdhip_stencil_table_set(None, b"group", 4) == _lib.ERR_ARGS
    o = _lib.TableOptions()
    hip.osdhip_table_options_init(ctypes.byref(o))
    assert (o.layout, o.keepCsr, o.group, o.kernel, o.order) == (0, 1, 8, 0, 0)


def test_device_table_cfg0_and_errors(hip):
    from opensubdiv_amd import HipStencilTable
    g = golden("cube_l3.npz")
    table = HipStencilTable.Create({k: g[k] for k in ("sizes", "offsets", "indices", "weights")})
    n = table.GetNumStencils()
    src = dev(g["src"])
    out = empty(n * 3)
    assert table_eval(hip, table.h, src, (0, 3, 3), [out], [(0, 3, 3)], 0, n) == 0
    np.testing.assert_array_equal(host(out, shape=(n, 3)), g["cpu_out"])
    # more outputs than weight arrays, length mismatch, bad output count
    assert table_eval(hip, table.h, src, (0, 3, 3), [out] * 3, [(0, 3, 3)] * 3, 0, n) == \
        _lib.ERR_ARGS
    assert table_eval(hip, table.h, src, (0, 3, 3), [out], [(0, 2, 3)], 0, n) == _lib.ERR_ARGS
    assert table_eval(hip, table.h, src, (0, 3, 3), [out, out], [(0, 3, 3)] * 2, 0, n) == \
        _lib.ERR_ARGS
    # the raw getters expose the Far layout on the device
    out2 = empty(n * 3)
    assert hip.osdhip_eval_stencils(ptr(src), D(0, 3, 3), ptr(out2), D(0, 3, 3),
                                    table.GetSizesBuffer(), table.GetOffsetsBuffer(),
                                    table.GetIndicesBuffer(), table.GetWeightsBuffer(), 0, n,
                                    stream()) == 0
    np.testing.assert_array_equal(host(out2, shape=(n, 3)), g["cpu_out"])


def _tiled_table(hip, t, group=8, order=0):
    from opensubdiv_amd import HipStencilTable
    return HipStencilTable.Create(t, layout=2, group=group, order=order)


@pytest.mark.parametrize("group,order", [(4, 0), (8, 0), (8, 1)])
def test_tiled_layout_mesh_tables_stage_in_lds(hip, group, order):
    """Locality-ordered slices on cube-sphere tables: with rows by dominant
    vertex every slice of the level-2 refinement table lists its control
    vertices (LDS path), the limit table mixes listed and gathering slices;
    rows in Far's order mix both; P / all six limit outputs, full tables
    and sub-ranges, equal the reference's CpuEvaluator outputs bit for
    bit."""
    g = golden("cubesphere_l2.npz")
    table = _tiled_table(hip, {k: g[k] for k in ("sizes", "offsets", "indices", "weights")},
                         group, order)
    info = table.layout()
    assert info["layout"] == 2 and 0 < info["listed_slices"] <= info["slices"]
    assert order == 1 or info["listed_slices"] == info["slices"]
    n = table.GetNumStencils()
    src = dev(g["src"])
    d7 = empty(n * 7)
    assert table_eval(hip, table.h, src, (0, 7, 7), [d7], [(0, 7, 7)], 0, n) == 0
    np.testing.assert_array_equal(host(d7, shape=(n, 7)), g["cpu_out7"])
    d3 = empty(n * 7, 0.0)
    assert table_eval(hip, table.h, src, (0, 3, 7), [d3], [(0, 3, 7)], 0, n) == 0
    o = host(d3, shape=(n, 7))
    np.testing.assert_array_equal(o[:, :3], g["cpu_out3"])
    assert np.all(o[:, 3:] == 0)
    for start, end in [(0, 1), (63, 65), (100, 1500), (1900, n)]:
        d = empty(n * 7)
        assert table_eval(hip, table.h, src, (0, 7, 7), [d], [(0, 7, 7)], start, end) == 0
        o = host(d, shape=(n, 7))
        np.testing.assert_array_equal(o[:end - start], g["cpu_out7"][start:end])
        assert np.all(np.isnan(o[end - start:]))

    g = golden("cubesphere_limit.npz")
    keys = ("weights", "du", "dv", "duu", "duv", "dvv")
    table = _tiled_table(hip, {k: g[k] for k in ("sizes", "offsets", "indices") + keys},
                         group, order)
    info = table.layout()
    # rows by dominant vertex: both paths in one table
    assert order == 1 or 0 < info["listed_slices"] < info["slices"]
    n = table.GetNumStencils()
    src = dev(g["src"])
    outs = [empty(n * 3) for _ in range(6)]
    assert table_eval(hip, table.h, src, (0, 3, 3), outs, [(0, 3, 3)] * 6, 0, n) == 0
    for k, name in enumerate(("cpu_p", "cpu_du", "cpu_dv", "cpu_duu", "cpu_duv", "cpu_dvv")):
        np.testing.assert_array_equal(host(outs[k], shape=(n, 3)), g[name], err_msg=name)


def _dominant_order_slices(t):
    """Host model of the dominant-vertex row order (stable by row; empty
    rows last): the record entries (64 x each slice's widest row) and the
    slices with at most 63 distinct control vertices."""
    sizes, offs, idx, w = (np.asarray(t[k]) for k in ("sizes", "offsets", "indices", "weights"))
    n = len(sizes)
    dom = np.full(n, 2 ** 32 - 1, np.int64)
    for r in np.nonzero(sizes)[0]:
        o = offs[r]
        dom[r] = idx[o + int(np.argmax(np.abs(w[o:o + sizes[r]])))]
    perm = np.argsort(dom, kind="stable")
    S = (n + 63) // 64
    entries, listed = 0, 0
    for s in range(S):
        rows = perm[64 * s:64 * s + 64]
        entries += 64 * int(sizes[rows].max())
        u = set()
        for r in rows:
            u.update(idx[offs[r]:offs[r] + sizes[r]].tolist())
        listed += len(u) <= 63
    return entries, listed


@pytest.mark.parametrize("n,level", [(12, 2), (8, 3), (6, 4)])
def test_tiled_rows_regrouped_by_size(hip, n, level):
    """Rows by dominant vertex, regrouped by size within 128- or 512-row
    windows when that leaves fewer record entries and no fewer listed
    slices: the device layout never streams more entries or lists fewer
    slices than the plain dominant order (modelled on the host), and the
    level-3 and level-4 refinement tables take a window; outputs equal the
    oracle's CpuEvaluator-order sums bit for bit."""
    from opensubdiv_amd import far, meshes
    m = meshes.cube_sphere(n, seed=3, radial_noise=0.01, crease_fraction=0.05, sharpness_max=4.0)
    topo = far.Topology(m)
    topo.refine_uniform(level)
    t = topo.stencil_table()
    table = _tiled_table(hip, {k: t[k] for k in ("sizes", "offsets", "indices", "weights")})
    info = table.layout()
    entries, listed = _dominant_order_slices(t)
    assert info["size_window"] in (0, 128, 512)
    assert info["record_entries"] <= entries and info["listed_slices"] >= listed
    if info["size_window"] == 0:
        assert info["record_entries"] == entries
    if level >= 3:
        assert info["size_window"] > 0 and info["record_entries"] < entries
    nr = len(t["sizes"])
    rng = np.random.default_rng(40 + level)
    ctrl = rng.standard_normal((m.nverts, 7)).astype(np.float32)
    out = empty(nr * 7)
    assert table_eval(hip, table.h, dev(ctrl.reshape(-1)), (0, 7, 7), [out], [(0, 7, 7)], 0,
                      nr) == 0
    want = oracle.apply_stencils(ctrl, t["sizes"], t["offsets"], t["indices"], t["weights"])
    np.testing.assert_array_equal(host(out, shape=(nr, 7)), want)
    # record slices (layout 0) hold 64 x the widest row of each run of 64 rows
    from opensubdiv_amd import HipStencilTable
    rec = HipStencilTable.Create({k: t[k] for k in ("sizes", "offsets", "indices", "weights")})
    sz = np.zeros(-(-nr // 64) * 64, np.int64)
    sz[:nr] = t["sizes"]
    assert rec.layout()["record_entries"] == 64 * int(sz.reshape(-1, 64).max(axis=1).sum())
    assert rec.layout()["size_window"] == 0
    # Far's order keeps its rows, order 2 the dominant order: no window
    far_order = _tiled_table(hip, {k: t[k] for k in ("sizes", "offsets", "indices", "weights")},
                             order=1)
    assert far_order.layout()["size_window"] == 0
    dom = _tiled_table(hip, {k: t[k] for k in ("sizes", "offsets", "indices", "weights")},
                       order=2)
    info = dom.layout()
    assert info["size_window"] == 0
    assert (info["record_entries"], info["listed_slices"]) == (entries, listed)
    out = empty(nr * 7)
    assert table_eval(hip, dom.h, dev(ctrl.reshape(-1)), (0, 7, 7), [out], [(0, 7, 7)], 0,
                      nr) == 0
    np.testing.assert_array_equal(host(out, shape=(nr, 7)), want)


@pytest.mark.parametrize("nw,L", [(1, 3), (1, 8), (3, 5), (6, 2)])
def test_tiled_layout_random_tables_both_paths(hip, nw, L):
    """Random tables over 40 control vertices (every slice listed, padding
    records on the zero slot) and over 5000 (most slices gather from
    memory, padding records skipped), with empty rows, against the oracle."""
    keys = ("weights", "du", "dv", "duu", "duv", "dvv")[:nw]
    for nsrc, listed in ((40, "all"), (5000, "few")):
        rng = np.random.default_rng(900 + 10 * nw + L + nsrc)
        n = 3001
        t = random_table(rng, n, nsrc, 14, nw)
        table = _tiled_table(hip, t)
        info = table.layout()
        if listed == "all":
            assert info["listed_slices"] == info["slices"]
        else:
            assert info["listed_slices"] < info["slices"] // 2
        src = rng.normal(size=(nsrc, L)).astype(np.float32)
        s_t = dev(src)
        for start, end in [(0, n), (17, 2222)]:
            outs = [empty(n * L) for _ in range(nw)]
            assert table_eval(hip, table.h, s_t, (0, L, L), outs, [(0, L, L)] * nw, start,
                              end) == 0
            for k in range(nw):
                want = oracle.apply_stencils(src, t["sizes"], t["offsets"], t["indices"],
                                             t[keys[k]])[start:end]
                got = host(outs[k], shape=(n, L))
                np.testing.assert_array_equal(got[:end - start], want,
                                              err_msg=f"{keys[k]} nsrc {nsrc} [{start},{end})")
                assert np.all(np.isnan(got[end - start:]))


def test_tiled_layout_edge_tables(hip):
    """Layout-2 builds of edge cases against the oracle: one row, a row
    repeating a control index, fewer rows than a slice, one row past a slice,
    every row empty, and exactly one full slice."""
    cases = []
    rng = np.random.default_rng(77)
    cases.append({"sizes": np.array([3], np.int32), "offsets": np.array([0], np.int32),
                  "indices": np.array([2, 0, 2], np.int32),
                  "weights": np.array([0.25, 0.5, -0.75], np.float32)})
    for n in (5, 63, 64, 65):
        t = random_table(rng, n, 30, 9)
        cases.append(t)
    blank = {"sizes": np.zeros(70, np.int32), "offsets": np.zeros(70, np.int32),
             "indices": np.zeros(0, np.int32), "weights": np.zeros(0, np.float32)}
    cases.append(blank)
    L = 5
    for k, t in enumerate(cases):
        table = _tiled_table(hip, t)
        n = table.GetNumStencils()
        info = table.layout()
        assert info["layout"] == 2 and info["slices"] == (n + 63) // 64, f"case {k}"
        src = rng.normal(size=(40, L)).astype(np.float32)
        s_t = dev(src)
        out = empty(max(1, n) * L)
        assert table_eval(hip, table.h, s_t, (0, L, L), [out], [(0, L, L)], 0, n) == 0
        if n == 0:
            continue
        want = oracle.apply_stencils(src, t["sizes"], t["offsets"], t["indices"], t["weights"])
        np.testing.assert_array_equal(host(out, shape=(n, L)), want, err_msg=f"case {k}")


@pytest.mark.parametrize("layout", [0, 2])
def test_hbr_regression_shapes_on_the_device(hip, layout):
    """The reference's regression shapes (regression/hbr_regression): level-3
    tables (every intermediate level) from the table compiler, evaluated on
    the device in float32 in place after the control points, equal the
    oracle's float32 sums bit for bit and Hbr's baseline positions within
    1e-5 of the bounding-box diagonal (matched one-to-one by nearest
    neighbour: Hbr numbers vertices in another order)."""
    from scipy.spatial import cKDTree
    from opensubdiv_amd import HipStencilTable, far
    g = golden("hbr_catmark.npz")
    names = sorted({k.split("/")[0] for k in g.files})
    for name in names:
        class Shape:
            pass
        m = Shape()
        for k in ("positions", "face_sizes", "face_verts", "crease_pairs", "crease_weights",
                  "corner_verts", "corner_weights"):
            setattr(m, k, g[f"{name}/{k}"])
        m.nverts, m.nfaces = len(m.positions), len(m.face_sizes)
        b, c = (int(x) for x in g[f"{name}/options"])
        t = far.Topology(m, boundary=b, creasing=c)
        t.refine_uniform(3)
        st = t.stencil_table(intermediate=True)
        n = len(st["sizes"])
        ctrl = np.ascontiguousarray(m.positions, np.float32)
        want = oracle.apply_stencils(ctrl, st["sizes"], st["offsets"], st["indices"],
                                     st["weights"])
        buf = np.full((m.nverts + n, 3), np.nan, np.float32)
        buf[:m.nverts] = ctrl
        d = dev(buf)
        h = HipStencilTable.Create(st, layout=layout)
        assert table_eval(hip, h.h, d, (0, 3, 3), [d], [(3 * m.nverts, 3, 3)], 0, n) == 0
        got = host(d, shape=(m.nverts + n, 3))
        np.testing.assert_array_equal(got[m.nverts:], want, err_msg=name)
        for level in ((1,) if name == "catmark_chaikin1" else (1, 2, 3)):
            nv = t.level_info(level)[0]
            lo = m.nverts + sum(t.level_info(k)[0] for k in range(1, level))
            base = g[f"{name}/level{level}"].astype(np.float64)
            diag = np.linalg.norm(base.max(0) - base.min(0))
            dist, nearest = cKDTree(base).query(got[lo:lo + nv].astype(np.float64))
            assert len(np.unique(nearest)) == nv, (name, level)
            assert dist.max() <= 1e-5 * diag, (name, level, dist.max())


@pytest.mark.parametrize("layout", [0, 2])
def test_loop_tables_from_the_compiler_on_the_device(hip, layout):
    """A Loop-scheme uniform table built by the table compiler (a triangulated,
    creased cube-sphere refined to level 3), evaluated on the device in the
    record and locality-ordered layouts, equals the oracle's CpuEvaluator-
    order sums bit for bit."""
    from opensubdiv_amd import HipStencilTable, far, meshes
    q = meshes.cube_sphere(6, seed=4, crease_fraction=0.1)
    tris = []
    for k, f in enumerate(np.asarray(q.face_verts).reshape(-1, 4)):
        a, b, c, d = (int(x) for x in f)
        tris += [(a, b, c), (a, c, d)] if k % 2 == 0 else [(a, b, d), (b, c, d)]

    class TriMesh:
        face_verts = np.asarray(tris, np.int32)
        nfaces = len(tris)
        nverts = q.nverts
        crease_pairs, crease_weights = q.crease_pairs, q.crease_weights
        corner_verts = np.zeros(0, np.int32)
        corner_weights = np.zeros(0, np.float32)

    topo = far.Topology(TriMesh, scheme=far.SCHEME_LOOP)
    topo.refine_uniform(3)
    t = topo.stencil_table()
    n = len(t["sizes"])
    assert n == topo.level_info(3)[0]
    table = HipStencilTable.Create({k: t[k] for k in ("sizes", "offsets", "indices", "weights")},
                                   layout=layout)
    rng = np.random.default_rng(17)
    ctrl = rng.standard_normal((q.nverts, 7)).astype(np.float32)
    out = empty(n * 7)
    assert table_eval(hip, table.h, dev(ctrl.reshape(-1)), (0, 7, 7), [out], [(0, 7, 7)], 0,
                      n) == 0
    want = oracle.apply_stencils(ctrl, t["sizes"], t["offsets"], t["indices"], t["weights"])
    np.testing.assert_array_equal(host(out, shape=(n, 7)), want)
