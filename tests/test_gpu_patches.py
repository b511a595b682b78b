"""GPU parity of HIP EvalPatches (C ABI) against the reference's golden
CpuEvaluator outputs and the CPU oracle: bit-exact (same basis forms and
accumulation order, no FMA)."""
import numpy as np
import pytest

import oracle
from conftest import golden
from gpu_util import dev, empty, host, ptr, stream
from opensubdiv_amd import _lib
from opensubdiv_amd.osd import PATCH_ARRAY_DTYPE, PATCH_COORD_DTYPE, PATCH_PARAM_DTYPE

pytestmark = pytest.mark.gpu
D = _lib.Desc
NODESC = D(0, 0, 0)


def eval_patches(lib, src, sd, outs, descs, n, coords, arrays, indices, params, sorted=False):
    args = []
    for k in range(6):
        if k < len(outs) and outs[k] is not None:
            args += [ptr(outs[k]), D(*descs[k])]
        else:
            args += [None, NODESC]
    fn = lib.osdhip_eval_patches_sorted if sorted else lib.osdhip_eval_patches
    return fn(ptr(src), D(*sd), *args, n, ptr(coords), ptr(arrays), ptr(indices), ptr(params),
              stream())


def golden_patches():
    g = golden("cubesphere_patches.npz")
    return (g, dev(g["src"]), dev(g["coords"]), dev(g["arrays"]), dev(g["indices"]),
            dev(g["params"]))


def test_golden_patches_p_du_dv_bit_exact(hip):
    g, src, coords, arrays, indices, params = golden_patches()
    n = len(g["coords"])
    outs = [empty(n * 3) for _ in range(3)]
    assert eval_patches(hip, src, (0, 3, 3), outs, [(0, 3, 3)] * 3, n, coords, arrays, indices,
                        params) == 0
    for o, k in zip(outs, ("cpu_p", "cpu_du", "cpu_dv")):
        np.testing.assert_array_equal(host(o, shape=(n, 3)), g[k], err_msg=k)


def test_golden_patches_position_only_and_second_derivatives(hip):
    g, src, coords, arrays, indices, params = golden_patches()
    n = len(g["coords"])
    p = empty(n * 3)
    assert eval_patches(hip, src, (0, 3, 3), [p], [(0, 3, 3)], n, coords, arrays, indices,
                        params) == 0
    np.testing.assert_array_equal(host(p, shape=(n, 3)), g["cpu_p_only"])
    outs = [empty(n * 3) for _ in range(6)]
    assert eval_patches(hip, src, (0, 3, 3), outs, [(0, 3, 3)] * 6, n, coords, arrays, indices,
                        params) == 0
    for o, k in zip(outs, ("cpu2_p", "cpu2_du", "cpu2_dv", "cpu2_duu", "cpu2_duv", "cpu2_dvv")):
        np.testing.assert_array_equal(host(o, shape=(n, 3)), g[k], err_msg=k)


def test_golden_patches_interleaved_outputs_and_null_p(hip):
    """P/du/dv interleaved in one buffer (stride 9); and du/dv without P
    (the reference CpuEvaluator crashes on a NULL dst; here it is skipped)."""
    g, src, coords, arrays, indices, params = golden_patches()
    n = len(g["coords"])
    buf = empty(n * 9)
    assert eval_patches(hip, src, (0, 3, 3), [buf, buf, buf], [(0, 3, 9), (3, 3, 9), (6, 3, 9)],
                        n, coords, arrays, indices, params) == 0
    o = host(buf, shape=(n, 9))
    np.testing.assert_array_equal(o[:, :3], g["cpu_p"])
    np.testing.assert_array_equal(o[:, 3:6], g["cpu_du"])
    np.testing.assert_array_equal(o[:, 6:], g["cpu_dv"])
    du, dv = empty(n * 3), empty(n * 3)
    assert eval_patches(hip, src, (0, 3, 3), [None, du, dv], [None, (0, 3, 3), (0, 3, 3)], n,
                        coords, arrays, indices, params) == 0
    np.testing.assert_array_equal(host(du, shape=(n, 3)), g["cpu_du"])
    np.testing.assert_array_equal(host(dv, shape=(n, 3)), g["cpu_dv"])


def synthetic_patch_table(rng, nsrc, per_array=400):
    """Random arrays of every supported type (plus a mixed regular/Gregory
    array), random params (depth 0-4, u/v, boundary masks), random CVs."""
    specs = [(6, 6), (9, 9), (3, 3), (6, 9)]
    arrays = np.zeros(len(specs), PATCH_ARRAY_DTYPE)
    params, indices = [], []
    base = prim = 0
    for a, (reg, irr) in enumerate(specs):
        stride = {6: 16, 9: 20, 3: 4}[irr] if reg == irr else 20
        arrays[a] = (reg, irr, per_array, base, stride, prim)
        for _ in range(per_array):
            depth = int(rng.integers(0, 5))
            u = int(rng.integers(0, 1 << depth))
            v = int(rng.integers(0, 1 << depth))
            regular = 1 if reg == 6 and (irr == 6 or rng.uniform() < 0.5) else 0
            boundary = int(rng.integers(0, 16)) if regular and rng.uniform() < 0.5 else 0
            f1 = (u << 22) | (v << 12) | (boundary << 7) | (regular << 5) | depth
            params.append((int(rng.integers(0, 1 << 20)), f1, 0.0))
            indices.append(rng.integers(0, nsrc, size=stride))
        base += per_array * stride
        prim += per_array
    params = np.array(params, dtype=PATCH_PARAM_DTYPE)
    return arrays, np.concatenate(indices).astype(np.int32), params


def synthetic_coords(rng, arrays, params, n):
    coords = np.zeros(n, PATCH_COORD_DTYPE)
    a = rng.integers(0, len(arrays), size=n)
    for i in range(n):
        arr = arrays[a[i]]
        pi = arr["primitiveIdBase"] + int(rng.integers(0, arr["numPatches"]))
        f1 = int(params[pi]["field1"])
        depth, u, v = f1 & 0xf, (f1 >> 22) & 0x3ff, (f1 >> 12) & 0x3ff
        s = (u + rng.uniform()) / (1 << depth)
        t = (v + rng.uniform()) / (1 << depth)
        coords[i] = (a[i], pi, 0, s, t)
    return coords


@pytest.mark.parametrize("L", [1, 3, 4, 7, 8, 11])
@pytest.mark.parametrize("deriv", [0, 1, 2])
def test_synthetic_tables_all_types_bit_exact(hip, L, deriv):
    rng = np.random.default_rng(1000 + 10 * L + deriv)
    nsrc, n = 900, 3000
    arrays, indices, params = synthetic_patch_table(rng, nsrc)
    coords = synthetic_coords(rng, arrays, params, n)
    sstride, soff = L + 2, 1
    src = rng.normal(size=(nsrc, sstride)).astype(np.float32)
    nout = {0: 1, 1: 3, 2: 6}[deriv]
    want = [np.zeros((n, L), np.float32) for _ in range(nout)]
    assert oracle.eval_patches(src, (soff, L, sstride), want, [(0, L, L)] * nout, coords, arrays,
                               indices, params)
    outs = [empty(n * (L + 1)) for _ in range(nout)]
    assert eval_patches(hip, dev(src), (soff, L, sstride), outs, [(0, L, L + 1)] * nout, n,
                        dev(coords), dev(arrays), dev(indices), dev(params)) == 0
    for k in range(nout):
        got = host(outs[k], shape=(n, L + 1))[:, :L]
        np.testing.assert_array_equal(got, want[k], err_msg=f"output {k}")


def test_unsupported_patch_type_gives_zeros(hip):
    """Types OsdEvaluatePatchBasis does not evaluate (POINTS, LINES, the
    legacy GREGORY / GREGORY_BOUNDARY): zero outputs, as CpuEvaluator's
    Clear() with no point added."""
    rng = np.random.default_rng(5)
    for ptype in (1, 2, 7, 8):
        arrays = np.array([(ptype, ptype, 4, 0, 16, 0)], PATCH_ARRAY_DTYPE)
        params = np.array([(0, 0, 0.0)] * 4, PATCH_PARAM_DTYPE)
        indices = np.zeros(64, np.int32)
        coords = np.array([(0, i, 0, 0.5, 0.5) for i in range(4)], PATCH_COORD_DTYPE)
        src = rng.normal(size=(2, 3)).astype(np.float32)
        p = dev(np.full(12, np.nan, np.float32))
        assert eval_patches(hip, dev(src), (0, 3, 3), [p], [(0, 3, 3)], 4, dev(coords),
                            dev(arrays), dev(indices), dev(params)) == 0
        assert np.all(host(p) == 0), ptype


LOOP_CASES = ("sphere", "grid", "strip", "uniform")


@pytest.mark.parametrize("name", LOOP_CASES)
def test_golden_loop_patches_bit_exact(hip, name):
    """Loop-scheme tables -- LOOP box-spline patches with boundary masks,
    GREGORY_TRIANGLE end caps, uniform TRIANGLES, rotated sub-triangles --
    against the reference CpuEvaluator's outputs, P only, P/du/dv and the
    2nd-derivative arity, and against the oracle with an interleaved
    stride-9 output."""
    g = golden("loop_patches.npz")
    src, coords, arrays = dev(g[name + "_src"]), dev(g[name + "_coords"]), dev(g[name + "_arrays"])
    indices, params = dev(g[name + "_indices"]), dev(g[name + "_params"])
    n = len(g[name + "_coords"])
    outs = [empty(n * 3) for _ in range(3)]
    assert eval_patches(hip, src, (0, 3, 3), outs, [(0, 3, 3)] * 3, n, coords, arrays, indices,
                        params) == 0
    for o, k in zip(outs, ("cpu_p", "cpu_du", "cpu_dv")):
        np.testing.assert_array_equal(host(o, shape=(n, 3)), g[f"{name}_{k}"], err_msg=k)
    outs = [empty(n * 3) for _ in range(6)]
    assert eval_patches(hip, src, (0, 3, 3), outs, [(0, 3, 3)] * 6, n, coords, arrays, indices,
                        params) == 0
    for o, k in zip(outs, ("cpu2_p", "cpu2_du", "cpu2_dv", "cpu2_duu", "cpu2_duv", "cpu2_dvv")):
        np.testing.assert_array_equal(host(o, shape=(n, 3)), g[f"{name}_{k}"], err_msg=k)
    p = empty(n * 3)
    assert eval_patches(hip, src, (0, 3, 3), [p], [(0, 3, 3)], n, coords, arrays, indices,
                        params) == 0
    np.testing.assert_array_equal(host(p, shape=(n, 3)), g[name + "_cpu_p"])
    inter = empty(n * 9)
    assert eval_patches(hip, src, (0, 3, 3), [inter, inter, inter],
                        [(0, 3, 9), (3, 3, 9), (6, 3, 9)], n, coords, arrays, indices,
                        params) == 0
    got = host(inter, shape=(n, 9))
    for j, k in enumerate(("cpu_p", "cpu_du", "cpu_dv")):
        np.testing.assert_array_equal(got[:, 3 * j:3 * j + 3], g[f"{name}_{k}"], err_msg=k)


@pytest.mark.parametrize("name", LOOP_CASES)
def test_golden_loop_patches_in_a_batch(hip, name):
    """The batched engine's phase B on Loop-scheme tables (no regular
    Catmark patches, so every coord takes the generic batch kernel), two
    frames (the second the first scaled by 2, a power of two, so its
    results are the first's scaled exactly): same bits as the reference
    CpuEvaluator's, run directly and as a graph launch."""
    import torch
    from opensubdiv_amd.batch import Batch
    g = golden("loop_patches.npz")
    src1 = g[name + "_src"]
    nsrc, n = len(src1), len(g[name + "_coords"])
    src = dev(np.stack([src1, 2 * src1], axis=1))              # (v, f, k), frames interleaved
    outs = [empty(2 * n * 3) for _ in range(3)]
    seg = dict(n=n, coords=dev(g[name + "_coords"]), arrays=dev(g[name + "_arrays"]),
               indices=dev(g[name + "_indices"]), params=dev(g[name + "_params"]), src=src,
               src_v=6, src_f=3, P=outs[0], dPdu=outs[1], dPdv=outs[2], out_v=3, out_f=n * 3,
               frames=2)
    b = Batch(3, [], [seg])
    assert nsrc > 0
    for run in (b.run, b.launch):
        for o in outs:
            o.fill_(float("nan"))
        run()
        torch.cuda.synchronize()
        for o, k in zip(outs, ("cpu_p", "cpu_du", "cpu_dv")):
            got = host(o, shape=(2, n, 3))
            np.testing.assert_array_equal(got[0], g[f"{name}_{k}"], err_msg=k)
            np.testing.assert_array_equal(got[1], 2 * g[f"{name}_{k}"], err_msg=k)
    b.close()


def test_patch_argument_errors(hip):
    g, src, coords, arrays, indices, params = golden_patches()
    p = empty(30)
    assert eval_patches(hip, None, (0, 3, 3), [p], [(0, 3, 3)], 10, coords, arrays, indices,
                        params) == _lib.ERR_ARGS
    assert eval_patches(hip, src, (0, 3, 3), [p], [(0, 2, 3)], 10, coords, arrays, indices,
                        params) == _lib.ERR_ARGS
    assert eval_patches(hip, src, (0, 3, 3), [], [], 10, coords, arrays, indices,
                        params) == _lib.ERR_ARGS
    assert eval_patches(hip, src, (0, 3, 3), [p], [(0, 3, 3)], 0, coords, arrays, indices,
                        params) == 0
    assert np.all(np.isnan(host(p)))


def test_python_mirror_eval_patches(hip):
    from opensubdiv_amd import BufferDescriptor, HipEvaluator, HipPatchTable
    g = golden("cubesphere_patches.npz")
    n = len(g["coords"])
    table = HipPatchTable.Create({"arrays": g["arrays"].view(PATCH_ARRAY_DTYPE).reshape(-1),
                                  "indices": g["indices"],
                                  "params": g["params"].view(PATCH_PARAM_DTYPE).reshape(-1)})
    src, coords = dev(g["src"]), dev(g["coords"])
    P, du, dv = empty(n * 3), empty(n * 3), empty(n * 3)
    d = BufferDescriptor(0, 3, 3)
    assert HipEvaluator.EvalPatches(src, d, P, d, du, d, dv, d, n, coords, table)
    assert HipEvaluator.Synchronize()
    np.testing.assert_array_equal(host(P, shape=(n, 3)), g["cpu_p"])
    np.testing.assert_array_equal(host(dv, shape=(n, 3)), g["cpu_dv"])
    # a face-varying channel holding the same arrays evaluates identically;
    # a channel that does not exist returns False
    arrays = g["arrays"].view(PATCH_ARRAY_DTYPE).reshape(-1)
    params = g["params"].view(PATCH_PARAM_DTYPE).reshape(-1)
    ft = HipPatchTable(arrays, g["indices"], params, fvar_channels=[(arrays, g["indices"], params)])
    assert ft.GetNumFVarChannels() == 1
    P2 = empty(n * 3)
    assert HipEvaluator.EvalPatchesFaceVarying(src, d, P2, d, n, coords, ft, 0)
    assert HipEvaluator.Synchronize()
    np.testing.assert_array_equal(host(P2, shape=(n, 3)), g["cpu_p"])
    assert not HipEvaluator.EvalPatchesFaceVarying(src, d, P2, d, n, coords, ft, 1)


@pytest.mark.parametrize("name", ["linear", "smooth"])
def test_golden_face_varying_channel_bit_exact(hip, name):
    """A real Far face-varying UV channel on the cube-sphere, end to end on
    the device as glEvalLimit runs it: the face-varying stencil table refines
    the UVs in a HipVertexBuffer (EvalStencils), then EvalPatchesFaceVarying
    evaluates the channel at PatchMap coords -- both bit-exact with the
    reference CpuEvaluator; linear QUADS patches (Far's defaults) and smooth
    B-spline / Gregory patches with face-varying local points."""
    from opensubdiv_amd import (BufferDescriptor, HipEvaluator, HipPatchTable, HipStencilTable,
                                HipVertexBuffer)
    g = golden("cubesphere_fvar.npz")
    nb, nt = len(g["uv"]), len(g[name + "_buffer"])
    vb = HipVertexBuffer.Create(2, nt)
    vb.UpdateData(g["uv"], 0, nb)
    st = HipStencilTable.Create({"sizes": g[name + "_st_sizes"], "offsets": g[name + "_st_offsets"],
                                 "indices": g[name + "_st_indices"],
                                 "weights": g[name + "_st_weights"]})
    assert HipEvaluator.EvalStencils(vb, BufferDescriptor(0, 2, 2), vb,
                                     BufferDescriptor(nb * 2, 2, 2), st)
    assert HipEvaluator.Synchronize()
    np.testing.assert_array_equal(vb.numpy(), g[name + "_buffer"])

    def view(key, dtype):
        return np.ascontiguousarray(g[key]).view(dtype).reshape(-1)

    table = HipPatchTable(view(name + "_vertex_arrays", PATCH_ARRAY_DTYPE),
                          g[name + "_vertex_indices"],
                          view(name + "_vertex_params", PATCH_PARAM_DTYPE),
                          fvar_channels=[(view(name + "_arrays", PATCH_ARRAY_DTYPE),
                                          g[name + "_indices"],
                                          view(name + "_params", PATCH_PARAM_DTYPE))])
    n = len(g[name + "_coords"])
    coords = dev(g[name + "_coords"])
    d = BufferDescriptor(0, 2, 2)
    outs = [empty(n * 2) for _ in range(6)]
    assert HipEvaluator.EvalPatchesFaceVarying(vb, d, outs[0], d, outs[1], d, outs[2], d,
                                               n, coords, table, 0)
    assert HipEvaluator.Synchronize()
    for o, k in zip(outs[:3], ("cpu_p", "cpu_du", "cpu_dv")):
        np.testing.assert_array_equal(host(o, shape=(n, 2)), g[name + "_" + k], err_msg=k)
    args = [x for o in outs[1:] for x in (o, d)]
    assert HipEvaluator.EvalPatchesFaceVarying(vb, d, outs[0], d, *args, n, coords, table, 0)
    assert HipEvaluator.Synchronize()
    for o, k in zip(outs, ("cpu2_p", "cpu2_du", "cpu2_dv", "cpu2_duu", "cpu2_duv", "cpu2_dvv")):
        np.testing.assert_array_equal(host(o, shape=(n, 2)), g[name + "_" + k], err_msg=k)


@pytest.mark.parametrize("deriv", [0, 1, 2])
def test_patch_sorted_order_bit_exact(hip, deriv):
    """Patch-ordered evaluation (osdhip_eval_patches_sorted: the coords
    binned by patch) returns every coord's result in its own row,
    bit-identical to caller order and to the oracle, on a small and a large
    set (coords drawn with repeats and in random order)."""
    g, src, _, arrays, indices, params = golden_patches()
    rng = np.random.default_rng(77 + deriv)
    base = g["coords"]
    nout = {0: 1, 1: 3, 2: 6}[deriv]
    for n in (1000, 70000):
        coords = base[rng.integers(0, len(base), size=n)]
        want = [np.zeros((n, 3), np.float32) for _ in range(nout)]
        assert oracle.eval_patches(g["src"], (0, 3, 3), want, [(0, 3, 3)] * nout, coords,
                                   g["arrays"], g["indices"], g["params"])
        dc = dev(coords)
        res = {}
        for srt in (True, False):
            outs = [empty(n * 3) for _ in range(nout)]
            assert eval_patches(hip, src, (0, 3, 3), outs, [(0, 3, 3)] * nout, n, dc,
                                arrays, indices, params, sorted=srt) == 0
            res[srt] = [host(o, shape=(n, 3)) for o in outs]
        for k in range(nout):
            np.testing.assert_array_equal(res[True][k], want[k], err_msg=f"sorted {n} {k}")
            np.testing.assert_array_equal(res[False][k], want[k], err_msg=f"caller {n} {k}")


def test_patch_sorted_order_beyond_the_bucket_count(hip):
    """More patches than the binning's 2^20 (patch p and p + 2^20 share a
    bin), coords drawn with repeats: still every coord's own result,
    bit-identical to caller order and to the oracle."""
    rng = np.random.default_rng(91)
    npatch, nsrc, n = (1 << 20) + 4096, 5000, 60000
    arrays = np.array([(3, 3, npatch, 0, 4, 0)], PATCH_ARRAY_DTYPE)
    params = np.zeros(npatch, PATCH_PARAM_DTYPE)                 # depth 0, not regular
    indices = rng.integers(0, nsrc, size=4 * npatch).astype(np.int32)
    pick = np.concatenate([rng.integers(0, npatch, size=n // 2),
                           (1 << 20) + rng.integers(0, 4096, size=n // 4),
                           rng.integers(0, 4096, size=n - n // 2 - n // 4)])
    coords = np.zeros(n, PATCH_COORD_DTYPE)
    coords["patchIndex"] = rng.permutation(pick)
    coords["s"], coords["t"] = rng.uniform(size=n), rng.uniform(size=n)
    src = rng.normal(size=(nsrc, 3)).astype(np.float32)
    want = [np.zeros((n, 3), np.float32) for _ in range(3)]
    assert oracle.eval_patches(src, (0, 3, 3), want, [(0, 3, 3)] * 3, coords, arrays, indices,
                               params)
    d = [dev(x) for x in (src, coords, arrays, indices, params)]
    for srt in (True, False):
        outs = [empty(n * 3) for _ in range(3)]
        assert eval_patches(hip, d[0], (0, 3, 3), outs, [(0, 3, 3)] * 3, n, *d[1:],
                            sorted=srt) == 0
        for k in range(3):
            np.testing.assert_array_equal(host(outs[k], shape=(n, 3)), want[k],
                                          err_msg=f"sorted={srt} {k}")


def normals(lib, du, dd, dv, vd, n, nd, count, normalize):
    return lib.osdhip_limit_normals(ptr(du), D(*dd), ptr(dv), D(*vd), ptr(n), D(*nd), count,
                                    int(normalize), stream())


@pytest.mark.parametrize("normalize", [False, True])
def test_limit_normals_after_golden_patches(hip, normalize):
    """EvalPatches on the golden cube-sphere coords, interleaved P/du/dv in one
    buffer, then limit normals into a padded buffer: bit-exact against the
    oracle's restatement of the reference tutorial's cross / normalize applied
    to the reference's own dPdu / dPdv."""
    g, src, coords, arrays, indices, params = golden_patches()
    n = len(g["coords"])
    inter = empty(n * 9)
    assert eval_patches(hip, src, (0, 3, 3), [inter, inter, inter], [(0, 3, 9), (3, 3, 9),
                                                                     (6, 3, 9)],
                        n, coords, arrays, indices, params) == 0
    nrm = empty(n * 4 + 1)
    assert normals(hip, inter, (3, 3, 9), inter, (6, 3, 9), nrm, (1, 3, 4), n, normalize) == 0
    want = np.full((n, 3), np.nan, np.float32)
    du, dv = np.ascontiguousarray(g["cpu_du"]), np.ascontiguousarray(g["cpu_dv"])
    assert oracle.limit_normals(du, (0, 3, 3), dv, (0, 3, 3), want, (0, 3, 3), n, normalize)
    got = host(nrm)[1:].reshape(n, 4)
    np.testing.assert_array_equal(got[:, :3], want)
    assert np.all(np.isnan(got[:, 3]))                    # the padding element is untouched


def test_limit_normals_random_and_errors(hip):
    """Random tangents with offsets and strides, normalised and not, with a
    zero normal (0 * (1 / sqrtf(0)) = NaN, as the reference tutorial's
    normalize gives); argument errors; the Python mirror."""
    rng = np.random.default_rng(5)
    n = 10007
    du = rng.normal(size=(n, 5)).astype(np.float32)
    dv = rng.normal(size=(n, 4)).astype(np.float32)
    du[17] = 0.0
    d_du, d_dv = dev(du), dev(dv)
    for normalize in (False, True):
        out = empty(n * 3)
        assert normals(hip, d_du, (1, 3, 5), d_dv, (0, 3, 4), out, (0, 3, 3), n, normalize) == 0
        want = np.zeros((n, 3), np.float32)
        assert oracle.limit_normals(du, (1, 3, 5), dv, (0, 3, 4), want, (0, 3, 3), n, normalize)
        np.testing.assert_array_equal(host(out, shape=(n, 3)), want)
        assert np.isnan(want[17]).all() == normalize
    out = empty(n * 3)
    assert normals(hip, d_du, (0, 2, 5), d_dv, (0, 3, 4), out, (0, 3, 3), n, True) == \
        _lib.ERR_ARGS
    assert normals(hip, None, (0, 3, 5), d_dv, (0, 3, 4), out, (0, 3, 3), n, True) == \
        _lib.ERR_ARGS
    assert normals(hip, d_du, (0, 3, 5), d_dv, (0, 3, 4), out, (0, 3, 3), 0, True) == 0
    assert np.isnan(host(out)).all()                      # count 0 writes nothing
    from opensubdiv_amd.osd import BufferDescriptor as BD, limit_normals
    assert limit_normals(d_du, BD(1, 3, 5), d_dv, BD(0, 3, 4), out, BD(0, 3, 3), n)
    np.testing.assert_array_equal(host(out, shape=(n, 3)), want)
    assert not limit_normals(d_du, BD(1, 3, 5), d_dv, BD(0, 2, 4), out, BD(0, 3, 3), n)


def test_invalid_phantom_indices_read_no_memory_outside_the_buffer(hip):
    """Boundary patches whose phantom points carry INDEX_INVALID (-1), as a
    PatchBuilder without fillMissingBoundaryPoints leaves them: the fold
    gives those points zero weight, and the kernel reads point 0 in their
    place, so the results equal the oracle's on the same table with any
    valid fill (here the random indices)."""
    rng = np.random.default_rng(77)
    nsrc, n, L = 900, 4000, 3
    arrays, indices, params = synthetic_patch_table(rng, nsrc)
    holes = indices.copy()
    phantom = {1: [0, 1, 2, 3], 2: [3, 7, 11, 15], 4: [12, 13, 14, 15], 8: [0, 4, 8, 12]}
    count = 0
    for arr in arrays:
        if arr["regDesc"] != 6:
            continue
        for p in range(arr["numPatches"]):
            pi = arr["primitiveIdBase"] + p
            f1 = int(params[pi]["field1"])
            if not (f1 >> 5) & 1:
                continue
            base = arr["indexBase"] + arr["stride"] * p
            for bit, pts in phantom.items():
                if (f1 >> 7) & bit:
                    holes[base + np.array(pts)] = -1
                    count += 1
    assert count > 100
    coords = synthetic_coords(rng, arrays, params, n)
    src = rng.normal(size=(nsrc, L)).astype(np.float32)
    want = [np.zeros((n, L), np.float32) for _ in range(3)]
    assert oracle.eval_patches(src, (0, L, L), want, [(0, L, L)] * 3, coords, arrays, indices,
                               params)
    outs = [empty(n * L) for _ in range(3)]
    assert eval_patches(hip, dev(src), (0, L, L), outs, [(0, L, L)] * 3, n, dev(coords),
                        dev(arrays), dev(holes), dev(params)) == 0
    for k in range(3):
        np.testing.assert_array_equal(host(outs[k], shape=(n, L)), want[k], err_msg=f"output {k}")


def _queries(rng, nptex, n):
    f = rng.integers(0, nptex, n).astype(np.int32)
    u, v = rng.random(n), rng.random(n)
    # medians of every depth (upper-half rule), the unit square's edges,
    # points just outside it, NaN and faces outside the map
    grid = np.array([0.0, 1.0, 0.5, 0.25, 0.75, 0.125, 0.0625, 1 - 2 ** -30, 2 ** -30])
    k = len(grid)
    u[:k * k], v[:k * k] = np.repeat(grid, k), np.tile(grid, k)
    u[k * k:k * k + 4] = [-1e-12, 1 + 1e-12, np.nan, 0.5]
    v[k * k:k * k + 4] = [0.5, 0.5, 0.5, -0.0]
    f[k * k + 4:k * k + 8] = [-1, nptex, nptex + 100, 2 ** 30]
    return f, u, v


def test_device_patch_map_matches_far_patch_map_records(hip):
    """Far::PatchMap's own records for the golden cube-sphere queries."""
    from opensubdiv_amd.osd import HipPatchMap
    g = golden("cubesphere_patches.npz")
    pm = HipPatchMap(g["arrays"].view(PATCH_ARRAY_DTYPE).reshape(-1),
                     g["params"].view(PATCH_PARAM_DTYPE).reshape(-1))
    n = len(g["ptex"])
    out = dev(np.full(n, -7, PATCH_COORD_DTYPE))
    assert pm.FindPatches(n, dev(g["ptex"].astype(np.int32)), dev(g["u"].astype(np.float64)),
                          dev(g["v"].astype(np.float64)), out)
    got = host(out, PATCH_COORD_DTYPE)
    want = g["coords"].view(PATCH_COORD_DTYPE).reshape(-1)
    for k in PATCH_COORD_DTYPE.names:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    pm.close()


@pytest.mark.parametrize("mesh", ["creased-sphere", "pentagonal-prism"])
def test_device_patch_map_matches_host_map(hip, mesh):
    """Tables of the table compiler (creased adaptive levels; an n-gon mesh
    whose sub-face patches are rooted at depth 1): the device map gives the
    host map's records for random queries, every median, the square's
    edges and points outside it, NaN and faces outside the map."""
    from opensubdiv_amd import far, meshes
    from opensubdiv_amd.osd import HipPatchMap
    if mesh == "creased-sphere":
        m = meshes.cube_sphere(5, seed=3, crease_fraction=0.2)
        nptex = m.nfaces
    else:
        from test_far_compiler import pentagonal_prism
        m = pentagonal_prism()
        nptex = int(sum(1 if k == 4 else k for k in m.face_sizes))
    pt = far.Topology(m).patch_table(3)
    rng = np.random.default_rng(21)
    f, u, v = _queries(rng, nptex, 20000)
    want = pt.find(f, u, v)
    pm = HipPatchMap(pt.arrays, pt.params)
    out = dev(np.zeros(len(f), PATCH_COORD_DTYPE))
    assert pm.FindPatches(len(f), dev(f), dev(u), dev(v), out)
    got = host(out, PATCH_COORD_DTYPE)
    for k in PATCH_COORD_DTYPE.names:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    assert (got["arrayIndex"] >= 0).sum() > 19000 and (got["arrayIndex"] == -1).sum() >= 7
    # argument errors; an empty query set is a no-op
    assert hip.osdhip_patch_map_find(pm.h, 0, None, None, None, None, stream()) == 0
    assert hip.osdhip_patch_map_find(pm.h, 5, None, ptr(out), ptr(out), ptr(out), stream()) == \
        _lib.ERR_ARGS
    assert hip.osdhip_patch_map_find(None, 5, ptr(out), ptr(out), ptr(out), ptr(out), stream()) \
        == _lib.ERR_ARGS
    pm.close()
    with pytest.raises(ValueError):
        HipPatchMap(np.zeros(0, PATCH_ARRAY_DTYPE), pt.params)
    bad = pt.arrays.copy()
    bad["numPatches"][0] += 1                      # arrays do not cover the params
    with pytest.raises(ValueError):
        HipPatchMap(bad, pt.params)


def test_device_tessellation_matches_bfr_and_feeds_eval_patches(hip):
    """Bfr's uniform quad pattern on the device for every ptex face of a
    creased cube-sphere at several rates: (u, v) bit-identical to the
    oracle's Bfr restatement, PatchCoords identical to the host patch map's
    for the same points, and EvalPatches on them identical to the oracle."""
    from opensubdiv_amd import far, meshes
    from opensubdiv_amd.osd import HipPatchMap
    m = meshes.cube_sphere(4, seed=5, crease_fraction=0.2)
    pt = far.Topology(m).patch_table(3)
    pm = HipPatchMap(pt.arrays, pt.params)
    faces = np.arange(m.nfaces, dtype=np.int32)[::-1].copy()      # any face order
    d_faces = dev(faces)
    for rate in (1, 2, 3, 8, 17):
        per = (rate + 1) ** 2
        n = len(faces) * per
        uv = empty(2 * n)
        coords = dev(np.zeros(n, PATCH_COORD_DTYPE))
        assert hip.osdhip_tess_uniform_quads(len(faces), ptr(d_faces), rate, ptr(uv), pm.h,
                                             ptr(coords), stream()) == 0
        want_uv = np.tile(oracle.tess_uniform_quad(rate), (len(faces), 1))
        got_uv = host(uv, shape=(n, 2))
        assert got_uv.tobytes() == want_uv.tobytes(), rate
        want = pt.find(np.repeat(faces, per), want_uv[:, 0].astype(np.float64),
                       want_uv[:, 1].astype(np.float64))
        got = host(coords, PATCH_COORD_DTYPE)
        assert got.tobytes() == want.tobytes(), rate
    # the last rate's coords straight into EvalPatches
    st = pt.point_stencils()
    ctrl = m.positions.astype(np.float32)
    pts = oracle.apply_stencils(ctrl, st["sizes"], st["offsets"], st["indices"], st["weights"])
    src = np.ascontiguousarray(np.concatenate([ctrl, pts]))
    outs = [empty(n * 3) for _ in range(3)]
    assert eval_patches(hip, dev(src), (0, 3, 3), outs, [(0, 3, 3)] * 3, n, coords,
                        dev(pt.arrays), dev(pt.indices), dev(pt.params)) == 0
    ref_out = [np.zeros((n, 3), np.float32) for _ in range(3)]
    assert oracle.eval_patches(src, (0, 3, 3), ref_out, [(0, 3, 3)] * 3, want, pt.arrays,
                               pt.indices, pt.params)
    for k in range(3):
        np.testing.assert_array_equal(host(outs[k], shape=(n, 3)), ref_out[k])
    # uv only (no map); argument errors
    assert hip.osdhip_tess_uniform_quads(len(faces), None, 4, ptr(uv), None, None, stream()) == 0
    assert hip.osdhip_tess_uniform_quads(len(faces), ptr(d_faces), 0, ptr(uv), None, None,
                                         stream()) == _lib.ERR_ARGS
    assert hip.osdhip_tess_uniform_quads(len(faces), ptr(d_faces), 4, None, None, None,
                                         stream()) == _lib.ERR_ARGS
    assert hip.osdhip_tess_uniform_quads(len(faces), None, 4, None, pm.h, ptr(coords),
                                         stream()) == _lib.ERR_ARGS
    assert hip.osdhip_tess_uniform_quads(len(faces), ptr(d_faces), 4, None, None, ptr(coords),
                                         stream()) == _lib.ERR_ARGS
    assert hip.osdhip_tess_uniform_quads(0, None, 4, None, None, None, stream()) == 0
    pm.close()


@pytest.mark.parametrize("mesh", ["grid", "disc5", "fan4"])
def test_boundary_patch_tables_bit_exact(hip, mesh):
    """Adaptive tables of open meshes (boundary B-spline patches with Far's
    boundary masks and filled phantom points, interior end caps): P, dPdu,
    dPdv at random and corner locations equal the oracle's bit for bit."""
    from opensubdiv_amd import far
    from test_far_compiler import PolyMesh, disc, open_grid
    m = {"grid": lambda: open_grid(5, creases=[(7, 8, 2.0)]), "disc5": lambda: disc(5),
         "fan4": lambda: PolyMesh([[0, 1, 2, 3], [0, 3, 4, 5], [0, 5, 6, 7], [0, 7, 8, 9]])}[mesh]()
    pt = far.Topology(m, boundary=2).patch_table(3)
    rng = np.random.default_rng(3)
    n = 4000
    f = rng.integers(0, m.nfaces, n).astype(np.int32)
    u, v = rng.random(n), rng.random(n)
    u[:4], v[:4] = [0.0, 1.0, 1.0, 0.0], [0.0, 0.0, 1.0, 1.0]
    coords = pt.find(f, u, v)
    assert (coords["arrayIndex"] >= 0).all()
    st = pt.point_stencils()
    ctrl = m.positions.astype(np.float32)
    pts = oracle.apply_stencils(ctrl, st["sizes"], st["offsets"], st["indices"], st["weights"])
    src = np.ascontiguousarray(np.concatenate([ctrl, pts]))
    want = [np.zeros((n, 3), np.float32) for _ in range(3)]
    assert oracle.eval_patches(src, (0, 3, 3), want, [(0, 3, 3)] * 3, coords, pt.arrays,
                               pt.indices, pt.params)
    outs = [empty(n * 3) for _ in range(3)]
    assert eval_patches(hip, dev(src), (0, 3, 3), outs, [(0, 3, 3)] * 3, n, dev(coords),
                        dev(pt.arrays), dev(pt.indices), dev(pt.params)) == 0
    for k in range(3):
        np.testing.assert_array_equal(host(outs[k], shape=(n, 3)), want[k])


@pytest.mark.parametrize("mesh", ["cube-with-triangles", "pentagonal-prism", "open-grid"])
def test_bilinear_patch_tables_bit_exact(hip, mesh):
    """The compiler's bilinear-scheme patch tables (QUADS patches, n-gon
    sub-faces at level 1) on the device: P, dPdu, dPdv at random locations
    of every ptex face equal the oracle's bit for bit."""
    from opensubdiv_amd import far
    from test_far_compiler import POLY_MESHES
    m = POLY_MESHES[mesh]()
    pt = far.Topology(m, scheme=far.SCHEME_BILINEAR).patch_table(2)
    assert all(a[0] == 3 for a in pt.arrays)                 # QUADS
    nptex = int(sum(1 if k == 4 else k for k in m.face_sizes))
    rng = np.random.default_rng(5)
    n = 3000
    f = rng.integers(0, nptex, n).astype(np.int32)
    u, v = rng.random(n), rng.random(n)
    coords = pt.find(f, u, v)
    assert (coords["arrayIndex"] >= 0).all()
    st = pt.point_stencils()
    ctrl = m.positions.astype(np.float32)
    pts = oracle.apply_stencils(ctrl, st["sizes"], st["offsets"], st["indices"], st["weights"]) \
        if len(st["sizes"]) else np.zeros((0, 3), np.float32)
    src = np.ascontiguousarray(np.concatenate([ctrl, pts]))
    want = [np.zeros((n, 3), np.float32) for _ in range(3)]
    assert oracle.eval_patches(src, (0, 3, 3), want, [(0, 3, 3)] * 3, coords, pt.arrays,
                               pt.indices, pt.params)
    outs = [empty(n * 3) for _ in range(3)]
    assert eval_patches(hip, dev(src), (0, 3, 3), outs, [(0, 3, 3)] * 3, n, dev(coords),
                        dev(pt.arrays), dev(pt.indices), dev(pt.params)) == 0
    for k in range(3):
        np.testing.assert_array_equal(host(outs[k], shape=(n, 3)), want[k])


@pytest.mark.parametrize("name", ["sphere", "prism", "grid"])
def test_bfr_face_queries_on_the_device_match_bfr_surfaces(hip, name):
    """Bfr-style per-face evaluation on the device: the compiler's table
    with Bfr's approximation options, its patch points refined by
    EvalStencils, the queries (base face, u, v in Bfr's parameterization of
    the face -- quads and n-gon sub-faces) located by
    osdhip_patch_map_find_faces, then EvalPatches -- against the reference's
    Bfr::Surface::Evaluate (fixture), within the 1e-5 bounding-box
    tolerance; the device's face -> ptex mapping equals the host's."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_far_compiler import FixtureMesh, bfr_to_ptex
    from conftest import bbox_rel_err
    from opensubdiv_amd import BufferDescriptor, HipEvaluator, HipPatchTable, HipStencilTable
    from opensubdiv_amd import HipVertexBuffer, far
    from opensubdiv_amd.osd import HipPatchMap
    g = golden("bfr_surfaces.npz")
    m = FixtureMesh(g, name)
    pt = far.Topology(m, boundary=int(g[name + "_boundary"])).bfr_patch_table()
    st = pt.point_stencils()
    nc, ns = m.nverts, len(st["sizes"])
    vb = HipVertexBuffer.Create(3, nc + ns)
    vb.UpdateData(m.positions.astype(np.float32), 0, nc)
    assert HipEvaluator.EvalStencils(vb, BufferDescriptor(0, 3, 3), vb,
                                     BufferDescriptor(nc * 3, 3, 3), HipStencilTable.Create(st))
    faces, u, v = g[name + "_faces"], g[name + "_u"], g[name + "_v"]
    n = len(faces)
    sizes = np.asarray(m.face_sizes, np.int32)
    base = np.concatenate([[0], np.cumsum(np.where(sizes == 4, 1, sizes))[:-1]]).astype(np.int32)
    pm = HipPatchMap(pt.arrays, pt.params)
    coords = dev(np.zeros((n, 5), np.int32))
    assert pm.FindFaceCoords(n, dev(faces), dev(u), dev(v), dev(sizes), dev(base), coords)
    assert HipEvaluator.Synchronize()
    got = host(coords).view(PATCH_COORD_DTYPE).reshape(-1)
    ptex, s, t = bfr_to_ptex(sizes, faces, u, v)
    want = pt.find(ptex, s, t)
    for k in ("arrayIndex", "patchIndex", "vertIndex", "s", "t"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    table = HipPatchTable(pt.arrays, pt.indices, pt.params)
    d = BufferDescriptor(0, 3, 3)
    P, du, dv = empty(n * 3), empty(n * 3), empty(n * 3)
    assert HipEvaluator.EvalPatches(vb, d, P, d, du, d, dv, d, n, coords, table)
    assert HipEvaluator.Synchronize()
    for o, k in zip((P, du, dv), ("P", "Du", "Dv")):
        assert bbox_rel_err(host(o, shape=(n, 3)).astype(np.float64), g[name + "_" + k]) < 1e-5, k
    pm.close()


def test_device_tessellation_of_every_parameterization_matches_bfr(hip):
    """osdhip_tess_coords for quads, Loop triangles and Catmark n-gon
    sub-faces at uniform, inner-only, outer and outer + inner rates: the
    coordinates of every face (three faces per pattern) are bit-identical to
    the reference's Bfr::Tessellation::GetCoords<float> (fixture)."""
    g = golden("tess_coords.npz")
    nf = 3
    for scheme, n, nr, r0, nc, c0 in g["configs"]:
        rates = np.ascontiguousarray(g["rates"][r0:r0 + nr], np.int32)
        uv = empty(nf * nc * 2)
        rc = hip.osdhip_tess_coords(int(scheme), int(n), int(nr), rates.ctypes.data, nf, ptr(uv),
                                    None, None, None, None, 0, None, stream())
        assert rc == 0, (scheme, n, list(rates))
        got = host(uv, shape=(nf, nc, 2))
        want = g["coords"][c0:c0 + nc]
        for f in range(nf):
            assert got[f].tobytes() == want.tobytes(), (scheme, n, list(rates), f)


def test_device_tessellation_to_patch_coords_of_base_faces(hip):
    """Tessellating base faces straight into PatchCoords: an n-gon mesh's
    quads and pentagons at per-edge rates through the device patch map,
    equal to osdhip_patch_map_find_faces of the same (face, u, v)."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_far_compiler import POLY_MESHES
    from opensubdiv_amd import far
    from opensubdiv_amd.osd import HipPatchMap
    m = POLY_MESHES["pentagonal-prism"]()
    pt = far.Topology(m).patch_table(3)
    sizes = np.asarray(m.face_sizes, np.int32)
    base = np.concatenate([[0], np.cumsum(np.where(sizes == 4, 1, sizes))[:-1]]).astype(np.int32)
    pm = HipPatchMap(pt.arrays, pt.params)
    for fs, rates in ((4, [3, 5, 2, 4]), (5, [4, 2, 3, 5, 1, 3]), (4, [6])):
        faces = np.nonzero(sizes == fs)[0].astype(np.int32)
        r = np.ascontiguousarray(rates, np.int32)
        nc = hip.osdhip_tess_count(1, fs, len(r), r.ctypes.data)
        assert nc > 0
        uv, coords = empty(len(faces) * nc * 2), dev(np.zeros((len(faces) * nc, 5), np.int32))
        dfaces, dsizes, dbase = dev(faces), dev(sizes), dev(base)
        assert hip.osdhip_tess_coords(1, fs, len(r), r.ctypes.data, len(faces), ptr(uv),
                                      ptr(dfaces), pm.h, ptr(dsizes), ptr(dbase), len(sizes),
                                      ptr(coords), stream()) == 0
        uvh = host(uv, shape=(len(faces) * nc, 2)).astype(np.float64)
        qf = np.repeat(faces, nc).astype(np.int32)
        want = dev(np.zeros((len(qf), 5), np.int32))
        assert pm.FindFaceCoords(len(qf), dev(qf), dev(np.ascontiguousarray(uvh[:, 0])),
                                 dev(np.ascontiguousarray(uvh[:, 1])), dsizes, dbase, want)
        got_h = host(coords).view(PATCH_COORD_DTYPE).reshape(-1)
        want_h = host(want).view(PATCH_COORD_DTYPE).reshape(-1)
        assert got_h.tobytes() == want_h.tobytes(), (fs, rates)
        assert (got_h["arrayIndex"] >= 0).all()
    pm.close()
