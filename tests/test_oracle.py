"""Pins the CPU oracle (oracle/osd_oracle.c) before it is trusted as the
checker of the HIP evaluator:

* golden vectors written by the real reference (Osd::CpuEvaluator, Far
  StencilTableReal<double>, Far PatchTable::EvaluateBasis<double>);
* closed forms computed by independent algorithms (Cox-de Boor recursion,
  Bernstein polynomials);
* invariants: partition of unity, zero-sum derivatives, finite differences,
  the boundary-fold reflection identity, Gregory -> Bezier reduction,
  PatchParam normalization / 2^depth scaling;
* brute force for the stencil gather.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import bbox_rel_err, golden
from opensubdiv_amd.osd import PATCH_ARRAY_DTYPE, PATCH_COORD_DTYPE, PATCH_PARAM_DTYPE

REGULAR, GREGORY, QUADS = 6, 9, 3
TRIANGLES, LOOP, GREGORY_TRIANGLE = 4, 5, 10
LOOP_CASES = ("sphere", "grid", "strip", "uniform")


# ------------------------------------------------------------ golden vectors

def test_stencils_cube_l3_bit_exact_vs_cpu_evaluator():
    g = golden("cube_l3.npz")
    out = oracle.apply_stencils(g["src"], g["sizes"], g["offsets"], g["indices"], g["weights"])
    assert len(g["sizes"]) == 386                      # uniform L3 of the 8-vertex cube
    np.testing.assert_array_equal(out, g["cpu_out"])


def test_stencils_cube_l3_fp64_vs_far_double_table():
    g = golden("cube_l3.npz")
    out = oracle.apply_stencils(g["src64"], g["sizes"], g["offsets"], g["indices"],
                                g["weights64"], np.float64)
    np.testing.assert_allclose(out, g["out64"], rtol=0, atol=1e-15)
    # float result within the contract of the double one
    assert bbox_rel_err(g["cpu_out"], g["out64"]) < 1e-6


def test_stencils_stride7_and_subrange_descriptors():
    g = golden("cubesphere_l2.npz")
    src = g["src"]
    n = len(g["sizes"])
    out7 = np.zeros((n, 7), np.float32)
    assert oracle.eval_stencils(src, (0, 7, 7), out7, (0, 7, 7), g["sizes"], g["offsets"],
                                g["indices"], g["weights"])
    np.testing.assert_array_equal(out7, g["cpu_out7"])
    out3 = np.zeros((n, 7), np.float32)
    assert oracle.eval_stencils(src, (0, 3, 7), out3, (0, 3, 7), g["sizes"], g["offsets"],
                                g["indices"], g["weights"])
    np.testing.assert_array_equal(out3[:, :3], g["cpu_out3"])
    assert np.all(out3[:, 3:] == 0)                    # untouched lanes of the stride


def test_limit_stencils_d1_d2_vs_cpu_evaluator():
    g = golden("cubesphere_limit.npz")
    n = len(g["sizes"])
    outs = [np.zeros((n, 3), np.float32) for _ in range(6)]
    assert oracle.eval_stencils_d2(g["src"], (0, 3, 3), outs, [(0, 3, 3)] * 6, g["sizes"],
                                   g["offsets"], g["indices"],
                                   [g[k] for k in ("weights", "du", "dv", "duu", "duv", "dvv")])
    for o, k in zip(outs, ("cpu_p", "cpu_du", "cpu_dv", "cpu_duu", "cpu_duv", "cpu_dvv")):
        np.testing.assert_array_equal(o, g[k])
    outs1 = [np.zeros((n, 3), np.float32) for _ in range(3)]
    assert oracle.eval_stencils_d1(g["src"], (0, 3, 3), outs1, [(0, 3, 3)] * 3, g["sizes"],
                                   g["offsets"], g["indices"], [g["weights"], g["du"], g["dv"]])
    for o, k in zip(outs1, ("cpu1_p", "cpu1_du", "cpu1_dv")):
        np.testing.assert_array_equal(o, g[k])


def _patch_fixture():
    g = golden("cubesphere_patches.npz")
    coords = np.ascontiguousarray(g["coords"]).view(PATCH_COORD_DTYPE).reshape(-1)
    arrays = np.ascontiguousarray(g["arrays"]).view(PATCH_ARRAY_DTYPE).reshape(-1)
    params = np.ascontiguousarray(g["params"]).view(PATCH_PARAM_DTYPE).reshape(-1)
    return g, coords, arrays, np.ascontiguousarray(g["indices"]), params


def test_patches_vs_cpu_evaluator():
    g, coords, arrays, idx, params = _patch_fixture()
    assert set(arrays["desc"].tolist()) == {REGULAR, GREGORY}
    n = len(coords)
    o = [np.zeros((n, 3), np.float32) for _ in range(3)]
    assert oracle.eval_patches(g["src"], (0, 3, 3), o, [(0, 3, 3)] * 3, coords, arrays, idx, params)
    # same polynomial forms and operation order as the reference: bit-exact
    for x, k in zip(o, ("cpu_p", "cpu_du", "cpu_dv")):
        np.testing.assert_array_equal(x, g[k], err_msg=k)
    o2 = [np.zeros((n, 3), np.float32) for _ in range(6)]
    assert oracle.eval_patches(g["src"], (0, 3, 3), o2, [(0, 3, 3)] * 6, coords, arrays, idx,
                               params)
    for x, k in zip(o2, ("cpu2_p", "cpu2_du", "cpu2_dv", "cpu2_duu", "cpu2_duv", "cpu2_dvv")):
        np.testing.assert_array_equal(x, g[k], err_msg=k)
    p_only = np.zeros((n, 3), np.float32)
    assert oracle.eval_patches(g["src"], (0, 3, 3), [p_only], [(0, 3, 3)], coords, arrays, idx,
                               params)
    np.testing.assert_array_equal(p_only, g["cpu_p_only"])


FVAR_CASES = ("linear", "smooth")


@pytest.mark.parametrize("name", FVAR_CASES)
def test_face_varying_channel_vs_cpu_evaluator(name):
    """A real Far face-varying UV channel (seams along the cube edges): the
    face-varying stencil table refines the UVs as Far's own table does, and
    EvalPatchesFaceVarying through the channel's flattened arrays reproduces
    CpuEvaluator bit for bit -- linear QUADS patches (Far's defaults) and
    smooth B-spline / Gregory patches with face-varying local points."""
    g = golden("cubesphere_fvar.npz")
    nb = len(g["uv"])
    buf = np.zeros_like(g[name + "_buffer"])
    buf[:nb] = g["uv"]
    assert oracle.eval_stencils(buf, (0, 2, 2), buf, (nb * 2, 2, 2), g[name + "_st_sizes"],
                                g[name + "_st_offsets"], g[name + "_st_indices"],
                                g[name + "_st_weights"])
    np.testing.assert_array_equal(buf, g[name + "_buffer"])
    arrays = np.ascontiguousarray(g[name + "_arrays"]).view(PATCH_ARRAY_DTYPE).reshape(-1)
    params = np.ascontiguousarray(g[name + "_params"]).view(PATCH_PARAM_DTYPE).reshape(-1)
    coords = np.ascontiguousarray(g[name + "_coords"]).view(PATCH_COORD_DTYPE).reshape(-1)
    types = set(arrays["regDesc"].tolist()) | set(arrays["desc"].tolist())
    assert types == ({QUADS} if name == "linear" else {REGULAR, GREGORY})
    n = len(coords)
    o = [np.zeros((n, 2), np.float32) for _ in range(6)]
    assert oracle.eval_patches(buf, (0, 2, 2), o[:3], [(0, 2, 2)] * 3, coords, arrays,
                               g[name + "_indices"], params)
    for x, k in zip(o[:3], ("cpu_p", "cpu_du", "cpu_dv")):
        np.testing.assert_array_equal(x, g[name + "_" + k], err_msg=k)
    assert oracle.eval_patches(buf, (0, 2, 2), o, [(0, 2, 2)] * 6, coords, arrays,
                               g[name + "_indices"], params)
    for x, k in zip(o, ("cpu2_p", "cpu2_du", "cpu2_dv", "cpu2_duu", "cpu2_duv", "cpu2_dvv")):
        np.testing.assert_array_equal(x, g[name + "_" + k], err_msg=k)


def test_linear_face_varying_patches_interpolate_the_face_values():
    """Closed form for the linear channel: a QUADS face-varying patch is the
    bilinear blend of its four refined face values, so at a patch corner the
    result is that corner's value, and on a chart the UVs stay inside the
    chart's atlas cell (seams do not blend across charts)."""
    g = golden("cubesphere_fvar.npz")
    arrays = np.ascontiguousarray(g["linear_arrays"]).view(PATCH_ARRAY_DTYPE).reshape(-1)
    params = np.ascontiguousarray(g["linear_params"]).view(PATCH_PARAM_DTYPE).reshape(-1)
    coords = np.ascontiguousarray(g["linear_coords"]).view(PATCH_COORD_DTYPE).reshape(-1)
    buf, idx = g["linear_buffer"], g["linear_indices"]
    p = g["linear_cpu_p"]
    # every coord's UV lies in the atlas cell of the cube face its ptex face is on
    ptex = params["field0"][coords["patchIndex"]] & 0xfffffff
    chart = ptex // 16
    cu, cv = chart % 3, chart // 3
    eps = 1e-6
    assert np.all((p[:, 0] >= cu / 3 - eps) & (p[:, 0] <= (cu + 1) / 3 + eps))
    assert np.all((p[:, 1] >= cv / 2 - eps) & (p[:, 1] <= (cv + 1) / 2 + eps))
    # bilinear closed form at the coords, in float64, within float rounding
    for k in range(0, len(coords), 97):
        c = coords[k]
        a = arrays[c["arrayIndex"]]
        base = a["indexBase"] + 4 * (c["patchIndex"] - a["primitiveIdBase"])
        cv4 = buf[idx[base:base + 4]].astype(np.float64)
        f1 = int(params["field1"][c["patchIndex"]])
        depth, u0, v0 = f1 & 15, (f1 >> 22) & 1023, (f1 >> 12) & 1023
        s = float(c["s"]) * (1 << depth) - u0
        t = float(c["t"]) * (1 << depth) - v0
        w = np.array([(1 - s) * (1 - t), s * (1 - t), s * t, (1 - s) * t])
        np.testing.assert_allclose(w @ cv4, p[k], rtol=0, atol=2e-7)


def test_patch_basis_f64_vs_far_evaluate_basis():
    g, coords, arrays, idx, params = _patch_fixture()
    for i in range(len(g["basis_n"])):
        c = coords[i]
        a = arrays[c["arrayIndex"]]
        p = params[c["patchIndex"]]
        ptype = a["regDesc"] if (p["field1"] >> 5) & 1 else a["desc"]
        n, wP, wDs, wDt, *_ = oracle.patch_basis(ptype, p["field0"], p["field1"],
                                                 float(c["s"]), float(c["t"]), 1, np.float64)
        assert n == g["basis_n"][i]
        np.testing.assert_allclose(wP[:n], g["basis_wP"][i, :n], atol=1e-13)
        np.testing.assert_allclose(wDs[:n], g["basis_wDs"][i, :n], atol=1e-11)
        np.testing.assert_allclose(wDt[:n], g["basis_wDt"][i, :n], atol=1e-11)


# --------------------------------------------- closed forms (independent)

def cox_de_boor(t, deriv=0):
    """Uniform cubic B-spline basis on [0,1] (knots ..., -3..4) by the
    Cox-de Boor recursion; derivatives by the standard derivative formula."""
    knots = np.arange(-3, 5, dtype=np.float64)

    def N(i, p, x):
        if p == 0:
            return 1.0 if knots[i] <= x < knots[i + 1] else 0.0
        a = (x - knots[i]) / (knots[i + p] - knots[i]) * N(i, p - 1, x)
        b = (knots[i + p + 1] - x) / (knots[i + p + 1] - knots[i + 1]) * N(i + 1, p - 1, x)
        return a + b

    def dN(i, p, x, k):
        if k == 0:
            return N(i, p, x)
        return p * (dN(i, p - 1, x, k - 1) / (knots[i + p] - knots[i]) -
                    dN(i + 1, p - 1, x, k - 1) / (knots[i + p + 1] - knots[i + 1]))

    x = min(max(t, 0.0), 1.0 - 1e-15)
    return np.array([dN(i, 3, x, deriv) for i in range(4)])


def bernstein(t, deriv=0):
    def b(k, n, x):
        if k < 0 or k > n:
            return 0.0
        return math.comb(n, k) * x ** k * (1 - x) ** (n - k)
    if deriv == 0:
        return np.array([b(k, 3, t) for k in range(4)])
    if deriv == 1:
        return np.array([3 * (b(k - 1, 2, t) - b(k, 2, t)) for k in range(4)])
    return np.array([6 * (b(k - 2, 1, t) - 2 * b(k - 1, 1, t) + b(k, 1, t)) for k in range(4)])


def field1(depth=0, u=0, v=0, boundary=0, regular=1, nonquad=0):
    return (u << 22) | (v << 12) | (boundary << 7) | (regular << 5) | (nonquad << 4) | depth


@pytest.mark.parametrize("s,t", [(0.0, 0.0), (0.3, 0.7), (0.999, 0.01), (0.5, 0.5), (0.125, 0.9)])
def test_bspline_basis_is_cox_de_boor_tensor_product(s, t):
    n, wP, wDs, wDt, wDss, wDst, wDtt = oracle.patch_basis(REGULAR, 0, field1(), s, t, 2,
                                                            np.float64)
    assert n == 16
    bs = [cox_de_boor(s, k) for k in range(3)]
    bt = [cox_de_boor(t, k) for k in range(3)]
    # point (row i, column j) = 4i + j, column along s
    np.testing.assert_allclose(wP[:16].reshape(4, 4), np.outer(bt[0], bs[0]), atol=1e-14)
    np.testing.assert_allclose(wDs[:16].reshape(4, 4), np.outer(bt[0], bs[1]), atol=1e-13)
    np.testing.assert_allclose(wDt[:16].reshape(4, 4), np.outer(bt[1], bs[0]), atol=1e-13)
    np.testing.assert_allclose(wDss[:16].reshape(4, 4), np.outer(bt[0], bs[2]), atol=1e-12)
    np.testing.assert_allclose(wDst[:16].reshape(4, 4), np.outer(bt[1], bs[1]), atol=1e-12)
    np.testing.assert_allclose(wDtt[:16].reshape(4, 4), np.outer(bt[2], bs[0]), atol=1e-12)


@pytest.mark.parametrize("s,t", [(0.2, 0.6), (0.0, 1.0), (0.75, 0.25)])
def test_bilinear_basis_closed_form(s, t):
    n, wP, wDs, wDt, wDss, wDst, wDtt = oracle.patch_basis(QUADS, 0, field1(), s, t, 2, np.float64)
    assert n == 4
    np.testing.assert_allclose(wP[:4], [(1 - s) * (1 - t), s * (1 - t), s * t, (1 - s) * t],
                               atol=1e-15)
    np.testing.assert_allclose(wDs[:4], [-(1 - t), 1 - t, t, -t], atol=1e-15)
    np.testing.assert_allclose(wDt[:4], [-(1 - s), -s, s, 1 - s], atol=1e-15)
    np.testing.assert_allclose(wDst[:4], [1, -1, 1, -1], atol=1e-15)
    np.testing.assert_allclose(wDss[:4], 0, atol=0)


GREG_BND = {0: (0, 0), 1: (1, 0), 7: (2, 0), 5: (3, 0), 2: (0, 1), 6: (3, 1), 16: (0, 2),
            12: (3, 2), 15: (0, 3), 17: (1, 3), 11: (2, 3), 10: (3, 3)}
GREG_INT = {(3, 4): (1, 1), (8, 9): (2, 1), (13, 14): (2, 2), (18, 19): (1, 2)}


@pytest.mark.parametrize("s,t", [(0.3, 0.4), (0.9, 0.05), (0.5, 0.5), (0.01, 0.99)])
def test_gregory_reduces_to_bicubic_bezier(s, t):
    """With each interior pair merged, the Gregory basis is the Bernstein
    tensor product (value and the approximate first/second derivatives)."""
    n, wP, wDs, wDt, wDss, wDst, wDtt = oracle.patch_basis(GREGORY, 0, field1(regular=0), s, t, 2,
                                                            np.float64)
    assert n == 20
    B = [bernstein(s, k) for k in range(3)]
    T = [bernstein(t, k) for k in range(3)]
    for (w, ds, dt) in [(wP, 0, 0), (wDs, 1, 0), (wDt, 0, 1), (wDss, 2, 0), (wDst, 1, 1),
                        (wDtt, 0, 2)]:
        for p, (c, r) in GREG_BND.items():
            assert abs(w[p] - B[ds][c] * T[dt][r]) < 1e-12
        for (p, m), (c, r) in GREG_INT.items():
            assert abs(w[p] + w[m] - B[ds][c] * T[dt][r]) < 1e-12
    # the rational blend is not trivial away from the corners
    assert 0.0 < wP[3] / (wP[3] + wP[4]) < 1.0


@pytest.mark.parametrize("ptype", [REGULAR, GREGORY, QUADS])
@pytest.mark.parametrize("depth,u,v", [(0, 0, 0), (2, 1, 3), (3, 7, 0)])
def test_partition_of_unity_and_zero_sum_derivatives(ptype, depth, u, v):
    rng = np.random.default_rng(depth * 31 + ptype)
    f1 = field1(depth, u, v, regular=int(ptype == REGULAR))
    for _ in range(8):
        s = (u + rng.uniform()) / (1 << depth)
        t = (v + rng.uniform()) / (1 << depth)
        n, wP, wDs, wDt, wDss, wDst, wDtt = oracle.patch_basis(ptype, 0, f1, s, t, 2, np.float64)
        assert abs(wP[:n].sum() - 1) < 1e-13
        for w in (wDs, wDt, wDss, wDst, wDtt):
            assert abs(w[:n].sum()) < 1e-9


@pytest.mark.parametrize("ptype", [REGULAR, QUADS])
@pytest.mark.parametrize("boundary", [0, 1, 2, 4, 8, 3, 12, 9])
def test_first_derivatives_match_finite_differences(ptype, boundary):
    """Exact derivatives for the polynomial bases, including the 2^depth
    scaling from PatchParam normalization (depth 2 here)."""
    depth, u, v = 2, 1, 2
    f1 = field1(depth, u, v, boundary=boundary if ptype == REGULAR else 0)
    h = 1e-6
    for s0, t0 in [(0.3, 0.6), (0.41, 0.52), (0.37, 0.74)]:
        s, t = s0, t0
        n, wP, wDs, wDt, *_ = oracle.patch_basis(ptype, 0, f1, s, t, 1, np.float64)
        _, a, *_ = oracle.patch_basis(ptype, 0, f1, s + h, t, 0, np.float64)
        _, b, *_ = oracle.patch_basis(ptype, 0, f1, s - h, t, 0, np.float64)
        np.testing.assert_allclose(wDs[:n], (a - b)[:n] / (2 * h), atol=1e-6)
        _, a, *_ = oracle.patch_basis(ptype, 0, f1, s, t + h, 0, np.float64)
        _, b, *_ = oracle.patch_basis(ptype, 0, f1, s, t - h, 0, np.float64)
        np.testing.assert_allclose(wDt[:n], (a - b)[:n] / (2 * h), atol=1e-6)


@pytest.mark.parametrize("boundary", [1, 2, 4, 8, 3, 6, 12, 9, 15])
def test_boundary_fold_is_reflection_of_phantom_points(boundary):
    """Folded weights on the 16 points == unfolded weights on points whose
    boundary row/column is replaced by the reflection 2*P1 - P2."""
    rng = np.random.default_rng(boundary)
    P = rng.normal(size=(4, 4, 3))              # [row(t), col(s)]
    Q = P.copy()
    if boundary & 1:
        Q[0, :] = 2 * Q[1, :] - Q[2, :]
    if boundary & 2:
        Q[:, 3] = 2 * Q[:, 2] - Q[:, 1]
    if boundary & 4:
        Q[3, :] = 2 * Q[2, :] - Q[1, :]
    if boundary & 8:
        Q[:, 0] = 2 * Q[:, 1] - Q[:, 2]
    if (boundary & 1) and (boundary & 2):       # corner: reflect the reflected row
        Q[0, 3] = 2 * Q[0, 2] - Q[0, 1]
    if (boundary & 4) and (boundary & 2):
        Q[3, 3] = 2 * Q[3, 2] - Q[3, 1]
    if (boundary & 4) and (boundary & 8):
        Q[3, 0] = 2 * Q[3, 1] - Q[3, 2]
    if (boundary & 1) and (boundary & 8):
        Q[0, 0] = 2 * Q[0, 1] - Q[0, 2]
    for s, t in [(0.2, 0.3), (0.7, 0.9), (0.5, 0.05)]:
        _, wf, dsf, dtf, *_ = oracle.patch_basis(REGULAR, 0, field1(boundary=boundary), s, t, 1,
                                                 np.float64)
        _, w0, ds0, dt0, *_ = oracle.patch_basis(REGULAR, 0, field1(), s, t, 1, np.float64)
        for a, b in [(wf, w0), (dsf, ds0), (dtf, dt0)]:
            lhs = np.einsum("k,kd->d", a[:16], P.reshape(16, 3))
            rhs = np.einsum("k,kd->d", b[:16], Q.reshape(16, 3))
            np.testing.assert_allclose(lhs, rhs, atol=1e-12)
        # folded weights vanish on the dropped row/column
        W = wf[:16].reshape(4, 4)
        if boundary & 1:
            assert np.all(W[0] == 0)
        if boundary & 8:
            assert np.all(W[:, 0] == 0)


def test_normalization_maps_ptex_space_to_patch_space():
    """A depth-d patch at (u,v) evaluated at ptex (s,t) equals the depth-0
    basis at (s 2^d - u, t 2^d - v), first derivatives x 2^d, second x 4^d."""
    d, u, v = 3, 5, 2
    for s_loc, t_loc in [(0.25, 0.75), (0.6, 0.1)]:
        s = (u + s_loc) / 8
        t = (v + t_loc) / 8
        a = oracle.patch_basis(REGULAR, 0, field1(d, u, v), s, t, 2, np.float64)
        b = oracle.patch_basis(REGULAR, 0, field1(), s_loc, t_loc, 2, np.float64)
        np.testing.assert_allclose(a[1], b[1], atol=1e-13)
        np.testing.assert_allclose(a[2], b[2] * 8, atol=1e-11)
        np.testing.assert_allclose(a[3], b[3] * 8, atol=1e-11)
        np.testing.assert_allclose(a[4], b[4] * 64, atol=1e-9)


def test_unsupported_patch_type_returns_zero_points():
    # POINTS, LINES and the legacy GREGORY / GREGORY_BOUNDARY types, which
    # OsdEvaluatePatchBasis does not evaluate
    for ptype in (0, 1, 2, 7, 8, 11):
        assert oracle.patch_basis(ptype, 0, field1(), 0.3, 0.3)[0] == 0


# ------------------------------------------------- triangular patch types

def _loop_fixture(name):
    g = golden("loop_patches.npz")
    coords = np.ascontiguousarray(g[name + "_coords"]).view(PATCH_COORD_DTYPE).reshape(-1)
    arrays = np.ascontiguousarray(g[name + "_arrays"]).view(PATCH_ARRAY_DTYPE).reshape(-1)
    params = np.ascontiguousarray(g[name + "_params"]).view(PATCH_PARAM_DTYPE).reshape(-1)
    return g, coords, arrays, np.ascontiguousarray(g[name + "_indices"]), params


@pytest.mark.parametrize("name", LOOP_CASES)
def test_loop_patches_vs_cpu_evaluator(name):
    """LOOP box-spline patches (boundary masks included), GREGORY_TRIANGLE
    end caps and uniform TRIANGLES: bit-exact against the reference's
    CpuEvaluator in all three arities."""
    g, coords, arrays, idx, params = _loop_fixture(name)
    types = set(arrays["desc"].tolist()) | set(arrays["regDesc"].tolist())
    assert types == ({TRIANGLES} if name == "uniform" else {LOOP, GREGORY_TRIANGLE})
    n = len(coords)
    o = [np.zeros((n, 3), np.float32) for _ in range(3)]
    assert oracle.eval_patches(g[name + "_src"], (0, 3, 3), o, [(0, 3, 3)] * 3, coords, arrays,
                               idx, params)
    for x, k in zip(o, ("cpu_p", "cpu_du", "cpu_dv")):
        np.testing.assert_array_equal(x, g[f"{name}_{k}"], err_msg=k)
    o2 = [np.zeros((n, 3), np.float32) for _ in range(6)]
    assert oracle.eval_patches(g[name + "_src"], (0, 3, 3), o2, [(0, 3, 3)] * 6, coords, arrays,
                               idx, params)
    for x, k in zip(o2, ("cpu2_p", "cpu2_du", "cpu2_dv", "cpu2_duu", "cpu2_duv", "cpu2_dvv")):
        np.testing.assert_array_equal(x, g[f"{name}_{k}"], err_msg=k)


def test_loop_fixture_covers_boundaries_and_rotated_triangles():
    masks, rotated = set(), 0
    for name in LOOP_CASES:
        g, coords, arrays, idx, params = _loop_fixture(name)
        f1 = params["field1"][coords["patchIndex"]].astype(np.int64)
        depth, u, v = f1 & 0xF, (f1 >> 22) & 0x3FF, (f1 >> 12) & 0x3FF
        masks |= set(((f1 >> 7) & 0x1F).tolist())
        rotated += int((u + v >= (1 << depth)).sum())
    # edge masks, vertex-only masks (upper bits 1) and edge + opposite
    # vertex masks (upper bits 2), and rotated sub-triangles
    assert {1, 2, 3, 4, 6} <= masks and {9, 10, 12} <= masks and {17, 18} <= masks
    assert rotated > 100


@pytest.mark.parametrize("name", LOOP_CASES)
def test_triangle_basis_f64_vs_far_evaluate_basis(name):
    g, coords, arrays, idx, params = _loop_fixture(name)
    for i in range(len(g[name + "_basis_n"])):
        c = coords[i]
        a = arrays[c["arrayIndex"]]
        p = params[c["patchIndex"]]
        ptype = a["regDesc"] if (p["field1"] >> 5) & 1 else a["desc"]
        n, wP, wDs, wDt, *_ = oracle.patch_basis(ptype, p["field0"], p["field1"],
                                                 float(c["s"]), float(c["t"]), 1, np.float64)
        assert n == g[name + "_basis_n"][i]
        np.testing.assert_allclose(wP[:n], g[name + "_basis_wP"][i, :n], atol=1e-13)
        np.testing.assert_allclose(wDs[:n], g[name + "_basis_wDs"][i, :n], atol=1e-11)
        np.testing.assert_allclose(wDt[:n], g[name + "_basis_wDt"][i, :n], atol=1e-11)


def bernstein_tri(i, j, u, v, du=0, dv=0):
    """Quartic Bernstein triangle polynomial 4!/(i! j! k!) u^i v^j w^k
    (k = 4 - i - j, w = 1 - u - v) and its u/v derivatives, by sympy-free
    differentiation of the monomial expansion in (u, v)."""
    k = 4 - i - j
    c = math.factorial(4) / (math.factorial(i) * math.factorial(j) * math.factorial(k))
    total = 0.0
    # w^k = sum_a sum_b k!/(a! b! (k-a-b)!) (-u)^a (-v)^b
    for a in range(k + 1):
        for b in range(k + 1 - a):
            m = math.factorial(k) / (math.factorial(a) * math.factorial(b) *
                                     math.factorial(k - a - b)) * (-1) ** (a + b)
            pu, pv = i + a, j + b
            if pu < du or pv < dv:
                continue
            fu = math.perm(pu, du) * u ** (pu - du)
            fv = math.perm(pv, dv) * v ** (pv - dv)
            total += m * fu * fv
    return c * total


# Gregory triangle point -> exponents (i of u, j of v) of the Bezier point it
# stands for: corner k's point, its two edge neighbours and its interior
# pair, for corners (0,0), (1,0), (0,1), then the three edge midpoints.
GREG_TRI_UV = {0: (0, 0), 1: (1, 0), 2: (0, 1), 5: (4, 0), 6: (3, 1), 7: (3, 0),
               10: (0, 4), 11: (0, 3), 12: (1, 3), 15: (2, 0), 16: (2, 2), 17: (0, 2)}
GREG_TRI_PAIRS = {(3, 4): (1, 1), (8, 9): (2, 1), (13, 14): (1, 2)}


@pytest.mark.parametrize("s,t", [(0.2, 0.3), (0.6, 0.1), (0.05, 0.9), (1 / 3, 1 / 3)])
def test_gregory_triangle_reduces_to_quartic_bezier_triangle(s, t):
    """Each Gregory boundary point carries its Bernstein weight and each
    interior pair sums to one (the approximate derivatives hold the rational
    factors fixed, so derivatives follow the same rule)."""
    n, *w = oracle.patch_basis(GREGORY_TRIANGLE, 0, field1(), s, t, 2, np.float64)
    assert n == 18
    for o, (du, dv) in enumerate([(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2)]):
        for p, (i, j) in GREG_TRI_UV.items():
            assert abs(w[o][p] - bernstein_tri(i, j, s, t, du, dv)) < 1e-11, (o, p)
        for (p, q), (i, j) in GREG_TRI_PAIRS.items():
            assert abs(w[o][p] + w[o][q] - bernstein_tri(i, j, s, t, du, dv)) < 1e-11, (o, p)
    assert 0.0 < w[0][3] / (w[0][3] + w[0][4]) < 1.0


def test_box_spline_corners_are_loop_limit_masks():
    """At the three corners of a regular patch the box spline is the Loop
    limit mask of a valence-6 vertex: 1/2 on the vertex, 1/12 on each of its
    six neighbours, 0 on the other five points."""
    corners = {}
    for s, t in [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)]:
        n, wP, *_ = oracle.patch_basis(LOOP, 0, field1(), s, t, 0, np.float64)
        assert n == 12
        w = wP[:12]
        center = int(np.argmax(w))
        assert abs(w[center] - 0.5) < 1e-15
        assert np.sum(np.abs(w - 1 / 12) < 1e-15) == 6
        assert np.sum(w == 0) == 5
        corners[(s, t)] = center
    assert len(set(corners.values())) == 3


@pytest.mark.parametrize("ptype", [LOOP, GREGORY_TRIANGLE, TRIANGLES])
@pytest.mark.parametrize("depth,u,v", [(0, 0, 0), (2, 1, 1), (2, 2, 1), (3, 3, 5)])
def test_triangle_bases_unity_and_ptex_space_derivatives(ptype, depth, u, v):
    """Partition of unity, zero-sum derivatives, and first and second
    derivatives equal to finite differences in ptex-face (s, t), for upright
    and rotated (u + v >= 2^depth) sub-triangles: pins the triangle
    normalization, the 2^depth / 4^depth scaling and the sign flip.  The
    Gregory triangle's approximate derivatives hold its rational factors
    fixed, so there the interior pairs are compared summed.

    Reference quirk kept: OsdEvaluatePatchBasis scales second derivatives by
    derivSign * d1Scale^2 (osd/patchBasis.h:1590-1596), so on a rotated
    triangle they are the negated ptex-space second derivatives (the chain
    rule gives +4^depth there); the golden vectors hold the reference's
    sign."""
    f1 = field1(depth, u, v, regular=int(ptype == LOOP))
    rotated = u + v >= (1 << depth)
    rng = np.random.default_rng(depth * 7 + u + ptype)
    h = 1e-5
    for _ in range(4):
        a, b = rng.uniform(0.1, 0.8), rng.uniform(0.05, 0.15)
        sl, tl = a * (1 - b) * 0.9 + 0.02, b                 # inside the local triangle
        if rotated:
            s, t = ((1 << depth) - u - sl) / (1 << depth), ((1 << depth) - v - tl) / (1 << depth)
        else:
            s, t = (u + sl) / (1 << depth), (v + tl) / (1 << depth)
        n, wP, wDs, wDt, wDss, wDst, wDtt = oracle.patch_basis(ptype, 0, f1, s, t, 2,
                                                                np.float64)
        assert abs(wP[:n].sum() - 1) < 1e-13
        for w in (wDs, wDt, wDss, wDst, wDtt):
            assert abs(w[:n].sum()) < 1e-8

        def merge(w):
            w = np.array(w[:n])
            if ptype == GREGORY_TRIANGLE:
                for p, q in GREG_TRI_PAIRS:
                    w[p], w[q] = w[p] + w[q], 0.0
            return w

        def P(ds=0.0, dt=0.0, k=0):
            return merge(oracle.patch_basis(ptype, 0, f1, s + ds, t + dt, 1, np.float64)[1 + k])
        wDs, wDt, wDss, wDst, wDtt = map(merge, (wDs, wDt, wDss, wDst, wDtt))
        scale = float(1 << depth)
        sign2 = -1.0 if rotated else 1.0
        np.testing.assert_allclose(wDs[:n], (P(h) - P(-h)) / (2 * h), atol=1e-5 * scale)
        np.testing.assert_allclose(wDt[:n], (P(0, h) - P(0, -h)) / (2 * h), atol=1e-5 * scale)
        np.testing.assert_allclose(wDss[:n], sign2 * (P(h, 0, 1) - P(-h, 0, 1)) / (2 * h),
                                   atol=1e-4 * scale * scale)
        np.testing.assert_allclose(wDst[:n], sign2 * (P(0, h, 1) - P(0, -h, 1)) / (2 * h),
                                   atol=1e-4 * scale * scale)
        np.testing.assert_allclose(wDtt[:n], sign2 * (P(0, h, 2) - P(0, -h, 2)) / (2 * h),
                                   atol=1e-4 * scale * scale)


def test_linear_triangle_closed_form():
    n, wP, wDs, wDt, *_ = oracle.patch_basis(TRIANGLES, 0, field1(regular=0), 0.2, 0.7, 1,
                                             np.float64)
    assert n == 3
    np.testing.assert_allclose(wP[:3], [0.1, 0.2, 0.7], atol=1e-15)
    np.testing.assert_array_equal(wDs[:3], [-1, 1, 0])
    np.testing.assert_array_equal(wDt[:3], [-1, 0, 1])


# ---------------------------------------------------------------- brute force

def random_csr(rng, n, nsrc, max_size=16):
    sizes = rng.integers(0, max_size + 1, size=n).astype(np.int32)
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
    nnz = int(sizes.sum())
    idx = rng.integers(0, nsrc, size=nnz).astype(np.int32)
    w = rng.uniform(-1, 1, size=nnz)
    return sizes, offsets, idx, w


@pytest.mark.parametrize("L", [1, 3, 4, 7, 8, 12])
def test_stencils_brute_force(L):
    rng = np.random.default_rng(L)
    sizes, offsets, idx, w = random_csr(rng, 300, 50)
    src = rng.normal(size=(50, L))
    ref = np.zeros((300, L))
    for i in range(300):
        for j in range(sizes[i]):
            ref[i] += w[offsets[i] + j] * src[idx[offsets[i] + j]]
    out64 = oracle.apply_stencils(src, sizes, offsets, idx, w, np.float64)
    np.testing.assert_allclose(out64, ref, atol=1e-12)
    out32 = oracle.apply_stencils(src, sizes, offsets, idx, w.astype(np.float32))
    assert bbox_rel_err(out32, ref) < 1e-5


def test_stencils_start_end_row_convention_and_errors():
    rng = np.random.default_rng(5)
    sizes, offsets, idx, w = random_csr(rng, 40, 10)
    w = w.astype(np.float32)
    src = rng.normal(size=(10, 3)).astype(np.float32)
    full = oracle.apply_stencils(src, sizes, offsets, idx, w)
    out = np.zeros((40, 3), np.float32)
    # CPU generic-path convention: stencils [start,end) -> rows 0..end-start-1
    assert oracle.eval_stencils(src, (0, 3, 3), out, (0, 3, 3), sizes, offsets, idx, w, 10, 25)
    np.testing.assert_array_equal(out[:15], full[10:25])
    assert np.all(out[15:] == 0)
    # no-op and length mismatch (osd/cpuEvaluator.cpp:46-47)
    assert oracle.eval_stencils(src, (0, 3, 3), out, (0, 3, 3), sizes, offsets, idx, w, 5, 5)
    assert not oracle.eval_stencils(src, (0, 3, 3), out, (0, 2, 3), sizes, offsets, idx, w, 0, 4)


@pytest.mark.parametrize("threads", [1, 2, 3, 8])
def test_omp_stencils_bit_exact_vs_cpu_evaluator(threads):
    """The OpenMP restatement (osd/ompKernel.cpp:70-122) against the real
    CpuEvaluator's vectors: same bits whatever the thread count, since every
    row still sums in index order."""
    g = golden("cube_l3.npz")
    n = len(g["sizes"])
    out = np.zeros((n, 3), np.float32)
    assert oracle.eval_stencils(g["src"], (0, 3, 3), out, (0, 3, 3), g["sizes"], g["offsets"],
                                g["indices"], g["weights"], threads=threads)
    np.testing.assert_array_equal(out, g["cpu_out"])
    g = golden("cubesphere_l2.npz")
    n = len(g["sizes"])
    out7 = np.zeros((n, 7), np.float32)
    assert oracle.eval_stencils(g["src"], (0, 7, 7), out7, (0, 7, 7), g["sizes"], g["offsets"],
                                g["indices"], g["weights"], threads=threads)
    np.testing.assert_array_equal(out7, g["cpu_out7"])
    # OMP writes stencil i to row i - start (osd/ompKernel.cpp:113)
    part = np.zeros((n, 7), np.float32)
    assert oracle.eval_stencils(g["src"], (0, 7, 7), part, (0, 7, 7), g["sizes"], g["offsets"],
                                g["indices"], g["weights"], 100, 900, threads=threads)
    np.testing.assert_array_equal(part[:800], g["cpu_out7"][100:900])
    assert np.all(part[800:] == 0)
    assert not oracle.eval_stencils(g["src"], (0, 7, 7), part, (0, 3, 7), g["sizes"],
                                    g["offsets"], g["indices"], g["weights"], threads=threads)


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_omp_patches_bit_exact_vs_cpu_evaluator(threads):
    """The OpenMP EvalPatches restatement (osd/ompEvaluator.cpp:163-350)
    against the real CpuEvaluator's P / dPdu / dPdv and P-only vectors."""
    g, coords, arrays, idx, params = _patch_fixture()
    n = len(coords)
    o = [np.zeros((n, 3), np.float32) for _ in range(3)]
    assert oracle.eval_patches(g["src"], (0, 3, 3), o, [(0, 3, 3)] * 3, coords, arrays, idx,
                               params, threads=threads)
    for x, k in zip(o, ("cpu_p", "cpu_du", "cpu_dv")):
        np.testing.assert_array_equal(x, g[k], err_msg=k)
    p_only = np.zeros((n, 3), np.float32)
    assert oracle.eval_patches(g["src"], (0, 3, 3), [p_only, None, None], [(0, 3, 3), None, None],
                               coords, arrays, idx, params, threads=threads)
    np.testing.assert_array_equal(p_only, g["cpu_p_only"])


def test_planted_mistake_is_caught_by_the_closed_form():
    """A plausible oracle bug (B-spline basis with columns reversed, i.e.
    transposed s/t roles) must fail the Cox-de Boor check."""
    n, wP, *_ = oracle.patch_basis(REGULAR, 0, field1(), 0.2, 0.7, 0, np.float64)
    wrong = wP[:16].reshape(4, 4).T
    exact = np.outer(cox_de_boor(0.7), cox_de_boor(0.2))
    assert np.allclose(wP[:16].reshape(4, 4), exact, atol=1e-14)
    assert not np.allclose(wrong, exact, atol=1e-6)


# ------------------------------------------------------------- limit normals

def _normals(du, dv, normalize, desc=(0, 3, 3)):
    n = len(du)
    out = np.zeros((n, 3), np.float32)
    assert oracle.limit_normals(du, desc, dv, desc, out, (0, 3, 3), n, normalize)
    return out


def test_limit_normals_closed_forms():
    """Right-handed basis (e1 x e2 = e3 and cyclic), scaling, unit length."""
    e = np.eye(3, dtype=np.float32)
    du = np.stack([e[0], e[1], e[2], 2 * e[1]]).astype(np.float32)
    dv = np.stack([e[1], e[2], e[0], 3 * e[2]]).astype(np.float32)
    np.testing.assert_array_equal(_normals(du, dv, False), [e[2], e[0], e[1], [6, 0, 0]])
    np.testing.assert_array_equal(_normals(du, dv, True), [e[2], e[0], e[1], e[0]])


def test_limit_normals_invariants_against_numpy_cross():
    """Against numpy's cross in float64: anti-commutative, orthogonal to both
    tangents, unit length after normalising; descriptor offsets and strides."""
    rng = np.random.default_rng(11)
    n = 4096
    du = rng.normal(size=(n, 5)).astype(np.float32)
    dv = rng.normal(size=(n, 5)).astype(np.float32)
    out = np.zeros((n, 3), np.float32)
    assert oracle.limit_normals(du, (2, 3, 5), dv, (1, 3, 5), out, (0, 3, 3), n, False)
    a, b = du[:, 2:5].astype(np.float64), dv[:, 1:4].astype(np.float64)
    ref = np.cross(a, b)
    assert np.allclose(out, ref, rtol=1e-5, atol=1e-5)
    rev = np.zeros_like(out)
    assert oracle.limit_normals(dv, (1, 3, 5), du, (2, 3, 5), rev, (0, 3, 3), n, False)
    np.testing.assert_array_equal(rev, -out)
    unit = np.zeros_like(out)
    assert oracle.limit_normals(du, (2, 3, 5), dv, (1, 3, 5), unit, (0, 3, 3), n, True)
    assert np.abs(np.linalg.norm(unit, axis=1) - 1).max() < 1e-6
    assert np.abs((unit * a).sum(1)).max() < 1e-5 * np.abs(a).max()
    assert np.abs((unit * b).sum(1)).max() < 1e-5 * np.abs(b).max()
    assert np.allclose(unit, ref / np.linalg.norm(ref, axis=1)[:, None], atol=1e-6)
    # length < 3 in any descriptor is rejected; count 0 is a no-op
    assert not oracle.limit_normals(du, (0, 2, 5), dv, (0, 3, 5), out, (0, 3, 3), n, True)
    assert not oracle.limit_normals(du, (0, 3, 5), dv, (0, 3, 5), out, (0, 2, 3), n, True)
    assert oracle.limit_normals(du, (0, 3, 5), dv, (0, 3, 5), out, (0, 3, 3), 0, True)


def test_limit_normals_point_outward_on_the_golden_sphere():
    """The reference's own dPdu / dPdv of the cube-sphere limit surface give
    unit normals within 10 degrees of the radial direction; a planted swap of
    the operands (dv x du) flips every one of them and fails this check."""
    g = golden("cubesphere_patches.npz")
    du, dv = np.ascontiguousarray(g["cpu_du"]), np.ascontiguousarray(g["cpu_dv"])
    p = g["cpu_p"].astype(np.float64)
    radial = p / np.linalg.norm(p, axis=1)[:, None]
    cos = (_normals(du, dv, True) * radial).sum(1)
    assert cos.min() > math.cos(math.radians(10))
    assert ((_normals(dv, du, True) * radial).sum(1) < 0).all()


# ------------------------------------------------ Bfr tessellation pattern

def test_tess_uniform_quad_matches_bfr_vectors():
    """The oracle's Bfr quad pattern against Bfr's own coordinates (fixture
    written by the reference's Bfr::Tessellation; live for more rates when
    the reference library is built), bit for bit, plus what the pattern
    fixes: (rate + 1)^2 points, the boundary first (4 rate points on the
    square's edges, counter-clockwise from the origin), every grid point
    i / rate once."""
    g = golden("bfr_tess_quad.npz")
    for key in g.files:
        rate = int(key[4:])
        uv = oracle.tess_uniform_quad(rate)
        assert uv.tobytes() == g[key].tobytes(), key
        assert len(uv) == (rate + 1) ** 2
        edge = uv[:4 * rate]
        assert np.all((edge[:, 0] == 0) | (edge[:, 0] == 1) | (edge[:, 1] == 0) |
                      (edge[:, 1] == 1))
        assert tuple(uv[0]) == (0.0, 0.0) and tuple(uv[rate]) == (1.0, 0.0)
        ij = np.rint(uv.astype(np.float64) * rate).astype(int)
        assert np.abs(uv * rate - ij).max() < 1e-4 * rate
        assert len({(a, b) for a, b in ij}) == len(uv)
    from oracle import ref
    if ref.available():
        for rate in range(1, 65):
            assert oracle.tess_uniform_quad(rate).tobytes() == \
                ref.tess_uniform_coords(4, rate).tobytes(), rate


def test_tess_planted_mistake_is_caught():
    """A plausible restatement error -- the iso-line coordinate computed as
    u0 + i * d instead of accumulated -- fails the bit comparison."""
    g = golden("bfr_tess_quad.npz")
    rate = 33
    d = np.float32(1.0) / np.float32(rate)
    wrong = (np.arange(rate, dtype=np.float32) * d).astype(np.float32)
    assert wrong.tobytes() != g["rate33"][:rate, 0].tobytes()


def _tess_fixture():
    g = golden("tess_coords.npz")
    for scheme, n, nr, r0, nc, c0 in g["configs"]:
        yield int(scheme), int(n), g["rates"][r0:r0 + nr], g["coords"][c0:c0 + nc]


def test_tess_coords_match_bfr_for_every_parameterization():
    """The oracle's Bfr::Tessellation restatement for quads, Loop triangles
    and Catmark n-gon sub-faces (3-12 sides) at uniform, inner-only, outer
    and outer + inner rates, bit for bit against the reference's own
    GetCoords<float> (fixture of 300 patterns; live for more when the
    reference library is built); plus what the pattern fixes: the uniform
    counts in closed form (quad (r+1)^2, triangle (r+1)(r+2)/2, n-gon
    sub-faces as Bfr counts them), the boundary first with the outer
    rates' points on the face's boundary, every coordinate inside the
    parameterization's domain."""
    for scheme, n, rates, want in _tess_fixture():
        uv = oracle.tess_coords(scheme, n, rates)
        assert uv is not None and uv.tobytes() == want.tobytes(), (scheme, n, list(rates))
        if len(rates) == 1:
            r = int(rates[0])
            if scheme != 2 and n == 4:
                assert len(uv) == (r + 1) ** 2
            elif scheme == 2:
                assert len(uv) == (r + 1) * (r + 2) // 2
        outer = list(rates) if len(rates) >= n else [int(rates[0])] * n
        if len(rates) == 2 and n == 4:
            outer = [int(rates[0]), int(rates[1])] * 2
        nb = sum(outer[:n])
        b = uv[:nb].astype(np.float64)
        if scheme == 2:
            assert np.all((np.abs(b[:, 0]) < 1e-6) | (np.abs(b[:, 1]) < 1e-6) |
                          (np.abs(b.sum(axis=1) - 1) < 1e-6))
            assert np.all(uv >= -1e-6) and np.all(uv.sum(axis=1) <= 1 + 1e-6)
        elif n == 4:
            assert ((b == 0) | (b == 1)).any(axis=1).all()
            assert np.all((uv >= -1e-6) & (uv <= 1 + 1e-6))
        else:
            udim = (2 + (n > 4)) if n < 10 else 1 + int(np.sqrt(n - 1))
            tile = np.floor(uv + 0.25)                  # sub-face tile of each coordinate
            assert np.all(tile[:, 0] < udim) and np.all(tile.min(axis=0) >= 0)
            assert np.all(uv - tile <= 0.5 + 1e-6)
    from oracle import ref
    if ref.available():
        rng = np.random.default_rng(3)
        for scheme, n in [(1, 4), (2, 3), (1, 3), (1, 5), (1, 8), (1, 11), (0, 6)]:
            for _ in range(40):
                m = int(rng.choice([1, n, n + 1]))
                rates = rng.integers(1, 20, m)
                a, b = oracle.tess_coords(scheme, n, rates), ref.tess_coords(scheme, n, rates)
                assert (a is None) == (b is None)
                assert a is None or a.tobytes() == b.tobytes(), (scheme, n, list(rates))


def test_tess_coords_invalid_arguments():
    """What Bfr rejects the oracle rejects: rates below 1, no rates, faces of
    fewer than 3 sides, Loop faces other than triangles."""
    assert oracle.tess_coords(1, 4, [0]) is None
    assert oracle.tess_coords(1, 4, []) is None
    assert oracle.tess_coords(1, 2, [3]) is None
    assert oracle.tess_coords(2, 4, [3]) is None
    assert oracle.tess_coords(1, 4, [3, 0, 2, 2]) is None
