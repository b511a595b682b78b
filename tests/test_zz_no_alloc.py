"""osdhip_eval_patches_sorted_noalloc (the allocation-free entry point: the
sort's workspace is the library's fixed per-GPU scratch array) against
osdhip_eval_patches_sorted (workspace from hipMallocAsync), called through
the C ABI: bit-identical outputs on two calls in a row, the second with
fewer coords over the same scratch, and rejection of a call beyond the
scratch limit."""
import numpy as np
import pytest

from conftest import golden
from gpu_util import dev, empty, host, ptr, stream
from opensubdiv_amd import _lib

pytestmark = pytest.mark.gpu
D = _lib.Desc
NODESC = D(0, 0, 0)
MAX_COORDS = 1 << 22            # OSDHIP_NOALLOC_MAX_PATCH_COORDS
ERR_TOO_LARGE = -3


def call(fn, src, outs, n, coords, arrays, indices, params):
    args = []
    for k in range(6):
        if k < len(outs):
            args += [ptr(outs[k]), D(0, 3, 3)]
        else:
            args += [None, NODESC]
    return fn(ptr(src), D(0, 3, 3), *args, n, ptr(coords), ptr(arrays), ptr(indices),
              ptr(params), stream())


def test_noalloc_sorted_patches_match_the_allocating_entry_point(hip):
    g = golden("cubesphere_patches.npz")
    src, arrays, indices, params = (dev(g["src"]), dev(g["arrays"]), dev(g["indices"]),
                                    dev(g["params"]))
    rng = np.random.default_rng(2024)
    base = g["coords"]
    # first call large (coords with repeats over every patch), second smaller
    for n, nout in ((300000, 3), (5000, 6)):
        coords = dev(base[rng.integers(0, len(base), size=n)])
        res = {}
        for name in ("osdhip_eval_patches_sorted", "osdhip_eval_patches_sorted_noalloc"):
            outs = [empty(n * 3) for _ in range(nout)]
            rc = call(getattr(hip, name), src, outs, n, coords, arrays, indices, params)
            assert rc == 0, (name, rc)
            res[name] = [host(o, shape=(n, 3)) for o in outs]
        for k in range(nout):
            a, b = res["osdhip_eval_patches_sorted"][k], res["osdhip_eval_patches_sorted_noalloc"][k]
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (n, k)
            assert not np.isnan(b).any()


def test_noalloc_rejects_more_coords_than_its_scratch(hip):
    g = golden("cubesphere_patches.npz")
    src = dev(g["src"])
    out = empty(3)
    rc = call(hip.osdhip_eval_patches_sorted_noalloc, src, [out], MAX_COORDS + 1,
              dev(g["coords"]), dev(g["arrays"]), dev(g["indices"]), dev(g["params"]))
    assert rc == ERR_TOO_LARGE
    assert np.isnan(host(out)).all()              # nothing was launched
