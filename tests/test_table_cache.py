"""Table files and the topology-keyed table cache (libosdfar, host only).

The hash is pinned against the xxhash package's XXH64 (an independent
implementation of the published algorithm); files are checked for exact
round trips of every section type, for rejection of anything stale,
truncated or corrupt, and for whole files under concurrent writers; meshes
loaded from the cache equal freshly compiled ones array for array."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from opensubdiv_amd import far, meshes, workloads
from opensubdiv_amd.osd import PATCH_COORD_DTYPE
from opensubdiv_amd.table_cache import TableCache, raw_table_arrays


@pytest.fixture(scope="module")
def lib():
    try:
        return far.load_far()
    except RuntimeError as e:
        pytest.skip(str(e))


def test_hash64_is_xxh64(lib):
    xxhash = pytest.importorskip("xxhash")
    rng = np.random.default_rng(0)
    for n in list(range(0, 70)) + [255, 256, 4097, 1 << 20]:
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        for seed in (0, 1, 2 ** 64 - 1):
            assert far.hash64(b, seed) == xxhash.xxh64(b, seed=seed).intdigest(), (n, seed)
    # the specification's value for the empty input
    assert far.hash64(b"") == 0xEF46DB3751D8E999


def _mesh(n=4, seed=3):
    return meshes.cube_sphere(n, seed=seed, radial_noise=0.01, crease_fraction=0.05,
                              sharpness_max=4.0)


def test_topology_key_tracks_topology_and_recipe(lib):
    m = _mesh()
    k = far.topology_key(m, "r")
    assert k == far.topology_key(_mesh(), "r") != 0            # deterministic
    assert k != far.topology_key(m, "r2")
    assert k != far.topology_key(m, "r", boundary=2)
    m2 = _mesh()
    m2.face_verts = m2.face_verts.copy()
    m2.face_verts[0, [0, 1]] = m2.face_verts[0, [1, 0]]
    assert far.topology_key(m2, "r") != k
    m3 = _mesh()
    m3.crease_weights = m3.crease_weights.copy()
    m3.crease_weights[0] += 0.5
    assert far.topology_key(m3, "r") != k
    # positions are not topology
    m4 = _mesh()
    m4.positions = m4.positions + 1.0
    assert far.topology_key(m4, "r") == k


def _arrays():
    rng = np.random.default_rng(1)
    coords = np.zeros(5, PATCH_COORD_DTYPE)
    coords["s"] = rng.random(5)
    return {"i32": rng.integers(-9, 9, 1000).astype(np.int32),
            "i64": rng.integers(-2 ** 40, 2 ** 40, 77),
            "f32": rng.normal(size=333).astype(np.float32),
            "f64": rng.normal(size=5),
            "empty": np.zeros(0, np.float32),
            "coords": coords}


def test_table_file_round_trip(lib, tmp_path):
    path = tmp_path / "t.osdt"
    arrays = _arrays()
    assert far.write_table_file(path, 42, arrays) == 0
    with far.TableFile.open(path, 42) as tf:
        assert tf.names() == list(arrays)
        for name, a in arrays.items():
            got = tf.array(name, a.dtype if name == "coords" else None)
            assert got.dtype == a.dtype and got.tobytes() == a.tobytes(), name
            if got.size:
                assert got.ctypes.data % 256 == 0
        with pytest.raises(KeyError):
            tf.array("nope")
    assert os.listdir(tmp_path) == ["t.osdt"]                 # no temporary left behind


def test_table_file_rejects_stale_truncated_and_corrupt(lib, tmp_path):
    path = tmp_path / "t.osdt"
    assert far.write_table_file(path, 7, _arrays()) == 0
    good = path.read_bytes()
    assert far.TableFile.open(path, 8) is None                 # other key
    assert far.TableFile.open(tmp_path / "missing.osdt", 7) is None

    def variant(data, name):
        p = tmp_path / name
        p.write_bytes(data)
        return p

    assert far.TableFile.open(variant(good[:-1], "trunc"), 7) is None
    assert far.TableFile.open(variant(good + b"\0" * 256, "long"), 7) is None
    bad = bytearray(good)
    bad[8] ^= 1                                                # version
    assert far.TableFile.open(variant(bytes(bad), "ver"), 7) is None
    bad = bytearray(good)
    bad[64 + 40] ^= 1                                          # a directory entry's count
    assert far.TableFile.open(variant(bytes(bad), "dir"), 7) is None
    bad = bytearray(good)
    at = good.find(_arrays()["f32"].tobytes())
    assert at > 0
    bad[at + 100] ^= 0x10                                      # a payload byte
    assert far.TableFile.open(variant(bytes(bad), "pay"), 7) is None
    tf = far.TableFile.open(variant(bytes(bad), "pay2"), 7, verify=False)
    assert tf is not None                                      # only the checksum differs
    tf.close()


def test_table_file_argument_errors(lib, tmp_path):
    a = np.zeros(3, np.int32)
    assert far.write_table_file(tmp_path / "x", 1, {"": a}) == -1
    assert far.write_table_file(tmp_path / "x", 1, {"n" * 32: a}) == -1
    secs = (far.Section * 2)(far.Section(b"a", 1, 3, a.ctypes.data),
                             far.Section(b"a", 1, 3, a.ctypes.data))
    assert lib.osdfar_table_file_write(str(tmp_path / "x").encode(), 1, 2, secs) == -1
    bad_type = (far.Section * 1)(far.Section(b"a", 9, 3, a.ctypes.data))
    assert lib.osdfar_table_file_write(str(tmp_path / "x").encode(), 1, 1, bad_type) == -1
    assert far.write_table_file(tmp_path / "no" / "such" / "dir", 1, {"a": a}) == -2
    assert not os.path.exists(tmp_path / "x")


def test_concurrent_writers_leave_one_whole_file(lib, tmp_path):
    path = tmp_path / "t.osdt"
    arrays = _arrays()
    with ThreadPoolExecutor(8) as ex:
        assert all(rc == 0 for rc in ex.map(lambda _: far.write_table_file(path, 5, arrays),
                                            range(32)))
    assert os.listdir(tmp_path) == ["t.osdt"]
    with far.TableFile.open(path, 5) as tf:
        assert tf.array("f32").tobytes() == arrays["f32"].tobytes()


def test_meshes_from_the_cache_equal_compiled_ones(lib, tmp_path):
    spec = workloads.MeshSpec(0, 6, 2, 2, 3000, 9)
    cache = TableCache(tmp_path)
    first = workloads.compile_mesh(spec, 7, 0, cache)
    assert (cache.hits, cache.misses) == (0, 1)
    got = workloads.compile_mesh(spec, 7, 0, cache)
    assert (cache.hits, cache.misses) == (1, 1) and got.source is not None
    ref = workloads.compile_mesh(spec, 7, 0)
    raw = raw_table_arrays(ref.refine, "refine")
    for k in ("sizes", "indices", "weights"):
        assert np.array_equal(getattr(got.refine, k), raw["refine." + k]), k
    for k in ("sizes", "offsets", "indices", "weights"):
        assert np.array_equal(got.points[k], ref.points[k]), k
    for k in ("arrays", "indices", "params", "coords"):
        x, y = getattr(got, k), getattr(ref, k)
        assert x.dtype == y.dtype and x.shape == y.shape and x.tobytes() == y.tobytes(), k
    assert np.array_equal(got.frames, ref.frames)
    # a corrupt file is a miss: recompiled and replaced
    got.release()
    first.release()
    path = cache.path(cache.key(_mesh(6, spec.seed), workloads.mesh_recipe(spec)))
    data = bytearray(open(path, "rb").read())
    data[-1000] ^= 1
    open(path, "wb").write(bytes(data))
    again = workloads.compile_mesh(spec, 7, 0, cache)
    assert (cache.hits, cache.misses) == (1, 2) and again.source is None
    again.release()
    assert workloads.compile_mesh(spec, 7, 0, cache).source is not None
    ref.release()
