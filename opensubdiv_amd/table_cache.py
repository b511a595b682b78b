"""Topology-keyed on-disk cache of compiled tables (libosdfar table files).

The reference rebuilds every table from the topology on each run; a process
that keeps hundreds of topologies resident spends most of its start-up in
that compiler.  ``TableCache`` stores what the compiler produced for a
topology under ``<directory>/<key as 16 hex digits>.osdt``, where the key
hashes the topology's arrays and a recipe string naming what was built, and
hands it back as numpy views of a read-only map.

    cache = TableCache("/var/cache/osd")
    key = cache.key(mesh, "refine L3; patches iso 3")
    got = cache.load(key)                # TableFile or None (missing / stale / corrupt)
    if got is None:
        cache.store(key, {"refine.sizes": ..., "refine.indices": ..., ...})

A file that does not match its key, format version or checksums is treated
as absent (and replaced by the next ``store``).  Writes go to a temporary
name and are renamed into place, so concurrent writers of one key leave one
whole file.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

from . import far


class MappedStencilTable:
    """A float stencil table read from a table file (sections
    ``<prefix>.sizes``, ``.indices``, ``.weights``): the same attributes as
    far.RawStencilTable, so HipStencilTable uploads it from the map without
    a host copy.  Valid while its TableFile is open."""

    def __init__(self, tf: far.TableFile, prefix: str, num_control: int):
        self.sizes = tf.array(prefix + ".sizes")
        self.indices = tf.array(prefix + ".indices")
        self.weights = tf.array(prefix + ".weights")
        if self.sizes.dtype != np.int32 or self.indices.dtype != np.int32 or \
                self.weights.dtype != np.float32 or len(self.indices) != len(self.weights) or \
                int(self.sizes.sum(dtype=np.int64)) != len(self.indices):
            raise ValueError(f"table file section {prefix!r} is not a float stencil table")
        self.num_stencils, self.nnz = len(self.sizes), len(self.indices)
        self.num_control = int(num_control)
        self.sizes_ptr = self.sizes.ctypes.data
        self.indices_ptr = self.indices.ctypes.data
        self.weights_ptr = self.weights.ctypes.data

    def close(self):
        pass


def raw_table_arrays(raw, prefix):
    """Sections of a far.RawStencilTable (views of the compiler's memory)."""
    def view(ptr, n, dt):
        if n == 0:
            return np.zeros(0, dt)
        return np.frombuffer((C.c_ubyte * (n * np.dtype(dt).itemsize)).from_address(ptr), dt)
    return {prefix + ".sizes": view(raw.sizes_ptr, raw.num_stencils, np.int32),
            prefix + ".indices": view(raw.indices_ptr, raw.nnz, np.int32),
            prefix + ".weights": view(raw.weights_ptr, raw.nnz, np.float32)}


class TableCache:
    SUFFIX = ".osdt"

    def __init__(self, directory, verify=True):
        self.directory = os.fspath(directory)
        self.verify = bool(verify)
        os.makedirs(self.directory, exist_ok=True)
        self.hits = self.misses = 0
        self._lock = threading.Lock()         # loads run on several host threads

    def count(self, hit, undo_hit=False):
        """Records a hit or a miss (undo_hit: a loaded file turned out
        unusable, so its hit becomes a miss)."""
        with self._lock:
            if undo_hit:
                self.hits -= 1
            if hit:
                self.hits += 1
            else:
                self.misses += 1

    @staticmethod
    def key(mesh, recipe, boundary=1):
        return far.topology_key(mesh, recipe, boundary)

    def path(self, key):
        return os.path.join(self.directory, f"{key:016x}{self.SUFFIX}")

    def load(self, key):
        tf = far.TableFile.open(self.path(key), key, self.verify)
        self.count(tf is not None)
        return tf

    def store(self, key, arrays):
        rc = far.write_table_file(self.path(key), key, arrays)
        if rc != 0:
            raise OSError(f"table file write failed ({rc}): {self.path(key)}")
