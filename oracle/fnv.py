"""FNV voxel keys and cuckoo-table sizing (oracle; test infrastructure only).

Follows csrc/common/hashmap_cuda.cu:248-261 (hash_kernel), :204-235 (kernel_hash_kernel) and
:341-360 (build_hash_table sizing).
"""
import math

import numpy as np

FNV_OFFSET = np.uint64(14695981039346656037)
FNV_PRIME = np.uint64(1099511628211)
MASK60 = np.uint64(0x0FFFFFFFFFFFFFFF)
BUCKET_SLOTS = 512


def fnv_fold60(coords) -> np.ndarray:
    """fold60(FNV-1a-64 over the uint32 words of each row), hashmap_cuda.cu:248-261.

    Each coordinate is cast to `unsigned int` (low 32 bits, two's complement), then
    h ^= w; h *= prime for every word, then h = (h >> 60) ^ (h & (2^60-1)).
    """
    c = np.asarray(coords)
    if c.ndim == 1:
        c = c[None, :]
    words = c.astype(np.int64).astype(np.uint64) & np.uint64(0xFFFFFFFF)
    h = np.full(c.shape[0], FNV_OFFSET, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for j in range(c.shape[1]):
            h ^= words[:, j]
            h *= FNV_PRIME
    return (h >> np.uint64(60)) ^ (h & MASK60)


def kernel_hash(coords, offsets) -> np.ndarray:
    """Keys of coords + offsets, shape (K, N), hashmap_cuda.cu:204-235 (out[k*N + i])."""
    c = np.asarray(coords, dtype=np.int64)
    o = np.asarray(offsets, dtype=np.int64)
    out = np.empty((o.shape[0], c.shape[0]), dtype=np.uint64)
    for k in range(o.shape[0]):
        q = c.copy()
        q[:, :3] += o[k]
        out[k] = fnv_fold60(q)
    return out


def table_sizing(n: int, enlarge: bool = False):
    """(table_size, num_buckets, evict_bound) of build_hash_table, hashmap_cuda.cu:349-360."""
    assert n > 0
    next_pow2 = int(2 ** math.ceil(math.log2(n)))
    table_size = 4 * next_pow2 if n < 2048 else 2 * next_pow2
    if enlarge:
        table_size = 4 * next_pow2
    table_size = max(table_size, 512)
    evict_bound = int(8 * math.ceil(math.log2(n)))
    num_buckets = -(-table_size // BUCKET_SLOTS)
    return table_size, num_buckets, evict_bound


def logical_map(keys, vals=None) -> dict:
    """The logical content of a built table: key -> value (default value = row index).

    The reference's slot layout is time-seeded and atomic-ordered (hashmap_cuda.cu:28,134), so the
    reproducible content of a table is this map (ext.py:74-84 query semantics: miss -> -1).
    """
    keys = np.asarray(keys, dtype=np.uint64)
    if vals is None:
        vals = np.arange(keys.shape[0], dtype=np.int64)
    return {int(k): int(v) for k, v in zip(keys, vals)}
