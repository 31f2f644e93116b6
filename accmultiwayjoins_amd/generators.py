"""Deterministic synthetic inputs (wrappers over libamj_gen, include/amj_gen.h).

Workloads (sizes as used by the tests and bench):
  * chain joins R1(a,b) x R2(b,c) x ... with int32 columns; join columns uniform or Zipf(s)
    over K keys with a seeded rank->key permutation;
  * R-MAT graphs (Graph500 style relabelled), symmetrised, deduplicated and degree-oriented;
  * the reference's generator shapes: columns of unique-modulo values, distinct uniform pairs,
    distinct pairs with one Zipf column, sequential pairs (the writer command lines of
    writers.py build .db files from them).
"""
from __future__ import annotations

import numpy as np

from ._native import gen


def _ptr(a: np.ndarray):
    return a.ctypes.data


def uniform(n: int, lo: int, hi: int, seed: int, stream: int) -> np.ndarray:
    out = np.empty(n, dtype=np.int32)
    gen().amj_gen_uniform(_ptr(out), n, lo, hi, seed, stream)
    return out


def zipf(n: int, k: int, s: float, seed: int, stream: int) -> np.ndarray:
    out = np.empty(n, dtype=np.int32)
    if gen().amj_gen_zipf(_ptr(out), n, k, float(s), seed, stream) != 0:
        raise ValueError("invalid Zipf key count")
    return out


def zipf_mapped(n: int, k: int, s: float, seed: int, stream: int, shuffle: bool) -> np.ndarray:
    """Zipf(s) over k values; shuffle=False maps rank r to value r-1 (value 0 most frequent)."""
    out = np.empty(n, dtype=np.int32)
    if gen().amj_gen_zipf_mapped(_ptr(out), n, k, float(s), seed, stream, int(shuffle)) != 0:
        raise ValueError("invalid Zipf key count")
    return out


def unique_column(n: int, value_range: int, seed: int, stream: int) -> np.ndarray:
    """j % value_range for j in [0, n), in a seeded random order."""
    out = np.empty(n, dtype=np.int32)
    if gen().amj_gen_unique_column(_ptr(out), n, value_range, seed, stream) != 0:
        raise ValueError("invalid range or size for a unique column")
    return out


def _pair_rc(rc: int, what: str):
    if rc == -1:
        raise ValueError(f"{what}: invalid ranges, or not fewer rows than distinct pairs")
    if rc != 0:
        raise ValueError(f"{what}: the ranges cannot hold that many distinct pairs")


def unique_pairs(n: int, range0, range1, seed: int = 42, shuffle: bool = False):
    """n distinct pairs uniform over range0 x range1 ((lo, hi) tuples), sorted unless shuffled."""
    k0 = np.empty(n, dtype=np.int32)
    k1 = np.empty(n, dtype=np.int32)
    _pair_rc(gen().amj_gen_unique_pairs(_ptr(k0), _ptr(k1), n, range0[0], range0[1], range1[0],
                                        range1[1], seed, int(shuffle)), "unique pairs")
    return k0, k1


def zipf_pairs(n: int, range0, range1, skew_col: int, z: float, seed: int = 42,
               shuffle: bool = False):
    """n distinct pairs: column skew_col Zipf(z), the other uniform and distinct per value."""
    k0 = np.empty(n, dtype=np.int32)
    k1 = np.empty(n, dtype=np.int32)
    _pair_rc(gen().amj_gen_zipf_pairs(_ptr(k0), _ptr(k1), n, range0[0], range0[1], range1[0],
                                      range1[1], seed, skew_col, float(z), int(shuffle)),
             "Zipf pairs")
    return k0, k1


def seq_pairs(n: int, range0):
    """(lo + i // p, i % p) with p = n // (hi - lo)."""
    k0 = np.empty(n, dtype=np.int32)
    k1 = np.empty(n, dtype=np.int32)
    if gen().amj_gen_seq_pairs(_ptr(k0), _ptr(k1), n, range0[0], range0[1]) != 0:
        raise ValueError("sequential pairs need at least one row per key0 value")
    return k0, k1


def chain_relations(sizes, key_range, seed=42, zipf_s=None, payload_range=None):
    """Relations R_i(attr i, attr i+1) of a chain join.

    sizes[i] rows for R_i. Join column values: R_i.(i+1) (probe side of the next join) uniform
    in [0, key_range); the build side R_{i+1}.(i+1) uniform as well, or Zipf(zipf_s) over
    key_range keys when zipf_s is given. The outer columns R_0.0 and R_last.last are uniform
    over payload_range (default: key_range). Seeds follow seed + 10*i + col.
    """
    n_rel = len(sizes)
    payload_range = payload_range or key_range
    rels = []
    for i, n in enumerate(sizes):
        cols = []
        for c in range(2):
            st = 10 * i + c
            is_first_col = c == 0
            if is_first_col and i > 0:  # build-side join column (hash column)
                if zipf_s is not None and zipf_s > 0:
                    cols.append(zipf(n, key_range, zipf_s, seed, st))
                else:
                    cols.append(uniform(n, 0, key_range, seed, st))
            elif (is_first_col and i == 0) or (not is_first_col and i == n_rel - 1):
                cols.append(uniform(n, 0, payload_range, seed, st))
            else:  # probe-side join column
                cols.append(uniform(n, 0, key_range, seed, st))
        rels.append((cols, [i, i + 1]))
    return rels


def rmat_edges(scale: int, edge_factor: int = 16, abc=(0.57, 0.19, 0.19), seed: int = 42):
    m = edge_factor << scale
    src = np.empty(m, dtype=np.int32)
    dst = np.empty(m, dtype=np.int32)
    gen().amj_gen_rmat(_ptr(src), _ptr(dst), scale, m, abc[0], abc[1], abc[2], seed)
    return src, dst


def orient(src: np.ndarray, dst: np.ndarray, nv: int):
    """Symmetrise, drop self-loops and duplicates, orient by (degree, id), sort by (src, dst)."""
    m = len(src)
    os_ = np.empty(m, dtype=np.int32)
    od = np.empty(m, dtype=np.int32)
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    k = gen().amj_graph_orient(_ptr(src), _ptr(dst), m, nv, _ptr(os_), _ptr(od))
    return os_[:k].copy(), od[:k].copy()


def rmat_oriented(scale: int, edge_factor: int = 16, seed: int = 42):
    s, d = rmat_edges(scale, edge_factor, seed=seed)
    return orient(s, d, 1 << scale)
