"""Deterministic synthetic inputs (wrappers over libamj_gen, include/amj_gen.h).

Workloads (sizes as used by the tests and bench):
  * chain joins R1(a,b) x R2(b,c) x ... with int32 columns; join columns uniform or Zipf(s)
    over K keys with a seeded rank->key permutation;
  * R-MAT graphs (Graph500 style relabelled), symmetrised, deduplicated and degree-oriented.
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
