"""ctypes access to the CPU oracle (oracle/liboracle.so) and brute-force restatements for tiny
inputs. Test infrastructure: the checker, never the thing measured."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from collections import defaultdict

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")


class ORel(C.Structure):
    _fields_ = [("ncols", C.c_uint32), ("rows", C.c_uint64), ("cols", C.c_void_p * 2),
                ("attr", C.c_uint32 * 2)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle")])
        _lib = C.CDLL(LIB)
        _lib.mhj_join.restype = C.c_int
        _lib.mhj_join.argtypes = [C.POINTER(ORel), C.c_uint32, C.c_int, C.c_int,
                                  C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_void_p), C.c_uint64,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
        for f in ("graph_triangles", "graph_four_cliques"):
            fn = getattr(_lib, f)
            fn.restype = C.c_uint64
            fn.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.POINTER(C.c_uint64)]
        _lib.oracle_tuple_hash.restype = C.c_uint64
        _lib.oracle_tuple_hash.argtypes = [C.c_void_p, C.c_int]
    return _lib


def mhj(relations, threads=0, general=False, materialize=False, cap=None):
    """Oracle chain/general multi-way join. relations: [(cols, attr_ids)]. Returns
    (count, checksum, tuples-or-None, partition_ms, probe_ms)."""
    keep = []
    arr = (ORel * len(relations))()
    attrs = set()
    for i, (cols, ids) in enumerate(relations):
        cs = [np.ascontiguousarray(c, dtype=np.int32) for c in cols]
        keep.append(cs)
        arr[i].ncols = len(cs)
        arr[i].rows = len(cs[0])
        for j, c in enumerate(cs):
            arr[i].cols[j] = c.ctypes.data
            arr[i].attr[j] = ids[j]
            attrs.add(ids[j])
    A = len(attrs)
    count = C.c_uint64()
    csum = C.c_uint64()
    pms, qms = C.c_double(), C.c_double()
    outp = None
    outs = None
    if materialize:
        cap = cap if cap is not None else 50_000_000
        outs = [np.zeros(cap, dtype=np.int32) for _ in range(A)]
        outp = (C.c_void_p * A)(*[o.ctypes.data for o in outs])
    rc = lib().mhj_join(arr, len(relations), threads, 1 if general else 0, C.byref(count),
                        C.byref(csum), outp, cap or 0, C.byref(pms), C.byref(qms))
    if rc < 0:
        raise RuntimeError(f"mhj_join failed: {rc}")
    tuples = None
    if materialize:
        n = min(count.value, cap)
        tuples = np.stack([o[:n] for o in outs], axis=1)
    return count.value, csum.value, tuples, pms.value, qms.value


def triangles(src, dst, threads=0):
    s = np.ascontiguousarray(src, dtype=np.int32)
    d = np.ascontiguousarray(dst, dtype=np.int32)
    cs = C.c_uint64()
    n = lib().graph_triangles(s.ctypes.data, d.ctypes.data, len(s), threads, C.byref(cs))
    return n, cs.value


def four_cliques(src, dst, threads=0):
    s = np.ascontiguousarray(src, dtype=np.int32)
    d = np.ascontiguousarray(dst, dtype=np.int32)
    cs = C.c_uint64()
    n = lib().graph_four_cliques(s.ctypes.data, d.ctypes.data, len(s), threads, C.byref(cs))
    return n, cs.value


# ---- pure-Python restatements (independent of the C oracle) for tiny inputs ----------------
M64 = (1 << 64) - 1


def mix64(z: int) -> int:
    """SplitMix64 finaliser (Steele, Lea & Flood, 2014)."""
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def tuple_hash(v) -> int:
    h = len(v)
    for i in range(0, len(v), 2):
        lo = v[i] & 0xFFFFFFFF
        hi = (v[i + 1] & 0xFFFFFFFF) if i + 1 < len(v) else 0
        h = mix64(h ^ (lo | (hi << 32)))
    return h


def checksum(tuples) -> int:
    s = 0
    for t in tuples:
        s = (s + tuple_hash([int(x) for x in t])) & M64
    return s


def brute_join(relations):
    """Nested-loop natural join of relations [(cols, attr_ids)] evaluated left to right with
    hash maps; returns the sorted list of result tuples in attribute order 0..A-1."""
    rows = []
    cols0, ids0 = relations[0]
    for r in range(len(cols0[0])):
        t = {}
        ok = True
        for c, a in zip(cols0, ids0):
            v = int(c[r])
            if a in t and t[a] != v:
                ok = False
            t[a] = v
        if ok:
            rows.append(t)
    for cols, ids in relations[1:]:
        index = defaultdict(list)
        for r in range(len(cols[0])):
            index[int(cols[0][r])].append(r)
        nxt = []
        for t in rows:
            for r in index.get(t[ids[0]], ()):
                u = dict(t)
                ok = True
                for c, a in zip(cols, ids):
                    v = int(c[r])
                    if a in u and u[a] != v:
                        ok = False
                        break
                    u[a] = v
                if ok:
                    nxt.append(u)
        rows = nxt
    A = len({a for _, ids in relations for a in ids})
    return sorted(tuple(t[a] for a in range(A)) for t in rows)


def brute_natural_join(relations):
    """Natural join of relations [(cols, attr_ids)] by nested loops in any attribute layout
    (products included); returns the sorted list of result tuples in attribute order 0..A-1."""
    rows = [dict()]
    for cols, ids in relations:
        tuples = [tuple(int(c[r]) for c in cols) for r in range(len(cols[0]))]
        nxt = []
        for t in rows:
            for vals in tuples:
                u = dict(t)
                ok = True
                for a, v in zip(ids, vals):
                    if a in u and u[a] != v:
                        ok = False
                        break
                    u[a] = v
                if ok:
                    nxt.append(u)
        rows = nxt
    A = len({a for _, ids in relations for a in ids})
    return sorted(tuple(t[a] for a in range(A)) for t in rows)


def sort_tuples(arr: np.ndarray) -> np.ndarray:
    if arr.size == 0:
        return arr
    return arr[np.lexsort(arr.T[::-1])]
