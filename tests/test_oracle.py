"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the GPU checker:
golden counts of the reference's own MHJ-CPU binary, brute force, closed forms, and the
reference's hand-made triangle fixture."""
import json
import os
from math import comb

import numpy as np
import pytest

import oracle_ffi as O
from accmultiwayjoins_amd import generators as G

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mhj_cpu_counts.json")

# LFTJ-GPU/uniform_writer.cu:18-19 (T_A/T_B), attributes per uniform_writer.cu:66-73:
# table 0 -> (0,1), table 1 -> (1,2), table 2 -> (2,0) stored as (0,2).
T_A = [[1, 1, 1, 2, 2, 9], [2, 2, 3, 3, 5, 12], [1, 1, 2, 2, 2, 22]]
T_B = [[2, 3, 5, 2, 5, 10], [3, 5, 2, 5, 5, 17], [2, 3, 2, 3, 5, 98]]
KAT_TRIANGLE = [(1, 2, 3), (1, 3, 2), (2, 2, 3), (2, 2, 5), (2, 5, 5)]


def kat_relations():
    return [([np.array(T_A[0]), np.array(T_B[0])], [0, 1]),
            ([np.array(T_A[1]), np.array(T_B[1])], [1, 2]),
            ([np.array(T_A[2]), np.array(T_B[2])], [0, 2])]


def test_splitmix64_published_vector():
    # SplitMix64 (Vigna) seeded with 1234567: first output of next() is mix64(seed)
    assert O.mix64(1234567) == 6457827717110365317
    assert O.mix64((1234567 + 0x9E3779B97F4A7C15) & O.M64) == 3203168211198807973


def test_tuple_hash_c_matches_python():
    rng = np.random.default_rng(0)
    for A in range(1, 7):
        v = rng.integers(-2**31, 2**31, size=A, dtype=np.int64).astype(np.int32)
        assert O.lib().oracle_tuple_hash(v.ctypes.data, A) == O.tuple_hash([int(x) for x in v])


def test_kat_triangle_brute_force():
    assert O.brute_join(kat_relations()) == KAT_TRIANGLE


def test_kat_triangle_oracle_general_compare():
    cnt, cs, tup, _, _ = O.mhj(kat_relations(), general=True, materialize=True, cap=100)
    assert cnt == 5
    assert [tuple(r) for r in O.sort_tuples(tup)] == KAT_TRIANGLE
    assert cs == O.checksum(KAT_TRIANGLE)


def test_single_key_probe_differs_on_triangle_kat():
    # The reference's single-join-key probe ignores the second join column of R2 (the
    # single_join_key=true branch used for every query by exec_AMHJ.cu:165-193), so on the KAT
    # it returns every 2-path instead of the 5 triangles.
    cnt, _, _, _, _ = O.mhj(kat_relations(), general=False)
    assert cnt > 5


@pytest.mark.parametrize("rec", json.load(open(GOLDEN))["datasets"], ids=lambda r: r["name"])
def test_oracle_matches_reference_mhj_cpu(rec):
    rels = G.chain_relations(rec["sizes"], rec["key_range"], seed=rec["seed"],
                             zipf_s=rec["zipf_s"], payload_range=rec["payload_range"])
    cnt, cs, _, _, _ = O.mhj(rels)
    assert cnt == rec["count"]
    cnt1, cs1, _, _, _ = O.mhj(rels, threads=1)
    assert (cnt1, cs1) == (cnt, cs)  # order-independent checksum


@pytest.mark.parametrize("case", range(12))
def test_oracle_vs_brute_force(case):
    rng = np.random.default_rng(100 + case)
    n_rel = 2 + case % 4
    sizes = [int(rng.integers(0 if case == 5 else 1, 60)) for _ in range(n_rel)]
    keys = int(rng.integers(1, 12))
    lo = -5 if case % 3 == 0 else 0  # negative keys too
    rels = []
    for i, n in enumerate(sizes):
        cols = [rng.integers(lo, lo + keys, size=n).astype(np.int32) for _ in range(2)]
        rels.append((cols, [i, i + 1]))
    want = O.brute_join(rels)
    cnt, cs, tup, _, _ = O.mhj(rels, materialize=True, cap=max(1, len(want)))
    assert cnt == len(want)
    assert [tuple(r) for r in O.sort_tuples(tup)] == want
    assert cs == O.checksum(want)


def test_oracle_general_compare_star_and_cycle():
    rng = np.random.default_rng(7)
    a = [rng.integers(0, 6, 40).astype(np.int32) for _ in range(2)]
    b = [rng.integers(0, 6, 40).astype(np.int32) for _ in range(2)]
    c = [rng.integers(0, 6, 40).astype(np.int32) for _ in range(2)]
    for rels in ([(a, [0, 1]), (b, [0, 2]), (c, [0, 3])],      # star on attribute 0
                 [(a, [0, 1]), (b, [1, 2]), (c, [2, 0])]):     # cycle; last table compares
        want = O.brute_join(rels)
        cnt, cs, tup, _, _ = O.mhj(rels, general=True, materialize=True, cap=max(1, len(want)))
        assert cnt == len(want)
        assert [tuple(r) for r in O.sort_tuples(tup)] == want


def _complete_oriented(n):
    s, d = zip(*[(i, j) for i in range(n) for j in range(i + 1, n)])
    return np.array(s, dtype=np.int32), np.array(d, dtype=np.int32)


@pytest.mark.parametrize("n", [3, 4, 7, 12])
def test_graph_oracle_closed_forms(n):
    s, d = _complete_oriented(n)
    assert O.triangles(s, d)[0] == comb(n, 3)
    assert O.four_cliques(s, d)[0] == comb(n, 4)


def _brute_cliques(s, d, k):
    adj = set(zip(s.tolist(), d.tolist()))
    verts = sorted(set(s.tolist()) | set(d.tolist()))
    out = []
    if k == 3:
        for (a, b) in adj:
            for c in verts:
                if (b, c) in adj and (a, c) in adj:
                    out.append((a, b, c))
    else:
        for (a, b) in adj:
            for c in verts:
                if (b, c) in adj and (a, c) in adj:
                    for e in verts:
                        if (c, e) in adj and (a, e) in adj and (b, e) in adj:
                            out.append((a, b, c, e))
    return sorted(out)


def test_graph_oracle_vs_brute_force_and_mhj_general():
    s0, d0 = G.rmat_edges(7, 8, seed=3)
    s, d = G.orient(s0, d0, 1 << 7)
    tri = _brute_cliques(s, d, 3)
    four = _brute_cliques(s, d, 4)
    n3, c3 = O.triangles(s, d)
    n4, c4 = O.four_cliques(s, d)
    assert (n3, c3) == (len(tri), O.checksum(tri))
    assert (n4, c4) == (len(four), O.checksum(four))
    # the same triangles through the general-compare hash join (A(0,1),B(1,2),C(0,2))
    rels = [([s, d], [0, 1]), ([s, d], [1, 2]), ([s, d], [0, 2])]
    cnt, cs, _, _, _ = O.mhj(rels, general=True)
    assert (cnt, cs) == (n3, c3)
    # and the 4-clique query A(0,1),B(1,2),C(2,3),D(0,3),E(0,2),F(1,3)
    rels4 = [([s, d], [0, 1]), ([s, d], [1, 2]), ([s, d], [2, 3]), ([s, d], [0, 3]),
             ([s, d], [0, 2]), ([s, d], [1, 3])]
    cnt4, cs4, _, _, _ = O.mhj(rels4, general=True)
    assert (cnt4, cs4) == (n4, c4)
