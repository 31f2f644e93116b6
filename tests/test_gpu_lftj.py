"""GPU parity of the leapfrog triejoin (ALFTJ DFS and BFS) and of the trie build.

Pins (no reference run): the triangle KAT of LFTJ-GPU/uniform_writer.cu:18-19 (5 results),
brute-force natural joins of small deduplicated relations, the CPU oracle's triangle and
4-clique counts and checksums on R-MAT graphs, closed forms C(n,3) / C(n,4) on K_n, the AMHJ
TRI/FOUR path on the same edges, and BFS == DFS.
"""
from math import comb

import numpy as np
import pytest

import oracle_ffi as O
from accmultiwayjoins_amd import AmjError
from accmultiwayjoins_amd import generators as G

pytestmark = pytest.mark.gpu

ALGOS = ["DFS", "BFS"]


def _tuples(r):
    return [tuple(int(x) for x in t) for t in O.sort_tuples(r.tuples())]


def _dedupe(cols):
    p = np.unique(np.stack([np.asarray(c, dtype=np.int64) for c in cols], axis=1), axis=0)
    return [p[:, 0].astype(np.int32), p[:, 1].astype(np.int32)]


def test_kat_triangle(ctx):
    from test_oracle import KAT_TRIANGLE, kat_relations
    tries = [ctx.trie(cols, ids) for cols, ids in kat_relations()]
    r = ctx.lftj(tries, [0, 1, 2], "DFS", fetch=True)
    assert r.count == 5 and r.path == "lftj-dfs"
    assert _tuples(r) == KAT_TRIANGLE
    assert r.checksum == O.checksum(KAT_TRIANGLE)
    b = ctx.lftj(tries, [0, 1, 2], "BFS")
    assert (b.count, b.checksum, b.path) == (5, r.checksum, "lftj-bfs")


def test_trie_levels_closed_form(ctx):
    rng = np.random.default_rng(3)
    n = 20000
    c0 = rng.integers(-50, 3000, n).astype(np.int32)  # negatives: no dense index
    c1 = rng.integers(-9, 40, n).astype(np.int32)
    c0[:500] = c0[500:1000]  # duplicate pairs
    c1[:500] = c1[500:1000]
    for cols in ([c0, c1], [np.abs(c0), c1]):  # second: dense level-0 index
        t = ctx.trie(cols)
        d0, off, d1 = t.levels()
        u = _dedupe(cols)
        want_d0, first = np.unique(u[0], return_index=True)
        assert np.array_equal(d0, want_d0)
        assert np.array_equal(off, np.append(first, len(u[0])).astype(np.uint32))
        assert np.array_equal(d1, u[1])
        assert t.max_children == int(np.diff(off).max())
        t.close()


def test_trie_empty_and_single(ctx):
    e = ctx.trie([np.zeros(0, np.int32), np.zeros(0, np.int32)])
    assert (e.level0, e.children) == (0, 0)
    one = ctx.trie([np.array([7], np.int32), np.array([7], np.int32)])
    assert (one.level0, one.children, one.max_children) == (1, 1, 1)
    for algo in ALGOS:
        r = ctx.lftj([e, one], [0, 1], algo)
        assert r.count == 0
        r = ctx.lftj([one, one, one], [0, 1, 2], algo,
                     trie_attrs=[(0, 1), (1, 2), (0, 2)])  # (7,7,7)
        assert r.count == 1 and r.checksum == O.checksum([(7, 7, 7)])


QUERIES = {
    "path": [(0, 1), (1, 2)],
    "triangle": [(0, 1), (1, 2), (0, 2)],
    "star": [(0, 1), (0, 2), (0, 3)],
    "cycle4": [(0, 1), (1, 2), (2, 3), (0, 3)],
    "clique4": [(0, 1), (1, 2), (2, 3), (0, 3), (0, 2), (1, 3)],
    "product": [(0, 1), (2, 3)],  # attribute 2 has no parent: level-0 candidates
    "two": [(0, 1), (0, 1)],
}


@pytest.mark.parametrize("name", sorted(QUERIES))
@pytest.mark.parametrize("algo", ALGOS)
def test_small_queries_equal_brute_force(ctx, name, algo):
    rng = np.random.default_rng(len(name))
    rels = []
    for x, y in QUERIES[name]:
        n = int(rng.integers(50, 300))
        cols = _dedupe([rng.integers(0, 12, n), rng.integers(0, 12, n)])
        rels.append((cols, [x, y]))
    want = O.brute_natural_join(rels)
    A = 1 + max(max(q) for q in QUERIES[name])
    tries = [ctx.trie(c, ids) for c, ids in rels]
    r = ctx.lftj(tries, list(range(A)), algo, fetch=True)
    assert r.count == len(want)
    assert r.checksum == O.checksum(want)
    if algo == "DFS":
        assert _tuples(r) == want


@pytest.mark.parametrize("scale", [8, 11, 14])
@pytest.mark.parametrize("algo", ALGOS)
def test_tri_four_match_graph_oracle(ctx, scale, algo):
    s, d = G.rmat_oriented(scale, 16, seed=42)
    n3, c3 = O.triangles(s, d)
    r3 = ctx.lftj_pattern(s, d, "TRI", algo, materialize=False)
    assert (r3.count, r3.checksum) == (n3, c3)
    if scale <= 11:
        n4, c4 = O.four_cliques(s, d)
        r4 = ctx.lftj_pattern(s, d, "FOUR", algo, materialize=False)
        assert (r4.count, r4.checksum) == (n4, c4)


def test_tri_materialized_and_amhj_agree(ctx):
    s, d = G.rmat_oriented(12, 16, seed=7)
    n3, c3 = O.triangles(s, d)
    r = ctx.lftj_pattern(s, d, "TRI", "DFS", fetch=True)
    assert r.count == n3 and r.rows_written == n3
    t = r.tuples()
    assert O.checksum(t) == c3
    h = ctx.pattern_join(s, d, "TRI", materialize=False)
    assert (h.count, h.checksum) == (n3, c3)
    h4 = ctx.pattern_join(s, d, "FOUR", materialize=False)
    l4 = ctx.lftj_pattern(s, d, "FOUR", "DFS", materialize=False)
    assert (h4.count, h4.checksum) == (l4.count, l4.checksum)


@pytest.mark.parametrize("n", [5, 30, 64, 65])
def test_complete_graph_closed_forms(ctx, n):
    a, b = np.triu_indices(n, 1)
    s, d = a.astype(np.int32), b.astype(np.int32)
    for algo in ALGOS:
        assert ctx.lftj_pattern(s, d, "TRI", algo, materialize=False).count == comb(n, 3)
        assert ctx.lftj_pattern(s, d, "FOUR", algo, materialize=False).count == comb(n, 4)


def test_duplicate_edges_are_ignored(ctx):
    s, d = G.rmat_oriented(9, 16, seed=1)
    n3, c3 = O.triangles(s, d)
    s2, d2 = np.concatenate([s, s[::3]]), np.concatenate([d, d[::3]])
    r = ctx.lftj_pattern(s2, d2, "TRI", "DFS", materialize=False)
    assert (r.count, r.checksum) == (n3, c3)


def test_attr_order_errors(ctx):
    t01 = ctx.trie([np.array([1, 2], np.int32), np.array([2, 3], np.int32)], (0, 1))
    t21 = ctx.trie([np.array([1, 2], np.int32), np.array([2, 3], np.int32)], (2, 1))
    with pytest.raises(AmjError, match="follow attr_order"):
        ctx.lftj([t01, t21], [0, 1, 2])
    with pytest.raises(AmjError, match="not in attr_order"):
        ctx.lftj([t01], [0, 2])
    with pytest.raises(AmjError, match="2..4 attributes"):
        ctx.lftj([t01], [0, 1, 2, 3, 4])
    with pytest.raises(AmjError, match="in no relation"):
        ctx.lftj([t01], [0, 1, 2])
    # the context stays usable after errors
    assert ctx.lftj([t01, t01], [0, 1]).count == 2


@pytest.mark.parametrize("query", ["TRI", "FOUR"])
@pytest.mark.parametrize("split", [0, 64], ids=["split1024", "split64"])
@pytest.mark.parametrize("mode", ["count", "write", "dro", "dro-overrun"])
def test_work_sharing_dfs_matches_oracle(ctx, query, split, mode):
    # ws=1: persistent waves, candidate runs longer than the split shared as tasks
    # (ALFTJ.cuh:464-499). Count + checksum equal the oracle's; written results are the same
    # multiset of tuples as the unit engine's (direct output: the row order is not fixed).
    s, d = G.rmat_oriented(11, 16, seed=3)
    want = O.triangles(s, d) if query == "TRI" else O.four_cliques(s, d)
    kw = dict(ws=1, ws_split_threshold=split)
    if mode == "count":
        r = ctx.lftj_pattern(s, d, query, "DFS", materialize=False, **kw)
        assert (r.count, r.checksum) == want and r.rows_written == 0
    else:
        if mode != "write":
            kw.update(dro=1, dro_capacity_rows=16 if mode == "dro-overrun" else 0)
        r = ctx.lftj_pattern(s, d, query, "DFS", fetch=True, **kw)
        assert (r.count, r.checksum) == want and r.rows_written == r.count
        base = ctx.lftj_pattern(s, d, query, "DFS", fetch=True)
        assert _tuples(r) == _tuples(base)
        if mode == "dro-overrun":
            assert r.passes == 2  # the first pass overran its 16 rows and counted everything
    assert r.path == "lftj-dfs"
    if split == 64:
        assert r.detail > 0  # tasks were shared between waves


def test_work_sharing_dfs_rmat18_matches_oracle(ctx):
    # the skewed R-MAT 18 graph (hub vertices with thousands of oriented edges): default split,
    # heaviest-first unit order, triangles counted and written, 4-cliques counted
    s, d = G.rmat_oriented(18, 16, seed=42)
    n3, c3 = O.triangles(s, d)
    r = ctx.lftj_pattern(s, d, "TRI", "DFS", materialize=False, ws=1)
    assert (r.count, r.checksum) == (n3, c3)
    assert r.detail > 0
    w = ctx.lftj_pattern(s, d, "TRI", "DFS", materialize=True, ws=1)
    assert (w.count, w.checksum, w.rows_written) == (n3, c3, n3)
    md5 = ctx.md5(w)
    ctx.free(w)
    u = ctx.lftj_pattern(s, d, "TRI", "DFS", materialize=True)
    assert ctx.md5(u) == md5  # the same multiset of tuples as the unit engine
    ctx.free(u)
    h4 = ctx.pattern_join(s, d, "FOUR", materialize=False)
    l4 = ctx.lftj_pattern(s, d, "FOUR", "DFS", materialize=False, ws=1)
    assert (l4.count, l4.checksum) == (h4.count, h4.checksum)


@pytest.mark.parametrize("query", ["TRI", "FOUR"])
@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("slice_rows", [37, 300, 5000])
def test_out_of_core_range_combinations(ctx, query, algo, slice_rows):
    s, d = G.rmat_oriented(10, 16, seed=5)
    want = O.triangles(s, d) if query == "TRI" else O.four_cliques(s, d)
    pattern = {"TRI": [(0, 1), (1, 2), (0, 2)],
               "FOUR": [(0, 1), (1, 2), (2, 3), (0, 3), (0, 2), (1, 3)]}[query]
    na = 3 if query == "TRI" else 4
    t = ctx.trie([s, d])
    # device trie sliced (opts.ooc), then the same trie offloaded to host memory
    r = ctx.lftj([t] * len(pattern), list(range(na)), algo, trie_attrs=pattern, ooc=1,
                 ooc_slice_rows=slice_rows, materialize=False)
    assert (r.count, r.checksum) == want and r.path == "lftj-ooc"
    t.offload()
    r2 = ctx.lftj([t] * len(pattern), list(range(na)), algo, trie_attrs=pattern,
                  ooc_slice_rows=slice_rows, materialize=False)
    assert (r2.count, r2.checksum) == want and r2.detail == r.detail
    assert r.detail > 1 or slice_rows >= t.level0  # several range combinations ran
    t3 = ctx.trie([s, d])
    t3.offload()
    r3 = ctx.lftj([t3] * len(pattern), list(range(na)), algo, trie_attrs=pattern,
                  materialize=False)  # ranges from the budget: everything fits, one combination
    assert (r3.count, r3.checksum) == want


def test_out_of_core_multiway_kat(ctx):
    from test_oracle import KAT_TRIANGLE, kat_relations
    tries = [ctx.trie(cols, ids) for cols, ids in kat_relations()]
    for t in tries:
        t.offload()
    r = ctx.lftj(tries, [0, 1, 2], "DFS", ooc_slice_rows=1, materialize=False)
    assert (r.count, r.checksum) == (5, O.checksum(KAT_TRIANGLE))


@pytest.mark.parametrize("query", ["TRI", "FOUR"])
@pytest.mark.parametrize("cap", [0, 16], ids=["sampled", "overrun-rerun"])
def test_dfs_direct_output_same_multiset(ctx, query, cap):
    # dro=1: one pass through per-wave buffers and atomic reservations; capacity 16 overruns
    # and runs again into an exact output. Same count, checksum and multiset of tuples.
    s, d = G.rmat_oriented(11, 16, seed=3)
    want = ctx.lftj_pattern(s, d, query, "DFS", fetch=True)
    r = ctx.lftj_pattern(s, d, query, "DFS", fetch=True, dro=1, dro_capacity_rows=cap)
    assert r.count > 16
    assert (r.count, r.checksum) == (want.count, want.checksum)
    assert r.rows_written == r.count
    assert _tuples(r) == _tuples(want)
    if cap:
        assert r.passes == 2  # the first pass overran its output


def test_heaviest_units_first_order_same_results(ctx):
    # R-MAT 18 has 2,716 units of 64 roots: enough for the DFS engine to launch its units in
    # descending weight order (count, write and direct passes). Results do not depend on it.
    s, d = G.rmat_oriented(18, 16, seed=42)
    assert len(np.unique(s)) > 2048 * 64
    n3, c3 = O.triangles(s, d)
    r = ctx.lftj_pattern(s, d, "TRI", "DFS", materialize=False)
    assert (r.count, r.checksum) == (n3, c3)
    w = ctx.lftj_pattern(s, d, "TRI", "DFS", materialize=True)
    assert (w.count, w.checksum, w.rows_written) == (n3, c3, n3)
    md5 = ctx.md5(w)
    ctx.free(w)
    x = ctx.lftj_pattern(s, d, "TRI", "DFS", materialize=True, dro=1)
    assert (x.count, x.checksum, x.rows_written) == (n3, c3, n3)
    assert ctx.md5(x) == md5  # the same multiset of tuples (an order-free fingerprint)
    ctx.free(x)
    h4 = ctx.pattern_join(s, d, "FOUR", materialize=False)
    l4 = ctx.lftj_pattern(s, d, "FOUR", "DFS", materialize=False)
    assert (l4.count, l4.checksum) == (h4.count, h4.checksum)
