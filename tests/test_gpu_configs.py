"""GPU parity at the benchmark's configured sizes, against the CPU oracle (count + the
order-independent checksum), not only engine against engine:

* the 4-way chain (R1 64 Mi rows, R2..R4 16 Mi rows, Zipf s = 1.0 over 4 Mi keys) through the
  frequent-item path with ws=1 fib=1 dro=1, materialised (2^32 result tuples);
* triangles on R-MAT scale 22 (64 M oriented edges) through ALFTJ DFS (unit engine and work
  sharing, counted and written), LFTJ BFS and AMHJ TRI;
* 4-cliques on R-MAT scale 20 through AMHJ FOUR (plain and work sharing) and ALFTJ DFS (unit
  engine and work sharing);
* the out-of-core 3-way chain of 1.5 G tuples per relation, probe streamed from pinned host
  memory under a 48 GiB budget, equal to the same join with the probe in HBM; and the
  out-of-core engine against the oracle on a 16 Mi-row chain cut into many slices.
"""
import numpy as np
import pytest

import oracle_ffi as O
from accmultiwayjoins_amd import generators as G

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(1200)]


def test_config2_chain4_full_size_zipf1(ctx):
    rels = G.chain_relations([64 << 20, 16 << 20, 16 << 20, 16 << 20], 4 << 20, seed=42,
                             zipf_s=1.0, payload_range=1 << 31)
    want_cnt, want_cs, _, _, _ = O.mhj(rels)
    assert want_cnt > 4_000_000_000
    r = ctx.chain_join(rels, materialize=True, checksum=True, ws=1, fib=1, dro=1)
    try:
        assert (r.count, r.checksum) == (want_cnt, want_cs)
        assert r.rows_written == want_cnt and r.path == "fib"
    finally:
        ctx.free(r)


@pytest.fixture(scope="module")
def rmat22():
    s, d = G.rmat_oriented(22, 16, seed=42)
    n3, c3 = O.triangles(s, d)
    return s, d, n3, c3


TRI = [(0, 1), (1, 2), (0, 2)]
FOUR = [(0, 1), (1, 2), (2, 3), (0, 3), (0, 2), (1, 3)]


def test_config3_triangles_rmat22_alftj(ctx, rmat22):
    s, d, n3, c3 = rmat22
    rel = ctx.relation([s, d], [0, 1])
    t = ctx.trie_from_relation(rel)
    try:
        for opts in (dict(), dict(ws=1)):
            r = ctx.lftj([t] * 3, [0, 1, 2], "DFS", trie_attrs=TRI, materialize=False, **opts)
            assert (r.count, r.checksum) == (n3, c3), opts
        for opts in (dict(dro=1), dict(ws=1), dict(ws=1, dro=1)):
            w = ctx.lftj([t] * 3, [0, 1, 2], "DFS", trie_attrs=TRI, materialize=True, **opts)
            try:
                assert (w.count, w.checksum, w.rows_written) == (n3, c3, n3), opts
            finally:
                ctx.free(w)
        b = ctx.lftj([t] * 3, [0, 1, 2], "BFS", trie_attrs=TRI, materialize=False)
        assert (b.count, b.checksum) == (n3, c3)
    finally:
        t.close()
        rel.close()


def test_config3_triangles_rmat22_amhj_tri(ctx, rmat22):
    s, d, n3, c3 = rmat22
    rel = ctx.relation([s, d], [0, 1])
    ht = ctx.hash_table(rel, 0, sort_vals=True)
    try:
        r = ctx.execute_pattern(rel, ht, "TRI", materialize=False)
        assert (r.count, r.checksum) == (n3, c3)
        w = ctx.execute_pattern(rel, ht, "TRI", materialize=True, dro=1)
        try:
            assert (w.count, w.checksum, w.rows_written) == (n3, c3, n3)
        finally:
            ctx.free(w)
    finally:
        ht.close()
        rel.close()


def test_config4_four_cliques_rmat20(ctx):
    s, d = G.rmat_oriented(20, 16, seed=42)
    n4, c4 = O.four_cliques(s, d)
    assert n4 > 10_000_000_000
    rel = ctx.relation([s, d], [0, 1])
    ht = ctx.hash_table(rel, 0, sort_vals=True)
    t = ctx.trie_from_relation(rel)
    try:
        for opts in (dict(), dict(ws=1)):
            a = ctx.execute_pattern(rel, ht, "FOUR", materialize=False, **opts)
            assert (a.count, a.checksum) == (n4, c4), ("amhj", opts)
            l4 = ctx.lftj([t] * 6, [0, 1, 2, 3], "DFS", trie_attrs=FOUR, materialize=False,
                          **opts)
            assert (l4.count, l4.checksum) == (n4, c4), ("alftj", opts)
    finally:
        t.close()
        ht.close()
        rel.close()


def test_config4_out_of_core_chain_1_5g_rows(ctx):
    from accmultiwayjoins_amd.amj import pinned_empty
    rows = 1_500_000_000
    rels = G.chain_relations([rows, rows, rows], rows, seed=42)
    pcols = []
    for c in rels[0][0]:
        p = pinned_empty(len(c))
        p[:] = c
        pcols.append(p)
    R0 = ctx.relation(pcols, rels[0][1], host=True)
    R = [ctx.relation(c, ids) for c, ids in rels[1:]]
    del rels
    H = [ctx.hash_table(r) for r in R]
    R0d = None
    try:
        budget = ctx.allocated_bytes() + (48 << 30)  # streamed slices beside the tables
        r = ctx.execute(R0, H, materialize=False, ooc=1, hbm_budget_bytes=budget)
        assert r.path == "ooc" and r.detail > 1  # several slices streamed
        R0d = ctx.relation(pcols, [0, 1])
        ref = ctx.execute(R0d, H, materialize=False)
        assert (r.count, r.checksum) == (ref.count, ref.checksum)
        assert abs(r.count - rows) < 0.01 * rows  # E[count] = rows for uniform keys
    finally:
        for x in (R0d, R0, *R, *H):
            if x is not None:
                x.close()


def test_out_of_core_chain_many_slices_matches_oracle(ctx):
    from accmultiwayjoins_amd.amj import pinned_empty
    n = 16 << 20
    rels = G.chain_relations([n, n, n], n, seed=7, zipf_s=0.7)
    want_cnt, want_cs, _, _, _ = O.mhj(rels)
    pcols = []
    for c in rels[0][0]:
        p = pinned_empty(len(c))
        p[:] = c
        pcols.append(p)
    R0 = ctx.relation(pcols, rels[0][1], host=True)
    R = [ctx.relation(c, ids) for c, ids in rels[1:]]
    H = [ctx.hash_table(r) for r in R]
    try:
        for opts in (dict(), dict(ws=1)):
            r = ctx.execute(R0, H, materialize=False, ooc=1, ooc_slice_rows=1 << 20, **opts)
            assert (r.count, r.checksum) == (want_cnt, want_cs), opts
            assert r.detail == 16  # 16 slices of 1 Mi rows
        w = ctx.execute(R0, H, materialize=True, ooc=1, ooc_slice_rows=3_000_000)
        try:
            assert (w.count, w.checksum, w.rows_written) == (want_cnt, want_cs, want_cnt)
        finally:
            ctx.free(w)
    finally:
        for x in (R0, *R, *H):
            x.close()
