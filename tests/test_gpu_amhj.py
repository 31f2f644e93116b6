"""GPU parity of the AMHJ multi-way hash join (libamj through its C ABI) against the CPU
oracle, the reference MHJ-CPU golden counts and brute force. Integer results are bit-exact:
count, order-independent checksum and, for small inputs, the sorted tuples."""
import json
import os

import numpy as np
import pytest

import oracle_ffi as O
from accmultiwayjoins_amd import AmjError
from accmultiwayjoins_amd import generators as G

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "mhj_cpu_counts.json")))

PATHS = [dict(), dict(fib=1), dict(fib=1, fib_suffix=1), dict(fib=1, fib_suffix=2),
         dict(fib=1, probe_order=2),
         dict(ws=1), dict(ws=1, ws_split_threshold=64), dict(dro=1),
         dict(dro=1, dro_capacity_rows=64), dict(ws=1, dro=1, ws_split_threshold=128),
         dict(engine="tpt"), dict(engine="open-addr"), dict(engine="pw")]
PATH_IDS = ["plain", "fib", "fib-nosuffix", "fib-deepsuffix", "fib-grouped", "ws", "ws-split64",
            "dro",
            "dro-overflow", "ws-dro", "thread-per-tuple", "open-addressing", "pairwise"]


def _rels(rec):
    return G.chain_relations(rec["sizes"], rec["key_range"], seed=rec["seed"],
                             zipf_s=rec["zipf_s"], payload_range=rec["payload_range"])


@pytest.mark.parametrize("opts", PATHS, ids=PATH_IDS)
@pytest.mark.parametrize("rec", GOLDEN["datasets"], ids=lambda r: r["name"])
def test_chain_matches_reference_counts_and_oracle_checksum(ctx, rec, opts):
    rels = _rels(rec)
    want_cnt, want_cs, _, _, _ = O.mhj(rels)
    assert want_cnt == rec["count"]
    heavy = (rec["zipf_s"] or 0) >= 1.0 and max(rec["sizes"]) >= 1 << 20
    if opts.get("engine") == "open-addr" and heavy:
        # linear probing walks a Zipf head key's whole cluster per probe row (about a minute
        # per pass on the 1 M-row set): count and checksum in one pass only
        r2 = ctx.chain_join(rels, materialize=False, **opts)
        assert (r2.count, r2.checksum) == (want_cnt, want_cs)
        return
    r = ctx.chain_join(rels, materialize=True, fetch=False, **opts)
    try:
        assert r.count == rec["count"]
        assert r.checksum == want_cs
        assert r.rows_written == r.count
    finally:
        ctx.free(r)
    r2 = ctx.chain_join(rels, materialize=False, **opts)
    assert (r2.count, r2.checksum) == (want_cnt, want_cs)
    assert r2.rows_written == 0


def _random_chain(rng, n_rel, max_rows, keys, lo=0, empty=None):
    rels = []
    for i in range(n_rel):
        n = 0 if empty == i else int(rng.integers(1, max_rows))
        cols = [rng.integers(lo, lo + keys, size=n).astype(np.int32) for _ in range(2)]
        rels.append((cols, [i, i + 1]))
    return rels


@pytest.mark.parametrize("opts", PATHS, ids=PATH_IDS)
@pytest.mark.parametrize("case", range(16))
def test_chain_small_tuples_equal_brute_force(ctx, case, opts):
    rng = np.random.default_rng(1000 + case)
    n_rel = 2 + case % 4  # 2..5 relations: at most 6 result attributes (MHJ-GPU/conf.h:20)
    rels = _random_chain(rng, n_rel, 300, keys=int(rng.integers(2, 40)),
                         lo=-20 if case % 4 == 1 else 0, empty=(1 if case == 7 else None))
    want = O.brute_join(rels)
    r = ctx.chain_join(rels, fetch=True, **opts)
    assert r.count == len(want)
    got = O.sort_tuples(r.tuples()) if r.count else np.zeros((0, n_rel + 1), np.int32)
    assert [tuple(int(x) for x in t) for t in got] == want
    assert r.checksum == O.checksum(want)


def test_fib_and_plain_write_identical_arrays(ctx):
    # both paths emit results in the same depth-first order over probe rows
    # 14 M results (Zipf 1.0 keys); each multiset comparison sorts them on the host
    rels = G.chain_relations([4000, 5000, 7000, 3000], 300, seed=3, zipf_s=1.0)
    a = ctx.chain_join(rels, fetch=True)
    a_sorted = O.sort_tuples(a.tuples())
    for suffix in (0, 1, 2):
        b = ctx.chain_join(rels, fetch=True, fib=1, out_chunk_rows=1000, fib_suffix=suffix,
                           probe_order=1)
        assert a.path == "plain" and b.path == "fib"
        assert a.count == b.count > 0
        for ca, cb in zip(a.cols, b.cols):
            assert np.array_equal(ca, cb)
        # grouped by join key the rows come out in another order, the same multiset
        c = ctx.chain_join(rels, fetch=True, fib=1, out_chunk_rows=1000, fib_suffix=suffix,
                           probe_order=2)
        assert (c.count, c.checksum) == (a.count, a.checksum)
        assert np.array_equal(O.sort_tuples(c.tuples()), a_sorted)


def test_window_writer_loops_over_windows(ctx):
    # 64-row output windows outnumber the writer's waves (its grid is capped), so every wave
    # writes many windows; the output must equal the unit engine's row for row as a multiset
    rels = G.chain_relations([1 << 22, 1 << 20, 1 << 20, 1 << 20], 1 << 18, seed=11,
                             zipf_s=0.5, payload_range=1 << 31)
    R = [ctx.relation(c, ids) for c, ids in rels]
    H = [ctx.hash_table(r) for r in R[1:]]
    want = ctx.execute(R[0], H, materialize=True)
    assert want.path == "plain" and want.count > 64 * 4 * 65536
    want_md5 = ctx.md5(want)
    ctx.free(want)
    for window in (64, 192):
        r = ctx.execute(R[0], H, fib=1, out_chunk_rows=window)
        assert r.path == "fib"
        assert (r.count, r.checksum) == (want.count, want.checksum)
        assert ctx.md5(r) == want_md5
        ctx.free(r)


def test_heavy_skew_single_key(ctx):
    # one build key holding every row: runs of 200k rows, results 50 * 200k * 3
    n = 200_000
    probe = ([np.arange(50, dtype=np.int32), np.full(50, 7, np.int32)], [0, 1])
    b1 = ([np.full(n, 7, np.int32), np.arange(n, dtype=np.int32) % 3], [1, 2])
    b2 = ([np.arange(3, dtype=np.int32), np.array([10, 11, 12], np.int32)], [2, 3])
    rels = [probe, b1, b2]
    want_cnt, want_cs, _, _, _ = O.mhj(rels)
    assert want_cnt == 50 * n
    for opts in PATHS:
        r = ctx.chain_join(rels, materialize=True, **opts)
        assert (r.count, r.checksum) == (want_cnt, want_cs)
        ctx.free(r)


def test_kat_triangle_through_chain_join(ctx):
    from test_oracle import KAT_TRIANGLE, kat_relations
    r = ctx.chain_join(kat_relations(), fetch=True)
    assert [tuple(int(x) for x in t) for t in O.sort_tuples(r.tuples())] == KAT_TRIANGLE


def test_star_and_cycle_general_compare(ctx):
    rng = np.random.default_rng(5)
    a = [rng.integers(0, 9, 500).astype(np.int32) for _ in range(2)]
    b = [rng.integers(0, 9, 400).astype(np.int32) for _ in range(2)]
    c = [rng.integers(0, 9, 300).astype(np.int32) for _ in range(2)]
    for rels in ([(a, [0, 1]), (b, [0, 2]), (c, [0, 3])],
                 [(a, [0, 1]), (b, [1, 2]), (c, [0, 2])]):
        want = O.brute_join(rels)
        r = ctx.chain_join(rels, fetch=True)
        assert r.count == len(want)
        assert [tuple(int(x) for x in t) for t in O.sort_tuples(r.tuples())] == want


@pytest.mark.parametrize("scale", [8, 11])
def test_pattern_tri_four_match_graph_oracle(ctx, scale):
    s, d = G.rmat_oriented(scale, 16, seed=42)
    n3, c3 = O.triangles(s, d)
    n4, c4 = O.four_cliques(s, d)
    r3 = ctx.pattern_join(s, d, "TRI", materialize=False)
    assert (r3.count, r3.checksum) == (n3, c3)
    r4 = ctx.pattern_join(s, d, "FOUR", materialize=False)
    assert (r4.count, r4.checksum) == (n4, c4)
    r3m = ctx.pattern_join(s, d, "TRI", fetch=True)
    assert r3m.count == n3
    assert O.checksum(r3m.tuples()) == c3


def test_thread_per_tuple_baseline_patterns_and_rules(ctx):
    # the baseline engine runs TRI / FOUR plans (compared columns) to the same results, and
    # refuses to be combined with the wave engines' switches
    s, d = G.rmat_oriented(10, 16, seed=42)
    n3, c3 = O.triangles(s, d)
    n4, c4 = O.four_cliques(s, d)
    r3 = ctx.pattern_join(s, d, "TRI", fetch=True, engine="tpt")
    assert r3.path == "tpt" and r3.count == n3 and r3.checksum == c3
    assert O.checksum(r3.tuples()) == c3
    r4 = ctx.pattern_join(s, d, "FOUR", materialize=False, engine="tpt")
    assert (r4.count, r4.checksum) == (n4, c4) and r4.passes == 1
    rels = G.chain_relations([5000, 3000, 3000], 500, seed=9)
    with pytest.raises(AmjError, match="runs? alone"):
        ctx.chain_join(rels, engine="tpt", ws=1)
    empty = [(c if i != 1 else [x[:0] for x in c], ids) for i, (c, ids) in enumerate(rels)]
    r = ctx.chain_join(empty, engine="tpt")
    assert (r.count, r.checksum, r.rows_written) == (0, 0, 0)


def test_open_addressing_and_pairwise_engines_rules(ctx):
    # open addressing runs TRI / FOUR (compared columns) to the oracle's results; the pairwise
    # engine joins chains only and says so; both refuse the wave engines' switches
    s, d = G.rmat_oriented(10, 16, seed=42)
    n3, c3 = O.triangles(s, d)
    n4, c4 = O.four_cliques(s, d)
    r3 = ctx.pattern_join(s, d, "TRI", fetch=True, engine="open-addr")
    assert r3.path == "open-addr" and (r3.count, r3.checksum) == (n3, c3)
    assert O.checksum(r3.tuples()) == c3
    r4 = ctx.pattern_join(s, d, "FOUR", materialize=False, engine="open-addr")
    assert (r4.count, r4.checksum) == (n4, c4)
    with pytest.raises(AmjError, match="pairwise"):
        ctx.pattern_join(s, d, "TRI", engine="pw")
    rels = G.chain_relations([5000, 3000, 3000, 2000], 500, seed=9, zipf_s=1.0)
    want_cnt, want_cs, _, _, _ = O.mhj(rels)
    r = ctx.chain_join(rels, engine="pw", fetch=True)
    assert r.path == "pw" and (r.count, r.checksum) == (want_cnt, want_cs)
    assert O.checksum(r.tuples()) == want_cs
    for eng in ("open-addr", "pw"):
        with pytest.raises(AmjError, match="runs? alone"):
            ctx.chain_join(rels, engine=eng, dro=1)
    empty = [(c if i != 2 else [x[:0] for x in c], ids) for i, (c, ids) in enumerate(rels)]
    for eng in ("open-addr", "pw"):
        r = ctx.chain_join(empty, engine=eng)
        assert (r.count, r.checksum, r.rows_written) == (0, 0, 0)


def test_resident_relations_and_tables_reused(ctx):
    rels = G.chain_relations([5000, 3000, 3000], 500, seed=9)
    want_cnt, want_cs, _, _, _ = O.mhj(rels)
    R = [ctx.relation(c, ids) for c, ids in rels]
    H = [ctx.hash_table(r) for r in R[1:]]
    for _ in range(3):
        r = ctx.execute(R[0], H, materialize=False)
        assert (r.count, r.checksum) == (want_cnt, want_cs)


def test_too_many_attributes_rejected(ctx):
    rng = np.random.default_rng(1)
    rels = _random_chain(rng, 8, 50, keys=5)  # 9 attributes > AMJ_MAX_ATTRS (8)
    with pytest.raises(AmjError) as e:
        ctx.chain_join(rels)
    assert "too many attributes" in str(e.value)


@pytest.mark.parametrize("n_rel", [6, 7, 8])
@pytest.mark.parametrize("opts", [dict(), dict(fib=1), dict(ws=1, dro=1, ws_split_threshold=64),
                                  dict(engine="tpt"), dict(engine="pw")],
                         ids=["plain", "fib", "ws-dro", "thread-per-tuple", "pairwise"])
def test_deep_chains_up_to_seven_build_tables(ctx, n_rel, opts):
    # the reference's TPC-H Q8 setting (MHJ-GPU/conf.h:16-20): 7 build tables, 8 attributes
    rels = G.chain_relations([3000] + [2000] * (n_rel - 1), 1500, seed=50 + n_rel, zipf_s=0.6)
    if n_rel == 8:  # 7 build tables within 8 attributes: the last table holds its key only
        cols, ids = rels[-1]
        rels[-1] = (cols[:1], ids[:1])
    n_attrs = len({a for _, ids in rels for a in ids})
    assert n_attrs <= 8
    want_cnt, want_cs, _, _, _ = O.mhj(rels)
    r = ctx.chain_join(rels, fetch=True, **opts)
    assert (r.count, r.checksum) == (want_cnt, want_cs) and r.count > 0
    assert len(r.cols) == n_attrs
    assert O.checksum(r.tuples()) == want_cs


def test_errors_are_reported_not_fatal(ctx):
    # hash attribute of the second build table is never bound before it
    rels = [([np.arange(4, dtype=np.int32)] * 2, [0, 1]),
            ([np.arange(4, dtype=np.int32)] * 2, [5, 6])]
    with pytest.raises(AmjError) as e:
        ctx.chain_join(rels)
    assert "not bound" in str(e.value)


@pytest.mark.parametrize("opts", [dict(ws=1), dict(ws=1, ws_split_threshold=64), dict(dro=1),
                                  dict(ws=1, dro=1, dro_capacity_rows=1000,
                                       ws_split_threshold=64)],
                         ids=["ws", "ws-split64", "dro", "ws-dro-overflow"])
def test_dynamic_paths_on_skew_and_patterns(ctx, opts):
    # one key of the probe side hits a long run in every table: tasks get shared
    n = 4096
    a = np.zeros(n, np.int32)
    a[::7] = np.arange(len(a[::7])) % 50
    rels = [([np.arange(n, dtype=np.int32), a], [0, 1]),
            ([np.zeros(3000, np.int32), np.arange(3000, dtype=np.int32) % 40], [1, 2]),
            ([np.arange(3000, dtype=np.int32) % 40, np.arange(3000, dtype=np.int32)], [2, 3])]
    want_cnt, want_cs, _, _, _ = O.mhj(rels)
    r = ctx.chain_join(rels, materialize=True, **opts)
    assert (r.count, r.checksum) == (want_cnt, want_cs)
    assert r.path in ("ws", "dro", "ws-dro")
    if opts.get("ws_split_threshold") == 64:
        assert r.detail > 0  # tasks were shared
    if opts.get("dro_capacity_rows"):
        assert r.passes == 2  # the estimate was overrun and the pass ran again
    ctx.free(r)
    s, d = G.rmat_oriented(10, 16, seed=42)
    n3, c3 = O.triangles(s, d)
    n4, c4 = O.four_cliques(s, d)
    r3 = ctx.pattern_join(s, d, "TRI", materialize=True, **opts)
    assert (r3.count, r3.checksum) == (n3, c3)
    ctx.free(r3)
    r4 = ctx.pattern_join(s, d, "FOUR", materialize=False, **opts)
    assert (r4.count, r4.checksum) == (n4, c4)


def test_ws_slow_pass_with_progress_is_not_a_stall(ctx):
    # One probe row opens 4,096 rows of R2 that each lead to 4,096 rows of R3: one wave walks
    # 16.7 M last-level results alone (the last level is never split) while every other wave
    # is idle. With a 50 ms stall limit the idle waves would give up if only completed units and
    # tasks counted as progress; the working wave's heartbeat keeps the pass alive, and its
    # count and checksum equal the oracle's.
    rng = np.random.default_rng(11)
    n = 20000
    b = rng.integers(100, 5000, n).astype(np.int32)
    b[1234] = 7
    r2b = np.concatenate([np.full(4096, 7, np.int32), rng.integers(100, 5000, 3000).astype(np.int32)])
    r2c = np.concatenate([np.full(4096, 9, np.int32), rng.integers(100, 5000, 3000).astype(np.int32)])
    r3c = np.concatenate([np.full(4096, 9, np.int32), rng.integers(100, 5000, 3000).astype(np.int32)])
    r3d = rng.integers(0, 1 << 30, len(r3c)).astype(np.int32)
    rels = [([rng.integers(0, 1 << 30, n).astype(np.int32), b], [0, 1]),
            ([r2b, r2c], [1, 2]), ([r3c, r3d], [2, 3])]
    want_cnt, want_cs, _, _, _ = O.mhj(rels)
    assert want_cnt > 4096 * 4096
    r = ctx.chain_join(rels, materialize=False, ws=1, ws_stall_ms=50)
    assert (r.count, r.checksum) == (want_cnt, want_cs) and r.path == "ws"


def test_autotune_choice_gives_the_same_result(ctx):
    rels = G.chain_relations([1 << 20, 1 << 18, 1 << 18, 1 << 18], 1 << 16, seed=42, zipf_s=1.0,
                             payload_range=1 << 31)
    R = [ctx.relation(c, ids) for c, ids in rels]
    H = [ctx.hash_table(r) for r in R[1:]]
    want = ctx.execute(R[0], H, materialize=False)
    tuned = ctx.autotune(R[0], H, materialize=True)
    opts = {k: v for k, v in tuned.items() if k != "estimated_ms"}
    r = ctx.execute(R[0], H, **opts)
    assert (r.count, r.checksum) == (want.count, want.checksum)
    ctx.free(r)
    again = ctx.autotune(R[0], H, materialize=True)  # remembered for this query shape
    assert again == tuned


def test_autotune_two_samples_of_a_large_probe(ctx):
    # above 4 M probe rows the candidates run on two strided samples (their engines' own
    # scratch, such as direct output's chunk tails, must leave the samples intact) and the
    # pick is the best extrapolated time; the picked options give the plain engine's result
    rels = G.chain_relations([9 << 20, 1 << 18, 1 << 18, 1 << 18], 1 << 19, seed=8, zipf_s=0.5,
                             payload_range=1 << 31)
    R = [ctx.relation(c, ids) for c, ids in rels]
    H = [ctx.hash_table(r) for r in R[1:]]
    want = ctx.execute(R[0], H, materialize=False)
    assert want.count > 0 and (want.count, want.checksum) == tuple(O.mhj(rels)[:2])
    for mat in (True, False):
        tuned = ctx.autotune(R[0], H, materialize=mat)
        assert 0 < tuned["estimated_ms"] < 1e6
        opts = {k: v for k, v in tuned.items() if k != "estimated_ms"}
        r = ctx.execute(R[0], H, **opts)
        assert (r.count, r.checksum) == (want.count, want.checksum)
        assert r.rows_written == (r.count if mat else 0)
        ctx.free(r)
