"""Batched and served execution: every result must be the same alone and in the mix (count,
checksum and, where materialised, the tuples), whatever the batch composition, the budget
splits and the submitting threads."""
import hashlib
import threading

import numpy as np
import pytest

import oracle_ffi as O
from accmultiwayjoins_amd import AmjError
from accmultiwayjoins_amd import generators as G
from accmultiwayjoins_amd.serving import Batch, Server

pytestmark = pytest.mark.gpu


def _pool():
    """Chain relations of several sizes and skews, plus one R-MAT edge relation."""
    chains = []
    for i, (sizes, keys, s) in enumerate([([3000, 800, 900], 300, None),
                                          ([20000, 5000, 5000, 5000], 2000, 0.8),
                                          ([1, 10, 10], 5, None),
                                          ([2000, 2000, 2000, 2000, 2000], 700, None),
                                          ([40000, 10000], 5000, 1.0)]):
        chains.append(G.chain_relations(sizes, keys, seed=100 + i, zipf_s=s,
                                        payload_range=1 << 20))
    src, dst = G.rmat_oriented(9, 16, seed=7)
    return chains, (src, dst)


def test_batch_equals_individual_queries(ctx):
    chains, (src, dst) = _pool()
    R = [[ctx.relation(c, ids) for c, ids in ch] for ch in chains]
    H = [[ctx.hash_table(r) for r in rs[1:]] for rs in R]
    E = ctx.relation([src, dst], [0, 1])
    HE = ctx.hash_table(E, 0, sort_vals=True)
    queries = []
    for k in range(len(R)):
        queries.append(("chain", k, True))
        queries.append(("chain", k, False))
    queries += [("TRI", None, True), ("FOUR", None, False), ("TRI", None, False)]
    want = []
    for kind, k, mat in queries:
        if kind == "chain":
            r = ctx.execute(R[k][0], H[k], fetch=True, materialize=mat)
        else:
            r = ctx.execute_pattern(E, HE, kind, fetch=True, materialize=mat)
        want.append(r)
    b = Batch(ctx)
    ids = []
    for kind, k, mat in queries:
        if kind == "chain":
            ids.append(b.add(R[k][0], H[k], materialize=mat))
        else:
            ids.append(b.add_pattern(E, HE, kind, materialize=mat))
    b.run()
    for (kind, k, mat), qid, w in zip(queries, ids, want):
        r = b.result(qid, fetch=True)
        assert r.path == "batch"
        assert (r.count, r.checksum) == (w.count, w.checksum), (kind, k, mat)
        if mat:
            # the batch writes every tuple at the position the unit engine gives it
            for a, c in zip(r.cols, w.cols):
                assert np.array_equal(a, c)
    b.close()
    n3, c3 = O.triangles(src, dst)
    assert (want[-3].count, want[-3].checksum) == (n3, c3)


def _sorted(cols):
    t = np.stack(cols, axis=1) if cols else np.zeros((0, 0), np.int32)
    return t[np.lexsort(t.T[::-1])] if len(t) else t


@pytest.mark.parametrize("capacity", [0, 64], ids=["sampled", "overrun-rerun"])
def test_direct_output_batch_equals_individual_queries(ctx, capacity):
    # dro=1 queries: no count pass; capacity 64 overruns every larger result, which is then
    # run again exactly. Same count, checksum and multiset of tuples as each query alone.
    chains, (src, dst) = _pool()
    R = [[ctx.relation(c, ids) for c, ids in ch] for ch in chains]
    H = [[ctx.hash_table(r) for r in rs[1:]] for rs in R]
    E = ctx.relation([src, dst], [0, 1])
    HE = ctx.hash_table(E, 0, sort_vals=True)
    queries = [("chain", k, mat) for k in range(len(R)) for mat in (True, False)]
    queries += [("TRI", None, True), ("FOUR", None, True), ("FOUR", None, False)]
    want = []
    for kind, k, mat in queries:
        if kind == "chain":
            want.append(ctx.execute(R[k][0], H[k], fetch=True, materialize=mat))
        else:
            want.append(ctx.execute_pattern(E, HE, kind, fetch=True, materialize=mat))
    b = Batch(ctx)
    ids = []
    for kind, k, mat in queries:
        if kind == "chain":
            ids.append(b.add(R[k][0], H[k], materialize=mat, dro=1, dro_capacity_rows=capacity))
        else:
            ids.append(b.add_pattern(E, HE, kind, materialize=mat, dro=1,
                                     dro_capacity_rows=capacity))
    b.run()
    for (kind, k, mat), qid, w in zip(queries, ids, want):
        r = b.result(qid, fetch=True)
        assert (r.count, r.checksum) == (w.count, w.checksum), (kind, k, mat)
        if mat:
            assert r.rows_written == w.count
            assert np.array_equal(_sorted(r.cols), _sorted(w.cols)), (kind, k)
    b.close()
    n4, c4 = O.four_cliques(src, dst)
    assert (want[-1].count, want[-1].checksum) == (n4, c4)


def test_direct_output_batch_admits_a_prefix(ctx):
    # with room for one result, the first dro query runs and the later ones are reported
    # as not admitted (the server runs them in a later batch)
    chains, _g = _pool()
    R = [ctx.relation(c, ids) for c, ids in chains[1]]
    H = [ctx.hash_table(r) for r in R[1:]]
    want = ctx.execute(R[0], H, materialize=True)
    need = want.count * len(want.attr_ids) * 4
    ctx.free(want)
    b = Batch(ctx)
    ids = [b.add(R[0], H, dro=1) for _ in range(3)]
    b.set_budget(need + need // 2)
    b.run()
    r = b.result(ids[0], fetch=True)
    assert (r.count, r.checksum) == (want.count, want.checksum)
    for i in ids[1:]:
        with pytest.raises(AmjError, match="not run"):
            b.result(i)
    b.close()


def _server_pool(srv):
    chains, (src, dst) = _pool()
    ids = []
    for ch in chains:
        ids.append([srv.add_relation(c, a) for c, a in ch])
    e = srv.add_relation([src, dst], (0, 1))
    reqs = [("chain", rid, True) for rid in ids] + [("chain", rid, False) for rid in ids]
    reqs += [("TRI", e, True), ("FOUR", e, False)]
    return reqs, chains, (src, dst)


def _submit(srv, req):
    kind, rid, mat = req
    if kind == "chain":
        return srv.submit(rid, materialize=mat)
    return srv.submit_pattern(rid, kind, materialize=mat)


def _key(out):
    """(count, checksum, digest of the sorted tuples)."""
    cnt, cs, _lat, t = out
    if t is None or not len(t):
        return cnt, cs, None
    ts = t[np.lexsort(t.T[::-1])]
    return cnt, cs, hashlib.md5(np.ascontiguousarray(ts).tobytes()).hexdigest()


@pytest.mark.parametrize("plain", [False, True], ids=["batched", "plain"])
def test_server_same_alone_and_in_mix(plain):
    with Server(0, max_batch=64, plain=plain) as srv:
        reqs, chains, (src, dst) = _server_pool(srv)
        alone = [_key(srv.wait(_submit(srv, q), fetch=True)) for q in reqs]
        # oracle for the chains
        for (kind, _rid, _mat), a, ch in zip(reqs, alone, chains + chains):
            if kind == "chain":
                cnt, cs, _, _, _ = O.mhj(ch)
                assert a[:2] == (cnt, cs)
        n3, c3 = O.triangles(src, dst)
        assert alone[-2][:2] == (n3, c3)
        # the same requests all at once, three times over, from one thread
        tickets = [_submit(srv, q) for q in reqs * 3]
        mix = [_key(srv.wait(t, fetch=True)) for t in tickets]
        assert mix == alone * 3
        st = srv.stats()
        assert st["served"] == 4 * len(reqs)
        if not plain:
            assert st["batches"] < st["served"]  # requests were batched


def test_server_triejoin_route_equals_hash_join_engines():
    # set_triejoin: TRI / FOUR requests run alone through the leapfrog triejoin over a trie
    # kept resident; the results (sorted tuples, count, checksum) equal the hash-join pattern
    # engines' in a plain server, alone and in a mix, and the oracle's
    with Server(0, max_batch=64, plain=True) as ref:
        reqs, _chains, (src, dst) = _server_pool(ref)
        want = [_key(ref.wait(_submit(ref, q), fetch=True)) for q in reqs]
    with Server(0, max_batch=64) as srv:
        srv.set_triejoin(1)
        reqs, _chains, _g = _server_pool(srv)
        alone = [_key(srv.wait(_submit(srv, q), fetch=True)) for q in reqs]
        assert alone == want
        tickets = [_submit(srv, q) for q in reqs * 2]
        assert [_key(srv.wait(t, fetch=True)) for t in tickets] == want * 2
        assert srv.tuning()["triejoin_alone"] == 6  # TRI and FOUR, three times each
    assert want[-2][:2] == O.triangles(src, dst)
    assert want[-1][:2] == O.four_cliques(src, dst)


@pytest.mark.parametrize("mode", ["batched", "plain", "routes"])
def test_server_threads_and_budget_splits(mode):
    # a budget that holds only a few results at once forces batch splits (batched) or waits
    # for releases (plain); every request still gets its result. "routes": larger chains run
    # alone with tuned options and the graph patterns through the triejoin, between batches;
    # requests still start in submission order, so clients that wait for their tickets in
    # order never wait for room held by a later result
    with Server(0, hbm_budget_bytes=16 << 20, max_batch=64, plain=mode == "plain") as srv:
        if mode == "routes":
            srv.set_autotune(10000)
            srv.set_triejoin(1)
        reqs, _chains, _g = _server_pool(srv)
        alone = [_key(srv.wait(_submit(srv, q), fetch=True)) for q in reqs]
        results = {}

        def client(tid):
            ts = [(i, _submit(srv, q)) for i, q in enumerate(reqs)]
            results[tid] = [(i, _key(srv.wait(t, fetch=True))) for i, t in ts]

        threads = [threading.Thread(target=client, args=(k,)) for k in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for tid in range(4):
            for i, k in results[tid]:
                assert k == alone[i]
        assert srv.stats()["bytes_held"] == 0


def test_batch_budget_overrun_reports_needed_bytes(ctx):
    ch = G.chain_relations([3000, 800, 900], 300, seed=100)
    R = [ctx.relation(c, a) for c, a in ch]
    H = [ctx.hash_table(r) for r in R[1:]]
    b = Batch(ctx)
    b.add(R[0], H)
    b.set_budget(1024)
    with pytest.raises(AmjError, match="budget"):
        b.run()
    b.close()


def test_server_request_larger_than_budget_fails_cleanly():
    with Server(0, hbm_budget_bytes=1 << 16, max_batch=8) as srv:
        ch = G.chain_relations([20000, 5000, 5000], 500, seed=3)
        ids = [srv.add_relation(c, a) for c, a in ch]
        t = srv.submit(ids, materialize=True)
        with pytest.raises(AmjError):
            srv.wait(t)
        # count-only requests hold no result memory and still run
        cnt, cs, _lat, _ = srv.wait(srv.submit(ids, materialize=False))
        assert (cnt, cs) == O.mhj(ch)[:2]


def test_direct_output_batch_heavy_first_order(ctx):
    # launches of at least 2,048 units take them heaviest first; every query's result is still
    # the one it has alone (count, checksum and the order-free fingerprint of its tuples)
    chains = [G.chain_relations([150_000, 40_000, 40_000], 20_000, seed=s, zipf_s=z)
              for s, z in ((11, 1.0), (12, 0.5))]
    R = [[ctx.relation(c, ids) for c, ids in ch] for ch in chains]
    H = [[ctx.hash_table(r) for r in rs[1:]] for rs in R]
    src, dst = G.rmat_oriented(17, 16, seed=5)
    E = ctx.relation([src, dst], [0, 1])
    HE = ctx.hash_table(E, 0, sort_vals=True)
    assert sum(-(-len(c[0][0][0]) // 64) for c in chains) >= 2048 and len(src) // 64 >= 2048
    want = [ctx.execute(R[k][0], H[k]) for k in range(2)]
    want.append(ctx.execute_pattern(E, HE, "TRI"))
    fps = [ctx.md5(w) for w in want]
    b = Batch(ctx)
    ids = [b.add(R[k][0], H[k], dro=1) for k in range(2)]
    ids.append(b.add_pattern(E, HE, "TRI", dro=1))
    b.run()
    for qid, w, fp in zip(ids, want, fps):
        r = b.result(qid)
        assert (r.count, r.checksum) == (w.count, w.checksum)
        assert r.rows_written == w.count
        assert ctx.md5(r) == fp
        ctx.free(r)
    b.close()
    for w in want:
        ctx.free(w)


@pytest.mark.parametrize("slab", [1 << 30, 200_000], ids=["fits", "small-slab-reruns"])
def test_captured_batch_replays_equal_eager_results(ctx, slab):
    # The batch's direct-output chain recorded into a HIP graph (sizes and places decided on the
    # device) and replayed three times, with eager work and synchronisations between the
    # replays: every replay gives each query's count, checksum and multiset of tuples as the
    # query run alone. A 200 KB slab cannot hold every output: those queries run again exactly.
    chains, (src, dst) = _pool()
    R = [[ctx.relation(c, ids) for c, ids in ch] for ch in chains]
    H = [[ctx.hash_table(r) for r in rs[1:]] for rs in R]
    E = ctx.relation([src, dst], [0, 1])
    HE = ctx.hash_table(E, 0, sort_vals=True)
    queries = [("chain", k, mat) for k in range(len(R)) for mat in (True, False)]
    queries += [("TRI", None, True), ("FOUR", None, False), ("FOUR", None, True)]
    want = []
    for kind, k, mat in queries:
        if kind == "chain":
            want.append(ctx.execute(R[k][0], H[k], fetch=True, materialize=mat))
        else:
            want.append(ctx.execute_pattern(E, HE, kind, fetch=True, materialize=mat))
    b = Batch(ctx)
    ids = []
    for j, (kind, k, mat) in enumerate(queries):
        cap = 64 if j == 2 else 0  # one query sized too small by the caller: it overruns
        if kind == "chain":
            ids.append(b.add(R[k][0], H[k], materialize=mat, dro=1, dro_capacity_rows=cap))
        else:
            ids.append(b.add_pattern(E, HE, kind, materialize=mat, dro=1, dro_capacity_rows=cap))
    b.capture(slab)
    with pytest.raises(AmjError, match="replay"):
        b.run()
    for rep in range(3):
        b.replay()
        for (kind, k, mat), qid, w in zip(queries, ids, want):
            r = b.result(qid, fetch=True)
            assert (r.count, r.checksum) == (w.count, w.checksum), (rep, kind, k, mat)
            if mat:
                assert r.rows_written == w.count
                assert np.array_equal(_sorted(r.cols), _sorted(w.cols)), (rep, kind, k)
        # eager work on the same stream between replays
        e = ctx.execute(R[1][0], H[1], materialize=False)
        assert (e.count, e.checksum) == (want[2].count, want[2].checksum)
    b.close()
    n4, c4 = O.four_cliques(src, dst)
    assert (want[-2].count, want[-2].checksum) == (n4, c4)
