"""amj_execute_noalloc (no device allocation) against amj_execute's count + checksum, through the
C ABI: two calls in a row, the second on a smaller probe relation, equal to the existing entry
point and to the CPU oracle, with the context's device bytes unchanged by the call; sizes past
the limit and options it does not support are rejected with a clear error."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
from accmultiwayjoins_amd import _native as N
from accmultiwayjoins_amd import amj as A
from accmultiwayjoins_amd import generators as G

pytestmark = pytest.mark.gpu


def _noalloc(ctx, probe, tables, **opts):
    hts = (C.c_void_p * len(tables))(*[t.h.value for t in tables])
    o = A._opts(materialize=False, **opts)
    res = N.AmjResult()
    rc = N.lib().amj_execute_noalloc(ctx.h, probe.h, hts, len(tables), C.byref(o), C.byref(res))
    return rc, res


def _existing(ctx, probe, tables):
    hts = (C.c_void_p * len(tables))(*[t.h.value for t in tables])
    o = A._opts(materialize=False)
    res = N.AmjResult()
    rc = N.lib().amj_execute(ctx.h, probe.h, hts, len(tables), C.byref(o), C.byref(res))
    assert rc == 0, N.lib().amj_last_error(ctx.h)
    return res


def test_noalloc_matches_execute_twice_then_smaller(ctx):
    rels = G.chain_relations([3_000_000, 400_000, 300_000, 200_000], 100_000, seed=42,
                             zipf_s=0.8, payload_range=1 << 30)
    small = [(cols, ids) for cols, ids in rels]
    small[0] = ([c[:12_345] for c in rels[0][0]], rels[0][1])
    R = [ctx.relation(c, ids) for c, ids in rels]
    H = [ctx.hash_table(r) for r in R[1:]]
    P_small = ctx.relation(*small[0])
    for probe, chain in ((R[0], rels), (P_small, small)):
        want_cnt, want_cs, _, _, _ = O.mhj(chain)
        before = ctx.allocated_bytes()
        rc, got = _noalloc(ctx, probe, H)
        assert rc == 0, N.lib().amj_last_error(ctx.h)
        assert ctx.allocated_bytes() == before  # no device bytes allocated by the call
        ref = _existing(ctx, probe, H)
        assert (got.count, got.checksum) == (ref.count, ref.checksum)
        assert (got.count, got.checksum) == (want_cnt, want_cs)
        assert got.rows_written == 0 and got.num_cols == 5
    # a repeated call on the same GPU reuses the array and gives the same totals
    rc, again = _noalloc(ctx, P_small, H)
    assert rc == 0 and (again.count, again.checksum) == (got.count, got.checksum)


def test_noalloc_rejects_what_it_cannot_do(ctx):
    rels = G.chain_relations([5000, 3000], 500, seed=7)
    R = [ctx.relation(c, ids) for c, ids in rels]
    H = [ctx.hash_table(R[1])]
    for bad in (dict(ws=1), dict(fib=1), dict(dro=1)):
        rc, _ = _noalloc(ctx, R[0], H, **bad)
        assert rc == -1
        assert b"count + checksum only" in N.lib().amj_last_error(ctx.h)
    n = (1 << 27) + 64  # one unit past the limit
    big = ctx.relation([np.zeros(n, np.int32), np.zeros(n, np.int32)], [0, 1])
    rc, _ = _noalloc(ctx, big, H)
    assert rc == -1
    assert b"exceeds its limit" in N.lib().amj_last_error(ctx.h)
    big.close()
    rc, res = _noalloc(ctx, R[0], H)  # still usable after the errors
    want_cnt, want_cs, _, _, _ = O.mhj(rels)
    assert rc == 0 and (res.count, res.checksum) == (want_cnt, want_cs)
