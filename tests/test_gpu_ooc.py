"""Out-of-core AMHJ: the probe relation streamed from host memory in slices must give the same
count, checksum and tuples as the in-HBM engines, for any slice size, with and without work
sharing, from pinned or pageable host memory."""
import numpy as np
import pytest

import oracle_ffi as O
from accmultiwayjoins_amd import AmjError
from accmultiwayjoins_amd import generators as G
from accmultiwayjoins_amd.amj import pinned_empty

pytestmark = pytest.mark.gpu


def _chain():
    return G.chain_relations([50000, 8000, 8000, 8000], 3000, seed=21, zipf_s=0.7,
                             payload_range=1 << 24)


def _pinned(cols):
    out = []
    for c in cols:
        p = pinned_empty(len(c))
        p[:] = c
        out.append(p)
    return out


@pytest.mark.parametrize("slice_rows", [64, 1000, 4096, 1 << 20])
@pytest.mark.parametrize("ws", [0, 1])
def test_ooc_equals_in_memory(ctx, slice_rows, ws):
    rels = _chain()
    want_cnt, want_cs, _, _, _ = O.mhj(rels)
    probe = ctx.relation(_pinned(rels[0][0]), rels[0][1], host=True)
    B = [ctx.relation(c, ids) for c, ids in rels[1:]]
    H = [ctx.hash_table(r) for r in B]
    r = ctx.execute(probe, H, materialize=False, ooc=1, ooc_slice_rows=slice_rows, ws=ws)
    assert (r.count, r.checksum) == (want_cnt, want_cs)
    assert r.path == "ooc"
    per = -(-slice_rows // 64) * 64  # slices are whole units of 64 rows
    assert r.detail == -(-50000 // per)
    m = ctx.execute(probe, H, fetch=True, ooc=1, ooc_slice_rows=slice_rows, ws=ws)
    assert m.count == want_cnt and m.passes == 2
    ref = ctx.execute(ctx.relation(*rels[0]), H, fetch=True)
    assert np.array_equal(O.sort_tuples(m.tuples()), O.sort_tuples(ref.tuples()))


def test_ooc_pageable_and_device_probe(ctx):
    rels = _chain()
    want_cnt, want_cs, _, _, _ = O.mhj(rels)
    H = [ctx.hash_table(ctx.relation(c, ids)) for c, ids in rels[1:]]
    host = ctx.relation(rels[0][0], rels[0][1], host=True)  # plain numpy memory
    r = ctx.execute(host, H, materialize=False, ooc_slice_rows=777)
    assert (r.count, r.checksum) == (want_cnt, want_cs)
    dev = ctx.relation(*rels[0])  # device-resident probe, sliced all the same
    r = ctx.execute(dev, H, materialize=False, ooc=1, ooc_slice_rows=4096)
    assert (r.count, r.checksum) == (want_cnt, want_cs) and r.detail == 13


def test_ooc_budget_and_errors(ctx):
    rels = _chain()
    probe = ctx.relation(_pinned(rels[0][0]), rels[0][1], host=True)
    H = [ctx.hash_table(ctx.relation(c, ids)) for c, ids in rels[1:]]
    with pytest.raises(AmjError, match="host-resident"):
        ctx.hash_table(probe)
    # a result that cannot fit the budget is refused; count + checksum still run
    with pytest.raises(AmjError, match="budget"):
        ctx.execute(probe, H, ooc=1, hbm_budget_bytes=ctx_allocated(ctx) + (1 << 20))
    r = ctx.execute(probe, H, ooc=1, materialize=False,
                    hbm_budget_bytes=ctx_allocated(ctx) + (64 << 20))
    assert r.count == O.mhj(rels)[0]


def ctx_allocated(ctx):
    return ctx.allocated_bytes()
