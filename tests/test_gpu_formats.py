"""GPU tests of the .tr trie files, CSR tries and the device-side result fingerprint."""
import struct

import numpy as np
import pytest

import oracle_ffi as O
from accmultiwayjoins_amd import AmjError
from accmultiwayjoins_amd import formats as F
from accmultiwayjoins_amd import generators as G

pytestmark = pytest.mark.gpu


def write_tr_reference_layout(path, levels, offsets, attrs):
    """IndexedTrie::serialization layout (LFTJ-GPU/IndexedTrie.cuh:122-155), written here
    without the library: u32 n, u64 data_len[n], u32 attr_list[n], int32 data per level,
    u64 offsets of every level but the last."""
    with open(path, "wb") as f:
        f.write(struct.pack("<I", len(levels)))
        f.write(np.array([len(l) for l in levels], dtype="<u8").tobytes())
        f.write(np.array(attrs, dtype="<u4").tobytes())
        for l in levels:
            f.write(np.asarray(l, dtype="<i4").tobytes())
        for o in offsets:
            f.write(np.asarray(o, dtype="<u8").tobytes())


def test_tr_roundtrip_and_external_file(ctx, tmp_path):
    s, d = G.rmat_oriented(10, 16, seed=3)
    n3, c3 = O.triangles(s, d)
    t = ctx.trie([s, d])
    p = str(tmp_path / "g.tr")
    t.save(p)
    d0, off, d1 = t.levels()
    raw = open(p, "rb").read()
    assert struct.unpack("<I", raw[:4])[0] == 2
    assert struct.unpack("<QQ", raw[4:20]) == (len(d0), len(d1))
    assert len(raw) == 4 + 16 + 8 + 4 * (len(d0) + len(d1)) + 8 * (len(d0) + 1)
    t2 = ctx.load_trie(p)
    for a, b in zip(t2.levels(), (d0, off, d1)):
        assert np.array_equal(a, b)
    # a file written independently in the reference layout loads to the same trie
    q = str(tmp_path / "ext.tr")
    write_tr_reference_layout(q, [d0, d1], [off.astype(np.uint64)], [0, 1])
    t3 = ctx.load_trie(q)
    for tr in (t2, t3):
        r = ctx.lftj([tr], [0, 1])
        assert r.count == len(d1)
    r = ctx.lftj_pattern(s, d, "TRI", "DFS", materialize=False)
    assert (r.count, r.checksum) == (n3, c3)


def test_tr_rejects_three_levels(ctx, tmp_path):
    q = str(tmp_path / "three.tr")
    write_tr_reference_layout(q, [[1], [2], [3]], [[0, 1], [0, 1]], [0, 1, 2])
    with pytest.raises(AmjError, match="two-level"):
        ctx.load_trie(q)


def test_csr_trie_validation_and_duplicates(ctx):
    with pytest.raises(AmjError, match="strictly increasing"):
        ctx.trie_from_csr([3, 1], [0, 1, 2], [5, 6])
    with pytest.raises(AmjError, match="sorted"):
        ctx.trie_from_csr([1, 3], [0, 2, 3], [6, 5, 1])
    with pytest.raises(AmjError, match="offsets"):
        ctx.trie_from_csr([1, 3], [0, 2, 4], [5, 6, 1])
    # duplicate children are kept: with R(0,1) = {(1,5),(1,5)} and S(1,.) = {(5,.)} the
    # driver list holds 5 twice
    t = ctx.trie_from_csr([1], [0, 2], [5, 5])
    assert (t.level0, t.children, t.max_children) == (1, 2, 2)


def test_result_md5_equals_host_fingerprint(ctx):
    rels = G.chain_relations([20000, 8000, 8000], 2000, seed=4)
    for opts in ({}, {"fib": 1, "ws": 1, "dro": 1}):
        r = ctx.chain_join(rels, materialize=True, **opts)
        dev = ctx.md5(r)
        ctx.free(r)
        h = ctx.chain_join(rels, fetch=True, **opts)
        assert dev == F.res_stat_md5(h.cols)
    e = ctx.chain_join([(c, ids) for c, ids in rels[:1]] + [([np.array([-1], np.int32),
                                                             np.array([0], np.int32)], [1, 2])])
    assert e.count == 0
    assert ctx.md5(e) == F.res_stat_md5([np.zeros(0, np.int32)] * 3)
    n = ctx.chain_join(rels, materialize=False)
    with pytest.raises(AmjError, match="materialised"):
        ctx.md5(n)
