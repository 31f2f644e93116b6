"""The TPC-H Q8 join shape (7 build tables, 8 attributes, the reference's Q8 setting of
MHJ-GPU/conf.h:16-20) on the GPU: the committed TPC-H-format fixture converted with tbl2db and
joined through every engine and through exec-AMHJ, against the CPU oracle and brute force."""
import os

import pytest

import oracle_ffi as O
from test_gpu_cli import field, run
from tpch_q8 import lineitem_rows, q8_relations

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("opts", [dict(), dict(fib=1), dict(ws=1, ws_split_threshold=64),
                                  dict(dro=1), dict(ws=1, dro=1, fib=1), dict(engine="tpt"),
                                  dict(engine="open-addr"), dict(engine="pw")],
                         ids=["plain", "fib", "ws", "dro", "ws-dro-fib", "thread-per-tuple",
                              "open-addressing", "pairwise"])
def test_q8_shape_all_engines(ctx, tmp_path, opts):
    rels = q8_relations(tmp_path)
    want = O.brute_join(rels)
    assert len(want) == 4 * lineitem_rows()
    r = ctx.chain_join(rels, fetch=True, **opts)
    assert (r.count, r.checksum) == (len(want), O.checksum(want))
    assert len(r.cols) == 8
    assert [tuple(int(x) for x in t) for t in O.sort_tuples(r.tuples())] == want


def test_q8_shape_through_exec_amhj(tmp_path):
    # the eight .db files in one directory (the two nation tables differ in their attribute
    # ids, so in their names); each file's position in the join order is given on stdin in
    # readdir order, as the reference asks for it (QueryProcessor.cuh:62-63)
    from accmultiwayjoins_amd import formats as F
    rels = q8_relations(tmp_path / "conv")
    d = tmp_path / "q8"
    d.mkdir()
    names = [os.path.basename(F.write_db(str(d), cols, ids)) for cols, ids in rels]
    assert len(set(names)) == len(rels)
    order = "\n".join(str(names.index(n)) for n in os.listdir(d))
    want = O.brute_join(rels)
    out = run(["exec-AMHJ", "NORMAL", str(d), "0", "0", "0", "0"], stdin=order + "\n")
    assert int(field(out, r"Total number of outputs: (\d+)")) == len(want)
    assert int(field(out, r"checksum=0x([0-9a-f]+)"), 16) == O.checksum(want)
