"""The TPC-H Q8 join shape (7 build tables, 8 attributes; MHJ-GPU/conf.h:16-20) on the CPU: the
committed .tbl fixture goes through tbl2db, and the oracle's count, checksum and tuples equal a
brute-force join and the closed form (4 results per lineitem row)."""
import oracle_ffi as O
from tpch_q8 import lineitem_rows, q8_relations


def test_q8_shape_oracle_matches_brute_force_and_closed_form(tmp_path):
    rels = q8_relations(tmp_path)
    assert len(rels) - 1 == 7 and len({a for _, ids in rels for a in ids}) == 8
    n, cs, tuples, _, _ = O.mhj(rels, materialize=True)
    want = O.brute_join(rels)
    assert n == len(want) == 4 * lineitem_rows()
    assert cs == O.checksum(want)
    assert [tuple(int(x) for x in t) for t in O.sort_tuples(tuples)] == want
