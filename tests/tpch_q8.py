"""The TPC-H Q8 join shape of the reference (MHJ-GPU/conf.h:16-20: 7 build tables, 8 result
attributes) over the committed TPC-H-format fixture (tests/golden/tpch), converted with tbl2db:

    lineitem(l_orderkey A0, l_partkey A1)            probe
    orders(o_orderkey A0, o_custkey A2)              customer(c_custkey A2, c_nationkey A3)
    nation n1(n_nationkey A3, n_regionkey A4)        region(r_regionkey A4)
    partsupp(ps_partkey A1, ps_suppkey A5)           supplier(s_suppkey A5, s_nationkey A6)
    nation n2(n_nationkey A6, n_regionkey A7)

Every foreign key of the fixture refers to an existing row and each part has four suppliers,
so the join has exactly 4 results per lineitem row."""
import os

from accmultiwayjoins_amd import convert, formats

TPCH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tpch")

# (table, .tbl fields, attribute ids) in join order
Q8 = [("lineitem", [0, 1], [0, 1]), ("orders", [0, 1], [0, 2]), ("customer", [0, 3], [2, 3]),
      ("nation", [0, 2], [3, 4]), ("region", [0], [4]), ("partsupp", [0, 1], [1, 5]),
      ("supplier", [0, 3], [5, 6]), ("nation", [0, 2], [6, 7])]


def q8_relations(tmpdir):
    """The Q8 relations, each written by tbl2db to its own directory and read back."""
    rels = []
    for k, (table, fields, ids) in enumerate(Q8):
        d = os.path.join(str(tmpdir), f"{k}_{table}")
        os.makedirs(d, exist_ok=True)
        path = convert.tbl_to_db(os.path.join(TPCH, table + ".tbl"), fields, ids, d)
        cols, attr_ids = formats.read_db(path)
        assert list(attr_ids) == ids
        rels.append((cols, ids))
    return rels


def lineitem_rows():
    with open(os.path.join(TPCH, "lineitem.tbl")) as f:
        return sum(1 for _ in f)
