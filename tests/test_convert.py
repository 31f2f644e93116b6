"""Edge-list ingest (exec_el_2_db / edgelist_to_csr): parsing with the reference's separator
set, the .db round trip, and (on the GPU) the .tr file of the device-built trie."""
import os

import numpy as np
import pytest

from accmultiwayjoins_amd import convert, formats


def _write(tmp_path, text, name="edges.txt"):
    p = tmp_path / name
    p.write_bytes(text)
    return str(p)


def test_separators_and_blank_lines(tmp_path):
    p = _write(tmp_path, b"1 2\n3\t4\n5|6\n\n7  |\t8\n")
    s, d = convert.read_edge_list(p)
    assert s.tolist() == [1, 3, 5, 7] and d.tolist() == [2, 4, 6, 8]
    assert s.dtype == np.int32 and d.dtype == np.int32


def test_bad_inputs_rejected(tmp_path):
    with pytest.raises(ValueError):
        convert.read_edge_list(_write(tmp_path, b"1 2 3\n", "odd.txt"))
    with pytest.raises(ValueError):
        convert.read_edge_list(_write(tmp_path, b"1 4294967296\n", "big.txt"))


def test_el_to_db_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    e = rng.integers(0, 1000, size=(500, 2))
    p = _write(tmp_path, "".join(f"{a}\t{b}\n" for a, b in e).encode())
    out = convert.el_to_db(p, str(tmp_path), attr_ids=(4, 7))
    assert os.path.basename(out) == "500_4_7.db"
    cols, ids = formats.read_db(out)
    assert ids == [4, 7]
    assert np.array_equal(cols[0], e[:, 0]) and np.array_equal(cols[1], e[:, 1])


def test_cli_usage(capsys):
    assert convert.main([]) == 2


@pytest.mark.gpu
def test_el_to_tr_levels(ctx, tmp_path):
    e = np.array([[3, 1], [1, 2], [3, 0], [1, 2], [0, 5], [3, 1]])
    p = _write(tmp_path, "".join(f"{a} {b}\n" for a, b in e).encode())
    out = convert.el_to_tr(p, str(tmp_path / "g.tr"), ctx=ctx)
    t = ctx.load_trie(out)
    d0, off, d1 = t.levels()
    # sorted distinct sources, their sorted distinct children (duplicate pairs dropped)
    assert d0.tolist() == [0, 1, 3]
    assert off.tolist() == [0, 1, 2, 4]
    assert d1.tolist() == [5, 2, 0, 1]
    t.close()


LINEITEM = (b"1|155190|7706|1|17|21168.23|0.04|0.02|N|O|1996-03-13|1996-02-12|1996-03-22|"
            b"DELIVER IN PERSON|TRUCK|egular courts above the|\n"
            b"1|67310|7311|2|36|45983.16|0.09|0.06|N|O|1996-04-12|1996-02-28|1996-04-20|"
            b"TAKE BACK RETURN|MAIL|ly final dependencies: slyly bold |\n"
            b"3|4297|1798|1|45|54058.05|0.06|0.00|R|F|1994-02-02|1994-01-04|1994-02-23|"
            b"NONE|AIR|ongside of the furiously brave acco|\n")


def test_tbl_to_db_fields_renamed_and_leading_integers(tmp_path):
    p = _write(tmp_path, LINEITEM, "lineitem.tbl")
    # l_orderkey, l_partkey, l_suppkey as attributes 5, 6, 7 (tbl_2_db's argument order)
    assert convert.main(["tbl2db", p, "0", "1", "2", "5", "6", "7", "--dir", str(tmp_path)]) == 0
    cols, ids = formats.read_db(os.path.join(str(tmp_path), "3_5_6_7.db"))
    assert ids == [5, 6, 7]
    assert [c.tolist() for c in cols] == [[1, 1, 3], [155190, 67310, 4297], [7706, 7311, 1798]]
    # output column order follows the argument order, not the field order; decimals and
    # dates give their leading integer
    out = convert.tbl_to_db(p, [10, 5, 3], [0, 1, 2], str(tmp_path))
    cols, ids = formats.read_db(out)
    assert [c.tolist() for c in cols] == [[1996, 1996, 1994], [21168, 45983, 54058], [1, 2, 1]]


def test_tbl_to_db_rejects_non_integer_fields_and_short_lines(tmp_path):
    p = _write(tmp_path, LINEITEM, "lineitem.tbl")
    with pytest.raises(ValueError, match="integer"):
        convert.read_tbl(p, [8])  # l_returnflag
    with pytest.raises(ValueError, match="fields"):
        convert.read_tbl(p, [40])
    assert convert.main(["tbl2db", p, "0", "1", "2"]) == 2  # fields without new ids
