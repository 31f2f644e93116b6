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
