"""Host-side formats and the result fingerprint (no GPU needed).

The res_stat fingerprint is pinned two ways: golden MD5s printed by the reference's own
res_stat (common-utils/res_stat.h with common-utils/md5.cpp, tests/golden/res_stat_md5.json),
and an independent restatement with numpy + hashlib.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from accmultiwayjoins_amd import formats as F
from accmultiwayjoins_amd import generators as G

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "res_stat_md5.json")


def md5_restated(cols):
    vals = [np.asarray(c, dtype=np.int64) & 0xFFFFF for c in cols]
    allv = np.concatenate(vals) if vals and sum(len(v) for v in vals) else np.zeros(0, np.int64)
    h = np.bincount(allv, minlength=1) if len(allv) else np.zeros(1, np.int64)
    text = "[" + ", ".join(str(int(x)) for x in h) + "]\n"
    return hashlib.md5(text.encode()).hexdigest()


@pytest.mark.parametrize("case", json.load(open(GOLDEN))["cases"],
                         ids=lambda c: "%dx%d" % (c["rows"], c["cols"]))
def test_res_stat_matches_reference_golden(case):
    cols = [G.uniform(case["rows"], case["lo"], case["hi"], case["seed"], c)
            for c in range(case["cols"])]
    assert F.res_stat_md5(cols) == case["md5"]
    assert md5_restated(cols) == case["md5"]


@pytest.mark.parametrize("n", [0, 1, 2, 9, 10, 11, 12, 13, 14, 20, 55, 100, 1000])
def test_md5_padding_boundaries(n):
    # histogram texts of many lengths cross the 55/56/64-byte padding cases of MD5
    cols = [np.arange(n, dtype=np.int32) % 7]
    assert F.res_stat_md5(cols) == md5_restated(cols)


def test_db_roundtrip_and_name(tmp_path):
    a = G.uniform(1234, -50, 50, 1, 0)
    b = G.uniform(1234, 0, 1 << 30, 1, 1)
    path = F.write_db(str(tmp_path), [a, b], [3, 7])
    assert os.path.basename(path) == "1234_3_7.db"
    assert os.path.getsize(path) == 1234 * 8
    # column 0 then column 1, little-endian int32 (read_rel_cols_mmap layout)
    raw = np.fromfile(path, dtype="<i4")
    assert np.array_equal(raw[:1234], a) and np.array_equal(raw[1234:], b)
    cols, ids = F.read_db(path)
    assert ids == [3, 7]
    assert np.array_equal(cols[0], a) and np.array_equal(cols[1], b)
    assert F.parse_db_name("/x/y/1000000_0_1.db") == (1000000, [0, 1])


@pytest.mark.parametrize("bad", ["x_0_1.db", "100.db", "100_0_1.txt", "100_a_1.db"])
def test_db_bad_names(bad):
    with pytest.raises(OSError):
        F.parse_db_name(bad)


def test_db_short_file(tmp_path):
    p = tmp_path / "100_0_1.db"
    p.write_bytes(b"\0" * 100)
    with pytest.raises(OSError, match="short read"):
        F.read_db(str(p))
