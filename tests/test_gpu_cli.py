"""End-to-end runs of the exec-AMHJ / exec-LFTJ command lines on .db / .tr datasets."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_ffi as O
from accmultiwayjoins_amd import formats as F
from accmultiwayjoins_amd import generators as G

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(__file__), "..", "accmultiwayjoins_amd", "bin")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mhj_cpu_counts.json")


def run(args, stdin=""):
    print("[cli]", " ".join(args), flush=True)
    p = subprocess.run([os.path.join(BIN, args[0])] + args[1:], input=stdin, capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return p.stdout


def field(out, pattern):
    m = re.search(pattern, out)
    assert m, out[-2000:]
    return m.group(1)


def golden(name):
    return next(r for r in json.load(open(GOLDEN))["datasets"] if r["name"] == name)


@pytest.mark.parametrize("flags", [["0", "0", "0", "0"], ["0", "1", "1", "1"],
                                   ["1", "0", "0", "0"], ["1", "1", "1", "0"]])
def test_exec_amhj_normal_matches_reference_count(tmp_path, flags):
    rec = golden("chain3_4096_keys1024")
    rels = G.chain_relations(rec["sizes"], rec["key_range"], seed=rec["seed"])
    paths = [F.write_db(str(tmp_path), c, ids) for c, ids in rels]
    # table order on stdin, one answer per file in readdir order (QueryProcessor.cuh:62-63)
    listing = [n for n in os.listdir(tmp_path) if n.endswith(".db")]
    order = "\n".join(str([os.path.basename(p) for p in paths].index(n)) for n in listing)
    out = run(["exec-AMHJ", "NORMAL", str(tmp_path)] + flags, stdin=order + "\n")
    cnt, cs, tup, _, _ = O.mhj(rels, materialize=True, cap=10 ** 7)
    assert cnt == rec["count"]
    assert int(field(out, r"Total number of outputs: (\d+)")) == cnt
    assert int(field(out, r"checksum=0x([0-9a-f]+)"), 16) == cs
    assert field(out, r"Md5sum of result histogram: ([0-9a-f]{32})") == \
        F.res_stat_md5([tup[:, a] for a in range(tup.shape[1])])
    # no order on stdin: inferred from the attribute chain
    out2 = run(["exec-AMHJ", "NORMAL", str(tmp_path)] + flags + ["--count-only"])
    assert int(field(out2, r"Total number of outputs: (\d+)")) == cnt
    assert int(field(out2, r"checksum=0x([0-9a-f]+)"), 16) == cs


def test_exec_amhj_tri_four(tmp_path):
    s, d = G.rmat_oriented(10, 16, seed=5)
    n3, c3 = O.triangles(s, d)
    n4, c4 = O.four_cliques(s, d)
    p = F.write_db(str(tmp_path), [s, d], [0, 1])
    out = run(["exec-AMHJ", "TRI", p, "0", "0", "0", "0"])
    assert int(field(out, r"Total number of outputs: (\d+)")) == n3
    assert int(field(out, r"checksum=0x([0-9a-f]+)"), 16) == c3
    out = run(["exec-AMHJ", "FOUR", p, "0", "1", "1", "1", "--count-only"])
    assert int(field(out, r"Total number of outputs: (\d+)")) == n4
    assert int(field(out, r"checksum=0x([0-9a-f]+)"), 16) == c4


@pytest.mark.parametrize("algo", ["0", "1"])
def test_exec_lftj_self_db_and_tr(ctx, tmp_path, algo):
    s, d = G.rmat_oriented(11, 16, seed=9)
    n3, c3 = O.triangles(s, d)
    a = tmp_path / "a"
    a.mkdir()
    F.write_db(str(a), [s, d], [0, 1])
    out = run(["exec-LFTJ", str(a), algo, "0", "0"])
    assert int(field(out, r"Output count: (\d+)")) == n3
    assert int(field(out, r"checksum=0x([0-9a-f]+)"), 16) == c3
    b = tmp_path / "b"
    b.mkdir()
    t = ctx.trie([s, d])
    t.save(str(b / ("%d_0_1.tr" % len(s))))
    out = run(["exec-LFTJ", str(b), algo, "0", "1"])
    assert "Read from Trie file" in out
    assert int(field(out, r"Output count: (\d+)")) == n3
    assert int(field(out, r"checksum=0x([0-9a-f]+)"), 16) == c3


def test_exec_lftj_dfs_from_tr_file(ctx, tmp_path):
    s, d = G.rmat_oriented(11, 16, seed=9)
    n3, c3 = O.triangles(s, d)
    t = ctx.trie([s, d])
    t.save(str(tmp_path / ("%d_0_1.tr" % len(s))))
    out = run(["exec-LFTJ", str(tmp_path), "1", "0", "0"])
    assert "Read from Trie file" in out
    assert int(field(out, r"Output count: (\d+)")) == n3
    assert int(field(out, r"checksum=0x([0-9a-f]+)"), 16) == c3


def test_exec_lftj_multiway_kat(tmp_path):
    from test_oracle import KAT_TRIANGLE, kat_relations
    for cols, ids in kat_relations():
        F.write_db(str(tmp_path), cols, ids)
    for algo in ("0", "1"):
        out = run(["exec-LFTJ", str(tmp_path), algo, "0", "0"])
        assert int(field(out, r"Output count: (\d+)")) == 5
        assert int(field(out, r"checksum=0x([0-9a-f]+)"), 16) == O.checksum(KAT_TRIANGLE)
