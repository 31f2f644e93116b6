"""The reference's generator shapes (distinct uniform pairs, Zipf pairs, unique-modulo columns,
sequential pairs) and its writer command lines: every invariant the reference's generators
guarantee, checked on the output, plus seed determinism and independence of the thread count."""
import os
import subprocess
import sys

import numpy as np
import pytest

from accmultiwayjoins_amd import formats
from accmultiwayjoins_amd import generators as G
from accmultiwayjoins_amd import writers as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _codes(k0, k1, r1):
    return k0.astype(np.int64) * r1 + k1


@pytest.mark.parametrize("n,r0,r1", [(1000, 100, 100), (9000, 100, 100), (9999, 100, 100),
                                     (200_000, 1 << 20, 1 << 20), (50, 1, 51)],
                         ids=["sparse", "dense", "all-but-one", "large", "one-row-key"])
def test_unique_pairs_distinct_sorted_in_range(n, r0, r1):
    k0, k1 = G.unique_pairs(n, (0, r0), (0, r1), seed=7)
    assert len(k0) == len(k1) == n
    assert k0.min() >= 0 and k0.max() < r0 and k1.min() >= 0 and k1.max() < r1
    c = _codes(k0, k1, r1)
    assert np.all(np.diff(c) > 0)  # strictly increasing: sorted and distinct
    s0, s1 = G.unique_pairs(n, (0, r0), (0, r1), seed=7, shuffle=True)
    assert np.array_equal(np.sort(_codes(s0, s1, r1)), c)  # the same pairs, shuffled
    if n > 100:
        assert not np.array_equal(_codes(s0, s1, r1), c)


def test_unique_pairs_offsets_and_uniformity():
    k0, k1 = G.unique_pairs(400_000, (-50, 50), (1000, 1000 + 50_000), seed=3)
    assert k0.min() == -50 and k0.max() == 49 and k1.min() >= 1000 and k1.max() < 51_000
    # every key0 value holds close to n / 100 pairs (chi-square, 99 degrees of freedom)
    h = np.bincount(k0 + 50, minlength=100)
    e = 400_000 / 100
    assert ((h - e) ** 2 / e).sum() < 200


def test_unique_pairs_rejects_impossible_requests():
    with pytest.raises(ValueError):
        G.unique_pairs(100, (0, 10), (0, 10))  # as many rows as pairs: the reference asserts
    with pytest.raises(ValueError):
        G.unique_pairs(5, (3, 3), (0, 10))


@pytest.mark.parametrize("skew_col", [0, 1])
def test_zipf_pairs_distinct_and_skewed(skew_col):
    n, rs, ru = 200_000, 1000, 100_000
    r0, r1 = ((0, rs), (0, ru)) if skew_col == 0 else ((0, ru), (0, rs))
    k0, k1 = G.zipf_pairs(n, r0, r1, skew_col, 1.0, seed=5)
    sk, un = (k0, k1) if skew_col == 0 else (k1, k0)
    assert sk.min() >= 0 and sk.max() < rs and un.min() >= 0 and un.max() < ru
    c = _codes(sk, un, ru)
    assert np.all(np.diff(c) > 0)  # ordered by (skewed, uniform), distinct
    # rank-frequency: the most frequent values follow n / (r * H_rs)
    h = np.sort(np.bincount(sk, minlength=rs))[::-1]
    harmonic = np.sum(1.0 / np.arange(1, rs + 1))
    for r in (1, 2, 4, 10):
        want = n / (r * harmonic)
        assert abs(h[r - 1] - want) < 5 * np.sqrt(want) + 0.02 * want, r
    s0, s1 = G.zipf_pairs(n, r0, r1, skew_col, 1.0, seed=5, shuffle=True)
    ssk, sun = (s0, s1) if skew_col == 0 else (s1, s0)
    assert np.array_equal(np.sort(_codes(ssk, sun, ru)), c)


def test_zipf_pairs_reports_a_value_that_cannot_be_made_distinct():
    # the top value of Zipf(1.0) over 1000 values occurs ~13 000 times in 100 000 rows, more
    # than a uniform range of 5000 values can hold distinct (the reference loops forever)
    with pytest.raises(ValueError, match="cannot hold"):
        G.zipf_pairs(100_000, (0, 1000), (0, 5000), 0, 1.0)
    k0, k1 = G.zipf_pairs(100_000, (0, 1000), (0, 5000), 0, 0.0)  # z = 0: uniform, fits
    assert np.all(np.diff(_codes(k0, k1, 5000)) > 0)


def test_unique_column_is_a_permutation_of_modulo_values():
    for n, r in ((1000, 300), (1000, 1000), (10, 1), (12345, 7)):
        c = G.unique_column(n, r, seed=1, stream=2)
        assert np.array_equal(np.sort(c), np.sort(np.arange(n) % r))
    c = G.unique_column(100_000, 100_000, seed=1, stream=2)
    assert not np.array_equal(c, np.arange(100_000))  # shuffled


def test_zipf_mapping_identity_and_shuffled():
    z = G.zipf_mapped(200_000, 50, 1.2, seed=9, stream=0, shuffle=False)
    h = np.bincount(z, minlength=50)
    assert np.argmax(h) == 0 and h[0] > h[1] > h[3] > h[10]
    assert np.array_equal(G.zipf_mapped(1000, 50, 1.2, 9, 0, True), G.zipf(1000, 50, 1.2, 9, 0))


def test_seq_pairs_closed_form():
    k0, k1 = G.seq_pairs(12, (5, 8))
    assert k0.tolist() == [5] * 4 + [6] * 4 + [7] * 4
    assert k1.tolist() == [0, 1, 2, 3] * 3
    with pytest.raises(ValueError):
        G.seq_pairs(2, (0, 3))


def test_generators_do_not_depend_on_the_thread_count():
    code = ("import hashlib, sys; sys.path.insert(0, %r)\n"
            "from accmultiwayjoins_amd import generators as G\n"
            "a, b = G.zipf_pairs(300000, (0, 1 << 16), (0, 1 << 20), 1, 0.8, seed=3)\n"
            "c, d = G.unique_pairs(300000, (0, 1 << 12), (0, 1 << 12), seed=3, shuffle=True)\n"
            "e = G.unique_column(300000, 1000, 3, 1)\n"
            "print(hashlib.md5(b''.join(x.tobytes() for x in (a, b, c, d, e))).hexdigest())\n"
            % ROOT)
    outs = set()
    for t in ("1", "3", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        outs.add(subprocess.run([sys.executable, "-c", code], env=env, check=True,
                                capture_output=True, text=True).stdout.strip())
    assert len(outs) == 1


def test_z_name_matches_the_reference_format():
    assert [W.z_name(z) for z in (1.0, 0.5, 1.25, 0.0, 2.0, 0.125)] == \
        ["1", "0.5", "1.25", "0", "2", "0.125"]


def test_uniform_and_skewed_writer_command_lines(tmp_path):
    d = str(tmp_path)
    assert W.main(["--seed", "4", "uniform", "1000", "2", "50", "3", "0", "200", "4", "1", d]) == 0
    cols, ids = formats.read_db(os.path.join(d, "1000_3_4.db"))
    assert ids == [3, 4] and cols[0].max() < 50 and cols[0].min() >= 0
    assert np.array_equal(np.sort(cols[1]), np.sort(np.arange(1000) % 200))
    assert W.main(["skewed", "5000", "2", "100", "1.0", "0", "70", "0", "1", d, "0"]) == 0
    cols, ids = formats.read_db(os.path.join(d, "5000_0_1.db"))
    assert ids == [0, 1] and np.argmax(np.bincount(cols[0])) == 0  # SHUFFLE 0: value 0 on top
    assert np.array_equal(np.sort(cols[1]), np.sort(np.arange(5000) % 70))
    assert W.main(["uniform", "10", "2", "5", "0", "0", d]) == 1  # wrong parameter count
    assert W.main([]) == 2


def test_triangle_writers_name_and_fill_their_directories(tmp_path):
    d = str(tmp_path)
    assert W.main(["triangle", "2000", "100", d]) == 0
    sub = os.path.join(d, "tri_2000_100_uu")
    assert sorted(os.listdir(sub)) == ["2000_0_1.db", "2000_0_2.db", "2000_1_2.db"]
    for f in os.listdir(sub):
        (k0, k1), _ = formats.read_db(os.path.join(sub, f))
        assert np.all(np.diff(_codes(k0, k1, 100)) > 0)
    assert W.main(["triangle", "2000", "100", d]) == 1  # the directory exists: refused
    assert W.main(["skewed-triangle", "3000", "2000", "0.5", "0", "1", d]) == 0
    sub = os.path.join(d, "tri_3000_2000_us_0.5")
    assert sorted(os.listdir(sub)) == ["3000_0_1.db", "3000_1_2.db", "3000_2_0.db"]
    (k0, k1), ids = formats.read_db(os.path.join(sub, "3000_2_0.db"))
    assert ids == [2, 0]
    assert np.all(np.diff(_codes(k1, k0, 2000)) > 0)  # ordered by the skewed column (key1)
    assert W.main(["skewed-triangle", "10", "10", "1", "1", "1", d]) == 1


@pytest.mark.gpu
def test_triangle_writer_tries_and_count(ctx, tmp_path):
    import oracle_ffi as O
    assert W.main(["--tries", "triangle", "3000", "120", str(tmp_path)]) == 0
    sub = os.path.join(str(tmp_path), "tri_3000_120_uu")
    rels, tries = [], []
    for a in ((0, 1), (1, 2), (0, 2)):
        cols, ids = formats.read_db(os.path.join(sub, f"3000_{a[0]}_{a[1]}.db"))
        rels.append((cols, ids))
        t = ctx.load_trie(os.path.join(sub, f"3000_{a[0]}_{a[1]}.tr"))
        d0, off, d1 = t.levels()
        # distinct sorted pairs: the trie holds them all, in order
        assert np.array_equal(np.repeat(d0, np.diff(off.astype(np.int64))), cols[0])
        assert np.array_equal(d1, cols[1])
        tries.append(t)
    want = O.brute_join(rels)
    r = ctx.lftj(tries, [0, 1, 2], fetch=True)
    assert r.count == len(want)
    assert [tuple(int(x) for x in t) for t in O.sort_tuples(r.tuples())] == want
    for t in tries:
        t.close()
