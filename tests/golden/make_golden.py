"""Writes tests/golden/mhj_cpu_counts.json: join counts printed by the reference's own CPU
multi-way hash join (MHJ-CPU/MHJ.cpp, built by `make -C oracle ref` into oracle/_ref/exec-MHJ)
on chain datasets produced by this package's deterministic generators; and
tests/golden/res_stat_md5.json: the reference's result fingerprint (res_stat,
common-utils/res_stat.h, built with a small driver into oracle/_ref/res_stat_ref) of columns
from the same generators.

Run in the build container only (needs the reference sources to build the binaries):
    python tests/golden/make_golden.py             # join counts
    python tests/golden/make_golden.py --res-stat  # result fingerprints

MHJ-CPU takes a directory of `<rows>_<attr0>_<attr1>.db` files (raw int32 columns, column 0
then column 1; common-utils/file_io.h:123-152) and uses readdir order as the table order
(MHJ.cpp:503-533), hashing every build table on its first column. The files are written to a
tmpfs directory and recreated until readdir returns them in chain order.
"""
from __future__ import annotations

import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
from accmultiwayjoins_amd import generators as G  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "exec-MHJ")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mhj_cpu_counts.json")

# name, relation sizes, key range, zipf s (None = uniform), payload range, seed
DATASETS = [
    ("c1_3x1M_uniform_seed42", [1_000_000] * 3, 1_000_000, None, None, 42),
    ("chain3_4096_keys1024", [4096] * 3, 1024, None, None, 42),
    ("chain3_20000_keys300", [20000, 5000, 7000], 300, None, None, 7),
    ("chain4_s0_1M_256K_keys64K", [1 << 20] + [1 << 18] * 3, 1 << 16, None, 1 << 31, 42),
    ("chain4_s05_1M_256K_keys64K", [1 << 20] + [1 << 18] * 3, 1 << 16, 0.5, 1 << 31, 42),
    ("chain4_s1_1M_256K_keys64K", [1 << 20] + [1 << 18] * 3, 1 << 16, 1.0, 1 << 31, 42),
    ("chain5_small_keys500", [3000, 2000, 2500, 1500, 1000], 500, None, None, 11),
    ("chain4_s1_small_keys100", [5000, 2000, 2000, 2000], 100, 1.0, 1 << 31, 5),
]


def write_db(dirname, rows, attrs, cols):
    path = os.path.join(dirname, "%d_%s.db" % (rows, "_".join(str(a) for a in attrs)))
    with open(path, "wb") as f:
        for c in cols:
            f.write(np.ascontiguousarray(c, dtype="<i4").tobytes())
    return os.path.basename(path)


def materialise_in_chain_order(rels, base):
    names = ["%d_%s.db" % (len(c[0]), "_".join(str(a) for a in ids)) for c, ids in rels]
    if len(set(names)) != len(names):
        raise ValueError("two relations would share a file name")
    for attempt in range(64):
        d = tempfile.mkdtemp(prefix="amj_golden_", dir=base)
        order = list(range(len(rels)))
        if attempt % 2 == 0:
            order.reverse()
        rng = np.random.default_rng(attempt)
        if attempt >= 2:
            rng.shuffle(order)
        for i in order:
            cols, ids = rels[i]
            write_db(d, len(cols[0]), ids, cols)
        if os.listdir(d) == names:
            return d
        shutil.rmtree(d)
    raise RuntimeError("could not obtain chain order from readdir")


def run_mhj(d, threads=8):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    p = subprocess.run([EXE, d], capture_output=True, text=True, env=env, timeout=1800)
    text = p.stdout + p.stderr
    m = re.search(r"Number of join results: (\d+)", text)
    if p.returncode != 0 or not m:
        raise RuntimeError("exec-MHJ failed:\n" + text[-2000:])
    return int(m.group(1))


RES_STAT_EXE = os.path.join(ROOT, "oracle", "_ref", "res_stat_ref")
RES_STAT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "res_stat_md5.json")
# rows, columns, value range [lo, hi), seed
RES_STAT_CASES = [
    (0, 2, 0, 10, 1), (1, 1, 5, 6, 1), (7, 3, -5, 5, 2), (1000, 2, 0, 100, 3),
    (100000, 4, -(1 << 31), (1 << 31) - 1, 4), (50000, 6, 0, 1 << 20, 5),
    (3000, 2, 1 << 20, (1 << 20) + 50, 6), (20000, 1, -1000, 0, 7),
]


def res_stat_columns(rows, ncols, lo, hi, seed):
    return [G.uniform(rows, lo, hi, seed, c) for c in range(ncols)]


def make_res_stat_golden():
    if not os.path.exists(RES_STAT_EXE):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    out = {"generator": "accmultiwayjoins_amd.generators.uniform(rows, lo, hi, seed, column)",
           "cases": []}
    with tempfile.TemporaryDirectory() as d:
        for rows, ncols, lo, hi, seed in RES_STAT_CASES:
            path = os.path.join(d, "cols.bin")
            with open(path, "wb") as f:
                for c in res_stat_columns(rows, ncols, lo, hi, seed):
                    f.write(np.ascontiguousarray(c, dtype="<i4").tobytes())
            p = subprocess.run([RES_STAT_EXE, path, str(rows), str(ncols)], capture_output=True,
                               text=True, timeout=600)
            m = re.search(r"Md5sum of result histogram: ([0-9a-f]{32})", p.stdout + p.stderr)
            if p.returncode != 0 or not m:
                raise RuntimeError("res_stat_ref failed:\n" + (p.stdout + p.stderr)[-2000:])
            rec = {"rows": rows, "cols": ncols, "lo": lo, "hi": hi, "seed": seed,
                   "md5": m.group(1)}
            print(rec, flush=True)
            out["cases"].append(rec)
    with open(RES_STAT_OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", RES_STAT_OUT)


def main():
    if "--res-stat" in sys.argv:
        make_res_stat_golden()
        return
    if not os.path.exists(EXE):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    base = "/dev/shm" if os.path.isdir("/dev/shm") else None
    out = {"generator": "accmultiwayjoins_amd.generators.chain_relations", "datasets": []}
    for name, sizes, keys, s, payload, seed in DATASETS:
        rels = G.chain_relations(sizes, keys, seed=seed, zipf_s=s, payload_range=payload)
        d = materialise_in_chain_order(rels, base)
        try:
            counts = {t: run_mhj(d, t) for t in (1, 8)}
        finally:
            shutil.rmtree(d)
        if len(set(counts.values())) != 1:
            raise RuntimeError(f"{name}: thread counts disagree {counts}")
        rec = {"name": name, "sizes": sizes, "key_range": keys, "zipf_s": s,
               "payload_range": payload, "seed": seed, "count": counts[1]}
        print(rec, flush=True)
        out["datasets"].append(rec)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
