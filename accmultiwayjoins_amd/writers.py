"""The reference's data-writer command lines, over this library's seeded generators.

Each writes .db relations (<rows>_<attr>_<attr>...db, raw int32 columns) in the reference's
naming, so the files load unchanged into exec-AMHJ / exec-LFTJ and into the reference. The
reference seeds these tools from the clock; here every output is a pure function of --seed.

    python -m accmultiwayjoins_amd.writers [--seed S] uniform TABLE_CARD NUM_COLS
        (COL_RANGE COL_ATTR GEN_TYPE)... BASE_ADDR
      MHJ-GPU/uniform_writer_.cu:98-132. GEN_TYPE 0: uniform in [0, COL_RANGE);
      1: the values j % COL_RANGE in random order.
    python -m accmultiwayjoins_amd.writers [--seed S] skewed TABLE_CARD NUM_COLS
        (COL_RANGE COL_ALPHA COL_ATTR)... BASE_ADDR SHUFFLE
      MHJ-GPU/skewed_writer_.cu:45-88. ALPHA 0: j % COL_RANGE in random order; ALPHA > 0:
      Zipf(ALPHA) over [0, COL_RANGE), ranks mapped to values by a seeded permutation when
      SHUFFLE is 1, rank r -> value r-1 when it is 0.
    python -m accmultiwayjoins_amd.writers [--seed S] [--tries] triangle CNT RANGE BASE_ADDR
      LFTJ-GPU/uniform_writer.cu:40-150: directory BASE_ADDR/tri_CNT_RANGE_uu with three
      tables of CNT distinct uniform pairs over [0, RANGE)^2, attributes (0,1), (1,2), (0,2);
      --tries also writes each table's trie as a .tr file (built on the device).
    python -m accmultiwayjoins_amd.writers [--seed S] skewed-triangle CNT RANGE Z KEY0_TYPE
        KEY1_TYPE BASE_ADDR
      LFTJ-GPU/skewed_writer.cu:19-69: one column Zipf(Z), the other uniform (type 1 =
      skewed, 0 = uniform; exactly one is skewed), CNT distinct pairs per table; directory
      BASE_ADDR/tri_CNT_RANGE_su_Z (key0 skewed) or _us_Z, attributes (0,1), (1,2), (2,0).
"""
from __future__ import annotations

import os
import sys
from typing import List, Sequence

from . import formats
from . import generators as G


def z_name(z: float) -> str:
    """A double as the reference prints it in directory names (helper.h:94-106): six decimals
    with trailing zeros removed, and no decimal point for whole numbers."""
    s = f"{z:.6f}"
    whole, frac = s.split(".")
    frac = frac.rstrip("0")
    return whole if not frac else f"{whole}.{frac}"


def uniform_table(card: int, ranges: Sequence[int], attrs: Sequence[int],
                  gen_types: Sequence[int], base_addr: str, seed: int = 42) -> str:
    """One table of len(ranges) columns (uniform_column_generator, uniform_writer_.cu:65-96)."""
    cols = []
    for i, (r, g) in enumerate(zip(ranges, gen_types)):
        if g == 0:
            cols.append(G.uniform(card, 0, r, seed, i))
        elif g == 1:
            cols.append(G.unique_column(card, r, seed, i))
        else:
            raise ValueError(f"wrong gen-type {g} for column {i} (0 uniform, 1 unique)")
    return formats.write_db(base_addr, cols, attrs)


def skewed_table(card: int, ranges: Sequence[int], alphas: Sequence[float], attrs: Sequence[int],
                 base_addr: str, shuffle: bool, seed: int = 42) -> str:
    """One table of Zipf / unique-modulo columns (skewed_writer, skewed_writer_.cu:15-43)."""
    cols = []
    for i, (r, a) in enumerate(zip(ranges, alphas)):
        if a < 0:
            raise ValueError(f"negative alpha for column {i}")
        if a == 0:
            cols.append(G.unique_column(card, r, seed, i))
        else:
            cols.append(G.zipf_mapped(card, r, a, seed, i, shuffle))
    return formats.write_db(base_addr, cols, attrs)


def _new_dir(path: str) -> str:
    # the reference stops when the directory exists (mkdir fails)
    os.mkdir(path)
    return path


def _save_trie(ctx, cols, attrs, path):
    t = ctx.trie(list(cols), attr_ids=tuple(attrs))
    try:
        t.save(path)
    finally:
        t.close()


def uniform_triangle(cnt: int, value_range: int, base_addr: str, seed: int = 42,
                     tries: bool = False) -> List[str]:
    """writeUniformTriangle (LFTJ-GPU/uniform_writer.cu:40-131): three tables of distinct
    uniform pairs; table i joins attributes i and (i+1) % 3, named with the smaller first."""
    d = _new_dir(os.path.join(base_addr, f"tri_{cnt}_{value_range}_uu"))
    out = []
    ctx = None
    if tries:
        from .amj import Context
        ctx = Context(0)
    try:
        for i in range(3):
            k0, k1 = G.unique_pairs(cnt, (0, value_range), (0, value_range), seed=seed + i)
            a = sorted((i, (i + 1) % 3))
            out.append(formats.write_db(d, [k0, k1], a))
            if ctx is not None:
                p = os.path.join(d, f"{cnt}_{a[0]}_{a[1]}.tr")
                _save_trie(ctx, [k0, k1], a, p)
                out.append(p)
    finally:
        if ctx is not None:
            ctx.close()
    return out


def skewed_triangle(cnt: int, value_range: int, z: float, key0_type: int, key1_type: int,
                    base_addr: str, seed: int = 42) -> List[str]:
    """writeSkewUSUniqueTriangle (LFTJ-GPU/skewed_writer.cu:19-57)."""
    if sorted((key0_type, key1_type)) != [0, 1]:
        raise ValueError("exactly one of KEY0_TYPE, KEY1_TYPE must be 1 (skewed)")
    skew_col = 0 if key0_type == 1 else 1
    tag = "su" if skew_col == 0 else "us"
    d = _new_dir(os.path.join(base_addr, f"tri_{cnt}_{value_range}_{tag}_{z_name(z)}"))
    out = []
    for i in range(3):
        k0, k1 = G.zipf_pairs(cnt, (0, value_range), (0, value_range), skew_col, z,
                              seed=seed + i)
        out.append(formats.write_db(d, [k0, k1], [i, (i + 1) % 3]))
    return out


def _usage() -> int:
    print(__doc__.strip(), file=sys.stderr)
    return 2


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    seed, tries = 42, False
    while argv and argv[0].startswith("--"):
        opt = argv.pop(0)
        if opt == "--seed" and argv:
            seed = int(argv.pop(0))
        elif opt == "--tries":
            tries = True
        else:
            return _usage()
    if not argv:
        return _usage()
    cmd, a = argv[0], argv[1:]
    try:
        if cmd == "uniform":
            card, nc = int(a[0]), int(a[1])
            if len(a) != 3 * nc + 3:
                raise ValueError("wrong parameters")
            spec = a[2:2 + 3 * nc]
            out = [uniform_table(card, [int(x) for x in spec[0::3]], [int(x) for x in spec[1::3]],
                                 [int(x) for x in spec[2::3]], a[-1], seed)]
        elif cmd == "skewed":
            card, nc = int(a[0]), int(a[1])
            if len(a) != 3 * nc + 4:
                raise ValueError("wrong parameters")
            spec = a[2:2 + 3 * nc]
            out = [skewed_table(card, [int(x) for x in spec[0::3]],
                                [float(x) for x in spec[1::3]], [int(x) for x in spec[2::3]],
                                a[-2], bool(int(a[-1])), seed)]
        elif cmd == "triangle":
            if len(a) != 3:
                raise ValueError("wrong parameters")
            out = uniform_triangle(int(a[0]), int(a[1]), a[2], seed, tries)
        elif cmd == "skewed-triangle":
            if len(a) != 6:
                raise ValueError("wrong parameters")
            out = skewed_triangle(int(a[0]), int(a[1]), float(a[2]), int(a[3]), int(a[4]), a[5],
                                  seed)
        else:
            return _usage()
    except (ValueError, IndexError, OSError) as e:
        print(f"error: {e}", file=sys.stderr)
        return 1
    for p in out:
        print(p)
    return 0


if __name__ == "__main__":
    sys.exit(main())
