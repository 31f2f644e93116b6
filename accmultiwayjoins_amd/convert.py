"""Edge-list ingest: the reference's converters exec_el_2_db (MHJ-GPU/exec_el_2_db.cu:22-87) and
edgelist_to_csr (LFTJ-GPU/edgelist_to_csr.cu:22-87), as library functions and a command line.

An edge list is a text file with one edge per line, two non-negative integers separated by
spaces, tabs or '|' (the reference's separator set " |\\t"). `el_to_db` writes the two columns
as a .db relation (<rows>_<a0>_<a1>.db, column 0 then column 1 as raw int32); `el_to_tr`
builds the sorted two-level trie on the device and writes it as a .tr file (the layout of
IndexedTrie::serialization, IndexedTrie.cuh:122-197).

`tbl_to_db` is the TPC-H converter tbl_2_db (MHJ-GPU/exec_tbl_2_db.cu:21-134): chosen fields
of a dbgen .tbl file (fields split on '|' and tabs, empty fields skipped as strtok does), each
read as the leading integer of the field (dates give their year, decimals their integer part),
written as <rows>_<new ids>.db.

    python -m accmultiwayjoins_amd.convert el2db <edge-list> [<directory>]
    python -m accmultiwayjoins_amd.convert el2tr <edge-list> [<output.tr>]
    python -m accmultiwayjoins_amd.convert tbl2db <file.tbl> <field>... <new id>... [--dir D]
"""
import os
import re
import sys
from typing import List, Sequence, Tuple

import numpy as np

from . import formats

_SEP = re.compile(rb"[ |\t]+")


def read_edge_list(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """(column 0, column 1) of a text edge list; blank lines are skipped."""
    with open(path, "rb") as f:
        data = f.read()
    vals = np.array(_SEP.sub(b" ", data).split(), dtype=np.int64)
    if len(vals) % 2:
        raise ValueError(f"{path}: odd number of values; expected two per line")
    if len(vals) and (vals.min() < 0 or vals.max() > np.iinfo(np.int32).max):
        raise ValueError(f"{path}: values must fit a non-negative int32")
    pairs = vals.reshape(-1, 2).astype(np.int32)
    return np.ascontiguousarray(pairs[:, 0]), np.ascontiguousarray(pairs[:, 1])


def el_to_db(path: str, directory: str = None, attr_ids: Sequence[int] = (0, 1)) -> str:
    """Write the edge list as <directory>/<rows>_<a0>_<a1>.db (default: beside the input)."""
    src, dst = read_edge_list(path)
    return formats.write_db(directory or os.path.dirname(os.path.abspath(path)), [src, dst],
                            attr_ids)


def el_to_tr(path: str, out_path: str = "output.tr", ctx=None) -> str:
    """Build the trie of the edge list on the device and write it as a .tr file."""
    from .amj import Context
    src, dst = read_edge_list(path)
    own = ctx is None
    ctx = ctx or Context(0)
    try:
        t = ctx.trie([src, dst])
        t.save(out_path)
        t.close()
    finally:
        if own:
            ctx.close()
    return out_path


_TBL_SEP = re.compile(rb"[\t|]+")
_LEADING_INT = re.compile(rb"\s*([+-]?\d+)")


def _leading_int(tok: bytes, line_no: int) -> int:
    m = _LEADING_INT.match(tok)
    if not m:
        raise ValueError(f"line {line_no}: field {tok[:40]!r} does not start with an integer")
    return int(m.group(1))


def read_tbl(path: str, fields: Sequence[int]) -> List[np.ndarray]:
    """The given fields (0-based) of every line of a .tbl file, as int32 columns. A value is
    the field's leading integer taken modulo 2^32 (the reference parses with stoul and stores
    an int)."""
    if not fields:
        raise ValueError("no fields to read")
    vals = [[] for _ in fields]
    want = sorted(set(fields))
    with open(path, "rb") as f:
        for n, line in enumerate(f, 1):
            toks = [t for t in _TBL_SEP.split(line.rstrip(b"\n")) if t]
            if toks and len(toks) <= want[-1]:
                raise ValueError(f"{path}:{n}: {len(toks)} fields, field {want[-1]} requested")
            if not toks:  # an empty line still counts as a row in the reference (wc -l)
                raise ValueError(f"{path}:{n}: empty line")
            got = {k: _leading_int(toks[k], n) for k in want}
            for j, k in enumerate(fields):
                vals[j].append(got[k])
    return [((np.array(v, dtype=np.int64) + (1 << 31)) % (1 << 32) - (1 << 31)).astype(np.int32)
            for v in vals]


def tbl_to_db(path: str, fields: Sequence[int], new_ids: Sequence[int],
              directory: str = ".") -> str:
    """Write fields of a .tbl file as <directory>/<rows>_<new ids>.db (tbl_2_db)."""
    if len(fields) != len(new_ids):
        raise ValueError("one new attribute id per retrieved field")
    return formats.write_db(directory, read_tbl(path, fields), list(new_ids))


def _usage() -> int:
    for line in __doc__.strip().splitlines()[-3:]:
        print(line.strip(), file=sys.stderr)
    return 2


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if len(argv) < 2 or argv[0] not in ("el2db", "el2tr", "tbl2db"):
        return _usage()
    if argv[0] == "el2db":
        out = el_to_db(argv[1], argv[2] if len(argv) > 2 else None)
    elif argv[0] == "el2tr":
        out = el_to_tr(argv[1], argv[2] if len(argv) > 2 else "output.tr")
    else:
        directory = "."
        if "--dir" in argv:
            i = argv.index("--dir")
            if i + 1 >= len(argv):
                return _usage()
            directory = argv[i + 1]
            del argv[i:i + 2]
        ids = argv[2:]
        if not ids or len(ids) % 2:
            return _usage()
        k = len(ids) // 2
        try:
            out = tbl_to_db(argv[1], [int(x) for x in ids[:k]], [int(x) for x in ids[k:]],
                            directory)
        except (ValueError, OSError) as e:
            print(f"error: {e}", file=sys.stderr)
            return 1
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
