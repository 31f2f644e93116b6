"""Edge-list ingest: the reference's converters exec_el_2_db (MHJ-GPU/exec_el_2_db.cu:22-87) and
edgelist_to_csr (LFTJ-GPU/edgelist_to_csr.cu:22-87), as library functions and a command line.

An edge list is a text file with one edge per line, two non-negative integers separated by
spaces, tabs or '|' (the reference's separator set " |\\t"). `el_to_db` writes the two columns
as a .db relation (<rows>_<a0>_<a1>.db, column 0 then column 1 as raw int32); `el_to_tr`
builds the sorted two-level trie on the device and writes it as a .tr file (the layout of
IndexedTrie::serialization, IndexedTrie.cuh:122-197).

    python -m accmultiwayjoins_amd.convert el2db <edge-list> [<directory>]
    python -m accmultiwayjoins_amd.convert el2tr <edge-list> [<output.tr>]
"""
import os
import re
import sys
from typing import Sequence, Tuple

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


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) < 2 or argv[0] not in ("el2db", "el2tr"):
        print(__doc__.strip().splitlines()[-2].strip(), file=sys.stderr)
        print(__doc__.strip().splitlines()[-1].strip(), file=sys.stderr)
        return 2
    if argv[0] == "el2db":
        out = el_to_db(argv[1], argv[2] if len(argv) > 2 else None)
    else:
        out = el_to_tr(argv[1], argv[2] if len(argv) > 2 else "output.tr")
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
