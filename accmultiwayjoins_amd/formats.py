"""The reference's file formats and result fingerprint, through libamj's host-side C ABI.

.db relation files: <rows>_<attr0>_<attr1>.db holding column 0 then column 1 as raw int32
(read_rel_cols_mmap, common-utils/file_io.h:90-121). res_stat: MD5 of the printed histogram
of the low 20 bits of every result value (common-utils/res_stat.h:32-133).
"""
import ctypes as C
import os
from typing import List, Sequence, Tuple

import numpy as np

from . import _native as N


def _err(rc: int):
    if rc != N.AMJ_OK:
        msg = N.lib().amj_file_error()
        raise OSError(msg.decode() if msg else "file error %d" % rc)


def parse_db_name(path: str) -> Tuple[int, List[int]]:
    """(rows, attribute ids) from a .db file name."""
    rows, nc = C.c_uint64(), C.c_uint32()
    ids = (C.c_uint32 * 8)()
    _err(N.lib().amj_db_parse_name(path.encode(), C.byref(rows), C.byref(nc), ids, 8))
    return rows.value, [ids[i] for i in range(min(nc.value, 8))]


def read_db(path: str) -> Tuple[List[np.ndarray], List[int]]:
    """Columns and attribute ids of a .db file."""
    rows, ids = parse_db_name(path)
    cols = [np.empty(rows, dtype=np.int32) for _ in ids]
    ptrs = (C.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
    _err(N.lib().amj_db_read(path.encode(), rows, len(cols), ptrs))
    return cols, ids


def write_db(directory: str, cols: Sequence[np.ndarray], attr_ids: Sequence[int]) -> str:
    """Write <directory>/<rows>_<attr ids>.db and return its path."""
    from .amj import _RelKeep
    keep = _RelKeep(cols, attr_ids)
    out = C.create_string_buffer(4096)
    _err(N.lib().amj_db_write(os.fspath(directory).encode(), C.byref(keep.rel), out, 4096))
    return out.value.decode()


def res_stat_md5(cols: Sequence[np.ndarray]) -> str:
    """res_stat fingerprint of host result columns."""
    arrs = [np.ascontiguousarray(c, dtype=np.int32) for c in cols]
    rows = len(arrs[0]) if arrs else 0
    ptrs = (C.c_void_p * max(len(arrs), 1))(*[a.ctypes.data for a in arrs])
    out = C.create_string_buffer(33)
    _err(N.lib().amj_res_stat_md5(ptrs, len(arrs), rows, out))
    return out.value.decode()
