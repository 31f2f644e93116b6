"""Batched and served execution of many join queries (amj_batch_* and amj_server_* of libamj).

Batch: queries over resident relations / hash tables run together, one count launch and one
write launch per group of equal depth (dro=1 queries: a sample launch and one direct-output
launch, rows placed by atomic reservation, so their row order is not the unit engine's). Server: a worker thread keeps relations resident and
serves submitted requests with continuous batching under an HBM budget for the results held
by clients; `plain=True` serves one query per launch instead (the comparison path).
"""
import ctypes as C
from typing import Sequence

import numpy as np

from . import _native as N
from .amj import AmjError, Context, HashTable, JoinResult, Relation, _opts, _RelKeep


class Batch:
    def __init__(self, ctx: Context):
        h = C.c_void_p()
        ctx._check(N.lib().amj_batch_create(ctx.h, C.byref(h)))
        self.ctx, self.h = ctx, h

    def add(self, probe: Relation, tables: Sequence[HashTable], materialize=True, dro=0,
            dro_capacity_rows=0) -> int:
        hts = (C.c_void_p * len(tables))(*[t.h.value for t in tables])
        o = _opts(materialize=materialize, dro=dro, dro_capacity_rows=dro_capacity_rows)
        qid = C.c_uint32()
        self.ctx._check(N.lib().amj_batch_add(self.h, probe.h, hts, len(tables), C.byref(o),
                                              C.byref(qid)))
        return qid.value

    def add_pattern(self, edges: Relation, table: HashTable, query="TRI", materialize=True,
                    dro=0, dro_capacity_rows=0) -> int:
        o = _opts(materialize=materialize, dro=dro, dro_capacity_rows=dro_capacity_rows)
        qid = C.c_uint32()
        self.ctx._check(N.lib().amj_batch_add_pattern(self.h, edges.h, table.h, N.QUERY[query],
                                                      C.byref(o), C.byref(qid)))
        return qid.value

    def set_budget(self, nbytes: int):
        self.ctx._check(N.lib().amj_batch_set_budget(self.h, int(nbytes)))

    def run(self) -> float:
        self.ctx._check(N.lib().amj_batch_run(self.h))
        ms = C.c_double()
        N.lib().amj_batch_info(self.h, None, C.byref(ms), None)
        return ms.value

    def result(self, qid: int, fetch: bool = False) -> JoinResult:
        res = N.AmjResult()
        self.ctx._check(N.lib().amj_batch_result(self.h, qid, C.byref(res)))
        return self.ctx._finish(res, fetch)

    def close(self):
        if self.h:
            N.lib().amj_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """amj_server: resident relations, continuous batching, HBM admission, latency per request."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, max_batch: int = 256,
                 plain: bool = False):
        h = C.c_void_p()
        rc = N.lib().amj_server_create(device, int(hbm_budget_bytes), max_batch, int(plain),
                                       C.byref(h))
        if rc != N.AMJ_OK:
            raise AmjError(rc, "cannot create a server on device %d" % device)
        self.h = h

    def _check(self, rc: int):
        if rc != N.AMJ_OK:
            msg = N.lib().amj_server_last_error(self.h)
            raise AmjError(rc, msg.decode() if msg else "")

    def add_relation(self, cols, attr_ids=(0, 1)) -> int:
        keep = _RelKeep(cols, attr_ids)
        rid = C.c_uint32()
        self._check(N.lib().amj_server_add_relation(self.h, C.byref(keep.rel), C.byref(rid)))
        return rid.value

    def submit(self, rel_ids: Sequence[int], materialize: bool = True) -> int:
        ids = (C.c_uint32 * len(rel_ids))(*rel_ids)
        t = C.c_uint64()
        self._check(N.lib().amj_server_submit(self.h, ids, len(rel_ids), int(materialize),
                                              C.byref(t)))
        return t.value

    def submit_pattern(self, edge_rel: int, query: str = "TRI", materialize: bool = True) -> int:
        t = C.c_uint64()
        self._check(N.lib().amj_server_submit_pattern(self.h, edge_rel, N.QUERY[query],
                                                      int(materialize), C.byref(t)))
        return t.value

    def wait(self, ticket: int, fetch: bool = False, release: bool = True):
        """(count, checksum, latency_ms, tuples or None); releases the result unless asked."""
        res = N.AmjResult()
        lat = C.c_double()
        self._check(N.lib().amj_server_wait(self.h, ticket, C.byref(res), C.byref(lat)))
        tuples = None
        if fetch:
            cols = []
            for c in range(res.num_cols):
                a = np.empty(res.rows_written, dtype=np.int32)
                if res.rows_written:
                    self._check(N.lib().amj_server_copy_col(self.h, C.byref(res), c, 0,
                                                            res.rows_written, a.ctypes.data))
                cols.append(a)
            tuples = np.stack(cols, axis=1) if cols else np.zeros((0, 0), np.int32)
        out = (res.count, res.checksum, lat.value, tuples)
        if release:
            self._check(N.lib().amj_server_release(self.h, C.byref(res)))
        return out

    def stats(self) -> dict:
        b, s, sp, held = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        busy = C.c_double()
        self._check(N.lib().amj_server_stats(self.h, C.byref(b), C.byref(s), C.byref(sp),
                                             C.byref(busy), C.byref(held)))
        return {"batches": b.value, "served": s.value, "splits": sp.value,
                "busy_ms": busy.value, "bytes_held": held.value}

    def close(self):
        if getattr(self, "h", None):
            N.lib().amj_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
