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

    def capture(self, slab_bytes: int = 0):
        """Record the direct-output chain into a HIP graph (every query added with dro=1)."""
        self.ctx._check(N.lib().amj_batch_capture(self.h, int(slab_bytes)))

    def replay(self) -> float:
        """Launch the captured graph; results as after run(). Returns the device time (ms)."""
        self.ctx._check(N.lib().amj_batch_replay(self.h))
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
                 plain=False):
        """plain: False (continuous batching), True or 1 (one query per launch, count + write),
        2 (one query per launch with direct output)."""
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

    def add_device_relation(self, cols: "DeviceColumns", attr_ids=(0, 1)) -> int:
        """Register device columns (amj_rel_t.on_device = 1): the server keeps their pointers,
        so they must stay allocated while the server runs."""
        ptrs = (C.c_void_p * len(cols.ptrs))(*cols.ptrs)
        ids = (C.c_uint32 * len(attr_ids))(*[int(x) for x in attr_ids])
        rel = N.AmjRel()
        rel.num_cols = len(cols.ptrs)
        rel.rows = cols.rows
        rel.cols = C.cast(ptrs, C.POINTER(C.c_void_p))
        rel.attr_ids = C.cast(ids, C.POINTER(C.c_uint32))
        rel.on_device = 1
        rid = C.c_uint32()
        self._check(N.lib().amj_server_add_relation(self.h, C.byref(rel), C.byref(rid)))
        return rid.value

    def set_autotune(self, min_probe_rows: int):
        """Chains probing at least this many rows run alone with autotuned engine options."""
        self._check(N.lib().amj_server_set_autotune(self.h, int(min_probe_rows)))

    def set_triejoin(self, min_edges: int):
        """TRI / FOUR over edge relations of at least this many rows run alone through the
        leapfrog triejoin (a trie built once per relation and kept resident); 0: off."""
        self._check(N.lib().amj_server_set_triejoin(self.h, int(min_edges)))

    def tuning(self) -> dict:
        a, k, t = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._check(N.lib().amj_server_tuning(self.h, C.byref(a), C.byref(k)))
        self._check(N.lib().amj_server_triejoin_runs(self.h, C.byref(t)))
        return {"tuned_alone": a.value, "classes": k.value, "triejoin_alone": t.value}

    def set_budget(self, hbm_budget_bytes: int):
        """Result budget, before the first request (the arena is reserved at the first one)."""
        self._check(N.lib().amj_server_set_budget(self.h, int(hbm_budget_bytes)))

    def memory(self) -> dict:
        r, b, f = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._check(N.lib().amj_server_memory(self.h, C.byref(r), C.byref(b), C.byref(f)))
        return {"resident_bytes": r.value, "budget_bytes": b.value, "device_free_bytes": f.value}

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

    def wait(self, ticket: int, fetch: bool = False, release: bool = True, timing=None):
        """(count, checksum, latency_ms, tuples or None); releases the result unless asked.
        A dict passed as `timing` gets the request's queue_ms (submission to the start of the
        launch or batch that served it)."""
        res = N.AmjResult()
        lat, qms = C.c_double(), C.c_double()
        self._check(N.lib().amj_server_wait_ex(self.h, ticket, C.byref(res), C.byref(lat),
                                               C.byref(qms)))
        if timing is not None:
            timing["queue_ms"] = qms.value
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


class DeviceColumns:
    """int32 columns in device memory owned by a context (amj_device_alloc), for relations
    registered with on_device = 1; filled by the device generators or copied from host arrays."""

    def __init__(self, ctx, rows: int, ncols: int = 2):
        self.ctx, self.rows, self.ptrs = ctx, int(rows), []
        self.alloc_rows = self.rows  # rows allocated (rows may shrink to those filled)
        for _ in range(ncols):
            p = C.c_void_p()
            ctx._check(N.lib().amj_device_alloc(ctx.h, 4 * max(self.rows, 1), C.byref(p)))
            self.ptrs.append(p.value)

    def uniform(self, col: int, lo: int, hi: int, seed: int, stream: int):
        self.ctx._check(N.lib().amj_gen_device_uniform(self.ctx.h, self.ptrs[col], self.rows,
                                                       int(lo), int(hi), seed, stream))

    def zipf(self, col: int, keys: int, s: float, seed: int, stream: int):
        self.ctx._check(N.lib().amj_gen_device_zipf(self.ctx.h, self.ptrs[col], self.rows,
                                                    int(keys), C.c_double(s), seed, stream))

    def to_host(self, col: int) -> np.ndarray:
        a = np.empty(self.rows, dtype=np.int32)
        self.ctx._check(N.lib().amj_memcpy(self.ctx.h, a.ctypes.data, self.ptrs[col],
                                           4 * self.rows))
        return a

    def close(self):
        for p in self.ptrs:
            N.lib().amj_device_free(self.ctx.h, p, 4 * max(self.alloc_rows, 1))
        self.ptrs = []


def device_rmat(ctx, scale: int, edge_factor: int = 16, seed: int = 42,
                abc=(0.57, 0.19, 0.19)) -> DeviceColumns:
    """An R-MAT graph generated on the device, oriented and sorted: the same edges as
    generators.rmat_oriented(scale, edge_factor, seed)."""
    m = edge_factor << scale
    cols = DeviceColumns(ctx, m)
    k = C.c_uint64()
    ctx._check(N.lib().amj_gen_device_rmat(ctx.h, scale, edge_factor, C.c_double(abc[0]),
                                           C.c_double(abc[1]), C.c_double(abc[2]), seed,
                                           cols.ptrs[0], cols.ptrs[1], C.byref(k)))
    cols.rows = k.value
    return cols

