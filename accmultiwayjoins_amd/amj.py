"""Host-side operator interface of the MI355X multi-way joins, over libamj's C ABI.

Mirrors the reference's operator boundary:
  * ``Context.chain_join``   -> exec-AMHJ NORMAL (MHJ-GPU/exec_AMHJ.cu:70-89, :163-203): the
    first relation probes, every other relation is hashed on its first column
    (QueryProcessor::build_hash_multiway, MHJ-GPU/QueryProcessor.cuh:76-110);
  * ``Context.pattern_join`` -> exec-AMHJ TRI / FOUR on one edge relation
    (MHJ-GPU/exec_AMHJ.cu:32-69, QueryProcessor::build_hash_single :112-169);
  * ``Context.execute``      -> OOC<...>::execute (MHJ-GPU/ooc.cuh:77-82) on resident relations
    and hash tables (``Context.relation`` / ``Context.hash_table``).
Switches keep the reference's names: ooc, ws, dro (the reference's `sbp` argument), fib.
Errors raise ``AmjError`` carrying the library's message instead of exiting the process.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _native as N


class AmjError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{N.STATUS.get(code, code)}: {msg}")
        self.code = code


@dataclass
class JoinResult:
    count: int
    checksum: int
    attr_ids: list
    path: str
    build_ms: float
    probe_ms: float
    probe_kernel_ms: float
    cols: Optional[list] = None  # numpy int32 columns when fetched
    rows_written: int = 0
    detail: int = 0   # fib: suffix-table depth; ws: tasks shared between waves
    passes: int = 0   # probe traversals run
    _res: Optional[N.AmjResult] = field(default=None, repr=False)

    def tuples(self) -> np.ndarray:
        """Materialised result as an (count, n_attrs) array (columns in attr_ids order)."""
        if self.cols is None:
            raise ValueError("result was not fetched")
        if not self.cols:
            return np.zeros((0, 0), dtype=np.int32)
        return np.stack(self.cols, axis=1)


def _opts(ooc=0, ws=0, dro=0, fib=0, materialize=True, checksum=True, hbm_budget_bytes=0,
          ooc_slice_rows=0, out_chunk_rows=0, dro_capacity_rows=0, probe_order=0,
          ws_split_threshold=0, fib_suffix=0, engine="wave", ws_stall_ms=0,
          ws_claim_units=0) -> N.AmjOpts:
    o = N.AmjOpts()
    N.lib().amj_default_opts(C.byref(o))
    o.ooc, o.ws, o.dro, o.fib = int(ooc), int(ws), int(dro), int(fib)
    o.materialize = 1 if materialize else 0
    o.checksum = 1 if checksum else 0
    o.hbm_budget_bytes = int(hbm_budget_bytes)
    o.ooc_slice_rows = int(ooc_slice_rows)
    o.out_chunk_rows = int(out_chunk_rows)
    o.dro_capacity_rows = int(dro_capacity_rows)
    o.ws_split_threshold = int(ws_split_threshold)
    o.fib_suffix = int(fib_suffix)
    o.probe_order = int(probe_order)
    o.ws_stall_ms = int(ws_stall_ms)
    o.ws_claim_units = int(ws_claim_units)
    engines = {"wave": 0, "tpt": 1, "open-addr": 2, "pw": 3}
    if engine not in engines:
        raise ValueError(f"engine must be one of {sorted(engines)}")
    o.engine = engines[engine]
    return o


def pinned_empty(n: int, dtype=np.int32) -> np.ndarray:
    """A numpy array in pinned host memory (amj_host_alloc); freed with the array."""
    import weakref
    itemsize = np.dtype(dtype).itemsize
    p = C.c_void_p()
    if N.lib().amj_host_alloc(max(1, n * itemsize), C.byref(p)) != N.AMJ_OK:
        raise MemoryError("amj_host_alloc(%d bytes) failed" % (n * itemsize))
    buf = (C.c_char * max(1, n * itemsize)).from_address(p.value)
    weakref.finalize(buf, N.lib().amj_host_free, p.value)
    return np.frombuffer(buf, dtype=dtype, count=n)


class _RelKeep:
    """Keeps the ctypes arrays behind an AmjRel alive."""

    def __init__(self, cols: Sequence[np.ndarray], attr_ids: Optional[Sequence[int]],
                 host: bool = False):
        self.arrays = [np.ascontiguousarray(c, dtype=np.int32) for c in cols]
        rows = len(self.arrays[0]) if self.arrays else 0
        if any(len(a) != rows for a in self.arrays):
            raise ValueError("columns of a relation must have equal length")
        self.ptrs = (C.c_void_p * len(self.arrays))(*[a.ctypes.data for a in self.arrays])
        self.ids = None
        if attr_ids is not None:
            self.ids = (C.c_uint32 * len(attr_ids))(*[int(x) for x in attr_ids])
        self.rel = N.AmjRel()
        self.rel.num_cols = len(self.arrays)
        self.rel.rows = rows
        self.rel.cols = C.cast(self.ptrs, C.POINTER(C.c_void_p))
        self.rel.attr_ids = C.cast(self.ids, C.POINTER(C.c_uint32)) if self.ids else None
        self.rel.on_device = N.AMJ_REL_HOST if host else 0


class Relation:
    """A relation (Relation<D,C>, MHJ-GPU/Relation.cuh:9-48): copied to HBM, or with
    host=True kept in host memory (pinned_empty arrays stream fastest) as an out-of-core
    probe relation."""

    def __init__(self, ctx: "Context", cols, attr_ids, host: bool = False):
        keep = _RelKeep(cols, attr_ids, host)
        h = C.c_void_p()
        ctx._check(N.lib().amj_relation_create(ctx.h, C.byref(keep.rel), C.byref(h)))
        self.ctx, self.h, self.rows = ctx, h, keep.rel.rows
        self._keep = keep if host else None  # host columns must outlive the relation

    def close(self):
        if self.h:
            N.lib().amj_relation_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HashTable:
    """Bucketed hash table over a relation (HashTable<D,C>, MHJ-GPU/Relation.cuh:50-77)."""

    def __init__(self, ctx: "Context", rel: Relation, key_col: int = 0, sort_vals: bool = False):
        h = C.c_void_p()
        ctx._check(N.lib().amj_hash_table_build(ctx.h, rel.h, key_col, int(sort_vals),
                                                C.byref(h)))
        self.ctx, self.h, self.rel = ctx, h, rel
        rows, buckets, ident, ms = C.c_uint64(), C.c_uint64(), C.c_int(), C.c_double()
        ctx._check(N.lib().amj_hash_table_info(ctx.h, h, C.byref(rows), C.byref(buckets),
                                               C.byref(ident), C.byref(ms)))
        self.rows, self.buckets = rows.value, buckets.value
        self.identity, self.build_ms = bool(ident.value), ms.value

    def close(self):
        if self.h:
            N.lib().amj_hash_table_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Trie:
    """Sorted two-level CSR trie of a two-column relation (IndexedTrie<D,C>,
    LFTJ-GPU/IndexedTrie.cuh:21-54); duplicate pairs are dropped."""

    def __init__(self, ctx: "Context", cols=None, attr_ids=(0, 1), handle=None):
        if handle is None:
            keep = _RelKeep(cols, attr_ids)
            h = C.c_void_p()
            ctx._check(N.lib().alftj_trie_build(ctx.h, C.byref(keep.rel), C.byref(h)))
        else:
            h = handle
        self.ctx, self.h, self.attr_ids = ctx, h, tuple(attr_ids)
        n0, n1, mx, ms = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_double()
        ctx._check(N.lib().alftj_trie_info(ctx.h, h, C.byref(n0), C.byref(n1), C.byref(mx),
                                           C.byref(ms)))
        self.level0, self.children = n0.value, n1.value
        self.max_children, self.build_ms = mx.value, ms.value

    def offload(self):
        """Move the trie to pinned host memory; queries over it then run out of core."""
        self.ctx._check(N.lib().alftj_trie_offload(self.ctx.h, self.h))

    def save(self, path: str):
        """Write the trie as a .tr file (IndexedTrie::serialization layout)."""
        self.ctx._check(N.lib().alftj_trie_save(self.ctx.h, self.h, path.encode()))

    def levels(self):
        """(level-0 values, level-0 offsets, level-1 values) as host arrays."""
        d0 = np.empty(self.level0, dtype=np.int32)
        off = np.empty(self.level0 + 1, dtype=np.uint32)
        d1 = np.empty(self.children, dtype=np.int32)
        self.ctx._check(N.lib().alftj_trie_copy_level(self.ctx.h, self.h, 0, d0.ctypes.data,
                                                      off.ctypes.data))
        self.ctx._check(N.lib().alftj_trie_copy_level(self.ctx.h, self.h, 1, d1.ctypes.data,
                                                      None))
        return d0, off, d1

    def close(self):
        if self.h:
            N.lib().alftj_trie_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One GPU, one HIP stream. Not shared between host threads."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = N.lib().amj_ctx_create(device, C.byref(h))
        if rc != N.AMJ_OK:
            raise AmjError(rc, "cannot create a context on device %d" % device)
        self.h = h

    # -- plumbing ---------------------------------------------------------------------------
    def _check(self, rc: int):
        if rc != N.AMJ_OK:
            msg = N.lib().amj_last_error(self.h)
            raise AmjError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            N.lib().amj_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_info(self):
        name = C.create_string_buffer(256)
        hbm = C.c_uint64()
        cus = C.c_int()
        self._check(N.lib().amj_device_info(self.h, name, 256, C.byref(hbm), C.byref(cus)))
        return name.value.decode(), hbm.value, cus.value

    def _finish(self, res: N.AmjResult, fetch: bool) -> JoinResult:
        out = JoinResult(
            count=res.count, checksum=res.checksum,
            attr_ids=[res.attr_ids[i] for i in range(res.num_cols)],
            path=N.PATHS.get(res.path, str(res.path)), build_ms=res.build_ms,
            probe_ms=res.probe_ms, probe_kernel_ms=res.probe_kernel_ms,
            rows_written=res.rows_written, detail=res.detail, passes=res.passes, _res=res)
        if fetch and res.rows_written:
            cols = []
            for c in range(res.num_cols):
                a = np.empty(res.rows_written, dtype=np.int32)
                self._check(N.lib().amj_result_copy_col(self.h, C.byref(res), c, 0,
                                                        res.rows_written, a.ctypes.data))
                cols.append(a)
            out.cols = cols
        elif fetch:
            out.cols = [np.empty(0, dtype=np.int32) for _ in range(res.num_cols)]
        if fetch:
            self.free(out)
        return out

    def free(self, r: JoinResult):
        if r._res is not None:
            self._check(N.lib().amj_result_free(self.h, C.byref(r._res)))
            r._res = None

    # -- operators --------------------------------------------------------------------------
    def chain_join(self, relations, fetch: bool = False, **opts) -> JoinResult:
        """exec-AMHJ NORMAL. relations: [(columns, attr_ids), ...], relations[0] probes."""
        keeps = [_RelKeep(cols, ids) for cols, ids in relations]
        arr = (N.AmjRel * len(keeps))(*[k.rel for k in keeps])
        o = _opts(**opts)
        res = N.AmjResult()
        self._check(N.lib().amj_chain_join(self.h, arr, len(keeps), C.byref(o), C.byref(res)))
        return self._finish(res, fetch)

    def pattern_join(self, src, dst, query: str = "TRI", fetch: bool = False,
                     **opts) -> JoinResult:
        """exec-AMHJ TRI / FOUR over one edge relation E(src, dst)."""
        keep = _RelKeep([src, dst], [0, 1])
        o = _opts(**opts)
        res = N.AmjResult()
        self._check(N.lib().amj_pattern_join(self.h, C.byref(keep.rel), N.QUERY[query],
                                             C.byref(o), C.byref(res)))
        return self._finish(res, fetch)

    def trie(self, cols, attr_ids=(0, 1)) -> Trie:
        return Trie(self, cols, attr_ids)

    def trie_from_relation(self, rel: Relation, attr_ids=(0, 1)) -> Trie:
        """A trie over a device-resident two-column relation."""
        h = C.c_void_p()
        self._check(N.lib().alftj_trie_build_resident(self.h, rel.h, C.byref(h)))
        return Trie(self, attr_ids=tuple(attr_ids), handle=h)

    def trie_from_csr(self, d0, off, d1, attr_ids=(0, 1)) -> Trie:
        """A trie from CSR arrays (level-0 values, u64 offsets, children)."""
        d0 = np.ascontiguousarray(d0, dtype=np.int32)
        off = np.ascontiguousarray(off, dtype=np.uint64)
        d1 = np.ascontiguousarray(d1, dtype=np.int32)
        if len(off) != len(d0) + 1:
            raise ValueError("offsets must have len(d0) + 1 entries")
        ids = (C.c_uint32 * 2)(*attr_ids)
        h = C.c_void_p()
        self._check(N.lib().alftj_trie_from_csr(self.h, len(d0), d0.ctypes.data, off.ctypes.data,
                                                len(d1), d1.ctypes.data, ids, C.byref(h)))
        return Trie(self, attr_ids=tuple(attr_ids), handle=h)

    def load_trie(self, path: str) -> Trie:
        """Read a two-level .tr file."""
        h = C.c_void_p()
        self._check(N.lib().alftj_trie_load(self.h, path.encode(), C.byref(h)))
        with open(path, "rb") as f:
            hdr = f.read(4 + 16 + 8)
        ids = tuple(np.frombuffer(hdr[20:28], dtype=np.uint32).tolist())
        return Trie(self, attr_ids=ids, handle=h)

    def md5(self, r: JoinResult) -> str:
        """res_stat fingerprint of a materialised result still held on the device."""
        if r._res is None:
            raise ValueError("the result has been freed (fetch=True frees it)")
        out = C.create_string_buffer(33)
        self._check(N.lib().amj_result_md5(self.h, C.byref(r._res), out))
        return out.value.decode()

    def lftj(self, tries: Sequence[Trie], attr_order: Sequence[int], algo: str = "DFS",
             fetch: bool = False, trie_attrs=None, **opts) -> JoinResult:
        """LFTJ_Base::evaluate over tries on attr_order; algo "DFS" or "BFS". trie_attrs
        ([(x, y), ...], one pair per trie) overrides the tries' own attribute ids."""
        ts = (C.c_void_p * len(tries))(*[t.h.value for t in tries])
        order = (C.c_uint32 * len(attr_order))(*attr_order)
        o = _opts(**opts)
        res = N.AmjResult()
        if trie_attrs is None:
            rc = N.lib().alftj_run(self.h, ts, len(tries), order, len(attr_order),
                                   N.ALGO[algo], C.byref(o), C.byref(res))
        else:
            flat = (C.c_uint32 * (2 * len(tries)))(*[int(a) for p in trie_attrs for a in p])
            rc = N.lib().alftj_run_aliased(self.h, ts, len(tries), flat, order, len(attr_order),
                                           N.ALGO[algo], C.byref(o), C.byref(res))
        self._check(rc)
        return self._finish(res, fetch)

    def lftj_pattern(self, src, dst, query: str = "TRI", algo: str = "DFS", fetch: bool = False,
                     **opts) -> JoinResult:
        """exec-LFTJ SELF mode: TRI (or FOUR) over one edge relation E(src, dst)."""
        keep = _RelKeep([src, dst], [0, 1])
        o = _opts(**opts)
        res = N.AmjResult()
        self._check(N.lib().alftj_pattern(self.h, C.byref(keep.rel), N.QUERY[query],
                                          N.ALGO[algo], C.byref(o), C.byref(res)))
        return self._finish(res, fetch)

    def autotune(self, probe: Relation, tables: Sequence[HashTable], **opts) -> dict:
        """Engine options chosen by timing candidates on two samples of the probe relation
        (amj_autotune); `estimated_ms` is the pick's device time extrapolated to the whole
        relation. Pass the options back to execute()."""
        hts = (C.c_void_p * len(tables))(*[t.h.value for t in tables])
        o = _opts(**opts)
        ms = C.c_double()
        self._check(N.lib().amj_autotune(self.h, probe.h, hts, len(tables), C.byref(o),
                                         C.byref(ms)))
        return {"ws": o.ws, "dro": o.dro, "fib": o.fib, "out_chunk_rows": o.out_chunk_rows,
                "fib_suffix": o.fib_suffix, "ws_split_threshold": o.ws_split_threshold,
                "ws_claim_units": o.ws_claim_units,
                "materialize": bool(o.materialize), "checksum": bool(o.checksum),
                "estimated_ms": ms.value}

    def execute_pattern(self, edges: Relation, table: HashTable, query: str = "TRI",
                        fetch: bool = False, **opts) -> JoinResult:
        """TRI / FOUR over a resident edge relation and its table (key 0, sort_vals=True)."""
        o = _opts(**opts)
        res = N.AmjResult()
        self._check(N.lib().amj_execute_pattern(self.h, edges.h, table.h, N.QUERY[query],
                                                C.byref(o), C.byref(res)))
        return self._finish(res, fetch)

    def relation(self, cols, attr_ids, host: bool = False) -> Relation:
        return Relation(self, cols, attr_ids, host)

    def allocated_bytes(self) -> int:
        """Device bytes held by this context (relations, tables, scratch, pooled results)."""
        a, p = C.c_uint64(), C.c_uint64()
        self._check(N.lib().amj_memory_info(self.h, C.byref(a), C.byref(p)))
        return a.value

    def hash_table(self, rel: Relation, key_col: int = 0, sort_vals: bool = False) -> HashTable:
        return HashTable(self, rel, key_col, sort_vals)

    def execute(self, probe: Relation, tables: Sequence[HashTable], fetch: bool = False,
                **opts) -> JoinResult:
        """OOC::execute over resident relations and hash tables."""
        hts = (C.c_void_p * len(tables))(*[t.h.value for t in tables])
        o = _opts(**opts)
        res = N.AmjResult()
        self._check(N.lib().amj_execute(self.h, probe.h, hts, len(tables), C.byref(o),
                                        C.byref(res)))
        return self._finish(res, fetch)
