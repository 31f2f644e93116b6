"""Benchmark: join output tuples/s of the AMHJ 4-way chain join on one MI355X.

Workload (one step): R1(a,b) x R2(b,c) x R3(c,d) x R4(d,e), int32 columns, R1 = 64M rows,
R2..R4 = 16M rows each, join columns over K = 4M keys: the build-side join column of every
join is Zipf(s) with a seeded rank->key permutation, the probe-side column uniform; the outer
columns a, e uniform over [0, 2^31). A step runs the three skews s = 0, 0.5, 1.0 back to back;
each join builds the three hash tables and probes with ws=1 dro=1 fib=1, materialising all
~2^32 result tuples (5 int32 columns) in HBM. Inputs are resident in HBM before timing.

A second leg times the triangle query (ALFTJ DFS, R-MAT scale 22) and reports it under
"triangle". Prints one JSON line (see the keys below). --gpus N > 1 exits: this build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E peak


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_inputs(G, scale, s, seed=42):
    n1 = int((64 << 20) * scale)
    nb = int((16 << 20) * scale)
    k = max(1, int((4 << 20) * scale))
    return G.chain_relations([n1, nb, nb, nb], k, seed=seed, zipf_s=s, payload_range=1 << 31)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "amhj_chain4_window_writer_pmc.md")


def measured_traffic():
    """HBM bytes per launch of the writer kernel at s=0 from the committed rocprofv3 PMC summary
    (FETCH_SIZE x the copy-calibrated read factor + WRITE_SIZE), or None without one."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def algorithmic_bytes(rels, count, n_attrs=5):
    """Compulsory HBM bytes of one chain join: every input column read once, every build
    table (key + value + bucket pointers) read once, every result tuple written once."""
    b = 0
    for i, (cols, _) in enumerate(rels):
        n = len(cols[0])
        b += 4 * len(cols) * n
        if i > 0:
            buckets = 1 << (n.bit_length() - 1)
            b += 8 * n + 4 * (buckets + 1)
    return b + 4 * n_attrs * count + 16


def cpu_baseline(rels, sample_rows, threads):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    sample = [([c[:sample_rows] for c in rels[0][0]], rels[0][1])] + list(rels[1:])
    t0 = time.perf_counter()
    cnt, cs, _, pms, qms = O.mhj(sample, threads=threads)
    wall = time.perf_counter() - t0
    return cnt, (pms + qms) / 1e3, wall


def config1_leg(ctx, G, threads, reps=5):
    """The CPU-runnable configuration at full size: the 3-way chain R1(a,b) x R2(b,c) x R3(c,d),
    1 M tuples of two int32 columns per relation, join keys uniform over [0, 1 M), seed 42.
    The oracle's MHJ-CPU restatement runs on the host cores (partition + probe, count +
    checksum), and plain AMHJ on the GPU (ooc=0 ws=0 dro=0 fib=0: count pass, scan, write pass)
    over relations already in HBM: hash builds + probe + materialised output, best of `reps`.
    Both must equal the MHJ-CPU count of the reference's own build recorded in
    tests/golden/mhj_cpu_counts.json."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    N = __import__("accmultiwayjoins_amd._native", fromlist=["lib"])
    rels = G.chain_relations([1_000_000] * 3, 1_000_000, seed=42)
    with open(os.path.join(ROOT, "tests", "golden", "mhj_cpu_counts.json")) as f:
        golden = next(d["count"] for d in json.load(f)["datasets"]
                      if d["name"] == "c1_3x1M_uniform_seed42")
    t0 = time.perf_counter()
    cnt, cs, _, pms, qms = O.mhj(rels, threads=threads)
    cpu_wall = time.perf_counter() - t0
    R = [ctx.relation(c, ids) for c, ids in rels]
    best = None
    try:
        for _ in range(reps):
            N.lib().amj_sync(ctx.h)
            t = time.perf_counter()
            H = [ctx.hash_table(r) for r in R[1:]]
            r = ctx.execute(R[0], H, materialize=True, checksum=True, ws=0, dro=0, fib=0,
                            ooc=0)
            N.lib().amj_sync(ctx.h)
            el = time.perf_counter() - t
            g = (r.count, r.checksum, r.rows_written, r.path)
            ctx.free(r)
            for h in H:
                h.close()
            best = el if best is None else min(best, el)
    finally:
        for x in R:
            x.close()
    if not (g[0] == cnt == golden and g[1] == cs and g[2] == cnt):
        raise SystemExit("bench: config 1 differs between the GPU, the oracle and the "
                         "reference's recorded count")
    log(f"[bench] config 1: {cnt} results; CPU {(pms + qms):.1f} ms ({threads} threads), "
        f"GPU plain {best * 1e3:.2f} ms")
    return {"count": cnt, "checksum": "%#018x" % cs, "reference_count": golden,
            "cpu": {"value": cnt / ((pms + qms) / 1e3), "unit": "tuples/s", "cores": threads,
                    "kind": "port", "ms": round(pms + qms, 3),
                    "wall_ms": round(cpu_wall * 1e3, 3)},
            "gpu_plain": {"value": cnt / best, "unit": "tuples/s", "ms": round(best * 1e3, 3),
                          "path": g[3],
                          "timed": "2 hash builds + count pass + scan + write pass, "
                                   "materialised, relations resident"},
            "workload": "3-way chain, 1 M tuples per relation, keys uniform over [0, 1 M), "
                        "seed 42 (full size)"}


def host_cpu():
    """(model name, logical CPUs, SMT active) of the host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    smt = None
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            smt = f.read().strip() == "1"
    except OSError:
        pass
    return model, os.cpu_count(), smt


def copy_bandwidth_gbs():
    """Device-to-device copy bandwidth (bytes read + written per second; 4 GiB copied, best of
    4, HIP events). Runs in the torch-only --copy-calibration child process."""
    import torch
    src = torch.ones(1 << 30, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    best = None
    for _ in range(4):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        b.synchronize()
        ms = a.elapsed_time(b)
        best = ms if best is None else min(best, ms)
    return 2 * 4 * (1 << 30) / (best / 1e3) / 1e9


def measured_copy_gbs():
    """The copy bandwidth measured by a child process (torch and libamj each bring their own
    HIP runtime, so the measurement does not share a process with the joins); None on failure."""
    import subprocess
    try:
        p = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-calibration"],
                           capture_output=True, text=True, timeout=300)
        return json.loads(p.stdout.strip().splitlines()[-1])["copy_gbs"]
    except Exception as e:  # noqa: BLE001 - reported as missing
        log(f"[bench] copy bandwidth not measured: {e!r}")
        return None


def triangle_leg(ctx, G, scale, steps, warmup):
    """Triangle query E(a,b) x E(b,c) x E(a,c) on an R-MAT graph (edge factor 16, parameters
    0.57/0.19/0.19/0.05, symmetrised, deduplicated, degree-oriented): one step = trie build +
    ALFTJ DFS with direct output (a sample launch sizes the output, one pass writes every
    triangle: 3 int32 columns), verified against the count/write DFS and BFS."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    t0 = time.perf_counter()
    src, dst = G.rmat_oriented(scale, 16, seed=42)
    log(f"[bench] triangle: R-MAT scale {scale}, {len(src)} oriented edges "
        f"({time.perf_counter() - t0:.1f} s)")
    rel = ctx.relation([src, dst], [0, 1])  # edges resident in HBM before timing

    tri = [(0, 1), (1, 2), (0, 2)]  # SELF mode: one trie aliased three times

    def run(algo, **opts):
        t = ctx.trie_from_relation(rel)
        r = ctx.lftj([t, t, t], [0, 1, 2], algo, trie_attrs=tri, **opts)
        r.build_ms = t.build_ms
        t.close()
        return r

    def one():
        r = run("DFS", materialize=True, checksum=False, dro=1)
        out = (r.count, r.build_ms, r.probe_ms, r.probe_kernel_ms)
        ctx.free(r)
        return out

    for _ in range(warmup):
        one()
    N = __import__("accmultiwayjoins_amd._native", fromlist=["lib"])
    N.lib().amj_sync(ctx.h)
    t = time.perf_counter()
    recs = [one() for _ in range(steps)]
    N.lib().amj_sync(ctx.h)
    el = time.perf_counter() - t
    cnt = recs[-1][0]
    chk = run("DFS", materialize=False)
    bfs = run("BFS", materialize=False)
    if (chk.count, chk.checksum) != (bfs.count, bfs.checksum) or chk.count != cnt:
        raise SystemExit("bench: triangle DFS / BFS disagree")
    d = run("DFS", materialize=True, checksum=True, dro=1)  # the timed engine, checksummed
    ctx.free(d)
    if (d.count, d.checksum) != (chk.count, chk.checksum):
        raise SystemExit("bench: triangle direct output disagrees with the count/write DFS")
    # AMHJ TRI on the same edges (value-sorted table built untimed), for comparison
    ht = ctx.hash_table(rel, 0, sort_vals=True)
    amhj_ms = []
    for k in range(warmup + steps):
        t = time.perf_counter()
        a = ctx.execute_pattern(rel, ht, "TRI", materialize=True, checksum=False, dro=1)
        N.lib().amj_sync(ctx.h)
        if k >= warmup:
            amhj_ms.append((time.perf_counter() - t) * 1e3)
        acount = a.count
        ctx.free(a)
    a = ctx.execute_pattern(rel, ht, "TRI", materialize=False)
    if (a.count, a.checksum) != (chk.count, chk.checksum) or acount != cnt:
        raise SystemExit("bench: triangle AMHJ / ALFTJ disagree")
    ht.close()
    rel.close()
    return {"value": cnt * steps / el, "unit": "tuples/s", "count": cnt,
            "checksum": "%#018x" % chk.checksum, "ms_per_step": el / steps * 1e3,
            "trie_build_ms": round(recs[-1][1], 3), "probe_ms": round(recs[-1][2], 3),
            "write_kernel_ms": round(recs[-1][3], 3), "bfs_ms": round(bfs.probe_ms, 3),
            "workload": f"ALFTJ DFS triangle, R-MAT scale {scale} edge factor 16, "
                        f"{len(src)} oriented edges, trie build + materialise (direct output: "
                        f"sampled size, one pass)",
            "verified_against": "ALFTJ count/write DFS and BFS (count + checksum)",
            "amhj_tri": {"value": cnt / (min(amhj_ms) / 1e3), "unit": "tuples/s",
                         "ms": round(min(amhj_ms), 3),
                         "workload": "AMHJ TRI (amj_execute_pattern, dro=1: sampled size, one "
                                     "materialising pass) on the same resident edges, table "
                                     "built untimed",
                         "verified_against": "ALFTJ DFS (count + checksum)"}}


def autotune_leg(ctx, R, skew, default):
    """The autotuner's pick (amj_autotune: engine, task and claim sizes, direct output, output
    window, suffix tables, timed on two strided samples of the probe relation and extrapolated
    to all of it) for the chain at one skew, then the whole join run with the picked options
    beside the default (untimed for `value`); both must give the default run's count and
    checksum."""
    H = [ctx.hash_table(r) for r in R[1:]]
    try:
        t0 = time.perf_counter()
        pick = ctx.autotune(R[0], H, materialize=True, checksum=True)
        tune_s = time.perf_counter() - t0
        opts = {k: pick[k] for k in ("ws", "dro", "fib", "out_chunk_rows", "fib_suffix",
                                     "ws_split_threshold", "ws_claim_units")}
        r = ctx.execute(R[0], H, materialize=True, checksum=False, **opts)  # as timed steps
        ms, path = r.probe_ms, r.path
        ctx.free(r)
        r = ctx.execute(R[0], H, materialize=True, checksum=True, **opts)
        got = (r.count, "%#018x" % r.checksum, ms, path)
        ctx.free(r)
    finally:
        for h in H:
            h.close()
    if (got[0], got[1]) != (default["count"], default["checksum"]):
        raise SystemExit(f"bench: the autotuned options gave {got[0]}/{got[1]}, the default "
                         f"{default['count']}/{default['checksum']}")
    log(f"[bench] autotune (s={skew}): {opts}, estimated {pick['estimated_ms']:.1f} ms; "
        f"whole join {got[2]:.1f} ms ({got[3]}) vs default {default['probe_ms']:.1f} ms")
    return {"skew": skew, "choice": opts, "path": got[3],
            "estimated_probe_ms": round(pick["estimated_ms"], 3), "tuning_s": round(tune_s, 3),
            "tuned_probe_ms": round(got[2], 3), "default_probe_ms": default["probe_ms"],
            "workload": "the chain of this skew; candidates timed on evenly strided 2 M- and "
                        "4 M-row samples of R1 (materialised), extrapolated to all of R1",
            "verified": "count + checksum equal to the default run"}


def ablation_leg(ctx, resident, s, checked, wave_probe_ms, G):
    """The comparison engines on the same resident chain at skew s, materialising, against the
    wave engine's probe time from the timed steps (counts and checksums must equal the verified
    ones): thread per tuple (one lane per probe row, MHJ.cuh) and the same walk over per-query
    open-addressing tables (AMHJ_open_addr.cuh; their build reported apart). The pairwise
    partitioned join (PW, hashJoinEPFL.cuh:174) materialises intermediates of up to 2^32 rows,
    so it runs the chain at a quarter of the size, beside the wave engine on that chain."""
    R = resident[s]
    H = [ctx.hash_table(r) for r in R[1:]]
    out = {"skew": s, "wave_probe_ms": wave_probe_ms}
    for eng, key in (("tpt", "thread_per_tuple"), ("open-addr", "open_addressing")):
        r = ctx.execute(R[0], H, engine=eng, materialize=True, checksum=True)
        ctx.free(r)
        if (r.count, r.checksum) != checked[:2]:
            raise SystemExit(f"bench: {eng} result differs at s={s}")
        out[key + "_probe_ms"] = round(r.probe_ms, 3)
        if eng == "open-addr":
            out["open_addressing_build_ms"] = round(r.build_ms, 3)
        log(f"[bench] ablation s={s}: {eng} probe {r.probe_ms:.1f} ms vs wave engine "
            f"{wave_probe_ms:.1f} ms")
    for h in H:
        h.close()
    out["speedup"] = round(out["thread_per_tuple_probe_ms"] / wave_probe_ms, 2)
    out["speedup_vs_open_addressing"] = round(out["open_addressing_probe_ms"] / wave_probe_ms, 2)
    # pairwise, on the chain at a quarter of the size (R1 16 Mi rows, R2..R4 4 Mi, 1 Mi keys)
    q = make_inputs(G, 0.25, s)
    Rq = [ctx.relation(c, ids) for c, ids in q]
    Hq = [ctx.hash_table(r) for r in Rq[1:]]
    w = ctx.execute(Rq[0], Hq, materialize=True, checksum=True, ws=1, dro=1, fib=1)
    ctx.free(w)
    pw = ctx.execute(Rq[0], Hq, engine="pw", materialize=True, checksum=True)
    ctx.free(pw)
    for x in (*Hq, *Rq):
        x.close()
    if (pw.count, pw.checksum) != (w.count, w.checksum):
        raise SystemExit(f"bench: pairwise result differs at s={s}")
    log(f"[bench] ablation s={s}, quarter chain: pairwise {pw.probe_ms:.1f} ms vs wave engine "
        f"{w.probe_ms:.1f} ms ({pw.count} tuples)")
    out["pairwise"] = {"chain": "R1 16 Mi rows, R2..R4 4 Mi rows, 1 Mi keys", "count": pw.count,
                       "pairwise_ms": round(pw.probe_ms, 3), "wave_probe_ms": round(w.probe_ms, 3),
                       "speedup": round(pw.probe_ms / w.probe_ms, 2)}
    out["verified"] = True
    return out


def ooc_leg(G, rows, budget_gb):
    """Out-of-core 3-way chain R1(a,b) x R2(b,c) x R3(c,d) with `rows` tuples per relation and
    join keys uniform over [0, rows): the probe relation stays in pinned host memory and is
    streamed in slices while R2 and R3 are hash tables in HBM; count + checksum under an
    artificial HBM budget. Verified against the same join with the probe resident in HBM."""
    from accmultiwayjoins_amd import Context
    from accmultiwayjoins_amd.amj import pinned_empty
    t0 = time.perf_counter()
    rels = G.chain_relations([rows, rows, rows], rows, seed=42)
    pcols = []
    for c in rels[0][0]:
        p = pinned_empty(len(c))
        p[:] = c
        pcols.append(p)
    log(f"[bench] ooc: {rows} rows per relation generated in {time.perf_counter() - t0:.1f} s")
    ctx = Context(0)
    R0 = ctx.relation(pcols, rels[0][1], host=True)
    R = [ctx.relation(c, ids) for c, ids in rels[1:]]
    del rels
    H = [ctx.hash_table(r) for r in R]
    build_ms = sum(h.build_ms for h in H)
    # the budget covers the streamed slices on top of the resident tables, as the reference's
    # OOC budget counts its input and output buffers only (MHJ-GPU/ooc.cuh:34-61)
    budget = ctx.allocated_bytes() + int(budget_gb * 2**30)
    r = ctx.execute(R0, H, materialize=False, ooc=1, hbm_budget_bytes=budget)
    log(f"[bench] ooc: {r.count} results, probe {r.probe_ms:.1f} ms (path {r.path})")
    R0d = ctx.relation(pcols, [0, 1])
    ref = ctx.execute(R0d, H, materialize=False)
    if (ref.count, ref.checksum) != (r.count, r.checksum):
        raise SystemExit("bench: out-of-core result differs from the in-HBM one")
    out = {"value": r.count / (r.probe_ms / 1e3), "unit": "tuples/s", "count": r.count,
           "checksum": "%#018x" % r.checksum, "probe_ms": round(r.probe_ms, 3),
           "build_ms": round(build_ms, 3), "in_hbm_probe_ms": round(ref.probe_ms, 3),
           "probe_gb_streamed": round(rows * 8 / 1e9, 2), "budget_gb": budget_gb,
           "path": r.path,
           "workload": f"3-way chain, {rows} int32x2 tuples per relation, keys uniform over "
                       f"[0, {rows}), probe streamed from pinned host memory, count + checksum, "
                       f"{budget_gb} GiB HBM budget for the streamed slices beside the resident "
                       f"tables; value = results / probe time (PCIe transfers included), builds "
                       f"reported apart",
           "verified_against": "the same join with the probe in HBM (count + checksum)"}
    for x in (R0d, R0, *R):
        x.close()
    ctx.close()
    return out


def four_leg(G, scale):
    """4-clique query (FOUR) on an R-MAT graph through AMHJ (amj_execute_pattern) and ALFTJ DFS
    over one trie aliased six times; count + checksum (at scale 20 the ~10^10 results do not
    fit one GPU's memory as tuples). The two engines must agree."""
    from accmultiwayjoins_amd import Context
    ctx = Context(0)
    src, dst = G.rmat_oriented(scale, 16, seed=42)
    rel = ctx.relation([src, dst], [0, 1])
    ht = ctx.hash_table(rel, 0, sort_vals=True)
    t = ctx.trie_from_relation(rel)
    four = [(0, 1), (1, 2), (2, 3), (0, 3), (0, 2), (1, 3)]
    res = {}
    for name, fn in (("amhj", lambda: ctx.execute_pattern(rel, ht, "FOUR", materialize=False)),
                     ("alftj_dfs", lambda: ctx.lftj([t] * 6, [0, 1, 2, 3], "DFS",
                                                    trie_attrs=four, materialize=False)),
                     ("alftj_dfs_ws", lambda: ctx.lftj([t] * 6, [0, 1, 2, 3], "DFS",
                                                       trie_attrs=four, materialize=False,
                                                       ws=1))):
        t0 = time.perf_counter()
        r = fn()
        N = __import__("accmultiwayjoins_amd._native", fromlist=["lib"])
        N.lib().amj_sync(ctx.h)
        el = time.perf_counter() - t0
        res[name] = (r.count, r.checksum, el, r.probe_ms)
        log(f"[bench] four: {name} {r.count} 4-cliques in {el * 1e3:.0f} ms")
    (ca, sa, ea, pa), (cl, sl, el_, pl) = res["amhj"], res["alftj_dfs"]
    cw, sw, _ew, pws = res["alftj_dfs_ws"]
    if not (ca, sa) == (cl, sl) == (cw, sw):
        raise SystemExit("bench: FOUR through AMHJ and ALFTJ disagree")
    t.close()
    ht.close()
    rel.close()
    ctx.close()
    return {"count": ca, "checksum": "%#018x" % sa, "edges": len(src),
            "amhj": {"value": ca / (pa / 1e3), "unit": "tuples/s", "probe_ms": round(pa, 3)},
            "alftj_dfs": {"value": cl / (pl / 1e3), "unit": "tuples/s", "probe_ms": round(pl, 3)},
            "alftj_dfs_ws": {"value": cw / (pws / 1e3), "unit": "tuples/s",
                             "probe_ms": round(pws, 3)},
            "workload": f"FOUR (4-clique) on R-MAT scale {scale} edge factor 16, {len(src)} "
                        f"oriented edges, count + checksum (hash table / trie built untimed)",
            "verified": "AMHJ, ALFTJ DFS and ALFTJ DFS with ws agree (count + checksum)"}


def _serve(mode, sources, stream, gaps, budget, fit_budget=False, tune_rows=0, trie_edges=0):
    """Serve `stream` through one amj_server: mode "served" (continuous batching), "plain" (one
    query per launch, count + write) or "plain-dro" (one query per launch, direct output).
    sources: relations (host arrays or DeviceColumns) registered once; stream: (kind, rel
    indices, materialize) per request, submitted after the gaps (Poisson arrivals). With
    fit_budget the result budget is what is left of `budget` and of the device after the
    resident relations and tables (set before the first request)."""
    import threading
    from accmultiwayjoins_amd.serving import DeviceColumns, Server
    plain = {"served": 0, "plain": 1, "plain-dro": 2}[mode]
    n = len(stream)
    # with fit_budget the server starts with a small arena and is given its result budget
    # once the resident relations and tables are in place
    with Server(0, hbm_budget_bytes=(1 << 30) if fit_budget else budget, max_batch=256,
                plain=plain) as srv:
        t_reg = time.perf_counter()
        ids = [srv.add_device_relation(c[0], c[1]) if isinstance(c[0], DeviceColumns)
               else srv.add_relation(c[0], c[1]) for c in sources]
        reg_s = time.perf_counter() - t_reg
        mem = srv.memory()
        if tune_rows and mode == "served":
            srv.set_autotune(tune_rows)
        if trie_edges and mode == "served":
            srv.set_triejoin(trie_edges)
        result_budget = budget
        if fit_budget:
            result_budget = max(1 << 30, min(budget - mem["resident_bytes"],
                                             mem["device_free_bytes"] - (8 << 30)))
            srv.set_budget(result_budget)
        tickets, lat, qw, outs, errors = [], [], [], [], []
        done = threading.Event()

        def waiter():
            try:
                for i in range(n):
                    while len(tickets) <= i:
                        time.sleep(0.0001)
                    tm = {}
                    cnt, cs, ms, _ = srv.wait(tickets[i], timing=tm)
                    lat.append(ms)
                    qw.append(tm["queue_ms"])
                    outs.append((cnt, cs))
            except Exception as e:  # noqa: BLE001 - reported below
                errors.append(repr(e))
            done.set()

        th = threading.Thread(target=waiter, daemon=True)
        th.start()
        t0 = time.perf_counter()
        due = t0
        for (kind, rel_idx, mat), gap in zip(stream, gaps):
            due += gap
            d = due - time.perf_counter()
            if d > 0:
                time.sleep(d)
            if kind == "chain":
                tickets.append(srv.submit([ids[j] for j in rel_idx], materialize=mat))
            else:
                tickets.append(srv.submit_pattern(ids[rel_idx[0]], kind, materialize=mat))
        arrivals_s = time.perf_counter() - t0
        deadline = time.perf_counter() + 600
        while not done.wait(timeout=20):  # progress lines while the stream drains
            log(f"[bench] serving ({mode}): {len(outs)}/{n} answered")
            if time.perf_counter() > deadline:
                raise SystemExit("bench: the serving leg did not finish within 600 s")
        el = time.perf_counter() - t0
        th.join()
        if errors:
            raise SystemExit(f"bench: a served query failed ({mode}): " + errors[0])
        st = srv.stats()
        tun = srv.tuning()
        log(f"[bench] serving ({mode}): {n} queries in {el:.2f} s, {st['batches']} launches "
            f"or batches, {tun['tuned_alone']} run alone with tuned options, "
            f"{tun['triejoin_alone']} through the triejoin")
    srt = sorted(lat)
    qs = sorted(qw)
    pick = lambda a, q: a[min(len(a) - 1, int(len(a) * q))]  # noqa: E731
    return {"queries_per_s": n / el, "p50_ms": pick(srt, 0.5), "p99_ms": pick(srt, 0.99),
            "queue_p50_ms": pick(qs, 0.5), "queue_p99_ms": pick(qs, 0.99),
            "batches": st["batches"], "seconds": el, "arrivals_s": round(arrivals_s, 3),
            "register_s": round(reg_s, 3),
            "resident_gb": round(mem["resident_bytes"] / 1e9, 2),
            "result_budget_gb": round(result_budget / 1e9, 2),
            "tuned_alone": tun["tuned_alone"], "tuned_classes": tun["classes"],
            "triejoin_alone": tun["triejoin_alone"],
            "busy_s": round(st["busy_ms"] / 1e3, 3)}, outs


def serving_config5_leg(n_queries, rate, seed=42, budget=200 << 30, scales=(14, 20),
                        max_rows=50_000_000, tune_rows=4 << 20, trie_edges=1 << 20):
    """The configured mixed stream: query i (seed 42 + i) is a chain of 3, 4 or 5 relations
    (60%, in equal thirds), a triangle query (20%) or a 4-clique query (20%). Chain relations
    have log-uniform sizes in [10^4, max_rows]; each join is foreign-key style: the build side's
    key column is Zipf(s) over as many keys as the build relation has rows (s ~ U[0, 1.2] per
    query), the probe side's uniform over the same keys, so a chain returns about as many tuples
    as its probe relation; outer columns uniform over [0, 2^30). Graph queries run on their own
    R-MAT graph, scale ~ U{14..20}, edge factor 16. All data is generated on the device. Chains
    and triangles are materialised, 4-cliques counted and checksummed (at scale 20 their 1.7e10
    tuples would take 273 GB). Poisson arrivals at `rate`; the same stream is served with
    continuous batching, one query per launch (count + write) and one query per launch with
    direct output; every result's count and checksum must agree, and a few small queries are
    checked against the CPU oracle."""
    import random
    from accmultiwayjoins_amd import Context
    from accmultiwayjoins_amd.serving import DeviceColumns, device_rmat
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    gctx = Context(0)  # owns the generated device columns
    sources, stream = [], []
    mix = {"chain": 0, "TRI": 0, "FOUR": 0}
    t0 = time.perf_counter()
    for i in range(n_queries):
        qs = seed + i
        rq = random.Random(qs)
        u = rq.random()
        if u < 0.6:
            k = 3 + rq.randrange(3)
            sizes = [int(10 ** rq.uniform(4, math.log10(max_rows))) for _ in range(k)]
            z = rq.uniform(0.0, 1.2)
            idx = []
            for j, nrow in enumerate(sizes):
                c = DeviceColumns(gctx, nrow)
                if j == 0:
                    c.uniform(0, 0, 1 << 30, qs, 10 * j)
                else:
                    c.zipf(0, sizes[j], z, qs, 10 * j)
                if j == k - 1:
                    c.uniform(1, 0, 1 << 30, qs, 10 * j + 1)
                else:
                    c.uniform(1, 0, sizes[j + 1], qs, 10 * j + 1)
                sources.append((c, (j, j + 1)))  # R_j(attr j, attr j + 1)
                idx.append(len(sources) - 1)
            stream.append(("chain", idx, True))
            mix["chain"] += 1
        else:
            kind = "TRI" if u < 0.8 else "FOUR"
            sc = rq.randint(scales[0], scales[1])
            sources.append((device_rmat(gctx, sc, 16, seed=qs), (0, 1)))
            stream.append((kind, [len(sources) - 1], kind == "TRI"))
            mix[kind] += 1
    rg = random.Random(seed)
    gaps = [rg.expovariate(rate) for _ in range(n_queries)]
    gen_s = time.perf_counter() - t0
    rows = sum(c.rows for c, _a in sources)
    log(f"[bench] serving: {n_queries} queries, {len(sources)} relations, {rows / 1e9:.2f} G "
        f"rows generated on the device in {gen_s:.1f} s")
    # a few small queries against the CPU oracle (count + checksum)
    chains = sorted((sum(sources[j][0].rows for j in idx), q) for q, (kind, idx, _m)
                    in enumerate(stream) if kind == "chain")
    oracle_checked = {}
    for _tot, q in chains[:3]:
        rels = [([sources[j][0].to_host(0), sources[j][0].to_host(1)], list(sources[j][1]))
                for j in stream[q][1]]
        cnt, cs, _, _, _ = O.mhj(rels)
        oracle_checked[q] = (cnt, cs)
    results, legs = {}, {}
    for mode in ("served", "plain", "plain-dro"):
        legs[mode], results[mode] = _serve(mode, sources, stream, gaps, budget, fit_budget=True,
                                           tune_rows=tune_rows, trie_edges=trie_edges)
    if not (results["served"] == results["plain"] == results["plain-dro"]):
        raise SystemExit("bench: served and plain results differ on the configured stream")
    for q, want in oracle_checked.items():
        if tuple(results["served"][q]) != want:
            raise SystemExit(f"bench: served query {q} differs from the CPU oracle")
    for c, _a in sources:
        c.close()
    gctx.close()
    return {"served": legs["served"], "plain": legs["plain"], "plain_dro": legs["plain-dro"],
            "speedup": legs["served"]["queries_per_s"] / legs["plain"]["queries_per_s"],
            "speedup_vs_plain_dro": (legs["served"]["queries_per_s"] /
                                     legs["plain-dro"]["queries_per_s"]),
            "workload": f"{n_queries} queries (seed 42+i): {mix['chain']} chains of 3/4/5 "
                        f"relations with log-uniform sizes in [1e4, {max_rows:.0e}] rows and "
                        f"Zipf s ~ U[0, 1.2] (materialised), {mix['TRI']} TRI (materialised) and "
                        f"{mix['FOUR']} FOUR (count + checksum) on their own R-MAT graphs of "
                        f"scale {scales[0]}..{scales[1]}; {len(sources)} relations, "
                        f"{rows / 1e9:.2f} G rows resident, generated on the device; Poisson "
                        f"arrivals at {rate} queries/s; {budget / 2**30:.0f} GiB HBM budget, of "
                        f"which what the resident relations and tables leave holds results",
            "engines": "served: continuous batching (each cycle runs what arrived during the "
                       "previous one as one batch; one direct-output launch per depth and "
                       f"kind), chains probing >= {tune_rows} rows alone with the engine options "
                       "the autotuner picks for their class, TRI / FOUR over graphs of >= "
                       f"{trie_edges} edges alone through the leapfrog triejoin (trie built once "
                       "per graph and kept resident); plain: one query per launch of the hash-"
                       "join engines, count + scan + write; plain_dro: one query per launch with "
                       "direct output",
            "distinct_queries": n_queries,
            "generation_s": round(gen_s, 2),
            "verified": "every result (count + checksum) equal in the three runs; "
                        f"{len(oracle_checked)} smallest chains equal to the CPU oracle"}


def graph_leg(G, reps=20, seed=7):
    """Host overhead of a batch: 64 small direct-output queries (chains of 3..5 relations of
    10^3..10^5 rows, triangles on R-MAT 10..12) run eagerly with amj_batch_run (a size-sample
    launch, a host read of the sizes, allocation, the direct-output launches, a host read of
    the totals) and as a captured HIP graph replayed with amj_batch_replay (one graph launch,
    one host read). Wall time per batch both ways; every replay's results equal the eager
    ones."""
    import random
    from accmultiwayjoins_amd import Context
    from accmultiwayjoins_amd.serving import Batch
    rng = random.Random(seed)
    ctx = Context(0)
    keep, qs = [], []
    for i in range(56):
        k = 3 + i % 3
        sizes = [int(10 ** rng.uniform(3, 5)) for _ in range(k)]
        rels = G.chain_relations(sizes, max(sizes[1:]), seed=seed + i, zipf_s=rng.uniform(0, 1),
                                 payload_range=1 << 30)
        R = [ctx.relation(c, ids) for c, ids in rels]
        H = [ctx.hash_table(r) for r in R[1:]]
        keep += R + H
        qs.append(("chain", R[0], H))
    for i in range(8):
        src, dst = G.rmat_oriented(10 + i % 3, 16, seed=seed + 100 + i)
        E = ctx.relation([src, dst], [0, 1])
        HE = ctx.hash_table(E, 0, sort_vals=True)
        keep += [E, HE]
        qs.append(("TRI", E, HE))

    def fill(b):
        return [b.add(p, h, materialize=True, dro=1) if kind == "chain"
                else b.add_pattern(p, h, kind, materialize=True, dro=1) for kind, p, h in qs]

    eager_ms, want = [], None
    for r in range(reps + 1):
        b = Batch(ctx)
        ids = fill(b)
        t0 = time.perf_counter()
        b.run()
        el = time.perf_counter() - t0
        got = [(x.count, x.checksum) for x in (b.result(q) for q in ids)]
        b.close()
        if r:
            eager_ms.append(el * 1e3)
        want = got if want is None else want
        if got != want:
            raise SystemExit("bench: eager batches differ between runs")
    b = Batch(ctx)
    ids = fill(b)
    t0 = time.perf_counter()
    b.capture(1 << 30)
    capture_ms = (time.perf_counter() - t0) * 1e3
    replay_ms = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        b.replay()
        el = time.perf_counter() - t0
        got = [(x.count, x.checksum) for x in (b.result(q) for q in ids)]
        if got != want:
            raise SystemExit("bench: a graph replay differs from the eager batch")
        if r:
            replay_ms.append(el * 1e3)
    b.close()
    for x in keep:
        x.close()
    ctx.close()
    eager_ms.sort()
    replay_ms.sort()
    return {"queries_per_batch": len(qs), "eager_ms_per_batch": round(eager_ms[len(eager_ms) // 2], 3),
            "replay_ms_per_batch": round(replay_ms[len(replay_ms) // 2], 3),
            "speedup": round(eager_ms[len(eager_ms) // 2] / replay_ms[len(replay_ms) // 2], 2),
            "capture_ms": round(capture_ms, 3), "reps": reps,
            "workload": "64 direct-output queries (56 chains of 3..5 relations of 1e3..1e5 rows, "
                        "8 TRI on R-MAT 10..12), materialised; median wall time per batch "
                        "(launches, host reads and allocation included)",
            "verified": "every replay's counts and checksums equal the eager batch's"}


def serving_leg(G, n_queries, max_rows, rate, seed=42, budget=200 << 30):
    """A smaller stream over a few resident relations (24 chains of 3/4/5 relations of 10^4 ..
    max_rows rows, R-MAT 12..16 graphs, 60% chain / 20% TRI / 20% FOUR, all materialised):
    launch overheads dominate, which is where batching gains most."""
    import random
    rng = random.Random(seed)
    chains = []
    for i in range(24):
        k = 3 + i % 3
        sizes = [int(10 ** rng.uniform(4, math.log10(max_rows))) for _ in range(k)]
        sizes[1:] = [max(1000, x // 4) for x in sizes[1:]]
        keys = max(sizes[1:])  # every join fans out at most 1 on average: |result| <~ |probe|
        chains.append(G.chain_relations(sizes, keys, seed=seed + i, zipf_s=rng.uniform(0, 1.2),
                                        payload_range=1 << 30))
    graphs = [G.rmat_oriented(sc, 16, seed=seed + sc) for sc in (12, 13, 14, 15, 16)]
    sources, cidx, gidx = [], [], []
    for ch in chains:
        cidx.append([])
        for c, a in ch:
            sources.append((c, a))
            cidx[-1].append(len(sources) - 1)
    for sd in graphs:
        sources.append(([sd[0], sd[1]], (0, 1)))
        gidx.append(len(sources) - 1)
    stream = []
    for q in range(n_queries):
        u = rng.random()
        if u < 0.6:
            stream.append(("chain", cidx[rng.randrange(len(chains))], True))
        else:
            stream.append(("TRI" if u < 0.8 else "FOUR", [gidx[rng.randrange(len(graphs))]],
                           True))
    gaps = [rng.expovariate(rate) for _ in range(n_queries)]
    served, a = _serve("served", sources, stream, gaps, budget)
    plain, b = _serve("plain", sources, stream, gaps, budget)
    if a != b:
        raise SystemExit("bench: served and plain results differ")
    return {"served": served, "plain": plain,
            "speedup": served["queries_per_s"] / plain["queries_per_s"],
            "workload": f"{n_queries} mixed queries over 24 resident chains (3/4/5 relations of "
                        f"1e4..{max_rows} rows) and 5 resident R-MAT graphs (scale 12..16), "
                        f"60% chain / 20% TRI / 20% FOUR, Poisson arrivals at {rate} queries/s, "
                        f"results materialised, {budget / 2**30:.0f} GiB HBM budget for held "
                        f"results",
            "distinct_queries": len(set((k, tuple(i)) for k, i, _m in stream)),
            "verified": "every result (count + checksum) equal in the served and plain runs"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0,
                    help="fraction of the full workload (1.0 = the configuration above)")
    ap.add_argument("--skews", default="0,0.5,1.0")
    ap.add_argument("--ws", type=int, default=1)
    ap.add_argument("--dro", type=int, default=1)
    ap.add_argument("--fib", type=int, default=1)
    ap.add_argument("--probe-order", type=int, default=0,
                    help="0: group probe rows by join key (default), 1: keep their order")
    ap.add_argument("--cpu-sample-rows", type=int, default=1 << 21)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-triangle", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-ablation", action="store_true",
                    help="skip the thread-per-tuple baseline run on the first skew")
    ap.add_argument("--serve-queries", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=100.0,
                    help="Poisson arrival rate of the configured stream (queries/s; above what "
                         "either path serves, so that both legs measure throughput, while "
                         "arrivals still overlap the batches on the GPU)")
    ap.add_argument("--serve-max-rows", type=int, default=50_000_000)
    ap.add_argument("--serve-small", action="store_true",
                    help="also serve the smaller stream over a few resident relations")
    ap.add_argument("--tri-scale", type=int, default=22)
    ap.add_argument("--ooc-rows", type=int, default=0,
                    help="also run the out-of-core chain leg with this many rows per relation "
                         "(1500000000 for the 1.5 B-tuple case; off by default: ~1 min)")
    ap.add_argument("--ooc-budget-gb", type=float, default=48.0)
    ap.add_argument("--four-scale", type=int, default=20,
                    help="also run FOUR through AMHJ and ALFTJ (with and without ws) on R-MAT of "
                         "this scale (the 4-clique case; 0: off)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device three times (torch), to scale the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiled run, then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copied 3 x 1 GiB")
        del src, dst
        print(json.dumps({"copy_gbs": copy_bandwidth_gbs()}), flush=True)
        return

    from accmultiwayjoins_amd import Context
    from accmultiwayjoins_amd import generators as G

    skews = [float(x) for x in args.skews.split(",")]
    t0 = time.perf_counter()
    inputs = {s: make_inputs(G, args.scale, s) for s in skews}
    log(f"[bench] generated inputs in {time.perf_counter() - t0:.1f} s")

    ctx = Context(0)
    dev_name, hbm, cus = ctx.device_info()
    log(f"[bench] device {dev_name}, {hbm / 2**30:.0f} GiB, {cus} CUs")
    resident = {}
    for s in skews:
        R = [ctx.relation(cols, ids) for cols, ids in inputs[s]]
        resident[s] = R
    opts = dict(ws=args.ws, dro=args.dro, fib=args.fib, materialize=True, checksum=False,
                probe_order=args.probe_order)

    def step(record):
        tot = 0
        for s in skews:
            R = resident[s]
            H = [ctx.hash_table(r) for r in R[1:]]
            r = ctx.execute(R[0], H, **opts)
            tot += r.count
            if record is not None:
                record.append((s, r.count, r.checksum, sum(h.build_ms for h in H), r.probe_ms,
                               r.probe_kernel_ms, r.path, 0))
            ctx.free(r)
            for h in H:
                h.close()
        return tot

    def verify():
        """Untimed: the materialising path with its checksum on must agree with the
        independent count-pass engine (count + checksum, no materialisation)."""
        out = {}
        for s in skews:
            R = resident[s]
            H = [ctx.hash_table(r) for r in R[1:]]
            a = ctx.execute(R[0], H, **dict(opts, checksum=True))
            ctx.free(a)
            b = ctx.execute(R[0], H, materialize=False)
            for h in H:
                h.close()
            if (a.count, a.checksum) != (b.count, b.checksum):
                raise SystemExit(f"bench: verification failed at s={s}: "
                                 f"{a.path} {a.count}/{a.checksum:#x} vs "
                                 f"{b.path} {b.count}/{b.checksum:#x}")
            out[s] = (a.count, a.checksum, a.path, b.path)
        return out

    for w in range(args.warmup):
        t = time.perf_counter()
        step(None)
        log(f"[bench] warmup {w} done in {time.perf_counter() - t:.2f} s")
    N = __import__("accmultiwayjoins_amd._native", fromlist=["lib"])
    N.lib().amj_sync(ctx.h)
    rec = []
    t_start = time.perf_counter()
    total = 0
    for k in range(args.steps):
        total += step(rec)
        log(f"[bench] step {k} done")
    N.lib().amj_sync(ctx.h)
    elapsed = time.perf_counter() - t_start
    value = total / elapsed

    # roofline of the dominant kernel (result materialisation), from HIP events on its stream
    last = rec[-len(skews):]
    k_ms = sum(x[5] for x in last)
    alg = sum(algorithmic_bytes(inputs[s], cnt) for (s, cnt, *_r) in last)
    achieved = alg / (k_ms / 1e3) / 1e9 if k_ms > 0 else 0.0
    checked = verify()
    log("[bench] verification: materialising path == count-pass engine for every skew")
    per_skew = {str(s): {"count": cnt, "checksum": "%#018x" % checked[s][1],
                         "build_ms": round(bms, 3), "probe_ms": round(pms, 3),
                         "write_kernel_ms": round(kms, 3), "path": path,
                         "verified_against": checked[s][3]}
                for (s, cnt, cs, bms, pms, kms, path, _x) in last}

    tuned = autotune_leg(ctx, resident[skews[0]], skews[0], per_skew[str(skews[0])])

    ablation = None
    if not args.no_ablation:
        ablation = ablation_leg(ctx, resident, skews[0], checked[skews[0]],
                                per_skew[str(skews[0])]["probe_ms"], G)

    tri = None
    if not args.no_triangle:
        tri = triangle_leg(ctx, G, args.tri_scale, args.steps, args.warmup)

    copy_gbs = measured_copy_gbs()
    log(f"[bench] device-to-device copy: {copy_gbs} GB/s (read + write)")

    cpu = None
    if not args.no_cpu:
        s0 = skews[-1]
        sample_rows = min(args.cpu_sample_rows, len(inputs[s0][0][0][0]))
        cnt, secs, wall = cpu_baseline(inputs[s0], sample_rows, args.cpu_threads)
        one_rows = sample_rows // 8
        cnt1, secs1, _ = cpu_baseline(inputs[s0], one_rows, 1)
        model, ncpu, smt = host_cpu()
        cpu = {"value": cnt / secs if secs > 0 else 0.0, "unit": "tuples/s",
               "cores": args.cpu_threads, "kind": "port",
               "sample": f"oracle MHJ-CPU restatement (partition + probe, count+checksum, no "
                         f"materialisation) on the s={s0} inputs with R1 cut to its first "
                         f"{sample_rows} rows; {cnt} results in {secs:.2f} s",
               "one_thread": {"value": cnt1 / secs1 if secs1 > 0 else 0.0,
                              "sample": f"R1 cut to {one_rows} rows; {cnt1} results in "
                                        f"{secs1:.2f} s"},
               "host": {"cpu_model": model, "logical_cpus": ncpu, "smt_active": smt}}
        log(f"[bench] cpu baseline: {cnt} tuples in {secs:.2f} s ({wall:.2f} s wall); "
            f"1 thread: {cnt1} in {secs1:.2f} s")
        cpu["config1"] = config1_leg(ctx, G, args.cpu_threads)

    out = {
        "metric": "join output tuples/s (4-way chain)",
        "value": value,
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "AMHJ NORMAL 4-way chain, R1 64M + R2..R4 16M int32x2 rows, "
                               "Zipf s in {0,0.5,1.0} over 4M keys, ws=1 dro=1 fib=1, "
                               "build + probe + materialise",
                   "scale": args.scale, "skews": skews, "ws": args.ws, "dro": args.dro,
                   "fib": args.fib, "per_skew": per_skew},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "measured_copy_gbs": copy_gbs,
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT)},
        "cpu_baseline": cpu,
        "ablation": ablation,
        "autotune": tuned,
        "triangle": tri,
        "serving": None,
    }
    for R in resident.values():
        for r in R:
            r.close()
    ctx.close()
    if args.ooc_rows:
        out["ooc"] = ooc_leg(G, args.ooc_rows, args.ooc_budget_gb)
    if args.four_scale:
        out["four"] = four_leg(G, args.four_scale)
    if not args.no_serving:  # its own server context, after the chain's memory is released
        serving = serving_config5_leg(args.serve_queries, args.serve_rate,
                                      max_rows=args.serve_max_rows)
        log(f"[bench] serving: {serving['served']['queries_per_s']:.1f} q/s served vs "
            f"{serving['plain']['queries_per_s']:.1f} q/s plain, "
            f"{serving['plain_dro']['queries_per_s']:.1f} q/s plain with direct output")
        out["serving"] = serving
    if not args.no_serving:
        gl = graph_leg(G)
        log(f"[bench] graph: eager {gl['eager_ms_per_batch']} ms vs replay "
            f"{gl['replay_ms_per_batch']} ms per batch of {gl['queries_per_batch']}")
        out["batch_graph"] = gl
    if args.serve_small:
        small = serving_leg(G, 512, 2_000_000, 4000.0)
        log(f"[bench] small stream: {small['speedup']:.2f}x")
        out["serving_small"] = small
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
