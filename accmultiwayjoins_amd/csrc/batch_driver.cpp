// Host driver of batched AMHJ execution (amj_batch_*): plans of many queries over resident
// relations and hash tables, executed together by the batch kernels (batch.hip) with two host
// round trips per batch whatever its size: counts -> one pooled buffer per materialised result
// -> writes and checksums.
#include <algorithm>
#include <cstring>

#include "runtime.h"

namespace amj {

hipError_t launch_batch_count(int L, const void* qs, uint32_t nq, uint64_t unit_begin,
                              uint64_t total, uint64_t* counts, hipStream_t s);
hipError_t launch_batch_write(int L, const void* qs, uint32_t nq, uint64_t unit_begin,
                              uint64_t total, const uint64_t* base, unsigned long long* csums,
                              hipStream_t s);
hipError_t launch_batch_totals(const void* qs, uint32_t nq, const uint64_t* base,
                               uint64_t* totals, hipStream_t s);
size_t batch_desc_bytes();
void batch_fill_desc(void* dst, const Plan& P, const ProbeOut& out, uint64_t unit0,
                     uint64_t units, uint64_t stride);
hipError_t launch_batch_weights(const void* qs, uint32_t nq, uint64_t unit_begin,
                                uint64_t total, uint64_t* keys, hipStream_t s);
hipError_t launch_batch_direct(int L, bool mat, const void* qs, uint32_t nq,
                               uint64_t unit_begin, uint64_t total, unsigned long long* rows,
                               unsigned long long* counts, unsigned long long* csums,
                               const uint64_t* order, hipStream_t s);
hipError_t launch_batch_place(const void* info, uint32_t nq, const uint64_t* sampled, void* dq,
                              int32_t* slab, uint64_t slab_words, uint64_t* cap_out,
                              uint64_t* off_out, uint32_t* admitted, hipStream_t s);
hipError_t launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s);
size_t batch_prep_info_bytes();
void batch_fill_prep_info(void* dst, uint64_t units, uint64_t sunits, uint64_t user_cap,
                          uint32_t n_attrs, bool mat);

}  // namespace amj

using namespace amj;

// Direct-output launches with at least this many units take them heaviest first.
constexpr uint64_t kBatchOrderMinUnits = 2048;

struct amj_batch {
  amj_ctx_t* ctx = nullptr;
  struct Query {
    Plan plan;
    std::vector<uint32_t> attr_ids;
    bool materialize = true;
    bool direct = false;    // dro: no count pass (output sized from a sample)
    uint64_t capacity = 0;  // dro: output rows given by the caller (0: from the sample)
    bool sampled = false;   // dro: the output size came from a sample launch
    int status = AMJ_OK;    // AMJ_ERR_NOMEM: not admitted (the budget was full before it)
    amj_result_t res{};
    bool taken = false;
  };
  std::vector<Query> qs;
  // A captured batch (amj_batch_capture): its chain recorded into a HIP graph over buffers the
  // batch owns, replayed by amj_batch_replay.
  struct Captured {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<void*, size_t>> bufs;  // device buffers owned by the graph
    std::vector<uint32_t> order;                 // launch order: query indices
    int32_t* slab = nullptr;                     // materialised results, placed on the device
    uint64_t* acc = nullptr;                     // rows, counts, checksums (3 x nq)
    uint64_t* caps = nullptr;                    // output rows of each query
    uint64_t* offs = nullptr;                    // its place in the slab (words)
    uint32_t* admitted = nullptr;                // it fitted the slab
    uint64_t replays = 0;
  } cap;
  bool captured = false;
  bool ran = false;
  double run_ms = 0.0;
  uint64_t budget = 0;  // bytes the batch's results may take (0: what the device has free)
  uint64_t needed = 0;  // bytes the materialised results take (known after the count pass)
};

extern "C" {

int amj_batch_create(amj_ctx_t* ctx, amj_batch_t** out) {
  if (!ctx || !out) return AMJ_ERR_ARG;
  *out = new amj_batch();
  (*out)->ctx = ctx;
  return AMJ_OK;
}

void amj_batch_destroy(amj_batch_t* b) {
  if (!b) return;
  for (auto& q : b->qs)
    if (b->ran && !q.taken) amj_result_free(b->ctx, &q.res);
  if (b->cap.exec) (void)hipGraphExecDestroy(b->cap.exec);
  if (b->cap.graph) (void)hipGraphDestroy(b->cap.graph);
  if (!b->cap.bufs.empty()) (void)hipStreamSynchronize(b->ctx->stream);
  for (auto& x : b->cap.bufs) dfree(b->ctx, x.first, x.second);
  delete b;
}

static int batch_add_plan(amj_batch_t* b, Plan& plan, std::vector<uint32_t>& ids,
                          const amj_opts_t* opts, uint32_t* id) {
  if (b->ran) {
    set_error(b->ctx, AMJ_ERR_STATE, "the batch has already run");
    return AMJ_ERR_STATE;
  }
  if (opts && opts->engine != AMJ_ENGINE_WAVE) {
    set_error(b->ctx, AMJ_ERR_ARG, "batches run the wave engine only (opts.engine = 0)");
    return AMJ_ERR_ARG;
  }
  amj_batch::Query q;
  q.plan = plan;
  q.attr_ids = ids;
  q.materialize = opts ? opts->materialize != 0 : true;
  q.direct = opts && opts->dro;
  q.capacity = opts && opts->dro ? opts->dro_capacity_rows : 0;
  b->qs.push_back(q);
  if (id) *id = (uint32_t)(b->qs.size() - 1);
  return AMJ_OK;
}

int amj_batch_add(amj_batch_t* b, const amj_relation_t* probe, amj_hash_table_t* const* hts,
                  uint32_t n_hts, const amj_opts_t* opts, uint32_t* id) {
  if (!b || !probe || !hts) return AMJ_ERR_ARG;
  if (probe->host) {
    set_error(b->ctx, AMJ_ERR_ARG, "batched queries need a device-resident probe relation");
    return AMJ_ERR_ARG;
  }
  ProbeUse pu{};
  pu.rel = probe;
  for (uint32_t c = 0; c < probe->num_cols; ++c) pu.attr_ids[c] = probe->attr_ids[c];
  std::vector<TableUse> uses;
  for (uint32_t i = 0; i < n_hts; ++i) {
    if (!hts[i]) return AMJ_ERR_ARG;
    TableUse u{};
    u.ht = hts[i];
    u.key_attr_id = hts[i]->rel->attr_ids[hts[i]->key_col];
    u.val_attr_id =
        hts[i]->rel->num_cols > 1 ? (int64_t)hts[i]->rel->attr_ids[1 - hts[i]->key_col] : -1;
    uses.push_back(u);
  }
  Plan plan;
  std::vector<uint32_t> ids;
  AMJ_TRY(make_plan(b->ctx, pu, uses, &plan, &ids, nullptr));
  return batch_add_plan(b, plan, ids, opts, id);
}

int amj_batch_add_pattern(amj_batch_t* b, const amj_relation_t* edges,
                          const amj_hash_table_t* ht, int query, const amj_opts_t* opts,
                          uint32_t* id) {
  if (!b) return AMJ_ERR_ARG;
  Plan plan;
  std::vector<uint32_t> ids;
  AMJ_TRY(make_pattern_plan(b->ctx, edges, ht, query, &plan, &ids));
  return batch_add_plan(b, plan, ids, opts, id);
}

}  // extern "C"

namespace {

// Output buffers of the materialised queries of one phase (rows[k] rows of qs[order[k]]), from
// the context's pool; all or nothing. Fails with AMJ_ERR_NOMEM, b->needed set, when their bytes
// exceed the batch budget or the device.
int alloc_outputs(amj_batch_t* b, const std::vector<uint32_t>& order,
                  const std::vector<uint64_t>& rows) {
  amj_ctx_t* ctx = b->ctx;
  uint64_t bytes = 0;
  for (size_t k = 0; k < order.size(); ++k)
    if (b->qs[order[k]].materialize && rows[k] > 0)
      bytes += rows[k] * (uint64_t)b->qs[order[k]].plan.n_attrs * 4;
  if (bytes == 0) return AMJ_OK;
  const uint64_t room = budget_left(ctx, nullptr);
  if (bytes > (b->budget ? std::min(b->budget, room) : room)) {
    b->needed = bytes;
    set_error(ctx, AMJ_ERR_NOMEM,
              "the batch's results (" + std::to_string(bytes) + " bytes) do not fit the budget");
    return AMJ_ERR_NOMEM;
  }
  for (size_t k = 0; k < order.size(); ++k) {
    auto& q = b->qs[order[k]];
    std::memset(&q.res, 0, sizeof(q.res));
    if (!q.materialize || rows[k] == 0) continue;
    const int rc = alloc_output(ctx, rows[k], q.plan.n_attrs, &q.res);
    if (rc != AMJ_OK) {
      for (size_t j = 0; j <= k; ++j) amj_result_free(ctx, &b->qs[order[j]].res);
      b->needed = bytes;
      return rc;
    }
  }
  b->needed += bytes;
  return AMJ_OK;
}

// Fill in the result of query q; its output buffer (if any) is already in q.res.
void set_result(amj_batch::Query& q, uint64_t count, uint64_t checksum) {
  q.res.count = count;
  q.res.checksum = checksum;
  q.res.num_cols = (uint32_t)q.plan.n_attrs;
  for (int a = 0; a < q.plan.n_attrs; ++a) q.res.attr_ids[a] = q.attr_ids[a];
  q.res.rows_written = q.res.cols[0] ? count : 0;
  q.res.passes = 2;
  q.res.path = AMJ_PATH_BATCH;
}

ProbeOut output_of(const amj_batch::Query& q, uint64_t capacity) {
  ProbeOut out{};
  if (q.res.cols[0]) {
    for (int a = 0; a < q.plan.n_attrs; ++a) out.cols[a] = q.res.cols[a];
    out.capacity = capacity;
  }
  return out;
}

// Queries sorted by depth (one launch per group of equal depth); calls f(k, e, L) per group.
template <typename F>
int for_depth_groups(amj_batch_t* b, const std::vector<uint32_t>& order, F f) {
  const uint32_t n = (uint32_t)order.size();
  for (uint32_t k = 0; k < n;) {
    uint32_t e = k;
    const int L = b->qs[order[k]].plan.n_levels;
    while (e < n && b->qs[order[e]].plan.n_levels == L) ++e;
    AMJ_TRY(f(k, e, L));
    k = e;
  }
  return AMJ_OK;
}

std::vector<uint32_t> depth_order(amj_batch_t* b, std::vector<uint32_t> idx) {
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) {
    return b->qs[x].plan.n_levels < b->qs[y].plan.n_levels;
  });
  return idx;
}

// Count pass over every unit (or over a strided sample of rows: stride[k] > 1), one scan;
// returns per-query totals. `desc` receives the descriptors (count-only outputs).
int batch_counts(amj_batch_t* b, const std::vector<uint32_t>& order,
                        const std::vector<uint64_t>& units, const std::vector<uint64_t>& stride,
                        std::vector<uint64_t>* unit0, std::vector<uint64_t>* cnt, void** ddesc_out,
                        void** counts_out) {
  amj_ctx_t* ctx = b->ctx;
  hipStream_t s = ctx->stream;
  const uint32_t nq = (uint32_t)order.size();
  unit0->assign(nq, 0);
  uint64_t total = 0;
  for (uint32_t k = 0; k < nq; ++k) {
    (*unit0)[k] = total;
    total += units[k];
  }
  const size_t db = batch_desc_bytes();
  std::vector<uint8_t> desc(db * nq);
  for (uint32_t k = 0; k < nq; ++k)
    batch_fill_desc(desc.data() + db * k, b->qs[order[k]].plan, ProbeOut{}, (*unit0)[k], units[k],
                    stride[k]);
  void *ddesc = nullptr, *counts = nullptr, *tmp = nullptr, *totals = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_MISC0, desc.size(), &ddesc));
  AMJ_TRY(scratch(ctx, SLOT_UNIT_COUNTS, (total + 1) * 8, &counts));
  AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(total + 1) * 8, &tmp));
  AMJ_TRY(scratch(ctx, SLOT_MISC3, (size_t)nq * 24, &totals));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(ddesc, desc.data(), desc.size(), hipMemcpyHostToDevice, s));
  AMJ_HIP_TRY(ctx, hipMemsetAsync(counts, 0, (total + 1) * 8, s));
  AMJ_TRY(for_depth_groups(b, order, [&](uint32_t k, uint32_t e, int L) -> int {
    const uint64_t gu = (*unit0)[e - 1] + units[e - 1] - (*unit0)[k];
    AMJ_HIP_TRY(ctx, launch_batch_count(L, (const uint8_t*)ddesc + db * k, e - k, (*unit0)[k],
                                        gu, (uint64_t*)counts, s));
    return AMJ_OK;
  }));
  AMJ_HIP_TRY(ctx, exclusive_scan_u64((uint64_t*)counts, (uint64_t*)counts, total + 1,
                                      (uint64_t*)tmp, s));
  AMJ_HIP_TRY(ctx, launch_batch_totals(ddesc, nq, (const uint64_t*)counts, (uint64_t*)totals, s));
  cnt->assign(nq, 0);
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(cnt->data(), totals, nq * 8, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  *ddesc_out = ddesc;
  if (counts_out) *counts_out = counts;
  return AMJ_OK;
}

// Exact execution: count pass, one scan, exact outputs, write pass (deterministic row order: every
// tuple lands where the unit engine puts it).
int run_exact(amj_batch_t* b, const std::vector<uint32_t>& idx) {
  amj_ctx_t* ctx = b->ctx;
  hipStream_t s = ctx->stream;
  const std::vector<uint32_t> order = depth_order(b, idx);
  const uint32_t nq = (uint32_t)order.size();
  std::vector<uint64_t> units(nq), stride(nq, 1), unit0, cnt;
  for (uint32_t k = 0; k < nq; ++k) units[k] = probe_units(b->qs[order[k]].plan);
  void *ddesc = nullptr, *counts = nullptr;
  AMJ_TRY(batch_counts(b, order, units, stride, &unit0, &cnt, &ddesc, &counts));
  AMJ_TRY(alloc_outputs(b, order, cnt));
  const size_t db = batch_desc_bytes();
  std::vector<uint8_t> desc(db * nq);
  for (uint32_t k = 0; k < nq; ++k) {
    auto& q = b->qs[order[k]];
    set_result(q, cnt[k], 0);
    batch_fill_desc(desc.data() + db * k, q.plan, output_of(q, cnt[k]), unit0[k], units[k], 1);
  }
  void* csums = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_MISC2, (size_t)nq * 8, &csums));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(ddesc, desc.data(), desc.size(), hipMemcpyHostToDevice, s));
  AMJ_HIP_TRY(ctx, hipMemsetAsync(csums, 0, nq * 8, s));
  AMJ_TRY(for_depth_groups(b, order, [&](uint32_t k, uint32_t e, int L) -> int {
    const uint64_t gu = unit0[e - 1] + units[e - 1] - unit0[k];
    AMJ_HIP_TRY(ctx, launch_batch_write(L, (const uint8_t*)ddesc + db * k, e - k, unit0[k], gu,
                                        (const uint64_t*)counts,
                                        (unsigned long long*)csums + k, s));
    return AMJ_OK;
  }));
  std::vector<uint64_t> cs(nq);
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(cs.data(), csums, nq * 8, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  for (uint32_t k = 0; k < nq; ++k) b->qs[order[k]].res.checksum = cs[k];
  return AMJ_OK;
}

// Direct output (dro): per-query capacities from a strided sample of units, one pass that
// reserves output rows with atomics. Materialised queries whose results overran their capacity
// are returned in *redo (their counts and checksums are exact, their rows incomplete).
constexpr uint64_t kSampleUnits = 1024;  // units sampled per query
constexpr double kSlack = 1.25;          // estimate -> output size

int run_direct(amj_batch_t* b, const std::vector<uint32_t>& idx,
               std::vector<uint32_t>* redo) {
  amj_ctx_t* ctx = b->ctx;
  hipStream_t s = ctx->stream;
  // 1. an output size for every query: exact for a query of at most kSampleUnits units, else
  //    the count of a strided sample of its units scaled up, with slack
  std::vector<uint64_t> capq(b->qs.size(), 0);
  {
    const std::vector<uint32_t> order = depth_order(b, idx);
    const uint32_t nq = (uint32_t)order.size();
    std::vector<uint64_t> units(nq), sunits(nq), stride(nq), unit0, sampled(nq, 0);
    bool need_sample = false;
    for (uint32_t k = 0; k < nq; ++k) {
      units[k] = probe_units(b->qs[order[k]].plan);
      sunits[k] = std::min(units[k], kSampleUnits);
      stride[k] = sunits[k] ? units[k] / sunits[k] : 1;
      need_sample |= b->qs[order[k]].materialize && !b->qs[order[k]].capacity;
    }
    void* ddesc = nullptr;
    if (need_sample)
      AMJ_TRY(batch_counts(b, order, sunits, stride, &unit0, &sampled, &ddesc, nullptr));
    for (uint32_t k = 0; k < nq; ++k) {
      auto& q = b->qs[order[k]];
      if (!q.materialize) continue;
      if (q.capacity) {
        capq[order[k]] = q.capacity;
      } else if (sunits[k] == units[k]) {
        capq[order[k]] = sampled[k];
      } else {
        const double est = (double)sampled[k] * (double)units[k] / (double)sunits[k];
        capq[order[k]] = (uint64_t)(est * kSlack) + 65536;
      }
      q.sampled = need_sample;
    }
  }
  // 2. admission in submission order: the longest prefix whose outputs fit the room; the
  //    rest is not run (amj_batch_result reports AMJ_ERR_NOMEM for it)
  std::vector<uint32_t> admitted;
  {
    const uint64_t room = budget_left(ctx, nullptr);
    const uint64_t lim = b->budget ? std::min(b->budget, room) : room;
    uint64_t bytes = 0;
    for (uint32_t i : idx) {
      const auto& q = b->qs[i];
      const uint64_t need = q.materialize ? capq[i] * (uint64_t)q.plan.n_attrs * 4 : 0;
      if (!admitted.empty() && bytes + need > lim) break;
      if (admitted.empty() && need > lim) {
        b->needed = need;
        set_error(ctx, AMJ_ERR_NOMEM, "the batch's results (" + std::to_string(need) +
                                          " bytes) do not fit the budget");
        return AMJ_ERR_NOMEM;
      }
      bytes += need;
      admitted.push_back(i);
    }
    for (size_t j = admitted.size(); j < idx.size(); ++j) b->qs[idx[j]].status = AMJ_ERR_NOMEM;
  }
  // 3. one direct-output pass over the admitted queries; inside a depth group the largest
  //    outputs come first, so their long units start early instead of ending the launch alone
  //    (one launch per depth and kind: materialising or count-only)
  std::vector<uint32_t> order = admitted;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    const int lx = b->qs[x].plan.n_levels, ly = b->qs[y].plan.n_levels;
    const bool mx = b->qs[x].materialize, my = b->qs[y].materialize;
    return lx != ly ? lx < ly : (mx != my ? mx : capq[x] > capq[y]);
  });
  const uint32_t nq = (uint32_t)order.size();
  std::vector<uint64_t> units(nq), cap(nq), unit0(nq);
  uint64_t total = 0;
  for (uint32_t k = 0; k < nq; ++k) {
    units[k] = probe_units(b->qs[order[k]].plan);
    cap[k] = capq[order[k]];
    unit0[k] = total;
    total += units[k];
  }
  AMJ_TRY(alloc_outputs(b, order, cap));
  const size_t db = batch_desc_bytes();
  std::vector<uint8_t> desc(db * nq);
  for (uint32_t k = 0; k < nq; ++k) {
    const auto& q = b->qs[order[k]];
    batch_fill_desc(desc.data() + db * k, q.plan, output_of(q, cap[k]), unit0[k], units[k], 1);
  }
  void *ddesc = nullptr, *acc = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_MISC0, desc.size(), &ddesc));
  AMJ_TRY(scratch(ctx, SLOT_MISC2, (size_t)nq * 24, &acc));
  unsigned long long* rows = (unsigned long long*)acc;
  unsigned long long* counts = rows + nq;
  unsigned long long* csums = rows + 2 * nq;
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(ddesc, desc.data(), desc.size(), hipMemcpyHostToDevice, s));
  AMJ_HIP_TRY(ctx, hipMemsetAsync(acc, 0, (size_t)nq * 24, s));
  for (uint32_t k = 0; k < nq;) {
    uint32_t e = k;
    const int L = b->qs[order[k]].plan.n_levels;
    const bool mat = b->qs[order[k]].materialize;
    while (e < nq && b->qs[order[e]].plan.n_levels == L && b->qs[order[e]].materialize == mat) ++e;
    const uint64_t gu = unit0[e - 1] + units[e - 1] - unit0[k];
    // heaviest units first when they outnumber the device's wave slots (the sort buffers are
    // sized for the whole batch, so no launch in flight loses its order to a reallocation)
    const uint64_t* uorder = nullptr;
    if (gu >= kBatchOrderMinUnits) {
      void *keys = nullptr, *alt = nullptr, *tmp = nullptr;
      AMJ_TRY(scratch(ctx, SLOT_SORT_KEYS, total * 8, &keys));
      AMJ_TRY(scratch(ctx, SLOT_SORT_ALT, total * 8, &alt));
      AMJ_TRY(scratch(ctx, SLOT_SORT_TEMP, radix_sort_temp_bytes(total), &tmp));
      AMJ_HIP_TRY(ctx, launch_batch_weights((const uint8_t*)ddesc + db * k, e - k, unit0[k], gu,
                                            (uint64_t*)keys, s));
      AMJ_HIP_TRY(ctx, radix_sort_u64((uint64_t*)keys, (uint64_t*)alt, gu, 0xf0u, tmp, s));
      uorder = (const uint64_t*)keys;
    }
    AMJ_HIP_TRY(ctx, launch_batch_direct(L, mat, (const uint8_t*)ddesc + db * k, e - k, unit0[k],
                                         gu, rows + k, counts + k, csums + k, uorder, s));
    k = e;
  }
  std::vector<uint64_t> h(3 * (size_t)nq);
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(h.data(), acc, (size_t)nq * 24, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  for (uint32_t k = 0; k < nq; ++k) {
    auto& q = b->qs[order[k]];
    const uint64_t cnt = h[nq + k], cs = h[2 * nq + k];
    if (q.materialize && cnt > cap[k]) {  // rows past the capacity were dropped: run again
      amj_result_free(ctx, &q.res);
      redo->push_back(order[k]);
      continue;
    }
    set_result(q, cnt, cs);
    q.res.passes = q.sampled ? 2 : 1;
  }
  return AMJ_OK;
}

}  // namespace

extern "C" {

int amj_batch_run(amj_batch_t* b) {
  if (!b) return AMJ_ERR_ARG;
  amj_ctx_t* ctx = b->ctx;
  if (b->captured) {
    set_error(ctx, AMJ_ERR_STATE, "a captured batch runs with amj_batch_replay");
    return AMJ_ERR_STATE;
  }
  if (b->ran) {
    set_error(ctx, AMJ_ERR_STATE, "the batch has already run");
    return AMJ_ERR_STATE;
  }
  hipStream_t s = ctx->stream;
  const uint32_t nq = (uint32_t)b->qs.size();
  b->ran = true;
  b->needed = 0;
  if (nq == 0) return AMJ_OK;
  std::vector<uint32_t> exact, direct;
  for (uint32_t i = 0; i < nq; ++i) (b->qs[i].direct ? direct : exact).push_back(i);
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  int rc = AMJ_OK;
  if (!direct.empty()) rc = run_direct(b, direct, &exact);
  if (rc == AMJ_OK && !exact.empty()) rc = run_exact(b, exact);
  if (rc != AMJ_OK) {  // nothing is handed out: give back what a phase produced
    for (auto& q : b->qs)
      if (q.res.cols[0]) amj_result_free(ctx, &q.res);
    for (auto& q : b->qs) std::memset(&q.res, 0, sizeof(q.res));
    b->ran = rc != AMJ_ERR_NOMEM;  // a batch over budget can run again with a larger one
    return rc;
  }
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  b->run_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  for (auto& q : b->qs) {
    q.res.probe_ms = b->run_ms;
    q.res.probe_kernel_ms = b->run_ms;
    q.res.probe_kernel_launches = 1;
  }
  return AMJ_OK;
}

int amj_batch_result(amj_batch_t* b, uint32_t id, amj_result_t* res) {
  if (!b || !res || id >= b->qs.size()) return AMJ_ERR_ARG;
  if (b->ran && b->qs[id].status != AMJ_OK) {
    set_error(b->ctx, b->qs[id].status,
              "query not run: the batch budget was full before it (run it in a later batch)");
    return b->qs[id].status;
  }
  if (!b->ran || b->qs[id].taken) {
    set_error(b->ctx, AMJ_ERR_STATE, "the batch has not run, or this result was taken");
    return AMJ_ERR_STATE;
  }
  *res = b->qs[id].res;
  b->qs[id].taken = true;
  return AMJ_OK;
}

int amj_batch_set_budget(amj_batch_t* b, uint64_t bytes) {
  if (!b) return AMJ_ERR_ARG;
  b->budget = bytes;
  return AMJ_OK;
}

int amj_batch_info(const amj_batch_t* b, uint32_t* queries, double* run_ms,
                   uint64_t* result_bytes) {
  if (!b) return AMJ_ERR_ARG;
  if (queries) *queries = (uint32_t)b->qs.size();
  if (run_ms) *run_ms = b->run_ms;
  if (result_bytes) *result_bytes = b->needed;
  return AMJ_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Captured batches.
// ------------------------------------------------------------------------------------------
namespace {

int cap_alloc(amj_batch_t* b, size_t bytes, void** p) {
  AMJ_TRY(dmalloc(b->ctx, p, bytes ? bytes : 8));
  b->cap.bufs.emplace_back(*p, bytes ? bytes : 8);
  return AMJ_OK;
}

// The launches of the chain, enqueued on the context's stream (recorded by the caller).
int enqueue_chain(amj_batch_t* b, const std::vector<uint64_t>& units,
                  const std::vector<uint64_t>& sunits, const std::vector<uint64_t>& unit0,
                  const std::vector<uint64_t>& sunit0, uint64_t stotal, void* sdesc, void* ddesc,
                  void* info, uint64_t* counts, uint64_t* scan_tmp, uint64_t* sampled,
                  uint64_t slab_words) {
  amj_ctx_t* ctx = b->ctx;
  hipStream_t s = ctx->stream;
  const auto& order = b->cap.order;
  const uint32_t nq = (uint32_t)order.size();
  const size_t db = batch_desc_bytes();
  // 1. size sample: strided units of every query that needs one, one scan, per-query totals
  if (stotal) {
    AMJ_HIP_TRY(ctx, launch_fill_u64(counts, stotal + 1, 0, s));
    for (uint32_t k = 0; k < nq;) {
      uint32_t e = k;
      const int L = b->qs[order[k]].plan.n_levels;
      while (e < nq && b->qs[order[e]].plan.n_levels == L) ++e;
      const uint64_t gu = sunit0[e - 1] + sunits[e - 1] - sunit0[k];
      if (gu)
        AMJ_HIP_TRY(ctx, launch_batch_count(L, (const uint8_t*)sdesc + db * k, e - k, sunit0[k],
                                            gu, counts, s));
      k = e;
    }
    AMJ_HIP_TRY(ctx, exclusive_scan_u64(counts, counts, stotal + 1, scan_tmp, s));
    AMJ_HIP_TRY(ctx, launch_batch_totals(sdesc, nq, counts, sampled, s));
  } else {
    AMJ_HIP_TRY(ctx, launch_fill_u64(sampled, nq, 0, s));
  }
  // 2. output sizes and places, on the device
  AMJ_HIP_TRY(ctx, launch_batch_place(info, nq, sampled, ddesc, b->cap.slab, slab_words,
                                      b->cap.caps, b->cap.offs, b->cap.admitted, s));
  // 3. one direct-output launch per depth and kind
  AMJ_HIP_TRY(ctx, launch_fill_u64(b->cap.acc, 3ull * nq, 0, s));
  unsigned long long* rows = (unsigned long long*)b->cap.acc;
  for (uint32_t k = 0; k < nq;) {
    uint32_t e = k;
    const int L = b->qs[order[k]].plan.n_levels;
    const bool mat = b->qs[order[k]].materialize;
    while (e < nq && b->qs[order[e]].plan.n_levels == L && b->qs[order[e]].materialize == mat) ++e;
    const uint64_t gu = unit0[e - 1] + units[e - 1] - unit0[k];
    AMJ_HIP_TRY(ctx, launch_batch_direct(L, mat, (const uint8_t*)ddesc + db * k, e - k, unit0[k],
                                         gu, rows + k, rows + nq + k, rows + 2 * nq + k, nullptr,
                                         s));
    k = e;
  }
  return AMJ_OK;
}

}  // namespace

extern "C" {

int amj_batch_capture(amj_batch_t* b, uint64_t slab_bytes) {
  if (!b) return AMJ_ERR_ARG;
  amj_ctx_t* ctx = b->ctx;
  if (b->ran || b->captured) {
    set_error(ctx, AMJ_ERR_STATE, "the batch has already run or been captured");
    return AMJ_ERR_STATE;
  }
  const uint32_t nq = (uint32_t)b->qs.size();
  if (nq == 0) {
    set_error(ctx, AMJ_ERR_ARG, "capture of an empty batch");
    return AMJ_ERR_ARG;
  }
  for (const auto& q : b->qs)
    if (!q.direct) {
      set_error(ctx, AMJ_ERR_ARG, "a captured batch holds direct-output queries (opts.dro = 1)");
      return AMJ_ERR_ARG;
    }
  if (slab_bytes == 0) slab_bytes = b->budget ? b->budget : (1ull << 30);
  hipStream_t s = ctx->stream;
  // launch order: by depth, then kind (materialising first), as the eager direct pass
  auto& order = b->cap.order;
  order.resize(nq);
  for (uint32_t i = 0; i < nq; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    const int lx = b->qs[x].plan.n_levels, ly = b->qs[y].plan.n_levels;
    const bool mx = b->qs[x].materialize, my = b->qs[y].materialize;
    return lx != ly ? lx < ly : (mx != my && mx);
  });
  std::vector<uint64_t> units(nq), sunits(nq), unit0(nq), sunit0(nq);
  uint64_t total = 0, stotal = 0;
  const size_t db = batch_desc_bytes(), pib = batch_prep_info_bytes();
  std::vector<uint8_t> sd(db * nq), dd(db * nq), inf(pib * nq);
  for (uint32_t k = 0; k < nq; ++k) {
    const auto& q = b->qs[order[k]];
    units[k] = probe_units(q.plan);
    sunits[k] = (q.materialize && !q.capacity) ? std::min(units[k], kSampleUnits) : 0;
    const uint64_t stride = sunits[k] ? units[k] / sunits[k] : 1;
    unit0[k] = total;
    sunit0[k] = stotal;
    total += units[k];
    stotal += sunits[k];
    batch_fill_desc(sd.data() + db * k, q.plan, ProbeOut{}, sunit0[k], sunits[k], stride);
    batch_fill_desc(dd.data() + db * k, q.plan, ProbeOut{}, unit0[k], units[k], 1);
    batch_fill_prep_info(inf.data() + pib * k, units[k], sunits[k], q.capacity,
                         (uint32_t)q.plan.n_attrs, q.materialize);
  }
  void *sdesc = nullptr, *ddesc = nullptr, *info = nullptr, *counts = nullptr, *tmp = nullptr,
       *sampled = nullptr, *p = nullptr;
  AMJ_TRY(cap_alloc(b, sd.size(), &sdesc));
  AMJ_TRY(cap_alloc(b, dd.size(), &ddesc));
  AMJ_TRY(cap_alloc(b, inf.size(), &info));
  AMJ_TRY(cap_alloc(b, (stotal + 1) * 8, &counts));
  AMJ_TRY(cap_alloc(b, scan_temp_elems(stotal + 1) * 8, &tmp));
  AMJ_TRY(cap_alloc(b, (size_t)nq * 8, &sampled));
  AMJ_TRY(cap_alloc(b, (size_t)nq * 24, &p));
  b->cap.acc = (uint64_t*)p;
  AMJ_TRY(cap_alloc(b, (size_t)nq * 8, &p));
  b->cap.caps = (uint64_t*)p;
  AMJ_TRY(cap_alloc(b, (size_t)nq * 8, &p));
  b->cap.offs = (uint64_t*)p;
  AMJ_TRY(cap_alloc(b, (size_t)nq * 4, &p));
  b->cap.admitted = (uint32_t*)p;
  const uint64_t slab_words = slab_bytes / 256 * 64;
  AMJ_TRY(cap_alloc(b, slab_words * 4, &p));
  b->cap.slab = (int32_t*)p;
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(sdesc, sd.data(), sd.size(), hipMemcpyHostToDevice, s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(ddesc, dd.data(), dd.size(), hipMemcpyHostToDevice, s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(info, inf.data(), inf.size(), hipMemcpyHostToDevice, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  AMJ_HIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_chain(b, units, sunits, unit0, sunit0, stotal, sdesc, ddesc, info,
                               (uint64_t*)counts, (uint64_t*)tmp, (uint64_t*)sampled,
                               slab_words);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);  // ends the capture on every path
  if (rc != AMJ_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  AMJ_HIP_TRY(ctx, e);
  b->cap.graph = g;
  AMJ_HIP_TRY(ctx, hipGraphInstantiate(&b->cap.exec, g, nullptr, nullptr, 0));
  b->captured = true;
  return AMJ_OK;
}

int amj_batch_replay(amj_batch_t* b) {
  if (!b) return AMJ_ERR_ARG;
  amj_ctx_t* ctx = b->ctx;
  if (!b->captured) {
    set_error(ctx, AMJ_ERR_STATE, "amj_batch_replay needs a captured batch (amj_batch_capture)");
    return AMJ_ERR_STATE;
  }
  hipStream_t s = ctx->stream;
  // results of the previous replay: views into the slab lapse, exact reruns not taken are freed
  if (b->ran)
    for (auto& q : b->qs)
      if (!q.taken) amj_result_free(ctx, &q.res);
  const auto& order = b->cap.order;
  const uint32_t nq = (uint32_t)order.size();
  std::vector<uint64_t> h(5 * (size_t)nq);
  std::vector<uint32_t> adm(nq);
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  AMJ_HIP_TRY(ctx, hipGraphLaunch(b->cap.exec, s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(h.data(), b->cap.acc, (size_t)nq * 24, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(h.data() + 3 * (size_t)nq, b->cap.caps, (size_t)nq * 8,
                                  hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(h.data() + 4 * (size_t)nq, b->cap.offs, (size_t)nq * 8,
                                  hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(adm.data(), b->cap.admitted, (size_t)nq * 4,
                                  hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  std::vector<uint32_t> redo;
  for (uint32_t k = 0; k < nq; ++k) {
    auto& q = b->qs[order[k]];
    std::memset(&q.res, 0, sizeof(q.res));
    q.status = AMJ_OK;
    q.taken = false;
    const uint64_t cnt = h[nq + k], cs = h[2 * (size_t)nq + k], capk = h[3 * (size_t)nq + k];
    if (q.materialize && (!adm[k] || cnt > capk)) {  // no room or overran: run again exactly
      redo.push_back(order[k]);
      continue;
    }
    if (q.materialize && cnt > 0)
      for (int a = 0; a < q.plan.n_attrs; ++a)
        q.res.cols[a] = b->cap.slab + h[4 * (size_t)nq + k] + (uint64_t)a * capk;
    set_result(q, cnt, cs);
    q.res.reserved[2] = 3;  // a view into the batch's slab: freeing it releases nothing
    q.res.passes = q.materialize && !q.capacity ? 2 : 1;
  }
  b->ran = true;
  if (!redo.empty()) {
    const int rc = run_exact(b, redo);
    if (rc != AMJ_OK) {
      for (uint32_t i : redo) b->qs[i].status = rc;
    }
  }
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  b->run_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  ++b->cap.replays;
  for (auto& q : b->qs) {
    q.res.probe_ms = b->run_ms;
    q.res.probe_kernel_ms = b->run_ms;
    q.res.probe_kernel_launches = 1;
  }
  return AMJ_OK;
}

}  // extern "C"
