// Host driver of batched AMHJ execution (amj_batch_*): plans of many queries over resident
// relations and hash tables, executed together by the batch kernels (batch.hip) with two host
// round trips per batch whatever its size: counts -> one slab for every materialised result
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
                     uint64_t units);

}  // namespace amj

using namespace amj;

struct amj_batch {
  amj_ctx_t* ctx = nullptr;
  struct Query {
    Plan plan;
    std::vector<uint32_t> attr_ids;
    bool materialize = true;
    amj_result_t res{};
    bool taken = false;
  };
  std::vector<Query> qs;
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
  delete b;
}

static int batch_add_plan(amj_batch_t* b, Plan& plan, std::vector<uint32_t>& ids,
                          const amj_opts_t* opts, uint32_t* id) {
  if (b->ran) {
    set_error(b->ctx, AMJ_ERR_STATE, "the batch has already run");
    return AMJ_ERR_STATE;
  }
  amj_batch::Query q;
  q.plan = plan;
  q.attr_ids = ids;
  q.materialize = opts ? opts->materialize != 0 : true;
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

int amj_batch_run(amj_batch_t* b) {
  if (!b) return AMJ_ERR_ARG;
  amj_ctx_t* ctx = b->ctx;
  if (b->ran) {
    set_error(ctx, AMJ_ERR_STATE, "the batch has already run");
    return AMJ_ERR_STATE;
  }
  hipStream_t s = ctx->stream;
  const uint32_t nq = (uint32_t)b->qs.size();
  b->ran = true;
  if (nq == 0) return AMJ_OK;
  // queries grouped by their number of build tables; units numbered consecutively
  std::vector<uint32_t> order(nq);
  for (uint32_t i = 0; i < nq; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    return b->qs[x].plan.n_levels < b->qs[y].plan.n_levels;
  });
  std::vector<uint64_t> unit0(nq), units(nq);
  uint64_t total = 0;
  for (uint32_t k = 0; k < nq; ++k) {
    const Plan& P = b->qs[order[k]].plan;
    unit0[k] = total;
    units[k] = probe_units(P);
    total += units[k];
  }
  const size_t db = batch_desc_bytes();
  std::vector<uint8_t> desc(db * nq);
  for (uint32_t k = 0; k < nq; ++k)
    batch_fill_desc(desc.data() + db * k, b->qs[order[k]].plan, ProbeOut{}, unit0[k], units[k]);
  void *ddesc = nullptr, *counts = nullptr, *tmp = nullptr, *totals = nullptr, *csums = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_MISC0, desc.size(), &ddesc));
  AMJ_TRY(scratch(ctx, SLOT_UNIT_COUNTS, (total + 1) * 8, &counts));
  AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(total + 1) * 8, &tmp));
  AMJ_TRY(scratch(ctx, SLOT_MISC3, (size_t)nq * 16, &totals));
  csums = (uint64_t*)totals + nq;
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(ddesc, desc.data(), desc.size(), hipMemcpyHostToDevice, s));
  AMJ_HIP_TRY(ctx, hipMemsetAsync(counts, 0, (total + 1) * 8, s));
  // count pass: one launch per group of equal depth
  for (uint32_t k = 0; k < nq;) {
    uint32_t e = k;
    const int L = b->qs[order[k]].plan.n_levels;
    while (e < nq && b->qs[order[e]].plan.n_levels == L) ++e;
    const uint64_t gu = unit0[e - 1] + units[e - 1] - unit0[k];
    AMJ_HIP_TRY(ctx, launch_batch_count(L, (const uint8_t*)ddesc + db * k, e - k, unit0[k], gu,
                                        (uint64_t*)counts, s));
    k = e;
  }
  AMJ_HIP_TRY(ctx, exclusive_scan_u64((uint64_t*)counts, (uint64_t*)counts, total + 1,
                                      (uint64_t*)tmp, s));
  AMJ_HIP_TRY(ctx, launch_batch_totals(ddesc, nq, (const uint64_t*)counts, (uint64_t*)totals, s));
  std::vector<uint64_t> cnt(nq);
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(cnt.data(), totals, nq * 8, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  // one slab for every materialised result
  uint64_t slab_bytes = 0;
  std::vector<uint64_t> off(nq, 0);
  for (uint32_t k = 0; k < nq; ++k) {
    const auto& q = b->qs[order[k]];
    if (!q.materialize || cnt[k] == 0) continue;
    off[k] = slab_bytes;
    slab_bytes += (cnt[k] * (uint64_t)q.plan.n_attrs * 4 + 255) / 256 * 256;
  }
  void* slab = nullptr;
  b->needed = slab_bytes;
  if (slab_bytes) {
    const uint64_t room = budget_left(ctx, nullptr);
    if (slab_bytes > (b->budget ? std::min(b->budget, room) : room)) {
      set_error(ctx, AMJ_ERR_NOMEM, "the batch's results (" + std::to_string(slab_bytes) +
                                        " bytes) do not fit the budget");
      return AMJ_ERR_NOMEM;
    }
    amj_result_t tmpres{};
    AMJ_TRY(alloc_output(ctx, slab_bytes / 4, 1, &tmpres));  // pooled, one column of words
    slab = tmpres.cols[0];
    amj_ctx::Slab sl;
    sl.bytes = (size_t)tmpres.reserved[0] | ((size_t)tmpres.reserved[1] << 32);
    sl.refs = 0;
    ctx->slabs[slab] = sl;
  }
  for (uint32_t k = 0; k < nq; ++k) {
    auto& q = b->qs[order[k]];
    ProbeOut out{};
    std::memset(&q.res, 0, sizeof(q.res));
    q.res.count = cnt[k];
    q.res.num_cols = (uint32_t)q.plan.n_attrs;
    for (int a = 0; a < q.plan.n_attrs; ++a) q.res.attr_ids[a] = q.attr_ids[a];
    if (q.materialize && cnt[k] > 0) {
      int32_t* base = (int32_t*)((uint8_t*)slab + off[k]);
      for (int a = 0; a < q.plan.n_attrs; ++a) {
        out.cols[a] = base + (uint64_t)a * cnt[k];
        q.res.cols[a] = out.cols[a];
      }
      out.capacity = cnt[k];
      q.res.rows_written = cnt[k];
      q.res.reserved[0] = (uint32_t)((uint64_t)(uintptr_t)slab & 0xffffffffu);
      q.res.reserved[1] = (uint32_t)((uint64_t)(uintptr_t)slab >> 32);
      q.res.reserved[2] = 2;
      ctx->slabs[slab].refs++;
    }
    batch_fill_desc(desc.data() + db * k, q.plan, out, unit0[k], units[k]);
  }
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(ddesc, desc.data(), desc.size(), hipMemcpyHostToDevice, s));
  AMJ_HIP_TRY(ctx, hipMemsetAsync(csums, 0, nq * 8, s));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
  for (uint32_t k = 0; k < nq;) {
    uint32_t e = k;
    const int L = b->qs[order[k]].plan.n_levels;
    while (e < nq && b->qs[order[e]].plan.n_levels == L) ++e;
    const uint64_t gu = unit0[e - 1] + units[e - 1] - unit0[k];
    AMJ_HIP_TRY(ctx, launch_batch_write(L, (const uint8_t*)ddesc + db * k, e - k, unit0[k], gu,
                                        (const uint64_t*)counts,
                                        (unsigned long long*)csums + k, s));
    k = e;
  }
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
  std::vector<uint64_t> cs(nq);
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(cs.data(), csums, nq * 8, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  b->run_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  const double kms = elapsed_ms(ctx->ev[3], ctx->ev[4]);
  for (uint32_t k = 0; k < nq; ++k) {
    auto& r = b->qs[order[k]].res;
    r.checksum = cs[k];
    r.probe_ms = b->run_ms;
    r.probe_kernel_ms = kms;
    r.probe_kernel_launches = 1;
    r.passes = 2;
    r.path = AMJ_PATH_BATCH;
  }
  return AMJ_OK;
}

int amj_batch_result(amj_batch_t* b, uint32_t id, amj_result_t* res) {
  if (!b || !res || id >= b->qs.size()) return AMJ_ERR_ARG;
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
