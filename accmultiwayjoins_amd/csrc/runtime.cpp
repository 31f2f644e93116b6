// Host runtime of libamj: context, device buffers, relations, hash build driver, query
// planning and the AMHJ execution driver (count/scan/write and frequent-item paths).
#include "runtime.h"
#include "tpt.h"

#include <algorithm>
#include <cstring>

namespace amj {

void set_error(amj_ctx_t* ctx, int code, const std::string& msg) {
  (void)code;
  if (ctx) ctx->last_error = msg;
}

uint32_t floor_log2_u64(uint64_t v) {
  uint32_t r = 0;
  while (v >>= 1) ++r;
  return r;
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
  return (double)ms;
}

int dmalloc(amj_ctx_t* ctx, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess && !ctx->out_pool.empty()) {  // parked result buffers give way first
    (void)hipGetLastError();
    pool_clear(ctx);
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error(ctx, AMJ_ERR_NOMEM,
              "hipMalloc(" + std::to_string(bytes) + " bytes) failed: " + hipGetErrorString(e));
    return AMJ_ERR_NOMEM;
  }
  ctx->allocated += bytes;
  ctx->peak = std::max(ctx->peak, ctx->allocated);
  return AMJ_OK;
}

void dfree(amj_ctx_t* ctx, void* p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  if (bytes == 0) bytes = 256;
  if (ctx) ctx->allocated -= std::min<uint64_t>(ctx->allocated, bytes);
}

int scratch(amj_ctx_t* ctx, int slot, size_t bytes, void** out) {
  auto& b = ctx->scratch[slot];
  if (b.bytes < bytes) {
    const size_t old = b.bytes;  // grow by at least half, so growing requests rarely reallocate
    if (b.p) {
      AMJ_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      dfree(ctx, b.p, b.bytes);
      b.p = nullptr;
      b.bytes = 0;
    }
    size_t want = std::max(bytes, old + old / 2);
    int rc = dmalloc(ctx, &b.p, want);
    if (rc != AMJ_OK) {
      rc = dmalloc(ctx, &b.p, bytes);
      if (rc != AMJ_OK) return rc;
      want = bytes;
    }
    b.bytes = want;
  }
  *out = b.p;
  return AMJ_OK;
}

// ------------------------------------------------------------------------------------------
// Hash build.
// ------------------------------------------------------------------------------------------
int build_hash_table(amj_ctx_t* ctx, const amj_relation* rel, uint32_t key_col, int sort_vals,
                     amj_hash_table** out) {
  *out = nullptr;
  if (!rel || key_col >= rel->num_cols) {
    set_error(ctx, AMJ_ERR_ARG, "hash build: bad relation or key column");
    return AMJ_ERR_ARG;
  }
  if (rel->host) {
    set_error(ctx, AMJ_ERR_ARG, "hash build: the relation is host-resident (only a probe "
                                "relation may stay on the host)");
    return AMJ_ERR_ARG;
  }
  if (rel->rows >= (1ull << 32)) {
    set_error(ctx, AMJ_ERR_ARG, "hash build: a build relation must have fewer than 2^32 rows");
    return AMJ_ERR_ARG;
  }
  const uint64_t n = rel->rows;
  // B = floorPowerOf2(N / BUC_RATIO) with BUC_RATIO = 1 (QueryProcessor.cuh:86, conf.h:25)
  const uint32_t log_b = n == 0 ? 0 : floor_log2_u64(n);
  const uint64_t buckets = 1ull << log_b;
  std::unique_ptr<amj_hash_table> ht(new amj_hash_table());
  ht->rel = rel;
  ht->key_col = key_col;
  ht->vals_sorted = sort_vals != 0;
  const int32_t* key = rel->cols[key_col];
  const int32_t* val = rel->num_cols > 1 ? rel->cols[1 - key_col] : nullptr;

  hipStream_t s = ctx->stream;
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[0], s));
  void *comp = nullptr, *alt = nullptr, *tmp = nullptr, *flag = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_SORT_KEYS, std::max<uint64_t>(n, 1) * 8, &comp));
  AMJ_TRY(scratch(ctx, SLOT_SORT_ALT, std::max<uint64_t>(n, 1) * 8, &alt));
  AMJ_TRY(scratch(ctx, SLOT_SORT_TEMP, radix_sort_temp_bytes(std::max<uint64_t>(n, 1)), &tmp));
  AMJ_TRY(scratch(ctx, SLOT_FLAG, 256, &flag));

  auto& d = ht->dev;
  d.buckets = buckets;
  AMJ_TRY(dmalloc(ctx, (void**)&d.ptrs, (buckets + 1) * sizeof(uint32_t)));
  AMJ_TRY(dmalloc(ctx, (void**)&d.keys, std::max<uint64_t>(n, 1) * sizeof(int32_t)));
  if (val) AMJ_TRY(dmalloc(ctx, (void**)&d.vals, std::max<uint64_t>(n, 1) * sizeof(int32_t)));

  uint32_t identity = 0;
  if (n > 0) {
    AMJ_HIP_TRY(ctx, launch_pack_composite(key, val, n, log_b, (uint64_t*)comp, s));
    // sort digits 4..7 (hash column) always, 0..3 (second column) when it is compared
    const uint32_t mask = sort_vals && val ? 0xffu : 0xf0u;
    AMJ_HIP_TRY(ctx, radix_sort_u64((uint64_t*)comp, (uint64_t*)alt, n, mask, tmp, s));
    AMJ_HIP_TRY(ctx, launch_unpack_composite((const uint64_t*)comp, n, log_b, d.keys, d.vals, s));
    AMJ_HIP_TRY(ctx, launch_bucket_ptrs((const uint64_t*)comp, n, log_b, buckets, d.ptrs, s));
    AMJ_HIP_TRY(ctx, launch_key_range(key, n, log_b, (uint32_t*)flag, s));
    uint32_t hflag = 1;
    AMJ_HIP_TRY(ctx, hipMemcpyAsync(&hflag, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
    identity = hflag == 0 ? 1u : 0u;
  } else {
    AMJ_HIP_TRY(ctx, hipMemsetAsync(d.ptrs, 0, (buckets + 1) * sizeof(uint32_t), s));
  }
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[1], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[1]));
  ht->build_ms = elapsed_ms(ctx->ev[0], ctx->ev[1]);

  d.view.ptrs = d.ptrs;
  d.view.keys = d.keys;
  d.view.vals = d.vals;
  d.view.pw = nullptr;
  d.view.n = n;
  d.view.log_b = log_b;
  d.view.identity = identity;
  d.view.vals_sorted = ht->vals_sorted ? 1u : 0u;
  *out = ht.release();
  return AMJ_OK;
}

// ------------------------------------------------------------------------------------------
// Planning: attribute numbering in first-occurrence order (exec_AMHJ.cu:76-88).
// ------------------------------------------------------------------------------------------
int make_plan(amj_ctx_t* ctx, const ProbeUse& probe, const std::vector<TableUse>& uses,
              Plan* plan, std::vector<uint32_t>* out_attr_ids, std::vector<int>* order) {
  std::memset(plan, 0, sizeof(Plan));
  const int L = (int)uses.size();
  if (L < 1 || L > kMaxLevels) {
    set_error(ctx, AMJ_ERR_ARG, "query needs 1.." + std::to_string(kMaxLevels) +
                                    " build tables, got " + std::to_string(L));
    return AMJ_ERR_ARG;
  }
  if (!probe.rel || probe.rel->num_cols < 1 || probe.rel->num_cols > AMJ_MAX_PROBE_COLS) {
    set_error(ctx, AMJ_ERR_ARG, "probe relation must have 1 or 2 columns");
    return AMJ_ERR_ARG;
  }
  std::vector<uint32_t> ids;  // renumbered attr -> original id
  auto find_id = [&](uint32_t id) -> int {
    for (size_t i = 0; i < ids.size(); ++i)
      if (ids[i] == id) return (int)i;
    return -1;
  };
  plan->probe_n = probe.rel->rows;
  plan->probe_on_host = probe.rel->host ? 1 : 0;
  plan->probe_ncols = (int)probe.rel->num_cols;
  for (uint32_t c = 0; c < probe.rel->num_cols; ++c) {
    plan->probe_cols[c] = probe.rel->cols[c];
    int a = find_id(probe.attr_ids[c]);
    if (a < 0) {
      ids.push_back(probe.attr_ids[c]);
      a = (int)ids.size() - 1;
    }
    plan->probe_attr[c] = a;
  }
  // Order: compare-only tables move to the earliest level where both attributes are bound.
  std::vector<int> remaining(L);
  for (int i = 0; i < L; ++i) remaining[i] = i;
  std::vector<int> ord;
  std::vector<uint32_t> bound_ids(ids.begin(), ids.end());
  auto is_bound = [&](uint32_t id) {
    return std::find(bound_ids.begin(), bound_ids.end(), id) != bound_ids.end();
  };
  while (!remaining.empty()) {
    int pick = -1;
    for (size_t i = 0; i < remaining.size(); ++i) {
      const TableUse& u = uses[remaining[i]];
      if (u.val_attr_id >= 0 && is_bound(u.key_attr_id) && is_bound((uint32_t)u.val_attr_id)) {
        pick = (int)i;
        break;
      }
    }
    if (pick < 0) pick = 0;
    const TableUse& u = uses[remaining[pick]];
    if (u.val_attr_id >= 0) bound_ids.push_back((uint32_t)u.val_attr_id);
    ord.push_back(remaining[pick]);
    remaining.erase(remaining.begin() + pick);
  }
  for (int t = 0; t < L; ++t) {
    const TableUse& u = uses[ord[t]];
    if (!u.ht) {
      set_error(ctx, AMJ_ERR_ARG, "null hash table");
      return AMJ_ERR_ARG;
    }
    const int ka = find_id(u.key_attr_id);
    if (ka < 0) {
      set_error(ctx, AMJ_ERR_ARG,
                "build table " + std::to_string(ord[t]) + ": hash attribute " +
                    std::to_string(u.key_attr_id) + " is not bound by an earlier relation");
      return AMJ_ERR_ARG;
    }
    plan->tab[t] = u.ht->dev.view;
    plan->key_attr[t] = ka;
    if (u.val_attr_id < 0 || !u.ht->dev.vals) {
      plan->val_attr[t] = -1;
      plan->val_mode[t] = VAL_NONE;
    } else {
      int va = find_id((uint32_t)u.val_attr_id);
      if (va >= 0) {
        if (!u.ht->vals_sorted) {
          set_error(ctx, AMJ_ERR_ARG,
                    "build table " + std::to_string(ord[t]) +
                        " compares its second column; build it with sort_vals=1");
          return AMJ_ERR_ARG;
        }
        plan->val_mode[t] = VAL_CHECK;
      } else {
        if ((int)ids.size() >= kMaxAttrs) {
          set_error(ctx, AMJ_ERR_ARG, "too many attributes in the result (max " +
                                          std::to_string(kMaxAttrs) + ")");
          return AMJ_ERR_ARG;
        }
        ids.push_back((uint32_t)u.val_attr_id);
        va = (int)ids.size() - 1;
        plan->val_mode[t] = VAL_BIND;
      }
      plan->val_attr[t] = va;
    }
  }
  plan->n_levels = L;
  plan->n_attrs = (int)ids.size();
  if (out_attr_ids) *out_attr_ids = ids;
  if (order) *order = ord;
  return AMJ_OK;
}

bool fib_eligible(const Plan& P) {
  for (int t = 0; t < P.n_levels; ++t) {
    if (P.val_mode[t] == VAL_CHECK) return false;
    if (t + 1 < P.n_levels) {
      if (P.val_mode[t] != VAL_BIND) return false;
      if (P.key_attr[t + 1] != P.val_attr[t]) return false;
    }
  }
  return true;
}

// Output columns: reuse a parked allocation when it is large enough.
// Freeing a large buffer and allocating another soon after can cost seconds on this device
// (the freed memory is cleared before it is handed out again), so released result buffers
// are parked and reused instead, up to a third of the device memory (more would starve other
// contexts on the device: a server reserves half of it for its results), or one buffer alone
// up to 2/5 of it (a 108 GB direct output of the 4-way chain with 2^32 results took 3.0 s per
// repeat when it was freed on release, 65 ms when parked); any allocation that fails gives the
// parked buffers back and tries again (dmalloc). Parking up to 3/5 left a server created next
// to such a context without room.
constexpr size_t kPoolCapBuffers = 256;

void pool_put(amj_ctx_t* ctx, void* p, size_t bytes) {
  if (!p) return;
  ctx->out_pool.push_back(amj_ctx::Buf{p, bytes});
  ctx->pool_bytes += bytes;
  const uint64_t cap_bytes = (uint64_t)ctx->props.totalGlobalMem / 3;
  const uint64_t alone_bytes = (uint64_t)ctx->props.totalGlobalMem / 5 * 2;
  auto over = [&] {
    if (ctx->out_pool.size() > kPoolCapBuffers) return true;
    if (ctx->out_pool.size() == 1) return ctx->pool_bytes > alone_bytes;
    return ctx->pool_bytes > cap_bytes;
  };
  while (!ctx->out_pool.empty() && over()) {
    const amj_ctx::Buf b = ctx->out_pool.front();
    ctx->out_pool.erase(ctx->out_pool.begin());
    ctx->pool_bytes -= b.bytes;
    dfree(ctx, b.p, b.bytes);
  }
}

void pool_clear(amj_ctx_t* ctx) {
  for (auto& b : ctx->out_pool) dfree(ctx, b.p, b.bytes);
  ctx->out_pool.clear();
  ctx->pool_bytes = 0;
}

int arena_drop(amj_ctx_t* ctx) {
  if (!ctx->arena.base) return AMJ_OK;
  if (ctx->arena.free_bytes != ctx->arena.size) {
    set_error(ctx, AMJ_ERR_STATE, "the result arena still holds results");
    return AMJ_ERR_STATE;
  }
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  dfree(ctx, ctx->arena.base, ctx->arena.size);
  ctx->arena.base = nullptr;
  ctx->arena.size = 0;
  ctx->arena.free.clear();
  ctx->arena.free_bytes = 0;
  return AMJ_OK;
}

int arena_reserve(amj_ctx_t* ctx, uint64_t bytes) {
  if (ctx->arena.base) {
    set_error(ctx, AMJ_ERR_STATE, "the context already has a result arena");
    return AMJ_ERR_STATE;
  }
  bytes = (bytes + 255) / 256 * 256;
  void* p = nullptr;
  AMJ_TRY(dmalloc(ctx, &p, bytes));
  ctx->arena.base = (char*)p;
  ctx->arena.size = bytes;
  ctx->arena.free.clear();
  ctx->arena.free[0] = bytes;
  ctx->arena.free_bytes = bytes;
  return AMJ_OK;
}

static bool arena_take(amj_ctx_t* ctx, uint64_t bytes, void** out) {
  auto& A = ctx->arena;
  bytes = (bytes + 255) / 256 * 256;
  for (auto it = A.free.begin(); it != A.free.end(); ++it) {
    if (it->second < bytes) continue;
    const uint64_t off = it->first, len = it->second;
    A.free.erase(it);
    if (len > bytes) A.free[off + bytes] = len - bytes;
    A.free_bytes -= bytes;
    *out = A.base + off;
    return true;
  }
  return false;
}

void arena_release(amj_ctx_t* ctx, void* p, uint64_t bytes) {
  auto& A = ctx->arena;
  bytes = (bytes + 255) / 256 * 256;
  uint64_t off = (uint64_t)((char*)p - A.base), len = bytes;
  A.free_bytes += bytes;
  auto next = A.free.lower_bound(off);
  if (next != A.free.end() && off + len == next->first) {  // merge with the range after
    len += next->second;
    next = A.free.erase(next);
  }
  if (next != A.free.begin()) {  // merge with the range before
    auto prev = std::prev(next);
    if (prev->first + prev->second == off) {
      prev->second += len;
      return;
    }
  }
  A.free[off] = len;
}

uint64_t arena_largest(const amj_ctx_t* ctx) {
  uint64_t m = 0;
  for (const auto& kv : ctx->arena.free) m = std::max(m, kv.second);
  return m;
}

// A result takes the smallest parked buffer that holds it when that buffer is at most twice
// its size (plus 64 MiB), so a small result never pins a large buffer; otherwise it is
// allocated, and only when the device is full is the pool given back first.
int alloc_output(amj_ctx_t* ctx, uint64_t rows, int n_attrs, amj_result_t* res) {
  const size_t bytes = std::max<uint64_t>(rows, 1) * sizeof(int32_t) * (size_t)n_attrs;
  void* base = nullptr;
  size_t got = 0;
  if (ctx->arena.base) {
    if (!arena_take(ctx, bytes, &base)) {
      set_error(ctx, AMJ_ERR_NOMEM, "result of " + std::to_string(bytes) +
                                        " bytes does not fit the result arena");
      return AMJ_ERR_NOMEM;
    }
    const uint64_t per_col = rows > 0 ? rows : 1;
    for (int a = 0; a < n_attrs; ++a) res->cols[a] = (int32_t*)base + (uint64_t)a * per_col;
    res->reserved[0] = (uint32_t)(bytes & 0xffffffffu);
    res->reserved[1] = (uint32_t)((uint64_t)bytes >> 32);
    res->reserved[2] = 2;  // cols[0] is the base of an arena range
    return AMJ_OK;
  }
  size_t best = 0;
  int pick = -1;
  for (size_t i = 0; i < ctx->out_pool.size(); ++i) {
    const size_t b = ctx->out_pool[i].bytes;
    if (b >= bytes && b <= 2 * bytes + (64u << 20) && (pick < 0 || b < best)) {
      pick = (int)i;
      best = b;
    }
  }
  if (pick >= 0) {
    base = ctx->out_pool[pick].p;
    got = ctx->out_pool[pick].bytes;
    ctx->out_pool.erase(ctx->out_pool.begin() + pick);
    ctx->pool_bytes -= got;
  } else {
    int rc = dmalloc(ctx, &base, bytes);
    if (rc == AMJ_ERR_NOMEM && !ctx->out_pool.empty()) {
      pool_clear(ctx);
      rc = dmalloc(ctx, &base, bytes);
    }
    if (rc != AMJ_OK) return rc;
    got = bytes;
  }
  const uint64_t per_col = got / sizeof(int32_t) / (size_t)n_attrs;
  for (int a = 0; a < n_attrs; ++a) res->cols[a] = (int32_t*)base + (uint64_t)a * per_col;
  res->reserved[0] = (uint32_t)(got & 0xffffffffu);
  res->reserved[1] = (uint32_t)(got >> 32);
  res->reserved[2] = 1;  // cols[0] is the base of a pooled allocation
  return AMJ_OK;
}

int finish_checksum(amj_ctx_t* ctx, const uint64_t* block_sums, uint64_t nb,
                           uint64_t* out) {
  void *sum = nullptr, *tmp = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_SUM, 64, &sum));
  AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(nb) * 8, &tmp));
  AMJ_HIP_TRY(ctx, reduce_sum_u64(block_sums, nb, (uint64_t*)sum, (uint64_t*)tmp, ctx->stream));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(out, sum, 8, hipMemcpyDeviceToHost, ctx->stream));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return AMJ_OK;
}

uint64_t budget_left(amj_ctx_t* ctx, const amj_opts_t* o) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  uint64_t left = free_b + ctx->pool_bytes;  // parked buffers are reusable
  if (ctx->arena.base) left = arena_largest(ctx);
  if (o && o->hbm_budget_bytes) {
    const uint64_t cap = o->hbm_budget_bytes > ctx->allocated ? o->hbm_budget_bytes - ctx->allocated : 0;
    left = std::min<uint64_t>(left, cap);
  }
  return left;
}

// Plain AMHJ: count pass -> exclusive scan -> write pass (AMHJ.cuh:314-377).
// Count + checksum without output: one pass that hashes every result and counts per block.
static int run_plain_count(amj_ctx_t* ctx, Plan& P, amj_result_t* res) {
  hipStream_t s = ctx->stream;
  const uint64_t nb = probe_write_blocks(P);
  void *bsum = nullptr, *bcnt = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_BLOCK_SUMS, nb * 8, &bsum));
  AMJ_TRY(scratch(ctx, SLOT_MISC1, nb * 8, &bcnt));
  uint64_t nb_used = 0;
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
  AMJ_HIP_TRY(ctx, launch_probe_write(P, nullptr, ProbeOut{}, true, (uint64_t*)bsum, &nb_used, s,
                                      (uint64_t*)bcnt));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
  AMJ_TRY(finish_checksum(ctx, (const uint64_t*)bsum, nb_used, &res->checksum));
  AMJ_TRY(finish_checksum(ctx, (const uint64_t*)bcnt, nb_used, &res->count));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->probe_kernel_ms = elapsed_ms(ctx->ev[3], ctx->ev[4]);
  res->probe_kernel_launches = 1;
  res->rows_written = 0;
  res->path = AMJ_PATH_PLAIN;
  res->passes = 1;
  return AMJ_OK;
}

static int run_plain(amj_ctx_t* ctx, Plan& P, const amj_opts_t* o, amj_result_t* res) {
  if (!o->materialize) return run_plain_count(ctx, P, res);
  hipStream_t s = ctx->stream;
  const uint64_t units = probe_units(P);
  void *counts = nullptr, *tmp = nullptr, *bsum = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_UNIT_COUNTS, (units + 1) * 8, &counts));
  AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(units + 1) * 8, &tmp));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  AMJ_HIP_TRY(ctx, hipMemsetAsync(counts, 0, (units + 1) * 8, s));
  AMJ_HIP_TRY(ctx, launch_probe_count(P, (uint64_t*)counts, s));
  AMJ_HIP_TRY(ctx, exclusive_scan_u64((uint64_t*)counts, (uint64_t*)counts, units + 1,
                                      (uint64_t*)tmp, s));
  uint64_t total = 0;
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(&total, (uint64_t*)counts + units, 8, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  res->count = total;

  ProbeOut out{};
  if (o->materialize && total > 0) {
    const uint64_t need = total * (uint64_t)P.n_attrs * 4;
    if (need > budget_left(ctx, o)) {
      set_error(ctx, AMJ_ERR_NOMEM, "result of " + std::to_string(total) +
                                        " tuples does not fit the HBM budget (use ooc=1 or "
                                        "materialize=0)");
      return AMJ_ERR_NOMEM;
    }
    AMJ_TRY(alloc_output(ctx, total, P.n_attrs, res));
    for (int a = 0; a < P.n_attrs; ++a) out.cols[a] = res->cols[a];
    out.capacity = total;
  }
  const uint64_t nb = probe_write_blocks(P);
  AMJ_TRY(scratch(ctx, SLOT_BLOCK_SUMS, nb * 8, &bsum));
  uint64_t nb_used = 0;
  const bool hash = o->checksum || !out.cols[0];
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
  AMJ_HIP_TRY(ctx, launch_probe_write(P, (const uint64_t*)counts, out, hash, (uint64_t*)bsum,
                                      &nb_used, s));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
  if (hash) AMJ_TRY(finish_checksum(ctx, (const uint64_t*)bsum, nb_used, &res->checksum));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->probe_kernel_ms = elapsed_ms(ctx->ev[3], ctx->ev[4]);
  res->probe_kernel_launches = 1;
  res->rows_written = out.cols[0] ? total : 0;
  res->path = AMJ_PATH_PLAIN;
  res->passes = 2;
  return AMJ_OK;
}

// Thread-per-tuple baseline: count pass, scan of the per-lane counts, write pass (the
// count pass also hashes when nothing is materialised).
// Open-addressing tables of every level (AMJ_ENGINE_OPEN_ADDR), built from the bucketed
// tables' key arrays: ceilPow2(2n) slots, filled on the device; freed with the holder.
struct OaHolder {
  amj_ctx_t* ctx;
  OaTables t{};
  std::vector<std::pair<void*, size_t>> bufs;
  explicit OaHolder(amj_ctx_t* c) : ctx(c) {}
  ~OaHolder() {
    if (!bufs.empty()) (void)hipStreamSynchronize(ctx->stream);
    for (auto& b : bufs) dfree(ctx, b.first, b.second);
  }
  int build(const Plan& P) {
    for (int l = 0; l < P.n_levels; ++l) {
      const uint64_t n = P.tab[l].n;
      if (n >= 0xffffffffull || n > (1ull << 30)) {
        set_error(ctx, AMJ_ERR_ARG, "open addressing: a build table exceeds 2^30 rows");
        return AMJ_ERR_ARG;
      }
      uint64_t cap = 64;
      while (cap < 2 * n) cap <<= 1;  // load factor at most 1/2 (OA_HT_RATIO, linear.cuh:11)
      void* p = nullptr;
      AMJ_TRY(dmalloc(ctx, &p, cap * 8));
      bufs.emplace_back(p, cap * 8);
      AMJ_HIP_TRY(ctx, hipMemsetAsync(p, 0xff, cap * 8, ctx->stream));
      AMJ_HIP_TRY(ctx, launch_oa_insert(P.tab[l].keys, n, (uint64_t*)p, (uint32_t)(cap - 1),
                                        ctx->stream));
      t.slots[l] = (const uint64_t*)p;
      t.mask[l] = (uint32_t)(cap - 1);
    }
    return AMJ_OK;
  }
};

static int run_tpt(amj_ctx_t* ctx, Plan& P, const amj_opts_t* o, amj_result_t* res) {
  hipStream_t s = ctx->stream;
  const bool open_addr = o->engine == AMJ_ENGINE_OPEN_ADDR;
  OaHolder oa(ctx);
  if (open_addr) {
    AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[0], s));
    AMJ_TRY(oa.build(P));
    AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[1], s));
    AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[1]));
    res->build_ms = elapsed_ms(ctx->ev[0], ctx->ev[1]);
  }
  const OaTables* oat = open_addr ? &oa.t : nullptr;
  const uint64_t nb = tpt_blocks(P), lanes = nb * kBlock;
  void *counts = nullptr, *tmp = nullptr, *bsum = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_UNIT_COUNTS, (lanes + 1) * 8, &counts));
  AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(lanes + 1) * 8, &tmp));
  AMJ_TRY(scratch(ctx, SLOT_BLOCK_SUMS, nb * 8, &bsum));
  const bool mat = o->materialize != 0;
  const bool hash_in_count = !mat;
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  AMJ_HIP_TRY(ctx, hipMemsetAsync((uint64_t*)counts + lanes, 0, 8, s));
  ProbeOut none{};
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
  AMJ_HIP_TRY(ctx, launch_tpt(P, false, hash_in_count, (uint64_t*)counts, nullptr, none,
                              (uint64_t*)bsum, s, oat));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
  AMJ_HIP_TRY(ctx, exclusive_scan_u64((uint64_t*)counts, (uint64_t*)counts, lanes + 1,
                                      (uint64_t*)tmp, s));
  uint64_t total = 0;
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(&total, (uint64_t*)counts + lanes, 8, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  res->count = total;
  double kernel_ms = elapsed_ms(ctx->ev[3], ctx->ev[4]);
  ProbeOut out{};
  if (mat && total > 0) {
    const uint64_t need = total * (uint64_t)P.n_attrs * 4;
    if (need > budget_left(ctx, o)) {
      set_error(ctx, AMJ_ERR_NOMEM, "result of " + std::to_string(total) +
                                        " tuples does not fit the HBM budget (use materialize=0)");
      return AMJ_ERR_NOMEM;
    }
    AMJ_TRY(alloc_output(ctx, total, P.n_attrs, res));
    for (int a = 0; a < P.n_attrs; ++a) out.cols[a] = res->cols[a];
    out.capacity = total;
    const bool hash = o->checksum != 0;
    AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
    AMJ_HIP_TRY(ctx, launch_tpt(P, true, hash, nullptr, (const uint64_t*)counts, out,
                                (uint64_t*)bsum, s, oat));
    AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
    if (hash) AMJ_TRY(finish_checksum(ctx, (const uint64_t*)bsum, nb, &res->checksum));
    AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[4]));
    kernel_ms += elapsed_ms(ctx->ev[3], ctx->ev[4]);
  } else if (hash_in_count) {
    AMJ_TRY(finish_checksum(ctx, (const uint64_t*)bsum, nb, &res->checksum));
  }  // else: materialising an empty result, checksum 0
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->probe_kernel_ms = kernel_ms;
  res->probe_kernel_launches = mat && total > 0 ? 2 : 1;
  res->rows_written = out.cols[0] ? total : 0;
  res->path = open_addr ? AMJ_PATH_OPEN_ADDR : AMJ_PATH_TPT;
  res->passes = mat && total > 0 ? 2 : 1;
  return AMJ_OK;
}

int execute_plan(amj_ctx_t* ctx, Plan& plan, const std::vector<uint32_t>& attr_ids,
                 const amj_opts_t* opts, amj_result_t* res) {
  amj_opts_t def;
  amj_default_opts(&def);
  const amj_opts_t* o = opts ? opts : &def;
  std::memset(res, 0, sizeof(*res));
  res->num_cols = (uint32_t)plan.n_attrs;
  for (int a = 0; a < plan.n_attrs; ++a) res->attr_ids[a] = attr_ids[a];
  if (o->engine == AMJ_ENGINE_TPT || o->engine == AMJ_ENGINE_OPEN_ADDR) {
    if (o->ooc || plan.probe_on_host || o->ws || o->dro || o->fib) {
      set_error(ctx, AMJ_ERR_ARG,
                "engines AMJ_ENGINE_TPT and AMJ_ENGINE_OPEN_ADDR run alone: clear ooc, ws, dro "
                "and fib");
      return AMJ_ERR_ARG;
    }
    return run_tpt(ctx, plan, o, res);
  }
  if (o->engine == AMJ_ENGINE_PW) {
    if (o->ooc || plan.probe_on_host || o->ws || o->dro || o->fib) {
      set_error(ctx, AMJ_ERR_ARG, "engine AMJ_ENGINE_PW runs alone: clear ooc, ws, dro and fib");
      return AMJ_ERR_ARG;
    }
    return run_pw(ctx, plan, o, res);
  }
  if (o->engine != AMJ_ENGINE_WAVE) {
    set_error(ctx, AMJ_ERR_ARG, "unknown engine " + std::to_string(o->engine));
    return AMJ_ERR_ARG;
  }
  if (o->ooc || plan.probe_on_host) return run_ooc(ctx, plan, o, res);
  if (o->fib && fib_eligible(plan) && window_plan_supported(plan)) {
    const int rc = run_fib(ctx, plan, o, res);
    if (rc != AMJ_ERR_STATE) return rc;
    std::memset(res, 0, sizeof(*res));  // extreme skew: fall back to the count/write path
    res->num_cols = (uint32_t)plan.n_attrs;
    for (int a = 0; a < plan.n_attrs; ++a) res->attr_ids[a] = attr_ids[a];
  }
  if (o->ws || o->dro) return run_dynamic(ctx, plan, o, res);
  return run_plain(ctx, plan, o, res);
}

}  // namespace amj
