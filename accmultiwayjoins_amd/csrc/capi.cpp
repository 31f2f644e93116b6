// extern "C" entry points of libamj (declared in include/amj.h).
#include <cstring>

#include "runtime.h"

using namespace amj;

namespace amj {

int make_pattern_plan(amj_ctx_t* ctx, const amj_relation* R, const amj_hash_table* H, int query,
                      Plan* plan, std::vector<uint32_t>* ids) {
  // exec_AMHJ.cu:42 A(0,1),B(1,2),C(0,2) and :59 A(0,1),B(1,2),C(2,3),D(0,3),E(0,2),F(1,3)
  // (make_plan moves each compare-only table to the first level where both its attributes
  // are bound, so FOUR checks E(0,2) right after B binds attribute 2)
  static const uint32_t tri[][2] = {{0, 1}, {1, 2}, {0, 2}};
  static const uint32_t four[][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {0, 2}, {1, 3}};
  const uint32_t(*q)[2] = query == AMJ_QUERY_TRI ? tri : four;
  const uint32_t nq = query == AMJ_QUERY_TRI ? 3 : 6;
  if (query != AMJ_QUERY_TRI && query != AMJ_QUERY_FOUR) {
    set_error(ctx, AMJ_ERR_ARG, "pattern query must be TRI or FOUR");
    return AMJ_ERR_ARG;
  }
  if (!R || !H || R->num_cols != 2 || H->rel != R || H->key_col != 0 || !H->vals_sorted) {
    set_error(ctx, AMJ_ERR_ARG,
              "TRI/FOUR need the edge relation's table on column 0 built with sort_vals=1");
    return AMJ_ERR_ARG;
  }
  ProbeUse pu{};
  pu.rel = R;
  pu.attr_ids[0] = q[0][0];
  pu.attr_ids[1] = q[0][1];
  std::vector<TableUse> uses;
  for (uint32_t i = 1; i < nq; ++i) {
    TableUse u{};
    u.ht = H;
    u.key_attr_id = q[i][0];
    u.val_attr_id = q[i][1];
    uses.push_back(u);
  }
  return make_plan(ctx, pu, uses, plan, ids, nullptr);
}

}  // namespace amj

static TableUse use_of(const amj_hash_table* ht) {
  TableUse u{};
  u.ht = ht;
  u.key_attr_id = ht->rel->attr_ids[ht->key_col];
  u.val_attr_id = ht->rel->num_cols > 1 ? (int64_t)ht->rel->attr_ids[1 - ht->key_col] : -1;
  return u;
}

namespace amj {

int make_execute_plan(amj_ctx_t* ctx, const amj_relation* probe, amj_hash_table* const* hts,
                      uint32_t n_hts, Plan* plan, std::vector<uint32_t>* ids) {
  ProbeUse pu{};
  pu.rel = probe;
  for (uint32_t c = 0; c < probe->num_cols; ++c) pu.attr_ids[c] = probe->attr_ids[c];
  std::vector<TableUse> uses;
  for (uint32_t i = 0; i < n_hts; ++i) {
    if (!hts[i]) {
      set_error(ctx, AMJ_ERR_ARG, "a hash table pointer is NULL");
      return AMJ_ERR_ARG;
    }
    uses.push_back(use_of(hts[i]));
  }
  return make_plan(ctx, pu, uses, plan, ids, nullptr);
}

}  // namespace amj

extern "C" {

const char* amj_version(void) { return "amj 0.1 (gfx950)"; }

void amj_default_opts(amj_opts_t* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->materialize = 1;
  o->checksum = 1;
}

int amj_ctx_create(int device, amj_ctx_t** out) {
  if (!out) return AMJ_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return AMJ_ERR_NODEV;
  }
  if (device < 0 || device >= n) return AMJ_ERR_ARG;
  std::unique_ptr<amj_ctx> ctx(new amj_ctx());
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) return AMJ_ERR_HIP;
  if (hipGetDeviceProperties(&ctx->props, device) != hipSuccess) return AMJ_ERR_HIP;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
    return AMJ_ERR_HIP;
  if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess)
    return AMJ_ERR_HIP;
  for (auto& e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) return AMJ_ERR_HIP;
  *out = ctx.release();
  return AMJ_OK;
}

void amj_ctx_destroy(amj_ctx_t* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  for (auto& kv : ctx->scratch) dfree(ctx, kv.second.p, kv.second.bytes);
  pool_clear(ctx);
  if (ctx->arena.base) dfree(ctx, ctx->arena.base, ctx->arena.size);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  delete ctx;
}

const char* amj_last_error(const amj_ctx_t* ctx) {
  return ctx ? ctx->last_error.c_str() : "no context";
}

int amj_host_alloc(size_t bytes, void** out) {
  if (!out) return AMJ_ERR_ARG;
  *out = nullptr;
  if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return AMJ_ERR_NOMEM;
  }
  return AMJ_OK;
}

void amj_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

int amj_memory_info(amj_ctx_t* ctx, uint64_t* allocated, uint64_t* peak) {
  if (!ctx) return AMJ_ERR_ARG;
  if (allocated) *allocated = ctx->allocated;
  if (peak) *peak = ctx->peak;
  return AMJ_OK;
}

int amj_sync(amj_ctx_t* ctx) {
  if (!ctx) return AMJ_ERR_ARG;
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return AMJ_OK;
}

int amj_device_info(amj_ctx_t* ctx, char* name, size_t name_len, uint64_t* hbm_bytes,
                    int* cu_count) {
  if (!ctx) return AMJ_ERR_ARG;
  if (name && name_len) {
    std::snprintf(name, name_len, "%s (%s)", ctx->props.name, ctx->props.gcnArchName);
  }
  if (hbm_bytes) *hbm_bytes = ctx->props.totalGlobalMem;
  if (cu_count) *cu_count = ctx->props.multiProcessorCount;
  return AMJ_OK;
}

int amj_relation_create(amj_ctx_t* ctx, const amj_rel_t* r, amj_relation_t** out) {
  if (!ctx || !r || !out) return AMJ_ERR_ARG;
  *out = nullptr;
  if (r->num_cols < 1 || r->num_cols > AMJ_MAX_PROBE_COLS || !r->cols) {
    set_error(ctx, AMJ_ERR_ARG, "relations must have 1 or 2 int32 columns");
    return AMJ_ERR_ARG;
  }
  std::unique_ptr<amj_relation> rel(new amj_relation());
  rel->num_cols = r->num_cols;
  rel->rows = r->rows;
  if (r->on_device == AMJ_REL_HOST) {  // stays in (preferably pinned) host memory
    for (uint32_t c = 0; c < r->num_cols; ++c) {
      rel->attr_ids[c] = r->attr_ids ? r->attr_ids[c] : c;
      rel->cols[c] = const_cast<int32_t*>(r->cols[c]);
    }
    rel->host = true;
    *out = rel.release();
    return AMJ_OK;
  }
  for (uint32_t c = 0; c < r->num_cols; ++c) {
    rel->attr_ids[c] = r->attr_ids ? r->attr_ids[c] : c;
    if (r->on_device) {
      rel->cols[c] = const_cast<int32_t*>(r->cols[c]);
    } else {
      AMJ_TRY(dmalloc(ctx, (void**)&rel->cols[c], std::max<uint64_t>(r->rows, 1) * 4));
      if (r->rows)
        AMJ_HIP_TRY(ctx, hipMemcpyAsync(rel->cols[c], r->cols[c], r->rows * 4,
                                        hipMemcpyHostToDevice, ctx->stream));
    }
  }
  rel->owned = !r->on_device;
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  *out = rel.release();
  return AMJ_OK;
}

void amj_relation_destroy(amj_ctx_t* ctx, amj_relation_t* rel) {
  if (!rel) return;
  if (rel->owned) {
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    for (uint32_t c = 0; c < rel->num_cols; ++c)
      dfree(ctx, rel->cols[c], std::max<uint64_t>(rel->rows, 1) * 4);
  }
  delete rel;
}

int amj_hash_table_build(amj_ctx_t* ctx, const amj_relation_t* rel, uint32_t key_col,
                         int sort_vals, amj_hash_table_t** out) {
  if (!ctx || !out) return AMJ_ERR_ARG;
  return build_hash_table(ctx, rel, key_col, sort_vals, out);
}

void amj_hash_table_destroy(amj_ctx_t* ctx, amj_hash_table_t* ht) {
  if (!ht) return;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  const uint64_t n = std::max<uint64_t>(ht->dev.view.n, 1);
  dfree(ctx, ht->dev.ptrs, (ht->dev.buckets + 1) * 4);
  dfree(ctx, ht->dev.keys, n * 4);
  dfree(ctx, ht->dev.vals, n * 4);
  delete ht;
}

int amj_hash_table_info(amj_ctx_t* ctx, const amj_hash_table_t* ht, uint64_t* rows,
                        uint64_t* buckets, int* identity, double* build_ms) {
  if (!ctx || !ht) return AMJ_ERR_ARG;
  if (rows) *rows = ht->dev.view.n;
  if (buckets) *buckets = ht->dev.buckets;
  if (identity) *identity = (int)ht->dev.view.identity;
  if (build_ms) *build_ms = ht->build_ms;
  return AMJ_OK;
}



int amj_execute(amj_ctx_t* ctx, const amj_relation_t* probe, amj_hash_table_t* const* hts,
                uint32_t n_hts, const amj_opts_t* opts, amj_result_t* res) {
  if (!ctx || !probe || !hts || !res) return AMJ_ERR_ARG;
  Plan plan;
  std::vector<uint32_t> ids;
  AMJ_TRY(make_execute_plan(ctx, probe, hts, n_hts, &plan, &ids));
  return execute_plan(ctx, plan, ids, opts, res);
}

int amj_chain_join(amj_ctx_t* ctx, const amj_rel_t* rels, uint32_t n, const amj_opts_t* opts,
                   amj_result_t* res) {
  if (!ctx || !rels || !res || n < 2) {
    if (ctx) set_error(ctx, AMJ_ERR_ARG, "chain join needs at least two relations");
    return AMJ_ERR_ARG;
  }
  std::memset(res, 0, sizeof(*res));
  std::vector<amj_relation_t*> R(n, nullptr);
  std::vector<amj_hash_table_t*> H(n - 1, nullptr);
  int rc = AMJ_OK;
  double build_ms = 0.0;
  for (uint32_t i = 0; i < n && rc == AMJ_OK; ++i) rc = amj_relation_create(ctx, &rels[i], &R[i]);
  // a build table needs value-sorted runs when its second attribute is bound earlier
  std::vector<uint32_t> seen;
  if (rc == AMJ_OK)
    for (uint32_t c = 0; c < R[0]->num_cols; ++c) seen.push_back(R[0]->attr_ids[c]);
  for (uint32_t i = 1; i < n && rc == AMJ_OK; ++i) {
    int sort_vals = 0;
    if (R[i]->num_cols > 1) {
      const uint32_t v = R[i]->attr_ids[1];
      sort_vals = std::find(seen.begin(), seen.end(), v) != seen.end() ? 1 : 0;
    }
    for (uint32_t c = 0; c < R[i]->num_cols; ++c) seen.push_back(R[i]->attr_ids[c]);
    rc = amj_hash_table_build(ctx, R[i], 0, sort_vals, &H[i - 1]);
    if (rc == AMJ_OK) build_ms += H[i - 1]->build_ms;
  }
  if (rc == AMJ_OK) rc = amj_execute(ctx, R[0], H.data(), n - 1, opts, res);
  if (rc == AMJ_OK) res->build_ms = build_ms;
  for (auto* h : H) amj_hash_table_destroy(ctx, h);
  for (auto* r : R) amj_relation_destroy(ctx, r);
  return rc;
}

int amj_pattern_join(amj_ctx_t* ctx, const amj_rel_t* edges, int query, const amj_opts_t* opts,
                     amj_result_t* res) {
  if (!ctx || !edges || !res) return AMJ_ERR_ARG;
  if (edges->num_cols != 2) {
    set_error(ctx, AMJ_ERR_ARG, "TRI/FOUR need a two-column edge relation");
    return AMJ_ERR_ARG;
  }
  std::memset(res, 0, sizeof(*res));
  if (query != AMJ_QUERY_TRI && query != AMJ_QUERY_FOUR) {
    set_error(ctx, AMJ_ERR_ARG, "pattern query must be TRI or FOUR");
    return AMJ_ERR_ARG;
  }
  amj_rel_t e = *edges;
  const uint32_t ids01[2] = {0, 1};
  e.attr_ids = ids01;
  amj_relation_t* R = nullptr;
  amj_hash_table_t* H = nullptr;
  int rc = amj_relation_create(ctx, &e, &R);
  // one table on the first column, aliased at every level (QueryProcessor.cuh:112-169)
  if (rc == AMJ_OK) rc = amj_hash_table_build(ctx, R, 0, 1, &H);
  if (rc == AMJ_OK) {
    Plan plan;
    std::vector<uint32_t> ids;
    rc = make_pattern_plan(ctx, R, H, query, &plan, &ids);
    if (rc == AMJ_OK) rc = execute_plan(ctx, plan, ids, opts, res);
    if (rc == AMJ_OK) res->build_ms = H->build_ms;
  }
  amj_hash_table_destroy(ctx, H);
  amj_relation_destroy(ctx, R);
  return rc;
}

int amj_execute_pattern(amj_ctx_t* ctx, const amj_relation_t* edges, const amj_hash_table_t* ht,
                        int query, const amj_opts_t* opts, amj_result_t* res) {
  if (!ctx || !edges || !ht || !res) return AMJ_ERR_ARG;
  std::memset(res, 0, sizeof(*res));
  Plan plan;
  std::vector<uint32_t> ids;
  AMJ_TRY(make_pattern_plan(ctx, edges, ht, query, &plan, &ids));
  return execute_plan(ctx, plan, ids, opts, res);
}

int amj_result_free(amj_ctx_t* ctx, amj_result_t* res) {
  if (!ctx || !res) return AMJ_ERR_ARG;
  if (res->reserved[2] == 1 && res->cols[0]) {
    pool_put(ctx, res->cols[0], (size_t)res->reserved[0] | ((size_t)res->reserved[1] << 32));
  } else if (res->reserved[2] == 2 && res->cols[0]) {
    arena_release(ctx, res->cols[0], (uint64_t)res->reserved[0] | ((uint64_t)res->reserved[1] << 32));
  }
  std::memset(res->cols, 0, sizeof(res->cols));
  res->reserved[2] = 0;
  res->rows_written = 0;
  return AMJ_OK;
}

int amj_result_copy_col(amj_ctx_t* ctx, const amj_result_t* res, uint32_t col, uint64_t row0,
                        uint64_t n, int32_t* host_out) {
  if (!ctx || !res || !host_out || col >= res->num_cols || !res->cols[col]) return AMJ_ERR_ARG;
  if (row0 + n > res->rows_written) {
    set_error(ctx, AMJ_ERR_ARG, "copy range beyond the materialised rows");
    return AMJ_ERR_ARG;
  }
  if (n == 0) return AMJ_OK;
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(host_out, res->cols[col] + row0, n * 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return AMJ_OK;
}

}  // extern "C"
