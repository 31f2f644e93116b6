// Autotuner: picks the engine options of a query by timing candidates on a sample of its probe
// relation, and remembers the choice per query shape in the context.
//
// Knobs (the reference fixes all of them at compile time, MHJ-GPU/conf.h): the engine (count /
// write unit engine, persistent ws engine, dro), the ws task size, and for frequent-item chains
// the output-window size and the suffix-table mode. Candidates run on an evenly strided sample
// of at most 4 M probe rows with the caller's materialise setting; the fastest (device time,
// best of two runs) wins.
#include <cstring>

#include "runtime.h"

namespace amj {

__global__ void gather_stride_kernel(const int32_t* __restrict__ src, uint64_t n_out,
                                     uint64_t stride, int32_t* __restrict__ dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_out;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i * stride];
}

}  // namespace amj

using namespace amj;

extern "C" {

int amj_autotune(amj_ctx_t* ctx, const amj_relation_t* probe, amj_hash_table_t* const* hts,
                 uint32_t n_hts, amj_opts_t* opts, double* tuned_ms) {
  if (!ctx || !probe || !hts || !opts || n_hts == 0 || n_hts > AMJ_MAX_BUILD_TABLES)
    return AMJ_ERR_ARG;
  if (probe->host) {
    set_error(ctx, AMJ_ERR_ARG, "autotune: the probe relation must be on the device");
    return AMJ_ERR_ARG;
  }
  std::vector<uint64_t> key{(uint64_t)(uintptr_t)probe, probe->rows,
                            (uint64_t)(opts->materialize != 0)};
  for (uint32_t i = 0; i < n_hts; ++i) key.push_back((uint64_t)(uintptr_t)hts[i]);
  auto& cache = ctx->tuned;
  auto hit = cache.find(key);
  if (hit != cache.end()) {
    const int mat = opts->materialize, cs = opts->checksum;
    *opts = hit->second.o;
    opts->materialize = mat;
    opts->checksum = cs;
    if (tuned_ms) *tuned_ms = hit->second.ms;
    return AMJ_OK;
  }
  // evenly strided sample of the probe relation
  const uint64_t n = probe->rows;
  const uint64_t stride = n > (4ull << 20) ? (n + (4ull << 20) - 1) / (4ull << 20) : 1;
  const uint64_t ns = (n + stride - 1) / stride;
  amj_relation sample = *probe;
  sample.owned = false;
  void* buf = nullptr;
  if (stride > 1) {
    AMJ_TRY(scratch(ctx, SLOT_MISC2, ns * 4 * probe->num_cols, &buf));
    for (uint32_t c = 0; c < probe->num_cols; ++c) {
      int32_t* dst = (int32_t*)buf + c * ns;
      hipLaunchKernelGGL(gather_stride_kernel, dim3(1024), dim3(256), 0, ctx->stream,
                         probe->cols[c], ns, stride, dst);
      sample.cols[c] = dst;
    }
    sample.rows = ns;
    AMJ_HIP_TRY(ctx, hipGetLastError());
  }
  // candidates
  std::vector<amj_opts_t> cands;
  amj_opts_t base = *opts;
  base.ooc = 0;
  auto add = [&](int ws, int dro, int fib, uint64_t window, int suffix, uint32_t split) {
    amj_opts_t c = base;
    c.ws = ws;
    c.dro = dro;
    c.fib = fib;
    c.out_chunk_rows = window;
    c.fib_suffix = suffix;
    c.ws_split_threshold = split;
    cands.push_back(c);
  };
  add(0, 0, 0, 0, 0, 0);                  // unit engine
  add(1, 0, 0, 0, 0, 1024);               // persistent engine, small tasks
  add(1, 0, 0, 0, 0, 8192);               // persistent engine, large tasks
  if (base.materialize) add(0, 1, 0, 0, 0, 0);  // direct output
  {
    ProbeUse pu{};
    pu.rel = &sample;
    for (uint32_t c = 0; c < sample.num_cols; ++c) pu.attr_ids[c] = sample.attr_ids[c];
    std::vector<TableUse> uses;
    for (uint32_t i = 0; i < n_hts; ++i) {
      TableUse u{};
      u.ht = hts[i];
      u.key_attr_id = hts[i]->rel->attr_ids[hts[i]->key_col];
      u.val_attr_id =
          hts[i]->rel->num_cols > 1 ? (int64_t)hts[i]->rel->attr_ids[1 - hts[i]->key_col] : -1;
      uses.push_back(u);
    }
    Plan P;
    std::vector<uint32_t> ids;
    AMJ_TRY(make_plan(ctx, pu, uses, &P, &ids, nullptr));
    if (fib_eligible(P) && window_plan_supported(P))
      for (uint64_t w : {1024ull, 4096ull, 16384ull})
        for (int sfx : {0, 1}) add(1, 1, 1, w, sfx, 0);
  }
  double best = 1e300;
  amj_opts_t pick = base;
  for (const auto& c : cands) {
    double t = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      amj_result_t r;
      const int rc = amj_execute(ctx, &sample, hts, n_hts, &c, &r);
      if (rc != AMJ_OK) {
        t = 1e300;
        break;
      }
      t = std::min(t, r.probe_ms);
      amj_result_free(ctx, &r);
    }
    if (t < best) {
      best = t;
      pick = c;
    }
  }
  if (best >= 1e300) {
    set_error(ctx, AMJ_ERR_STATE, "autotune: no candidate ran");
    return AMJ_ERR_STATE;
  }
  pick.materialize = opts->materialize;
  pick.checksum = opts->checksum;
  *opts = pick;
  cache[key] = amj_ctx::Tuned{pick, best};
  if (tuned_ms) *tuned_ms = best;
  return AMJ_OK;
}

void amj_autotune_forget(amj_ctx_t* ctx) {
  if (ctx) ctx->tuned.clear();
}

}  // extern "C"
