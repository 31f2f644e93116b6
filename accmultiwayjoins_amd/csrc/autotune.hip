// Autotuner: picks the engine options of a query by timing candidates on a sample of its probe
// relation, and remembers the choice per query shape in the context.
//
// Knobs (the reference fixes all of them at compile time, MHJ-GPU/conf.h): the engine (count /
// write unit engine, persistent ws engine, dro), the ws task size (how many waves share a heavy
// key's run) and claim size (units a wave takes at once), and for frequent-item chains the
// output-window size (rows per wave step) and the suffix-table mode. Candidates run on an evenly
// strided sample of at most 4 M probe rows with the caller's materialise setting; the fastest
// (device time, best of two runs) wins and is remembered per query shape. The hash tables'
// bucket count stays the reference's rule (B = floorPow2(rows), QueryProcessor.cuh:86), which
// the tables were built with before the query.
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
  Plan full;
  std::vector<uint32_t> full_ids;
  AMJ_TRY(make_execute_plan(ctx, probe, hts, n_hts, &full, &full_ids));
  // the query shape: depth, attributes, and per level the size class of the table, its bucket
  // count, whether it is an identity table and what the level does with the second column,
  // with the probe's size class; queries of one shape share a choice (relations freed and
  // reallocated at the same address are not mistaken for the ones tuned before)
  auto log2c = [](uint64_t x) {
    uint64_t b = 0;
    while ((1ull << b) < x && b < 63) ++b;
    return b;
  };
  std::vector<uint64_t> key{(uint64_t)(opts->materialize != 0), (uint64_t)full.n_levels,
                            (uint64_t)full.n_attrs, log2c(full.probe_n)};
  for (int t = 0; t < full.n_levels; ++t) {
    key.push_back(log2c(full.tab[t].n));
    key.push_back(full.tab[t].log_b);
    key.push_back(full.tab[t].identity ? 1 : 0);
    key.push_back((uint64_t)full.val_mode[t]);
  }
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
  // candidates: the engine, the persistent engine's task size (waves per heavy key) and claim
  // size (units per claim), direct output, and for frequent-item chains the output window
  // (rows per wave step) and the suffix tables
  std::vector<amj_opts_t> cands;
  amj_opts_t base = *opts;
  base.ooc = 0;
  auto add = [&](int ws, int dro, int fib, uint64_t window, int suffix, uint32_t split,
                 uint32_t claim) {
    amj_opts_t c = base;
    c.ws = ws;
    c.dro = dro;
    c.fib = fib;
    c.out_chunk_rows = window;
    c.fib_suffix = suffix;
    c.ws_split_threshold = split;
    c.ws_claim_units = claim;
    cands.push_back(c);
  };
  add(0, 0, 0, 0, 0, 0, 0);                    // unit engine
  for (uint32_t split : {1024u, 8192u})
    for (uint32_t claim : {4u, 64u}) add(1, 0, 0, 0, 0, split, claim);  // persistent engine
  if (base.materialize) add(0, 1, 0, 0, 0, 0, 0);  // direct output
  if (fib_eligible(full) && window_plan_supported(full))
    for (uint64_t w : {1024ull, 4096ull, 16384ull})
      for (int sfx : {0, 1}) add(1, 1, 1, w, sfx, 0, 0);
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
