// Autotuner: picks the engine options of a query by timing candidates on a sample of its probe
// relation, and remembers the choice per query shape in the context.
//
// Knobs (the reference fixes all of them at compile time, MHJ-GPU/conf.h): the engine (count /
// write unit engine, persistent ws engine, dro), the ws task size (how many waves share a heavy
// key's run) and claim size (units a wave takes at once), and for frequent-item chains the
// output-window size (rows per wave step) and the suffix-table mode. Candidates run with the
// caller's materialise setting on two evenly strided samples of the probe relation (at most 2 M
// and 4 M rows); the device time (best of two runs on each) extrapolated to the whole relation
// decides, and the pick is remembered per query shape. The hash tables'
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
  // Two evenly strided samples of the probe relation, of at most 2 M and 4 M rows (the
  // relation itself when it holds at most 4 M rows). A candidate's device time is taken as
  // affine in the probe rows: costs that scale with the build tables (the suffix tables of the
  // frequent-item path) plus a cost per probe row. Timed on both samples, it is extrapolated
  // to the whole relation, so a fixed cost counts once, as it does in the real run.
  constexpr uint64_t kLarge = 4ull << 20;
  const uint64_t n = probe->rows;
  const bool two = n > kLarge;
  const uint64_t s_large = two ? (n + kLarge - 1) / kLarge : 1, s_small = 2 * s_large;
  const uint64_t n_large = (n + s_large - 1) / s_large, n_small = (n + s_small - 1) / s_small;
  amj_relation large = *probe, small = *probe;
  large.owned = small.owned = false;
  if (two) {
    void* buf = nullptr;
    AMJ_TRY(scratch(ctx, SLOT_TUNE_SAMPLE, (n_large + n_small) * 4 * probe->num_cols, &buf));
    int32_t* dst = (int32_t*)buf;
    for (uint32_t c = 0; c < probe->num_cols; ++c) {
      hipLaunchKernelGGL(gather_stride_kernel, dim3(1024), dim3(256), 0, ctx->stream,
                         probe->cols[c], n_large, s_large, dst);
      large.cols[c] = dst;
      dst += n_large;
      hipLaunchKernelGGL(gather_stride_kernel, dim3(1024), dim3(256), 0, ctx->stream,
                         probe->cols[c], n_small, s_small, dst);
      small.cols[c] = dst;
      dst += n_small;
    }
    large.rows = n_large;
    small.rows = n_small;
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
  // best of two runs (the first may also allocate scratch)
  auto time_on = [&](const amj_relation& rel, const amj_opts_t& c) {
    double t = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      amj_result_t r;
      if (amj_execute(ctx, &rel, hts, n_hts, &c, &r) != AMJ_OK) return 1e300;
      t = std::min(t, r.probe_ms);
      amj_result_free(ctx, &r);
    }
    return t;
  };
  double best = 1e300;
  amj_opts_t pick = base;
  for (const auto& c : cands) {
    const double t_large = time_on(large, c);
    if (t_large >= 1e300) continue;
    double t = t_large;
    if (two) {
      const double t_small = time_on(small, c);
      if (t_small >= 1e300) continue;
      t += std::max(0.0, t_large - t_small) * (double)(n - n_large) / (double)(n_large - n_small);
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
