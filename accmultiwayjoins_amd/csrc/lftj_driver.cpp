// Host driver of the leapfrog triejoin: trie objects, query plans over them, and the DFS / BFS
// executions behind the alftj_* entry points.
//
// Reference: exec_LFTJ.cu:62-254 (query_processing: load or build the tries, SELF mode aliasing
// trie 0 for the triangle, attr_order 0..n-1) and ALFTJ::evaluate (ALFTJ.cuh:1112-1304: build
// the attribute -> relation map from the 2-level tries, pick the smallest level-0 relation,
// count, allocate, write).
#include <algorithm>
#include <cstring>

#include "lftj.h"
#include "probe_common.h"
#include "runtime.h"
#include "ws_queue.h"

namespace amj {

struct LUse {
  const amj_trie* t;
  uint32_t x_id, y_id;
};

// Attribute i of the plan is attr_order[i]; relation r joins positions x[r] < y[r].
static int make_lplan(amj_ctx_t* ctx, const std::vector<LUse>& uses, const uint32_t* order,
                      uint32_t na, LPlan* P) {
  std::memset(P, 0, sizeof(*P));
  if (na < 2 || na > (uint32_t)kMaxLftjAttrs) {
    set_error(ctx, AMJ_ERR_ARG, "LFTJ queries need 2..4 attributes");
    return AMJ_ERR_ARG;
  }
  if (uses.empty() || uses.size() > (size_t)kMaxLftjRels) {
    set_error(ctx, AMJ_ERR_ARG, "LFTJ queries need 1..6 relations");
    return AMJ_ERR_ARG;
  }
  auto pos_of = [&](uint32_t id) -> int {
    for (uint32_t i = 0; i < na; ++i)
      if (order[i] == id) return (int)i;
    return -1;
  };
  for (uint32_t i = 0; i < na; ++i)
    for (uint32_t j = i + 1; j < na; ++j)
      if (order[i] == order[j]) {
        set_error(ctx, AMJ_ERR_ARG, "attr_order repeats an attribute");
        return AMJ_ERR_ARG;
      }
  P->n_attrs = (int32_t)na;
  P->n_rels = (int32_t)uses.size();
  P->off32 = 1;
  uint64_t widest = 0;  // longest candidate list any frontier entry can take
  for (size_t r = 0; r < uses.size(); ++r) {
    const int px = pos_of(uses[r].x_id), py = pos_of(uses[r].y_id);
    if (px < 0 || py < 0) {
      set_error(ctx, AMJ_ERR_ARG, "a trie attribute is not in attr_order");
      return AMJ_ERR_ARG;
    }
    if (px >= py) {
      set_error(ctx, AMJ_ERR_ARG,
                "trie levels must follow attr_order (build the trie with its columns swapped)");
      return AMJ_ERR_ARG;
    }
    P->rel[r] = uses[r].t->dev.view;
    P->x[r] = px;
    P->y[r] = py;
    P->root_mask[px] |= 1u << r;
    P->child_mask[py] |= 1u << r;
    widest = std::max<uint64_t>(widest, uses[r].t->dev.max_run);
    const TrieDev& d = uses[r].t->dev;
    if ((uint64_t)d.n0 + 1 >= (1ull << 30) || d.n1 >= (1ull << 30) || d.vrange_n >= (1u << 29))
      P->off32 = 0;
  }
  if (P->root_mask[0] == 0u) {
    set_error(ctx, AMJ_ERR_ARG, "no relation starts at the first attribute of attr_order");
    return AMJ_ERR_ARG;
  }
  for (uint32_t i = 1; i < na; ++i) {
    if ((P->root_mask[i] | P->child_mask[i]) == 0u) {
      set_error(ctx, AMJ_ERR_ARG, "an attribute of attr_order is in no relation");
      return AMJ_ERR_ARG;
    }
    if (P->child_mask[i] == 0u)  // candidates come from a level-0 list
      for (size_t r = 0; r < uses.size(); ++r)
        if ((P->root_mask[i] >> r) & 1u) widest = std::max<uint64_t>(widest, P->rel[r].n0);
  }
  // a wave's frontier level is indexed with 32 bits: 64 lists of at most 2^26 entries
  if (widest >= (1ull << 26)) {
    set_error(ctx, AMJ_ERR_ARG, "a candidate list exceeds 2^26 entries");
    return AMJ_ERR_ARG;
  }
  // level 0 enumerates the smallest level-0 list among the relations rooted at attribute 0
  // (ALFTJ.cuh:1188-1196)
  int best = -1;
  for (size_t r = 0; r < uses.size(); ++r)
    if (((P->root_mask[0] >> r) & 1u) && (best < 0 || P->rel[r].n0 < P->rel[best].n0))
      best = (int)r;
  P->driver0 = best;
  P->roots = P->rel[best].n0;
  return AMJ_OK;
}

constexpr uint64_t kLftjSampleRoots = 65536;

// DFS with direct output (opts.dro, materialising): the output is sized from the counts of an
// evenly strided sample of at most kLftjSampleRoots roots (x1.5 + 64 Ki rows; exact when every
// root is sampled, or opts.dro_capacity_rows), one pass writes through per-wave LDS buffers with one
// atomic reservation per 256 rows, and a pass that overran its output runs once more into an
// exact one. Count and checksum are those of the count/write engine; the row order is not.
// Launch order of the DFS units, heaviest first, when there are more units than wave slots
// (otherwise every unit starts at once and the order cannot shorten the tail): a unit whose
// roots open long candidate lists otherwise starts late and runs alone at the end.
constexpr uint64_t kLftjOrderMinUnits = 2048;

static int unit_order(amj_ctx_t* ctx, const LPlan& P, const uint64_t** order) {
  *order = nullptr;
  const uint64_t units = lftj_units(P);
  if (units < kLftjOrderMinUnits) return AMJ_OK;
  hipStream_t s = ctx->stream;
  void *keys = nullptr, *alt = nullptr, *tmp = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_SORT_KEYS, units * 8, &keys));
  AMJ_TRY(scratch(ctx, SLOT_SORT_ALT, units * 8, &alt));
  AMJ_TRY(scratch(ctx, SLOT_SORT_TEMP, radix_sort_temp_bytes(units), &tmp));
  AMJ_HIP_TRY(ctx, launch_lftj_unit_weights(P, (uint64_t*)keys, s));
  // the weight bytes only: the sort is stable, so equal weights keep unit order
  AMJ_HIP_TRY(ctx, radix_sort_u64((uint64_t*)keys, (uint64_t*)alt, units, 0xf0u, tmp, s));
  *order = (const uint64_t*)keys;
  return AMJ_OK;
}

// Output size of a direct-output pass from a strided sample of kLftjSampleRoots roots, each
// root's candidates for attribute 1 counted in windows of kLftjSampleWindow by one wave each (a
// launch lasts as long as its slowest wave: counted whole, 4 roots per wave, the heaviest
// sampled root of R-MAT 22 held the triangle sample at 6.3 ms; in windows of 256 candidates
// the direct path took 3.7 ms less, of 64 another 1.25 ms less). x1.5 + 64 Ki rows; exact when
// every root is sampled.
static int sample_capacity(amj_ctx_t* ctx, const LPlan& P, uint64_t* capacity) {
  hipStream_t s = ctx->stream;
  const uint64_t roots = P.roots;
  *capacity = 0;
  if (roots == 0) return AMJ_OK;
  const uint64_t sr = std::min<uint64_t>(roots, kLftjSampleRoots);
  const uint64_t stride = roots / sr;
  void *woff = nullptr, *counts = nullptr, *tmp = nullptr, *sum = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_LFTJ_SAMPLE, (sr + 1) * 8, &woff));
  AMJ_TRY(scratch(ctx, SLOT_SUM, 64, &sum));
  AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(sr + 1) * 8, &tmp));
  AMJ_HIP_TRY(ctx, launch_lftj_sample_widths(P, (uint64_t*)woff, sr, stride, stride / 2, s));
  AMJ_HIP_TRY(ctx, exclusive_scan_u64((const uint64_t*)woff, (uint64_t*)woff, sr + 1,
                                      (uint64_t*)tmp, s));
  uint64_t nw = 0;
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(&nw, (const uint64_t*)woff + sr, 8, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  uint64_t sampled = 0;
  if (nw > 0) {
    AMJ_TRY(scratch(ctx, SLOT_UNIT_COUNTS, nw * 8, &counts));
    AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(nw) * 8, &tmp));
    AMJ_HIP_TRY(ctx, launch_lftj_sample_windows(P, (const uint64_t*)woff, sr, nw, stride,
                                                stride / 2, (uint64_t*)counts, s));
    AMJ_HIP_TRY(ctx, reduce_sum_u64((const uint64_t*)counts, nw, (uint64_t*)sum,
                                    (uint64_t*)tmp, s));
    AMJ_HIP_TRY(ctx, hipMemcpyAsync(&sampled, sum, 8, hipMemcpyDeviceToHost, s));
    AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  }
  *capacity = sr == roots
                  ? sampled
                  : (uint64_t)((double)sampled * (double)roots / (double)sr * 1.5) + 65536;
  return AMJ_OK;
}

static int run_lftj_direct(amj_ctx_t* ctx, const LPlan& P, const amj_opts_t* o,
                           amj_result_t* res) {
  hipStream_t s = ctx->stream;
  const uint64_t units = lftj_units(P);
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  uint64_t expected = o->dro_capacity_rows;
  uint32_t passes = 0;
  bool slack = false;  // a sampled size gets room for the unused chunk tails
  if (expected == 0 && units > 0) {
    AMJ_TRY(sample_capacity(ctx, P, &expected));
    ++passes;
    slack = true;
  }
  const uint64_t* order = nullptr;
  AMJ_TRY(unit_order(ctx, P, &order));
  const uint64_t nb = units == 0 ? 1 : (units + kWavesPerBlock - 1) / kWavesPerBlock;
  void *bsum = nullptr, *bcnt = nullptr, *ctl = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_BLOCK_SUMS, nb * 8, &bsum));
  AMJ_TRY(scratch(ctx, SLOT_MISC1, nb * 8, &bcnt));
  AMJ_TRY(scratch(ctx, SLOT_FLAG, 256, &ctl));
  const bool hash = o->checksum != 0;
  uint64_t reserved = 0, n = 0, nb_used = 0, capacity = 0;
  uint32_t chunk = 256;
  double kms = 0.0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    chunk = std::max<uint32_t>(direct_chunk_rows(expected, units), 256);
    const uint64_t want = expected + (slack ? units * (uint64_t)chunk : 0);
    capacity = (want + chunk - 1) / chunk * chunk;
    const uint64_t n_chunks = capacity / chunk;
    void *tail = nullptr, *tpre = nullptr, *stmp = nullptr;
    AMJ_TRY(scratch(ctx, SLOT_MISC2, std::max<uint64_t>(n_chunks, 1) * 8, &tail));
    LOut out{};
    if (capacity > 0) {
      if (capacity * (uint64_t)P.n_attrs * 4 > budget_left(ctx, o)) {
        set_error(ctx, AMJ_ERR_NOMEM, "result of " + std::to_string(capacity) +
                                          " tuples does not fit the HBM budget (use "
                                          "materialize=0)");
        return AMJ_ERR_NOMEM;
      }
      AMJ_TRY(alloc_output(ctx, capacity, P.n_attrs, res));
      for (int a = 0; a < P.n_attrs; ++a) out.c[a] = res->cols[a];
      AMJ_HIP_TRY(ctx, hipMemsetAsync(tail, 0, n_chunks * 8, s));
    }
    AMJ_HIP_TRY(ctx, hipMemsetAsync(ctl, 0, 8, s));
    AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
    AMJ_HIP_TRY(ctx, launch_lftj_direct(P, out, capacity, chunk, hash, (unsigned long long*)ctl,
                                        (uint64_t*)tail, (uint64_t*)bsum, (uint64_t*)bcnt,
                                        &nb_used, order, s));
    AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
    AMJ_HIP_TRY(ctx, hipMemcpyAsync(&reserved, ctl, 8, hipMemcpyDeviceToHost, s));
    AMJ_TRY(finish_checksum(ctx, (const uint64_t*)bcnt, nb_used, &n));
    kms += elapsed_ms(ctx->ev[3], ctx->ev[4]);
    ++passes;
    if (reserved <= capacity) {
      // the results are the first n rows once the rows past n fill the tails below it
      const uint64_t used_chunks = reserved / chunk;
      if (reserved > n) {
        AMJ_TRY(scratch(ctx, SLOT_MISC3, (used_chunks + 1) * 8, &tpre));
        AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(used_chunks) * 8, &stmp));
        AMJ_HIP_TRY(ctx, exclusive_scan_u64((const uint64_t*)tail, (uint64_t*)tpre, used_chunks,
                                            (uint64_t*)stmp, s));
        ProbeOut cols{};
        for (int a = 0; a < P.n_attrs; ++a) cols.cols[a] = out.c[a];
        AMJ_HIP_TRY(ctx, launch_chunk_fill(cols, P.n_attrs, (const uint64_t*)tail,
                                           (const uint64_t*)tpre, used_chunks, chunk, n,
                                           reserved, s));
      }
      break;
    }
    // overran: the pass counted every result; run again with room for them and the tails
    amj_result_free(ctx, res);
    expected = n;
    slack = true;
  }
  if (reserved > capacity) {
    set_error(ctx, AMJ_ERR_STATE, "direct output overran an allocation sized for it");
    return AMJ_ERR_STATE;
  }
  if (hash) AMJ_TRY(finish_checksum(ctx, (const uint64_t*)bsum, nb_used, &res->checksum));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->count = n;
  res->rows_written = n;
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->probe_kernel_ms = kms;
  res->probe_kernel_launches = passes;
  res->passes = passes;
  res->path = AMJ_PATH_LFTJ_DFS;
  return AMJ_OK;
}

// Count + checksum without output: one pass that hashes every result and counts per block.
static int run_lftj_count(amj_ctx_t* ctx, const LPlan& P, amj_result_t* res) {
  hipStream_t s = ctx->stream;
  const uint64_t units = lftj_units(P);
  const uint64_t nb = units == 0 ? 1 : (units + kWavesPerBlock - 1) / kWavesPerBlock;
  void *bsum = nullptr, *bcnt = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_BLOCK_SUMS, nb * 8, &bsum));
  AMJ_TRY(scratch(ctx, SLOT_MISC1, nb * 8, &bcnt));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  const uint64_t* order = nullptr;
  AMJ_TRY(unit_order(ctx, P, &order));
  uint64_t nb_used = 0;
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
  AMJ_HIP_TRY(ctx, launch_lftj_write(P, nullptr, LOut{}, true, (uint64_t*)bsum, &nb_used, order,
                                     s, (uint64_t*)bcnt));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
  AMJ_TRY(finish_checksum(ctx, (const uint64_t*)bsum, nb_used, &res->checksum));
  AMJ_TRY(finish_checksum(ctx, (const uint64_t*)bcnt, nb_used, &res->count));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->probe_kernel_ms = elapsed_ms(ctx->ev[3], ctx->ev[4]);
  res->probe_kernel_launches = 1;
  res->rows_written = 0;
  res->passes = 1;
  res->path = AMJ_PATH_LFTJ_DFS;
  return AMJ_OK;
}

// Work sharing (opts.ws, ALFTJ.cuh:464-499): persistent waves over the units, long candidate
// runs split into tasks of `split` candidates (opts.ws_split_threshold, default the reference's
// BUC_THRESHOLD of 1,024). Without output one pass counts and hashes. Materialising, the
// output is sized by a counting pass (or, with dro, from the sample of the direct-output path)
// and written in one pass with direct output; an overrun runs once more into an exact size.
constexpr uint64_t kLftjSplit = 1024;        // ALFTJ.cuh:22 BUC_THRESHOLD
constexpr uint64_t kLftjTaskSlots = 8192;    // task records of 5,440 bytes
constexpr uint64_t kLftjStallMs = 20000;     // no progress for this long: give up

static int lftj_ws_pass(amj_ctx_t* ctx, const LPlan& P, const LOut& out, int mode,
                        const WsArgs& A, const uint64_t* order, unsigned grid, WsCtl* hc,
                        double* kms, uint32_t chunk = 0, uint64_t* tails = nullptr) {
  hipStream_t s = ctx->stream;
  AMJ_HIP_TRY(ctx, hipMemsetAsync(A.ctl, 0, sizeof(WsCtl), s));
  if (A.split) AMJ_HIP_TRY(ctx, hipMemsetAsync(A.ready, 0, A.cap * 4, s));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
  if (A.units)
    AMJ_HIP_TRY(ctx, launch_lftj_ws(P, out, mode, &A, order, grid, s, chunk, tails));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(hc, A.ctl, sizeof(WsCtl), hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  *kms += elapsed_ms(ctx->ev[3], ctx->ev[4]);
  if (hc->flags & kFlagSpin) {
    set_error(ctx, AMJ_ERR_STATE, "ALFTJ work sharing: no progress within the time limit");
    return AMJ_ERR_STATE;
  }
  return AMJ_OK;
}

static int run_lftj_ws(amj_ctx_t* ctx, const LPlan& P_in, const amj_opts_t* o,
                       amj_result_t* res) {
  hipStream_t s = ctx->stream;
  // units of 64 roots: the waves claim them dynamically and split long runs into tasks, so
  // they need not be small
  LPlan P = P_in;
  P.rpu = kWave;
  const uint64_t units = lftj_units(P);
  WsArgs A{};
  void *ctl = nullptr, *tasks = nullptr, *ready = nullptr;
  A.units = units;
  A.split = o->ws_split_threshold ? (uint64_t)(o->ws_split_threshold + 63) / 64 * 64 : kLftjSplit;
  A.cap = kLftjTaskSlots;
  AMJ_TRY(scratch(ctx, SLOT_DYN_CTL, sizeof(WsCtl), &ctl));
  AMJ_TRY(scratch(ctx, SLOT_DYN_TASKS, A.cap * lftj_ws_task_bytes(), &tasks));
  AMJ_TRY(scratch(ctx, SLOT_DYN_READY, A.cap * 4, &ready));
  A.ctl = (WsCtl*)ctl;
  A.tasks = (uint32_t*)tasks;
  A.ready = (uint32_t*)ready;
  int clock_khz = 0;
  if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, ctx->device) !=
          hipSuccess || clock_khz <= 0)
    clock_khz = 100000;
  A.stall_ticks = (o->ws_stall_ms ? (uint64_t)o->ws_stall_ms : kLftjStallMs) * (uint64_t)clock_khz;
  A.max_claim = o->ws_claim_units;
  const bool mat = o->materialize != 0;
  const bool hash = o->checksum || !mat;
  const int final_mode = mat ? (hash ? MODE_WRITE_HASH : MODE_WRITE) : MODE_HASH;
  // persistent grid: the blocks the final pass's kernel fits on the device, at most one wave
  // per unit
  const uint64_t fit = (uint64_t)lftj_ws_blocks_per_cu(P, final_mode) *
                       (uint64_t)ctx->props.multiProcessorCount;
  const unsigned grid = (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>((units + kWavesPerBlock - 1) / kWavesPerBlock, fit));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  const uint64_t* order = nullptr;
  AMJ_TRY(unit_order(ctx, P, &order));
  WsCtl hc{};
  double kms = 0.0;
  uint32_t passes = 0;
  uint64_t pushed = 0;
  if (!mat) {
    AMJ_TRY(lftj_ws_pass(ctx, P, LOut{}, MODE_HASH, A, order, grid, &hc, &kms));
    ++passes;
    pushed = hc.pushed;
    res->count = hc.count;
    res->checksum = hc.csum;
  } else {
    uint64_t capacity = o->dro_capacity_rows;
    if (capacity == 0 && o->dro) {
      AMJ_TRY(sample_capacity(ctx, P, &capacity));
    } else if (capacity == 0) {  // counting pass, then an exact allocation
      AMJ_TRY(lftj_ws_pass(ctx, P, LOut{}, MODE_COUNT, A, order, grid, &hc, &kms));
      ++passes;
      pushed += hc.pushed;
      capacity = hc.count;
    }
    // large outputs are reserved in chunks: room for the unused tail of every wave's last
    // chunk, then the fill-up pass (as in run_lftj_direct)
    const uint64_t waves = (uint64_t)grid * kWavesPerBlock;
    bool user_sized = o->dro_capacity_rows != 0;
    uint64_t rows = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
      LOut out{};
      const uint32_t chunk = direct_chunk_rows(capacity, waves);
      uint64_t alloc = capacity;
      if (chunk) {
        alloc = capacity + (user_sized ? 0 : waves * (uint64_t)chunk);
        alloc = (alloc + chunk - 1) / chunk * chunk;
      }
      void* tails = nullptr;
      if (alloc > 0) {
        if (alloc * (uint64_t)P.n_attrs * 4 > budget_left(ctx, o)) {
          set_error(ctx, AMJ_ERR_NOMEM, "result of " + std::to_string(alloc) +
                                            " tuples does not fit the HBM budget (use "
                                            "materialize=0)");
          return AMJ_ERR_NOMEM;
        }
        AMJ_TRY(alloc_output(ctx, alloc, P.n_attrs, res));
        for (int a = 0; a < P.n_attrs; ++a) out.c[a] = res->cols[a];
        if (chunk) {
          AMJ_TRY(scratch(ctx, SLOT_MISC2, alloc / chunk * 8, &tails));
          AMJ_HIP_TRY(ctx, hipMemsetAsync(tails, 0, alloc / chunk * 8, s));
        }
      }
      A.out_cap = alloc;
      AMJ_TRY(lftj_ws_pass(ctx, P, out, final_mode, A, order, grid, &hc, &kms,
                           tails ? chunk : 0u, (uint64_t*)tails));
      ++passes;
      pushed += hc.pushed;
      rows = hc.out_rows;
      if (rows <= alloc) {
        if (tails && rows > hc.count) {
          const uint64_t used_chunks = rows / chunk;
          void *tpre = nullptr, *stmp = nullptr;
          AMJ_TRY(scratch(ctx, SLOT_MISC3, (used_chunks + 1) * 8, &tpre));
          AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(used_chunks) * 8, &stmp));
          AMJ_HIP_TRY(ctx, exclusive_scan_u64((const uint64_t*)tails, (uint64_t*)tpre,
                                              used_chunks, (uint64_t*)stmp, s));
          ProbeOut cols{};
          for (int a = 0; a < P.n_attrs; ++a) cols.cols[a] = out.c[a];
          AMJ_HIP_TRY(ctx, launch_chunk_fill(cols, P.n_attrs, (const uint64_t*)tails,
                                             (const uint64_t*)tpre, used_chunks, chunk,
                                             hc.count, rows, s));
          AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
        }
        break;
      }
      // overran: the pass counted every result; run again with room for them
      amj_result_free(ctx, res);
      capacity = hc.count;
      user_sized = false;
    }
    if (rows > A.out_cap) {
      set_error(ctx, AMJ_ERR_STATE, "direct output overran an allocation sized for it");
      return AMJ_ERR_STATE;
    }
    res->count = hc.count;
    res->rows_written = hc.count;
    res->checksum = hash ? hc.csum : 0;
  }
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->probe_kernel_ms = kms;
  res->probe_kernel_launches = passes;
  res->passes = passes;
  res->detail = (uint32_t)std::min<uint64_t>(pushed, 0xffffffffull);
  res->path = AMJ_PATH_LFTJ_DFS;
  return AMJ_OK;
}

static int run_lftj_dfs(amj_ctx_t* ctx, const LPlan& P, const amj_opts_t* o, amj_result_t* res) {
  if (o->ws) return run_lftj_ws(ctx, P, o, res);
  if (o->dro && o->materialize) return run_lftj_direct(ctx, P, o, res);
  if (!o->materialize) return run_lftj_count(ctx, P, res);
  hipStream_t s = ctx->stream;
  const uint64_t units = lftj_units(P);
  void *counts = nullptr, *tmp = nullptr, *bsum = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_UNIT_COUNTS, (units + 1) * 8, &counts));
  AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(units + 1) * 8, &tmp));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  const uint64_t* order = nullptr;
  AMJ_TRY(unit_order(ctx, P, &order));
  AMJ_HIP_TRY(ctx, hipMemsetAsync(counts, 0, (units + 1) * 8, s));
  AMJ_HIP_TRY(ctx, launch_lftj_count(P, (uint64_t*)counts, order, s));
  AMJ_HIP_TRY(ctx, exclusive_scan_u64((uint64_t*)counts, (uint64_t*)counts, units + 1,
                                      (uint64_t*)tmp, s));
  uint64_t total = 0;
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(&total, (uint64_t*)counts + units, 8, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  res->count = total;
  LOut out{};
  if (o->materialize && total > 0) {
    const uint64_t need = total * (uint64_t)P.n_attrs * 4;
    if (need > budget_left(ctx, o)) {
      set_error(ctx, AMJ_ERR_NOMEM, "result of " + std::to_string(total) +
                                        " tuples does not fit the HBM budget (use "
                                        "materialize=0)");
      return AMJ_ERR_NOMEM;
    }
    AMJ_TRY(alloc_output(ctx, total, P.n_attrs, res));
    for (int a = 0; a < P.n_attrs; ++a) out.c[a] = res->cols[a];
  }
  const uint64_t nb = units == 0 ? 1 : (units + kWavesPerBlock - 1) / kWavesPerBlock;
  AMJ_TRY(scratch(ctx, SLOT_BLOCK_SUMS, nb * 8, &bsum));
  const bool hash = o->checksum || !out.c[0];
  uint64_t nb_used = 0;
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
  AMJ_HIP_TRY(ctx, launch_lftj_write(P, (const uint64_t*)counts, out, hash, (uint64_t*)bsum,
                                     &nb_used, order, s));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
  if (hash) AMJ_TRY(finish_checksum(ctx, (const uint64_t*)bsum, nb_used, &res->checksum));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->probe_kernel_ms = elapsed_ms(ctx->ev[3], ctx->ev[4]);
  res->probe_kernel_launches = 1;
  res->rows_written = out.c[0] ? total : 0;
  res->path = AMJ_PATH_LFTJ_DFS;
  return AMJ_OK;
}

int lftj_execute(amj_ctx_t* ctx, const std::vector<LUse>& uses_in, const uint32_t* order,
                 uint32_t na, int algo, const amj_opts_t* opts, amj_result_t* res) {
  amj_opts_t def;
  amj_default_opts(&def);
  const amj_opts_t* o = opts ? opts : &def;
  std::memset(res, 0, sizeof(*res));
  // relations ordered by the positions of their attributes in attr_order (the result set does
  // not depend on it), so that a triangle or a 4-clique is recognised whatever the trie order
  std::vector<LUse> uses = uses_in;
  auto pos_key = [&](const LUse& u) {
    uint32_t px = na, py = na;
    for (uint32_t i = 0; i < na; ++i) {
      if (order[i] == u.x_id) px = i;
      if (order[i] == u.y_id) py = i;
    }
    return px * 64u + py;
  };
  std::stable_sort(uses.begin(), uses.end(),
                   [&](const LUse& a, const LUse& b) { return pos_key(a) < pos_key(b); });
  LPlan P;
  AMJ_TRY(make_lplan(ctx, uses, order, na, &P));
  res->num_cols = na;
  for (uint32_t a = 0; a < na; ++a) res->attr_ids[a] = order[a];
  if (algo != ALFTJ_DFS && algo != ALFTJ_BFS) {
    set_error(ctx, AMJ_ERR_ARG, "algo must be ALFTJ_BFS (0) or ALFTJ_DFS (1)");
    return AMJ_ERR_ARG;
  }
  bool any_host = false;
  for (const auto& u : uses) any_host = any_host || u.t->host;
  if (o->ooc || any_host) {
    // out of core: level-0 arrays readable from the host for the range cuts
    std::vector<HostView> hv(uses.size());
    std::vector<std::vector<int32_t>> d0s(uses.size());
    std::vector<std::vector<uint32_t>> offs(uses.size());
    for (size_t r = 0; r < uses.size(); ++r) {
      const amj_trie* t = uses[r].t;
      if (t->host) {
        hv[r] = HostView{t->h_d0, t->h_off, t->dev.n0};
        continue;
      }
      d0s[r].resize(t->dev.n0);
      offs[r].resize((size_t)t->dev.n0 + 1);
      if (t->dev.n0)
        AMJ_HIP_TRY(ctx, hipMemcpyAsync(d0s[r].data(), t->dev.d0, (size_t)t->dev.n0 * 4,
                                        hipMemcpyDeviceToHost, ctx->stream));
      AMJ_HIP_TRY(ctx, hipMemcpyAsync(offs[r].data(), t->dev.off, ((size_t)t->dev.n0 + 1) * 4,
                                      hipMemcpyDeviceToHost, ctx->stream));
      hv[r] = HostView{d0s[r].data(), offs[r].data(), t->dev.n0};
    }
    AMJ_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    uint64_t parts = 0;
    if (o->ooc_slice_rows) {  // level-0 values per range
      uint64_t mx = 0;
      for (const auto& h : hv) mx = std::max<uint64_t>(mx, h.n0);
      parts = std::max<uint64_t>(1, (mx + o->ooc_slice_rows - 1) / o->ooc_slice_rows);
    }
    return run_lftj_ooc(ctx, P, hv.data(), o, parts, res, nullptr);
  }
  if (algo == ALFTJ_DFS) {
    // the unit engine runs one wave per unit: with few roots, fewer roots per unit, so that a
    // launch has many more units than the device has wave slots (a unit of 64 roots of a
    // skewed graph can run far longer than the average one)
    P.rpu = kWave;
    const uint64_t target = (uint64_t)ctx->props.multiProcessorCount * 128;
    while (P.rpu > 1 && P.roots / P.rpu < target) P.rpu >>= 1;
    return run_lftj_dfs(ctx, P, o, res);
  }
  // BFS: count and checksum (LFTJ_BFS::evaluate reports a count only, LFTJ_BFS.cuh:457-500)
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
  AMJ_TRY(run_lftj_bfs(ctx, P, budget_left(ctx, o), &res->count, &res->checksum));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->path = AMJ_PATH_LFTJ_BFS;
  return AMJ_OK;
}

}  // namespace amj

using namespace amj;

extern "C" {

int alftj_trie_build_resident(amj_ctx_t* ctx, const amj_relation_t* R, amj_trie_t** out) {
  if (!ctx || !R || !out) return AMJ_ERR_ARG;
  *out = nullptr;
  if (R->num_cols != 2 || R->host) {
    set_error(ctx, AMJ_ERR_ARG, "a trie is built from a two-column device relation");
    return AMJ_ERR_ARG;
  }
  std::unique_ptr<amj_trie> t(new amj_trie());
  t->attr_ids[0] = R->attr_ids[0];
  t->attr_ids[1] = R->attr_ids[1];
  const int rc = build_trie(ctx, R->cols[0], R->cols[1], R->rows, &t->dev);
  if (rc != AMJ_OK) {
    free_trie(ctx, &t->dev);
    return rc;
  }
  *out = t.release();
  return AMJ_OK;
}

int alftj_trie_build(amj_ctx_t* ctx, const amj_rel_t* rel, amj_trie_t** out) {
  if (!ctx || !rel || !out) return AMJ_ERR_ARG;
  *out = nullptr;
  if (rel->num_cols != 2) {
    set_error(ctx, AMJ_ERR_ARG, "a trie is built from a two-column relation");
    return AMJ_ERR_ARG;
  }
  amj_relation_t* R = nullptr;
  AMJ_TRY(amj_relation_create(ctx, rel, &R));
  const int rc = alftj_trie_build_resident(ctx, R, out);
  amj_relation_destroy(ctx, R);
  return rc;
}

void alftj_trie_destroy(amj_ctx_t* ctx, amj_trie_t* t) {
  if (!t) return;
  if (t->host) {
    (void)hipHostFree(t->h_d0);
    (void)hipHostFree(t->h_off);
    (void)hipHostFree(t->h_d1);
    t->dev = TrieDev{};
  } else {
    free_trie(ctx, &t->dev);
  }
  delete t;
}

int alftj_trie_offload(amj_ctx_t* ctx, amj_trie_t* t) {
  if (!ctx || !t) return AMJ_ERR_ARG;
  if (t->host) return AMJ_OK;
  const uint64_t n0 = t->dev.n0, n1 = t->dev.n1;
  int32_t *d0 = nullptr, *d1 = nullptr;
  uint32_t* off = nullptr;
  if (hipHostMalloc((void**)&d0, std::max<uint64_t>(n0, 1) * 4, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&off, (n0 + 1) * 4, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&d1, std::max<uint64_t>(n1, 1) * 4, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipHostFree(d0);
    (void)hipHostFree(off);
    (void)hipHostFree(d1);
    set_error(ctx, AMJ_ERR_NOMEM, "pinned host allocation for the trie failed");
    return AMJ_ERR_NOMEM;
  }
  hipStream_t s = ctx->stream;
  if (n0) AMJ_HIP_TRY(ctx, hipMemcpyAsync(d0, t->dev.d0, n0 * 4, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(off, t->dev.off, (n0 + 1) * 4, hipMemcpyDeviceToHost, s));
  if (n1) AMJ_HIP_TRY(ctx, hipMemcpyAsync(d1, t->dev.d1, n1 * 4, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  const uint32_t max_run = t->dev.max_run;
  const double ms = t->dev.build_ms;
  free_trie(ctx, &t->dev);
  t->host = true;
  t->h_d0 = d0;
  t->h_off = off;
  t->h_d1 = d1;
  t->dev.n0 = (uint32_t)n0;
  t->dev.n1 = (uint32_t)n1;
  t->dev.max_run = max_run;
  t->dev.build_ms = ms;
  t->dev.view = LRel{d0, off, d1, nullptr, (uint32_t)n0, 0};
  return AMJ_OK;
}

int alftj_trie_info(amj_ctx_t* ctx, const amj_trie_t* t, uint64_t* level0, uint64_t* children,
                    uint64_t* max_children, double* build_ms) {
  if (!ctx || !t) return AMJ_ERR_ARG;
  if (level0) *level0 = t->dev.n0;
  if (children) *children = t->dev.n1;
  if (max_children) *max_children = t->dev.max_run;
  if (build_ms) *build_ms = t->dev.build_ms;
  return AMJ_OK;
}

int alftj_trie_copy_level(amj_ctx_t* ctx, const amj_trie_t* t, int level, int32_t* values,
                          uint32_t* offsets) {
  if (!ctx || !t || (level != 0 && level != 1)) return AMJ_ERR_ARG;
  hipStream_t s = ctx->stream;
  if (level == 0) {
    if (values && t->dev.n0)
      AMJ_HIP_TRY(ctx, hipMemcpyAsync(values, t->dev.view.d0, (uint64_t)t->dev.n0 * 4,
                                      hipMemcpyDefault, s));
    if (offsets)
      AMJ_HIP_TRY(ctx, hipMemcpyAsync(offsets, t->dev.view.off, ((uint64_t)t->dev.n0 + 1) * 4,
                                      hipMemcpyDefault, s));
  } else if (values && t->dev.n1) {
    AMJ_HIP_TRY(ctx, hipMemcpyAsync(values, t->dev.view.d1, (uint64_t)t->dev.n1 * 4,
                                    hipMemcpyDefault, s));
  }
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  return AMJ_OK;
}

int alftj_run(amj_ctx_t* ctx, amj_trie_t* const* tries, uint32_t n, const uint32_t* attr_order,
              uint32_t n_attrs, int algo, const amj_opts_t* opts, amj_result_t* res) {
  if (!ctx || !tries || !attr_order || !res) return AMJ_ERR_ARG;
  std::vector<LUse> uses;
  for (uint32_t r = 0; r < n; ++r) {
    if (!tries[r]) return AMJ_ERR_ARG;
    uses.push_back(LUse{tries[r], tries[r]->attr_ids[0], tries[r]->attr_ids[1]});
  }
  return lftj_execute(ctx, uses, attr_order, n_attrs, algo, opts, res);
}

int alftj_run_aliased(amj_ctx_t* ctx, amj_trie_t* const* tries, uint32_t n,
                      const uint32_t* trie_attrs, const uint32_t* attr_order, uint32_t n_attrs,
                      int algo, const amj_opts_t* opts, amj_result_t* res) {
  if (!ctx || !tries || !trie_attrs || !attr_order || !res) return AMJ_ERR_ARG;
  std::vector<LUse> uses;
  for (uint32_t r = 0; r < n; ++r) {
    if (!tries[r]) return AMJ_ERR_ARG;
    uses.push_back(LUse{tries[r], trie_attrs[2 * r], trie_attrs[2 * r + 1]});
  }
  return lftj_execute(ctx, uses, attr_order, n_attrs, algo, opts, res);
}

int alftj_pattern(amj_ctx_t* ctx, const amj_rel_t* edges, int query, int algo,
                  const amj_opts_t* opts, amj_result_t* res) {
  if (!ctx || !edges || !res) return AMJ_ERR_ARG;
  std::memset(res, 0, sizeof(*res));
  // SELF mode (exec_LFTJ.cu:198-222): one trie, aliased as (0,1), (1,2), (0,2) for the
  // triangle; the 4-clique adds (2,3), (0,3), (1,3)
  static const uint32_t tri[][2] = {{0, 1}, {1, 2}, {0, 2}};
  static const uint32_t four[][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {0, 2}, {1, 3}};
  const uint32_t(*q)[2] = nullptr;
  uint32_t nq = 0, na = 0;
  if (query == AMJ_QUERY_TRI) {
    q = tri;
    nq = 3;
    na = 3;
  } else if (query == AMJ_QUERY_FOUR) {
    q = four;
    nq = 6;
    na = 4;
  } else {
    set_error(ctx, AMJ_ERR_ARG, "pattern query must be TRI or FOUR");
    return AMJ_ERR_ARG;
  }
  amj_rel_t e = *edges;
  const uint32_t ids01[2] = {0, 1};
  e.attr_ids = ids01;
  amj_trie_t* T = nullptr;
  AMJ_TRY(alftj_trie_build(ctx, &e, &T));
  std::vector<LUse> uses;
  for (uint32_t i = 0; i < nq; ++i) uses.push_back(LUse{T, q[i][0], q[i][1]});
  const uint32_t order[4] = {0, 1, 2, 3};
  const int rc = lftj_execute(ctx, uses, order, na, algo, opts, res);
  if (rc == AMJ_OK) res->build_ms = T->dev.build_ms;
  alftj_trie_destroy(ctx, T);
  return rc;
}

}  // extern "C"
