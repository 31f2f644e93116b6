// Frequent-item path of the AMHJ probe for linear chains: subtree counts of every join value,
// suffix tables, and the output-window writer.
//
// Reference: AMHJ_WS_FIB + SubfixBuffer (MHJ-GPU/Joins/AMHJ_WS_FIB.cuh:29-916,
// MHJ-GPU/SubfixBuffer.cuh:21-98) buffers the 256 most frequent values of every join column:
// the first warp to reach one publishes its result count and the position of the result tuples
// it wrote, and later warps reuse the count or copy those suffix tuples. Here the buffer holds
// every join value:
//   1. bottom-up, W_t[r] = number of results below row r of build table t and its prefix PW_t
//      (one hash lookup per build row), plus a 16-byte record per row (its value and the run it
//      leads to one table down);
//   2. while they stay small, the suffix tables S_t (all tuples below every row of table t, row r
//      owning rows [PW_t[r], PW_t[r+1])) are materialised bottom-up, each with the writer below
//      using table t as the probe side and S_{t+1} as the suffix;
//   3. the result is written by the output-window writer over the remaining top levels, whose
//      last hop copies contiguous runs of the deepest suffix table.
#include <algorithm>
#include <cstring>

#include "runtime.h"

namespace amj {

constexpr uint64_t kDefaultWindow = 4096;  // results per window of the output-window writer

// One launch of the output-window writer over plan Q; returns AMJ_ERR_STATE on overflow.
static int window_pass(amj_ctx_t* ctx, const Plan& Q, const uint64_t* off, uint64_t total,
                       const ProbeOut& out, bool hash, uint64_t window, uint64_t* csum,
                       double* kernel_ms) {
  hipStream_t s = ctx->stream;
  if (total == 0) {
    if (csum) *csum = 0;
    return AMJ_OK;
  }
  const uint64_t n_windows = (total + window - 1) / window;
  const uint64_t nb = window_blocks(total, window);
  void *bsum = nullptr, *first = nullptr, *flag = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_BLOCK_SUMS, (nb + 1) * 8, &bsum));
  AMJ_TRY(scratch(ctx, SLOT_MISC0, (n_windows + 1) * 8, &first));
  AMJ_TRY(scratch(ctx, SLOT_FLAG, 256, &flag));
  AMJ_HIP_TRY(ctx, hipMemsetAsync(flag, 0, 4, s));
  uint64_t nb_used = 0;
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
  AMJ_HIP_TRY(ctx, launch_window_write(Q, off, (uint64_t*)first, total, window, out, hash,
                                       (uint64_t*)bsum, (uint32_t*)flag, &nb_used, s));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
  uint32_t overflow = 0;
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(&overflow, flag, 4, hipMemcpyDeviceToHost, s));
  if (hash && csum) {
    AMJ_TRY(finish_checksum(ctx, (const uint64_t*)bsum, nb_used, csum));
  } else {
    AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  }
  if (kernel_ms) *kernel_ms = elapsed_ms(ctx->ev[3], ctx->ev[4]);
  if (overflow) {
    set_error(ctx, AMJ_ERR_STATE, "a frontier level exceeded 2^32 children");
    return AMJ_ERR_STATE;
  }
  return AMJ_OK;
}

int run_fib(amj_ctx_t* ctx, const Plan& plan, const amj_opts_t* o, amj_result_t* res) {
  hipStream_t s = ctx->stream;
  // a local copy: the probe columns are redirected to the grouped copy and the tables get their
  // prefix arrays, and the caller's plan stays as it was for a fallback engine
  Plan P = plan;
  const int L = P.n_levels;
  const int lastv = P.val_mode[L - 1] == VAL_BIND ? 1 : 0;
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));

  // 0. probe rows grouped by join key: rows that share a key are then neighbours, so their
  // lookups, run records and the suffix rows they copy are served from cache (one stable
  // radix sort on the key bytes, the other column carried along)
  const uint64_t np = P.probe_n;
  if (np > 0 && (o->probe_order == 2 || (o->probe_order == 0 && np >= (1ull << 20)))) {
    int kcol = 0;
    for (int c = 0; c < P.probe_ncols; ++c)
      if (P.probe_attr[c] == P.key_attr[0]) kcol = c;
    const int other = P.probe_ncols == 2 ? 1 - kcol : -1;
    void *comp = nullptr, *alt = nullptr, *tmp = nullptr, *cols = nullptr;
    AMJ_TRY(scratch(ctx, SLOT_SORT_KEYS, np * 8, &comp));
    AMJ_TRY(scratch(ctx, SLOT_SORT_ALT, np * 8, &alt));
    AMJ_TRY(scratch(ctx, SLOT_SORT_TEMP, radix_sort_temp_bytes(np), &tmp));
    AMJ_TRY(scratch(ctx, SLOT_PROBE_SORTED, np * 4 * P.probe_ncols, &cols));
    int32_t* sk = (int32_t*)cols;
    int32_t* so = other >= 0 ? sk + np : nullptr;
    AMJ_HIP_TRY(ctx, launch_pack_composite(P.probe_cols[kcol],
                                           other >= 0 ? P.probe_cols[other] : nullptr, np, 0,
                                           (uint64_t*)comp, s));
    AMJ_HIP_TRY(ctx, radix_sort_u64((uint64_t*)comp, (uint64_t*)alt, np, 0xf0u, tmp, s));
    AMJ_HIP_TRY(ctx, launch_unpack_composite((const uint64_t*)comp, np, 0, sk, so, s));
    P.probe_cols[kcol] = sk;
    if (other >= 0) P.probe_cols[other] = so;
  }

  // 1. weights, prefixes and records, bottom-up (weights of table t need PW of table t+1)
  RunRef* rec[kMaxLevels] = {nullptr};
  uint64_t* pw[kMaxLevels] = {nullptr};
  void* w = nullptr;
  void* tmp = nullptr;
  for (int t = L - 2; t >= 0; --t) {
    const uint64_t n = P.tab[t].n;
    void *p = nullptr, *nx = nullptr;
    AMJ_TRY(scratch(ctx, SLOT_PW0 + t, (n + 1) * 8, &p));
    AMJ_TRY(scratch(ctx, SLOT_NEXT0 + t, std::max<uint64_t>(n, 1) * sizeof(RunRef), &nx));
    AMJ_TRY(scratch(ctx, SLOT_FIB_W, (n + 1) * 8, &w));
    AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(n + 1) * 8, &tmp));
    AMJ_HIP_TRY(ctx, launch_fib_weights(P, t, (uint64_t*)w, (RunRef*)nx, s));
    AMJ_HIP_TRY(ctx, exclusive_scan_u64((uint64_t*)w, (uint64_t*)p, n + 1, (uint64_t*)tmp, s));
    pw[t] = (uint64_t*)p;
    rec[t] = (RunRef*)nx;
    P.tab[t].pw = pw[t];
  }
  const uint64_t n0 = P.probe_n;
  void *off = nullptr, *pnx = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_PROBE_OFF, (n0 + 1) * 8, &off));
  AMJ_TRY(scratch(ctx, SLOT_PROBE_NEXT, std::max<uint64_t>(n0, 1) * sizeof(RunRef), &pnx));
  AMJ_TRY(scratch(ctx, SLOT_FIB_W, (n0 + 1) * 8, &w));
  AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(n0 + 1) * 8, &tmp));
  AMJ_HIP_TRY(ctx, launch_fib_probe_weights(P, (uint64_t*)w, (RunRef*)pnx, s));
  AMJ_HIP_TRY(ctx, exclusive_scan_u64((uint64_t*)w, (uint64_t*)off, n0 + 1, (uint64_t*)tmp, s));

  // sizes: total results and |S_t| for every table
  uint64_t total = 0;
  uint64_t srows[kMaxLevels] = {0};
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(&total, (uint64_t*)off + n0, 8, hipMemcpyDeviceToHost, s));
  for (int t = 0; t <= L - 2; ++t)
    AMJ_HIP_TRY(ctx, hipMemcpyAsync(&srows[t], pw[t] + P.tab[t].n, 8, hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  srows[L - 1] = P.tab[L - 1].n;
  res->count = total;

  // 2. suffix depth k: S_t for t in [k, L-2] are materialised (S_{L-1} is the last table)
  int ncols[kMaxLevels];
  for (int t = 0; t < L; ++t) ncols[t] = (L - 1 - t) + lastv;
  const uint64_t out_bytes = total * (uint64_t)P.n_attrs * 4;
  const uint64_t room = budget_left(ctx, o);
  uint64_t sfx_budget = std::max<uint64_t>(out_bytes / 4, 1ull << 30);
  sfx_budget = std::min<uint64_t>(sfx_budget, room / 4);
  const int mode = o->fib_suffix;  // 0 size-based, 1 none, 2 as deep as memory allows
  if (mode == 2) sfx_budget = room / 2;
  int k = L - 1;
  uint64_t used = 0;
  if (mode != 1) {
    for (int t = L - 2; t >= 0; --t) {
      const uint64_t bytes = srows[t] * (uint64_t)ncols[t] * 4;
      if (srows[t] >= (1ull << 32) || used + bytes > sfx_budget) break;
      used += bytes;
      k = t;
    }
  }

  // 3. records that lead into a materialised suffix table address its rows
  for (int t = L - 3; t >= 0 && t >= k - 1; --t)
    AMJ_HIP_TRY(ctx, launch_records_to_suffix(rec[t], P.tab[t].n, pw[t + 1], s));
  if (k == 0 && L >= 2)
    AMJ_HIP_TRY(ctx, launch_records_to_suffix((RunRef*)pnx, n0, pw[0], s));

  // 4. suffix tables, bottom-up
  const int32_t* sbase[kMaxLevels];
  uint64_t sstride[kMaxLevels];
  sbase[L - 1] = P.tab[L - 1].vals;
  sstride[L - 1] = P.tab[L - 1].n;
  // windows are whole staging blocks of the writer
  const uint64_t window = o->out_chunk_rows
                              ? (o->out_chunk_rows + kWindowAlign - 1) / kWindowAlign * kWindowAlign
                              : kDefaultWindow;
  for (int t = L - 2; t >= k; --t) {
    void* buf = nullptr;
    AMJ_TRY(scratch(ctx, SLOT_SFX0 + t, std::max<uint64_t>(srows[t], 1) * ncols[t] * 4, &buf));
    Plan Q;
    std::memset(&Q, 0, sizeof(Q));
    Q.n_levels = 1;
    Q.n_attrs = ncols[t];
    Q.probe_ncols = 1;
    Q.probe_n = P.tab[t].n;
    Q.probe_cols[0] = P.tab[t].vals;
    Q.probe_attr[0] = 0;
    Q.probe_next = rec[t];
    Q.sfx_base = sbase[t + 1];
    Q.sfx_stride = sstride[t + 1];
    Q.sfx_cols = ncols[t + 1];
    ProbeOut po{};
    for (int a = 0; a < ncols[t]; ++a) po.cols[a] = (int32_t*)buf + (uint64_t)a * srows[t];
    po.capacity = srows[t];
    AMJ_TRY(window_pass(ctx, Q, pw[t], srows[t], po, false, window, nullptr, nullptr));
    sbase[t] = (const int32_t*)buf;
    sstride[t] = srows[t];
  }

  // 5. the result: top levels 0..k-1 traversed, last hop into S_k
  Plan F = P;
  F.n_levels = k + 1;
  F.probe_next = (const RunRef*)pnx;
  for (int t = 0; t < k; ++t) F.next[t] = rec[t];
  F.sfx_base = sbase[k];
  F.sfx_stride = sstride[k];
  F.sfx_cols = ncols[k];
  ProbeOut out{};
  if (o->materialize && total > 0) {
    if (out_bytes > budget_left(ctx, o)) {
      set_error(ctx, AMJ_ERR_NOMEM, "result of " + std::to_string(total) +
                                        " tuples does not fit the HBM budget (use ooc=1 or "
                                        "materialize=0)");
      return AMJ_ERR_NOMEM;
    }
    AMJ_TRY(alloc_output(ctx, total, P.n_attrs, res));
    for (int a = 0; a < P.n_attrs; ++a) out.cols[a] = res->cols[a];
    out.capacity = total;
  }
  const bool hash = o->checksum || !out.cols[0];
  double kms = 0.0;
  const int rc = window_pass(ctx, F, (const uint64_t*)off, total, out, hash, window,
                             hash ? &res->checksum : nullptr, &kms);
  if (rc != AMJ_OK) {
    amj_result_free(ctx, res);
    return rc;
  }
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->probe_kernel_ms = kms;
  res->probe_kernel_launches = 1;
  res->rows_written = out.cols[0] ? total : 0;
  res->path = AMJ_PATH_FIB;
  res->detail = (uint32_t)k;  // suffix depth used
  res->passes = 1;
  return AMJ_OK;
}

}  // namespace amj
