// Out-of-core leapfrog triejoin: tries kept in host memory (or too large for the budget) are
// evaluated per combination of level-0 value ranges.
//
// Reference: OOCWrapper + BudgetPlanner (LFTJ-GPU/ooc_wrapper.cuh:14-427) split the budget over
// the attributes, binary-search level-0 value ranges so the slices fill the budget, and run the
// join once per combination of ranges (:263-371), on views of the tries.
//
// Here every attribute that is the level-0 attribute of some relation is cut into P value
// ranges of about equal trie bytes; a combination picks one range per such attribute, and every
// relation contributes the slice of its trie whose level-0 values lie in the range of its
// level-0 attribute (contiguous in d0, off and d1). A result tuple is found in exactly the one
// combination whose ranges hold its values: there every lookup it needs succeeds as in the
// whole trie, and in every other combination one of its level-0 lookups fails. The slices of the
// next combination are copied on the side stream into the other half of a double buffer while
// the current one is joined (count + checksum, summed over the combinations).
#include <algorithm>
#include <cstring>

#include "lftj.h"
#include "runtime.h"

namespace amj {

__global__ void rebase_off_kernel(uint32_t* off, uint32_t n, uint32_t base) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    off[i] -= base;
}

namespace {

uint32_t lower_row(const HostView& h, int64_t v) {  // first row with d0 >= v
  return (uint32_t)(std::lower_bound(h.d0, h.d0 + h.n0, v,
                                     [](int32_t a, int64_t b) { return (int64_t)a < b; }) -
                    h.d0);
}

}  // namespace

// hv[r]: host-readable level-0 values and offsets of relation r, for the range cuts and the
// slice bounds; the slices themselves are copied from P.rel[r] (host or device memory).
int run_lftj_ooc(amj_ctx_t* ctx, const LPlan& P, const HostView* hv, const amj_opts_t* o,
                 uint64_t parts_hint, amj_result_t* res, uint64_t* combos_out) {
  hipStream_t s = ctx->stream, side = ctx->side;
  const int R = P.n_rels, A = P.n_attrs;
  // bytes of each relation (the slices of a combination share the budget)
  std::vector<uint64_t> total(R);
  uint64_t all = 0;
  for (int r = 0; r < R; ++r) {
    total[r] = (uint64_t)hv[r].n0 * 8 + (uint64_t)hv[r].off[hv[r].n0] * 4;
    all += total[r];
  }
  uint64_t parts = parts_hint;
  if (!parts) {
    const uint64_t room = std::max<uint64_t>(budget_left(ctx, o) / 4, 1 << 20);
    parts = std::max<uint64_t>(1, (all + room - 1) / room);
  }
  // value cuts per root attribute, from the first relation rooted there (by trie bytes)
  std::vector<std::vector<int64_t>> cuts(A);  // ranges [cuts[k], cuts[k+1])
  for (int i = 0; i < A; ++i) {
    if (!P.root_mask[i]) continue;
    int r0 = -1;
    for (int r = 0; r < R; ++r)
      if ((P.root_mask[i] >> r) & 1u) {
        r0 = r;
        break;
      }
    const HostView& h = hv[r0];
    std::vector<int64_t> c{INT64_MIN};
    const uint64_t step = std::max<uint64_t>(1, total[r0] / parts);
    uint64_t acc = 0, next = step;
    for (uint32_t j = 0; j < h.n0 && c.size() < parts; ++j) {
      acc += 8 + 4ull * (h.off[j + 1] - h.off[j]);
      if (acc >= next && j + 1 < h.n0) {
        c.push_back((int64_t)h.d0[j + 1]);
        next += step;
      }
    }
    c.push_back(INT64_MAX);
    cuts[i] = c;
  }
  // device double buffer: per relation d0, off, d1 slices
  uint64_t cap0 = 0, cap1 = 0;  // largest level-0 slice and child slice over all relations
  for (int r = 0; r < R; ++r) {
    const int x = P.x[r];
    for (size_t k = 0; k + 1 < cuts[x].size(); ++k) {
      const uint32_t lo = lower_row(hv[r], cuts[x][k]), hi = lower_row(hv[r], cuts[x][k + 1]);
      cap0 = std::max<uint64_t>(cap0, hi - lo);
      cap1 = std::max<uint64_t>(cap1, hv[r].off[hi] - hv[r].off[lo]);
    }
  }
  const uint64_t per_rel = (cap0 * 2 + 1 + cap1) * 4 + 256;
  void* buf[2] = {nullptr, nullptr};
  AMJ_TRY(scratch(ctx, SLOT_OOC_A, per_rel * R, &buf[0]));
  AMJ_TRY(scratch(ctx, SLOT_OOC_B, per_rel * R, &buf[1]));
  hipEvent_t copied[2], done[2];
  for (int b = 0; b < 2; ++b) {
    AMJ_HIP_TRY(ctx, hipEventCreateWithFlags(&copied[b], hipEventDisableTiming));
    AMJ_HIP_TRY(ctx, hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
    AMJ_HIP_TRY(ctx, hipEventRecord(done[b], s));
  }
  struct Cleanup {
    hipEvent_t* e;
    ~Cleanup() {
      for (int k = 0; k < 4; ++k) (void)hipEventDestroy(e[k]);
    }
  };
  hipEvent_t evs[4] = {copied[0], copied[1], done[0], done[1]};
  Cleanup cleanup{evs};
  // per-combination unit counts, scan and block sums
  const uint64_t max_units = (cap0 + kWave - 1) / kWave + 1;
  void *counts = nullptr, *tmp = nullptr, *bsum = nullptr, *csum_dev = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_UNIT_COUNTS, (max_units + 1) * 8, &counts));
  AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(max_units + 1) * 8, &tmp));
  AMJ_TRY(scratch(ctx, SLOT_BLOCK_SUMS, (max_units / kWavesPerBlock + 2) * 8, &bsum));
  AMJ_TRY(scratch(ctx, SLOT_SUM, 64, &csum_dev));

  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  std::vector<size_t> idx(A, 0);  // odometer over the root attributes
  uint64_t combos = 0, count = 0, csum = 0;
  for (;;) {
    const int b = (int)(combos & 1);
    LPlan Q = P;
    AMJ_HIP_TRY(ctx, hipStreamWaitEvent(side, done[b], 0));
    uint8_t* base = (uint8_t*)buf[b];
    for (int r = 0; r < R; ++r) {
      const int x = P.x[r];
      const HostView& h = hv[r];
      const uint32_t lo = lower_row(h, cuts[x][idx[x]]), hi = lower_row(h, cuts[x][idx[x] + 1]);
      const uint32_t c0 = h.off[lo], c1 = h.off[hi];
      int32_t* d0 = (int32_t*)(base + per_rel * r);
      uint32_t* off = (uint32_t*)(d0 + cap0);
      int32_t* d1 = (int32_t*)(off + cap0 + 1);
      if (hi > lo)
        AMJ_HIP_TRY(ctx, hipMemcpyAsync(d0, P.rel[r].d0 + lo, (uint64_t)(hi - lo) * 4,
                                        hipMemcpyDefault, side));
      AMJ_HIP_TRY(ctx, hipMemcpyAsync(off, P.rel[r].off + lo, (uint64_t)(hi - lo + 1) * 4,
                                      hipMemcpyDefault, side));
      if (c1 > c0)
        AMJ_HIP_TRY(ctx, hipMemcpyAsync(d1, P.rel[r].d1 + c0, (uint64_t)(c1 - c0) * 4,
                                        hipMemcpyDefault, side));
      if (c0)
        hipLaunchKernelGGL(rebase_off_kernel, dim3(std::min<uint32_t>((hi - lo + 256) / 256, 1024u)),
                           dim3(256), 0, side, off, hi - lo + 1, c0);
      Q.rel[r] = LRel{d0, off, d1, nullptr, hi - lo, 0};
    }
    AMJ_HIP_TRY(ctx, hipEventRecord(copied[b], side));
    AMJ_HIP_TRY(ctx, hipStreamWaitEvent(s, copied[b], 0));
    Q.roots = Q.rel[Q.driver0].n0;
    const uint64_t units = lftj_units(Q);
    if (units) {
      // count + checksum of this combination: one checksum-only pass that also counts
      uint64_t nb = 0;
      AMJ_HIP_TRY(ctx, launch_lftj_write(Q, nullptr, LOut{}, true, (uint64_t*)bsum, &nb, nullptr,
                                         s, (uint64_t*)counts));
      AMJ_HIP_TRY(ctx, reduce_sum_u64((const uint64_t*)bsum, nb, (uint64_t*)csum_dev,
                                      (uint64_t*)tmp, s));
      AMJ_HIP_TRY(ctx, reduce_sum_u64((const uint64_t*)counts, nb, (uint64_t*)csum_dev + 1,
                                      (uint64_t*)tmp, s));
      uint64_t hc[2] = {0, 0};
      AMJ_HIP_TRY(ctx, hipMemcpyAsync(&hc[1], csum_dev, 8, hipMemcpyDeviceToHost, s));
      AMJ_HIP_TRY(ctx, hipMemcpyAsync(&hc[0], (uint64_t*)csum_dev + 1, 8, hipMemcpyDeviceToHost,
                                      s));
      AMJ_HIP_TRY(ctx, hipEventRecord(done[b], s));
      AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
      count += hc[0];
      csum += hc[1];
    } else {
      AMJ_HIP_TRY(ctx, hipEventRecord(done[b], s));
    }
    ++combos;
    int i = A - 1;  // advance the odometer
    for (; i >= 0; --i) {
      if (!P.root_mask[i]) continue;
      if (++idx[i] + 1 < cuts[i].size()) break;
      idx[i] = 0;
    }
    if (i < 0) break;
  }
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(side));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->count = count;
  res->checksum = csum;
  res->rows_written = 0;
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->passes = 1;
  res->detail = (uint32_t)std::min<uint64_t>(combos, 0xffffffffull);
  res->path = AMJ_PATH_LFTJ_OOC;
  if (combos_out) *combos_out = combos;
  return AMJ_OK;
}

}  // namespace amj
