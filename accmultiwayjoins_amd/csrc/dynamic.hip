// Persistent AMHJ probe with work sharing (ws) and direct result output (dro).
//
// Reference: AMHJ_WS (MHJ-GPU/Joins/AMHJ_WS.cuh:32-557) keeps persistent warps that claim probe
// tuples with an atomic counter and, when a warp meets a long bucket, appends the tail as
// 1024-entry tasks to a TaskBook (TaskBook.cuh:13-109) served through a global queue
// (GCQueue.cuh:30-107) until an idle-warp barrier (CBarrier.cuh:9-46) sees every warp idle.
// AMHJ_DRO (Joins/AMHJ_DRO.cuh:18-707) drops the count pass: it sizes the output from a sample
// and writes into chunks in a single pass, falling back to residual passes when chunks run out.
//
// Here persistent waves run the frontier DFS of probe_engine.h:
//   * a wave claims units of 64 probe rows from a global counter, in guided chunks;
//   * when a level of its frontier has more than `split` children left, it keeps the next
//     `split` and publishes the rest as tasks of `split` children each; a task is the 64-lane
//     frontier of that level (bindings, runs, prefix) plus its child range, so the wave that
//     takes it continues exactly where the owner would have;
//   * results gather in a per-wave LDS buffer and are placed with one atomic reservation per
//     full buffer (not per tuple, nor per last-level frontier);
//   * a wave leaves when no unit is left, the queue is empty and no wave holds work (a wave
//     holding work is running, so it finishes and, having pushed tasks, takes them itself if
//     nobody else does). Every wait is bounded: a wave that sees no unit or task completed
//     anywhere for the time limit raises an error flag and leaves, so the kernel always ends.
// With dro the output is allocated from a sampled estimate before the single pass; if the
// reservations overrun it, the pass knows the exact total and is run once more into an exact
// allocation.
#include "probe_engine.h"
#include "runtime.h"
#include "ws_queue.h"

namespace amj {

using DynCtl = WsCtl;
using DynArgs = WsArgs;

// Per-wave state of the persistent engine besides the accumulators.
struct DynWave {
  DirectStage st;  // staged output rows (materialising passes)
  WsWave q;        // sharing state
};

// Task record: 16 header words then 10 words per lane (bindings, lo, len, excl).
constexpr int kTaskHdr = 16;
constexpr int kTaskLaneWords = kMaxAttrs + 4;
constexpr int kTaskWords = kTaskHdr + kTaskLaneWords * kWave;

// Publish the level frontier `l` with child range [c0, c1) into slot `slot`.
template <int t, int NA>
__device__ __forceinline__ void write_task(const DynArgs& A, uint64_t slot, const LvN<NA>& l,
                                           uint64_t c0, uint64_t c1) {
  uint32_t* rec = A.tasks + slot * (uint64_t)kTaskWords;
  const int lane = lane_id();
  uint32_t* lw = rec + kTaskHdr;
  sfor<0, NA>([&](auto a) { lw[a * kWave + lane] = (uint32_t)l.b[a]; });
  lw[(kMaxAttrs + 0) * kWave + lane] = l.lo;
  lw[(kMaxAttrs + 1) * kWave + lane] = l.len;
  lw[(kMaxAttrs + 2) * kWave + lane] = (uint32_t)l.excl;
  lw[(kMaxAttrs + 3) * kWave + lane] = (uint32_t)(l.excl >> 32);
  if (lane == 0) {
    rec[0] = (uint32_t)t;
    rec[1] = (uint32_t)c0;
    rec[2] = (uint32_t)(c0 >> 32);
    rec[3] = (uint32_t)c1;
    rec[4] = (uint32_t)(c1 >> 32);
    rec[5] = (uint32_t)l.T;
    rec[6] = (uint32_t)(l.T >> 32);
  }
  publish_task(A, slot);
}

template <int NA>
__device__ __forceinline__ void read_task(const DynArgs& A, uint64_t slot, int& t, uint64_t& c0,
                                          uint64_t& c1, LvN<NA>& l) {
  const uint32_t* rec = A.tasks + slot * (uint64_t)kTaskWords;
  const int lane = lane_id();
  t = (int)uniform(rec[0]);
  c0 = uniform(((uint64_t)rec[2] << 32) | rec[1]);
  c1 = uniform(((uint64_t)rec[4] << 32) | rec[3]);
  l.T = uniform(((uint64_t)rec[6] << 32) | rec[5]);
  const uint32_t* lw = rec + kTaskHdr;
  sfor<0, NA>([&](auto a) { l.b[a] = (int32_t)lw[a * kWave + lane]; });
  l.lo = lw[(kMaxAttrs + 0) * kWave + lane];
  l.len = lw[(kMaxAttrs + 1) * kWave + lane];
  l.excl = ((uint64_t)lw[(kMaxAttrs + 3) * kWave + lane] << 32) | lw[(kMaxAttrs + 2) * kWave + lane];
}

template <int L, int t, int MODE>
__device__ __forceinline__ void dyn_descend(const Plan& P, const LvL<L>& cur, uint64_t c0, uint64_t c1,
                            const ProbeOut& out, const DynArgs& A, Acc& acc, DynWave& W) {
  const auto step = [&]() { ws_step(A, W.q); };
  if constexpr (t == L - 1) {
    // the last level is never split: its children are the results, [0, T)
    if constexpr (MODE == MODE_COUNT) {
      acc.count += cur.T;
    } else if constexpr (!mode_mat<MODE>()) {
      emit<false, mode_hash<MODE>()>(P, t, cur, 0, out, acc, step);
      acc.count += cur.T;
    } else {
      direct_emit<true, mode_hash<MODE>()>(P, t, cur, out, acc, W.st, &A.ctl->out_rows, step);
      acc.count += cur.T;
    }
  } else {
    for (uint64_t c = c0; c < c1; c += kWave) {
      step();
      if (!W.q.qfull && ws_should_split(c, c1, A.split)) {
        // keep [c, c + split), publish the remaining pieces (the last ones if the queue is short)
        const uint64_t k = ws_pieces(c, c1, A.split);
        uint64_t got = 0;
        bool full = false;
        const uint64_t slot = reserve_tasks(A, k, &got, &full);
        if (full) W.q.qfull = true;  // slots are never reused within a launch: stop asking
        if (got) {
          for (uint64_t j = 0; j < got; ++j) {
            uint64_t p0 = 0, p1 = 0;
            ws_piece(c, c1, A.split, k, got, j, &p0, &p1);
            write_task<t>(A, slot + j, cur, p0, p1);
          }
          if (lane_id() == 0) add_relaxed(&A.ctl->pushed, (unsigned long long)got);
          c1 = ws_keep_end(c, c1, A.split, k, got);
        }
      }
      LvL<L> ch;
      make_child<t>(P, cur, c, c1, ch);
      if (ch.T) dyn_descend<L, t + 1, MODE>(P, ch, 0, ch.T, out, A, acc, W);
    }
  }
}

template <int L, int MODE>
__global__ __launch_bounds__(kBlock) void dyn_kernel(const Plan P, const ProbeOut out,
                                                     const DynArgs A) {
  constexpr bool MAT = mode_mat<MODE>();
  __shared__ int32_t stage_mem[MAT ? direct_stage_words<L>() : 1];
  DynWave W{{(lds_i32*)&stage_mem[MAT ? (threadIdx.x / kWave) * na_of(L) * kDirectRows : 0], 0u},
            {false, 0u}};
  Acc acc{};
  ws_serve(
      A,
      [&](uint64_t u, uint64_t ue) {
        for (uint64_t v = u; v < ue; ++v) {
          LvL<L> l0;
          make_level0(P, v * kWave, l0);
          if (l0.T) dyn_descend<L, 0, MODE>(P, l0, 0, l0.T, out, A, acc, W);
        }
      },
      [&](uint64_t slot) {
        int t = 0;
        uint64_t c0 = 0, c1 = 0;
        LvL<L> l;
        read_task(A, slot, t, c0, c1, l);
        sfor<0, L - 1>([&](auto tt) {
          if ((int)tt == t) dyn_descend<L, tt, MODE>(P, l, c0, c1, out, A, acc, W);
        });
      });
  if (MAT) {
    direct_flush<na_of(L)>(out, P.n_attrs, W.st, &A.ctl->out_rows);
    direct_tail(out, W.st);
  }
  ws_accumulate(A, acc.count, acc.csum);
}

// Results of `samples` units spread evenly over the probe relation (the output estimate).
template <int L>
__global__ __launch_bounds__(kBlock) void dyn_sample_kernel(const Plan P, uint64_t units,
                                                            uint64_t samples,
                                                            unsigned long long* sum) {
  const uint64_t s = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (s >= samples) return;  // no barriers in this kernel
  const uint64_t u = s * units / samples;
  LvL<L> l0;
  make_level0(P, u * kWave, l0);
  Acc acc{};
  const DynArgs none{};  // no sharing, no output
  DynWave W{{nullptr, 0u}, {true, 0u}};
  if (l0.T) dyn_descend<L, 0, MODE_COUNT>(P, l0, 0, l0.T, ProbeOut{}, none, acc, W);
  if (lane_id() == 0) add_relaxed(sum, (unsigned long long)acc.count);
}

hipError_t launch_dyn(const Plan& P, const ProbeOut& out, bool hash, const void* args_v,
                      unsigned grid, hipStream_t s) {
  const DynArgs& A = *(const DynArgs*)args_v;
  const int mode = out.cols[0] ? (hash ? MODE_WRITE_HASH : MODE_WRITE) : (hash ? MODE_HASH
                                                                              : MODE_COUNT);
#define LAUNCH_M(LV, M) \
  hipLaunchKernelGGL((dyn_kernel<LV, M>), dim3(grid), dim3(kBlock), 0, s, P, out, A)
#define LAUNCH(LV)                                                 \
  if (mode == MODE_WRITE) LAUNCH_M(LV, MODE_WRITE);                \
  else if (mode == MODE_WRITE_HASH) LAUNCH_M(LV, MODE_WRITE_HASH); \
  else if (mode == MODE_HASH) LAUNCH_M(LV, MODE_HASH);             \
  else LAUNCH_M(LV, MODE_COUNT)
  AMJ_DISPATCH_L(P, LAUNCH)
#undef LAUNCH
#undef LAUNCH_M
  return hipGetLastError();
}

hipError_t launch_dyn_sample(const Plan& P, uint64_t samples, unsigned long long* sum,
                             hipStream_t s) {
  const uint64_t units = probe_units(P);
  if (units == 0 || samples == 0) return hipSuccess;
  const unsigned grid = (unsigned)((samples + kWavesPerBlock - 1) / kWavesPerBlock);
#define LAUNCH(LV)                                                                       \
  hipLaunchKernelGGL(dyn_sample_kernel<LV>, dim3(grid), dim3(kBlock), 0, s, P, units, \
                     samples, sum)
  AMJ_DISPATCH_L(P, LAUNCH)
#undef LAUNCH
  return hipGetLastError();
}

// Persistent grid: as many workgroups as fit on the device at once for the instantiation that
// runs (its registers and, when materialising, its staging buffer bound the occupancy).
template <int L, int MODE>
static int occupancy_of() {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dyn_kernel<L, MODE>, kBlock, 0) !=
          hipSuccess || per_cu <= 0)
    per_cu = 2;
  return per_cu;
}

int dyn_occupancy_grid(const Plan& P, bool mat, bool hash, int cus) {
  const int mode = mat ? (hash ? MODE_WRITE_HASH : MODE_WRITE) : (hash ? MODE_HASH : MODE_COUNT);
  int per_cu = 2;
#define OCC_M(LV, M) per_cu = occupancy_of<LV, M>()
#define OCC(LV)                                                \
  if (mode == MODE_WRITE) OCC_M(LV, MODE_WRITE);                \
  else if (mode == MODE_WRITE_HASH) OCC_M(LV, MODE_WRITE_HASH); \
  else if (mode == MODE_HASH) OCC_M(LV, MODE_HASH);             \
  else OCC_M(LV, MODE_COUNT)
  switch (P.n_levels) {
    case 1: OCC(1); break;
    case 2: OCC(2); break;
    case 3: OCC(3); break;
    case 4: OCC(4); break;
    case 5: OCC(5); break;
    case 6: OCC(6); break;
    case 7: OCC(7); break;
    default: break;
  }
#undef OCC
#undef OCC_M
  return per_cu * cus;
}

// ------------------------------------------------------------------------------------------
// Host driver.
// ------------------------------------------------------------------------------------------
constexpr uint64_t kDefaultSplit = 4096;      // children per shared task
constexpr uint64_t kTaskSlots = 1 << 14;      // 16384 tasks x 2.6 KB
constexpr uint64_t kStallMs = 20000;          // no completed work for this long: give up

static int dyn_pass(amj_ctx_t* ctx, const Plan& P, const ProbeOut& out, bool hash,
                    DynArgs A, unsigned grid, DynCtl* host_ctl, double* kernel_ms) {
  hipStream_t s = ctx->stream;
  AMJ_HIP_TRY(ctx, hipMemsetAsync(A.ctl, 0, sizeof(DynCtl), s));
  if (A.split) AMJ_HIP_TRY(ctx, hipMemsetAsync(A.ready, 0, A.cap * 4, s));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
  AMJ_HIP_TRY(ctx, launch_dyn(P, out, hash, &A, grid, s));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(host_ctl, A.ctl, sizeof(DynCtl), hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  if (kernel_ms) *kernel_ms += elapsed_ms(ctx->ev[3], ctx->ev[4]);
  if (host_ctl->flags & kFlagSpin) {
    set_error(ctx, AMJ_ERR_STATE, "persistent probe: no work completed within the time limit");
    return AMJ_ERR_STATE;
  }
  return AMJ_OK;
}

int run_dynamic(amj_ctx_t* ctx, Plan& P, const amj_opts_t* o, amj_result_t* res) {
  hipStream_t s = ctx->stream;
  const uint64_t units = probe_units(P);
  DynArgs A{};
  void *ctl = nullptr, *tasks = nullptr, *ready = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_DYN_CTL, 256, &ctl));
  A.ctl = (DynCtl*)ctl;
  A.units = units;
  int clock_khz = 0;
  if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, ctx->device) !=
          hipSuccess || clock_khz <= 0)
    clock_khz = 100000;
  A.stall_ticks = (o->ws_stall_ms ? (uint64_t)o->ws_stall_ms : kStallMs) * (uint64_t)clock_khz;
  A.max_claim = o->ws_claim_units;
  if (o->ws) {
    A.split = o->ws_split_threshold ? (uint64_t)(o->ws_split_threshold + 63) / 64 * 64
                                    : kDefaultSplit;
    A.cap = kTaskSlots;
    AMJ_TRY(scratch(ctx, SLOT_DYN_TASKS, A.cap * kTaskWords * 4, &tasks));
    AMJ_TRY(scratch(ctx, SLOT_DYN_READY, A.cap * 4, &ready));
    A.tasks = (uint32_t*)tasks;
    A.ready = (uint32_t*)ready;
  }
  const bool mat = o->materialize != 0;
  // sized for the final pass (a count pass before it fits at least as many waves)
  const unsigned grid = (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>((units + kWavesPerBlock - 1) / kWavesPerBlock,
                            (uint64_t)dyn_occupancy_grid(P, mat, o->checksum || !mat,
                                                         ctx->props.multiProcessorCount)));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  DynCtl hc{};
  double kms = 0.0;
  uint64_t capacity = 0;
  uint32_t passes = 0;
  if (!o->dro || !mat) {
    // count pass, then an exact allocation (WS without DRO, AMHJ_WS.cuh:478-557)
    if (mat) {
      A.out_cap = 0;
      AMJ_TRY(dyn_pass(ctx, P, ProbeOut{}, false, A, grid, &hc, &kms));
      ++passes;
      capacity = hc.count;
    }
  } else {
    // DRO: output sized from a sample of units (AMHJ_DRO.cuh:590-640)
    const uint64_t samples = std::min<uint64_t>(units, 4096);
    void* sum = nullptr;
    AMJ_TRY(scratch(ctx, SLOT_SUM, 64, &sum));
    AMJ_HIP_TRY(ctx, hipMemsetAsync(sum, 0, 8, s));
    AMJ_HIP_TRY(ctx, launch_dyn_sample(P, samples, (unsigned long long*)sum, s));
    uint64_t sampled = 0;
    AMJ_HIP_TRY(ctx, hipMemcpyAsync(&sampled, sum, 8, hipMemcpyDeviceToHost, s));
    AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
    const double est = samples ? (double)sampled * (double)units / (double)samples : 0.0;
    capacity = (uint64_t)(est * 1.25) + 65536;
    if (o->dro_capacity_rows) capacity = o->dro_capacity_rows;
  }
  // outputs of many rows are reserved in chunks (see DirectStage): the allocation then has room
  // for the unused tail of every wave's last chunk, and a fill-up pass makes the rows dense
  const uint64_t waves = (uint64_t)grid * kWavesPerBlock;
  bool user_sized = o->dro && mat && o->dro_capacity_rows;
  for (int attempt = 0; attempt < 2; ++attempt) {
    ProbeOut out{};
    const uint32_t chunk = mat ? direct_chunk_rows(capacity, waves) : 0u;
    uint64_t alloc = capacity;
    if (chunk) {
      alloc = capacity + (user_sized ? 0 : waves * (uint64_t)chunk);
      alloc = (alloc + chunk - 1) / chunk * chunk;
    }
    if (mat && alloc > 0) {
      const uint64_t need = alloc * (uint64_t)P.n_attrs * 4;
      if (need > budget_left(ctx, o)) {
        set_error(ctx, AMJ_ERR_NOMEM, "result of " + std::to_string(alloc) +
                                          " tuples does not fit the HBM budget (use ooc=1 or "
                                          "materialize=0)");
        return AMJ_ERR_NOMEM;
      }
      AMJ_TRY(alloc_output(ctx, alloc, P.n_attrs, res));
      for (int a = 0; a < P.n_attrs; ++a) out.cols[a] = res->cols[a];
      out.capacity = alloc;
      if (chunk) {
        void* tails = nullptr;
        AMJ_TRY(scratch(ctx, SLOT_MISC2, alloc / chunk * 8, &tails));
        AMJ_HIP_TRY(ctx, hipMemsetAsync(tails, 0, alloc / chunk * 8, s));
        out.chunk = chunk;
        out.tails = (uint64_t*)tails;
      }
    }
    A.out_cap = out.capacity;
    const bool hash = o->checksum || !mat;
    AMJ_TRY(dyn_pass(ctx, P, out, hash, A, grid, &hc, &kms));
    ++passes;
    if (hc.out_rows <= out.capacity) {
      if (out.chunk && hc.out_rows > hc.count) {
        const uint64_t used_chunks = hc.out_rows / out.chunk;
        void *tpre = nullptr, *stmp = nullptr;
        AMJ_TRY(scratch(ctx, SLOT_MISC3, (used_chunks + 1) * 8, &tpre));
        AMJ_TRY(scratch(ctx, SLOT_SCAN_TEMP, scan_temp_elems(used_chunks) * 8, &stmp));
        AMJ_HIP_TRY(ctx, exclusive_scan_u64(out.tails, (uint64_t*)tpre, used_chunks,
                                            (uint64_t*)stmp, s));
        AMJ_HIP_TRY(ctx, launch_chunk_fill(out, P.n_attrs, out.tails, (const uint64_t*)tpre,
                                           used_chunks, out.chunk, hc.count, hc.out_rows, s));
      }
      break;
    }
    // the estimate was short: the pass counted every result, rerun into an allocation with
    // room for them (and the chunk tails)
    amj_result_free(ctx, res);
    capacity = hc.count;
    user_sized = false;
  }
  if (mat && hc.out_rows > A.out_cap) {
    set_error(ctx, AMJ_ERR_STATE, "direct output overran an allocation sized for it");
    return AMJ_ERR_STATE;
  }
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->count = hc.count;
  res->checksum = hc.csum;
  res->rows_written = mat ? hc.count : 0;
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->probe_kernel_ms = kms;
  res->probe_kernel_launches = passes;
  res->passes = passes;
  res->detail = (uint32_t)std::min<unsigned long long>(hc.pushed, 0xffffffffull);
  res->path = o->dro && mat ? (o->ws ? AMJ_PATH_WS_DRO : AMJ_PATH_DRO) : AMJ_PATH_WS;
  return AMJ_OK;
}


// ------------------------------------------------------------------------------------------
// Out of core (ooc.cuh:166-306): the probe relation stays in host memory and is streamed in
// slices through two device buffers. Slice i+1 is copied on the side stream while slice i is
// probed on the main stream (events order the reuse of a buffer); the persistent engine of
// this file probes each slice and keeps accumulating the count, the checksum and the output
// reservations in its control block, so the whole stream is enqueued without a host round
// trip. Materialising takes a counting stream first (exact allocation), then a writing one.
// ------------------------------------------------------------------------------------------
namespace {

struct OocEvents {
  hipEvent_t copied[2]{}, done[2]{};
  ~OocEvents() {
    for (int b = 0; b < 2; ++b) {
      if (copied[b]) (void)hipEventDestroy(copied[b]);
      if (done[b]) (void)hipEventDestroy(done[b]);
    }
  }
};

}  // namespace

static int ooc_stream(amj_ctx_t* ctx, const Plan& P, const ProbeOut& out, bool hash, DynArgs A,
                      uint64_t slice, int32_t* const* bufs, unsigned grid, bool reset_out,
                      DynCtl* hc, uint64_t* n_slices) {
  hipStream_t s = ctx->stream, side = ctx->side;
  OocEvents ev;
  for (int b = 0; b < 2; ++b) {
    AMJ_HIP_TRY(ctx, hipEventCreateWithFlags(&ev.copied[b], hipEventDisableTiming));
    AMJ_HIP_TRY(ctx, hipEventCreateWithFlags(&ev.done[b], hipEventDisableTiming));
  }
  if (reset_out) AMJ_HIP_TRY(ctx, hipMemsetAsync(A.ctl, 0, sizeof(DynCtl), s));
  // the side stream must not overwrite a buffer before the main stream has caught up
  AMJ_HIP_TRY(ctx, hipEventRecord(ev.done[0], s));
  AMJ_HIP_TRY(ctx, hipEventRecord(ev.done[1], s));
  const uint64_t n = P.probe_n;
  uint64_t k = 0;
  for (uint64_t r0 = 0; r0 < n; r0 += slice, ++k) {
    const int b = (int)(k & 1);
    const uint64_t rows = n - r0 < slice ? n - r0 : slice;
    AMJ_HIP_TRY(ctx, hipStreamWaitEvent(side, ev.done[b], 0));
    for (int c = 0; c < P.probe_ncols; ++c)
      if (rows)
        AMJ_HIP_TRY(ctx, hipMemcpyAsync(bufs[b] + (uint64_t)c * slice, P.probe_cols[c] + r0,
                                        rows * 4, hipMemcpyDefault, side));
    AMJ_HIP_TRY(ctx, hipEventRecord(ev.copied[b], side));
    AMJ_HIP_TRY(ctx, hipStreamWaitEvent(s, ev.copied[b], 0));
    // fresh unit and task counters for the slice; count, checksum and reservations carry on
    AMJ_HIP_TRY(ctx, hipMemsetAsync(A.ctl, 0, 3 * sizeof(unsigned long long), s));
    if (A.split) AMJ_HIP_TRY(ctx, hipMemsetAsync(A.ready, 0, A.cap * 4, s));
    Plan Q = P;
    Q.probe_on_host = 0;
    Q.probe_n = rows;
    for (int c = 0; c < P.probe_ncols; ++c) Q.probe_cols[c] = bufs[b] + (uint64_t)c * slice;
    DynArgs Ai = A;
    Ai.units = (rows + kWave - 1) / kWave;
    if (Ai.units) AMJ_HIP_TRY(ctx, launch_dyn(Q, out, hash, &Ai, grid, s));
    AMJ_HIP_TRY(ctx, hipEventRecord(ev.done[b], s));
  }
  AMJ_HIP_TRY(ctx, hipMemcpyAsync(hc, A.ctl, sizeof(DynCtl), hipMemcpyDeviceToHost, s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(s));
  AMJ_HIP_TRY(ctx, hipStreamSynchronize(side));
  *n_slices = k;
  if (hc->flags & kFlagSpin) {
    set_error(ctx, AMJ_ERR_STATE, "persistent probe: no work completed within the time limit");
    return AMJ_ERR_STATE;
  }
  return AMJ_OK;
}

int run_ooc(amj_ctx_t* ctx, Plan& P, const amj_opts_t* o, amj_result_t* res) {
  hipStream_t s = ctx->stream;
  const int nc = P.probe_ncols;
  DynArgs A{};
  void *ctl = nullptr, *tasks = nullptr, *ready = nullptr;
  AMJ_TRY(scratch(ctx, SLOT_DYN_CTL, 256, &ctl));
  A.ctl = (DynCtl*)ctl;
  int clock_khz = 0;
  if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, ctx->device) !=
          hipSuccess || clock_khz <= 0)
    clock_khz = 100000;
  A.stall_ticks = (o->ws_stall_ms ? (uint64_t)o->ws_stall_ms : kStallMs) * (uint64_t)clock_khz;
  A.max_claim = o->ws_claim_units;
  if (o->ws) {
    A.split = o->ws_split_threshold ? (uint64_t)(o->ws_split_threshold + 63) / 64 * 64
                                    : kDefaultSplit;
    A.cap = kTaskSlots;
    AMJ_TRY(scratch(ctx, SLOT_DYN_TASKS, A.cap * kTaskWords * 4, &tasks));
    AMJ_TRY(scratch(ctx, SLOT_DYN_READY, A.cap * 4, &ready));
    A.tasks = (uint32_t*)tasks;
    A.ready = (uint32_t*)ready;
  }
  // slice: the option, else a quarter of the room left split over the two buffers
  uint64_t slice = o->ooc_slice_rows;
  if (!slice) {
    const uint64_t room = budget_left(ctx, o);
    slice = room / 4 / (2ull * nc * 4);
    slice = std::min<uint64_t>(slice, 1ull << 26);
  }
  slice = std::max<uint64_t>((slice + kWave - 1) / kWave * kWave, kWave);
  void* buf[2] = {nullptr, nullptr};
  AMJ_TRY(scratch(ctx, SLOT_OOC_A, slice * nc * 4, &buf[0]));
  AMJ_TRY(scratch(ctx, SLOT_OOC_B, slice * nc * 4, &buf[1]));
  int32_t* bufs[2] = {(int32_t*)buf[0], (int32_t*)buf[1]};
  const bool mat = o->materialize != 0;
  const unsigned grid = (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>((slice / kWave + kWavesPerBlock - 1) / kWavesPerBlock,
                            (uint64_t)dyn_occupancy_grid(P, mat, o->checksum || !mat,
                                                         ctx->props.multiProcessorCount)));
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
  DynCtl hc{};
  uint64_t n_slices = 0;
  uint32_t passes = 0;
  ProbeOut out{};
  if (mat) {  // counting stream, then the exact allocation
    A.out_cap = 0;
    AMJ_TRY(ooc_stream(ctx, P, ProbeOut{}, false, A, slice, bufs, grid, true, &hc, &n_slices));
    ++passes;
    const uint64_t total = hc.count;
    if (total > 0) {
      const uint64_t need = total * (uint64_t)P.n_attrs * 4;
      if (need > budget_left(ctx, o)) {
        set_error(ctx, AMJ_ERR_NOMEM, "result of " + std::to_string(total) +
                                          " tuples does not fit the HBM budget (use "
                                          "materialize=0)");
        return AMJ_ERR_NOMEM;
      }
      AMJ_TRY(alloc_output(ctx, total, P.n_attrs, res));
      for (int a = 0; a < P.n_attrs; ++a) out.cols[a] = res->cols[a];
      out.capacity = total;
    }
    A.out_cap = total;
  }
  const bool hash = o->checksum || !mat;
  AMJ_TRY(ooc_stream(ctx, P, out, hash, A, slice, bufs, grid, true, &hc, &n_slices));
  ++passes;
  if (hc.out_rows > out.capacity) {
    set_error(ctx, AMJ_ERR_STATE, "out-of-core probe overran its exact allocation");
    return AMJ_ERR_STATE;
  }
  AMJ_HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
  AMJ_HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
  res->count = hc.count;
  res->checksum = hash ? hc.csum : 0;
  res->rows_written = out.cols[0] ? hc.count : 0;
  res->probe_ms = elapsed_ms(ctx->ev[2], ctx->ev[5]);
  res->probe_kernel_ms = 0.0;
  res->probe_kernel_launches = n_slices * passes;
  res->passes = passes;
  res->detail = (uint32_t)std::min<uint64_t>(n_slices, 0xffffffffull);
  res->path = AMJ_PATH_OOC;
  return AMJ_OK;
}

}  // namespace amj
