// ALFTJ with work sharing (opts.ws on the DFS): persistent waves, long intersections split into
// tasks that idle waves take over.
//
// Reference: ALFTJ_count / ALFTJ_write with WS (LFTJ-GPU/ALFTJ.cuh:464-499): when the smaller of
// the two lists a warp intersects at an attribute is longer than BUC_THRESHOLD (1024,
// ALFTJ.cuh:22), the warp cuts it into BUC_THRESHOLD-element pieces, bounds each piece's range
// in the larger list with galloping searches, and pushes (piece, bound range, bindings, level)
// to a TaskBook served through a GCQueue; idle warps pop the tasks and continue the DFS from
// that level.
//
// Here a wave's level is the flattened candidate lists of its 64 frontier entries (each entry's
// shortest participating list, lftj_dfs.h), so a long intersection is a long run of candidates
// of that level. A wave at candidate c of a level with more than `split` + 64 candidates left
// keeps the next `split` and publishes the rest as tasks of `split` candidates each: the task is
// the level's 64-lane frontier (bindings, every relation's child range, the candidate list of
// each entry and its prefix) plus the candidate range, so the wave that takes it checks exactly
// the candidates the owner would have. For a query of no compiled shape every level can be
// split, the last one included (the last attribute's candidates still have to be intersected).
// For a triangle or a 4-clique (compiled shapes, lftj_common.h) the levels that the unit engine
// runs cooperatively (the last attribute; for a 4-clique the last two) take the same
// cooperative paths here on the task's candidate range, and the levels above them are split.
// Each candidate is checked against the other lists with a search over that list's whole child
// range rather than a galloping-bounded sub-range. The queue, the ready flags, the unit claims
// and the termination are the protocol of the persistent AMHJ engine (ws_queue.h); results are
// counted and hashed, or written through per-wave LDS buffers with chunked reservations (direct
// output, DirectSink).
#include <type_traits>

#include "lftj_dfs.h"
#include "ws_queue.h"

namespace amj {

// Task record: kLTaskHdr header words (level, first and end candidate, level total), then
// kLTaskLaneWords words per lane: bindings, the child range [lo, hi) of every relation, the
// candidate list (pointer, length, relation) and its exclusive prefix.
constexpr int kLTaskHdr = 16;
constexpr int kLTaskLaneWords = kMaxLftjAttrs + 2 * kR + 5;
constexpr int kLTaskWords = kLTaskHdr + kLTaskLaneWords * kWave;
constexpr int kLwB = 0, kLwLo = kMaxLftjAttrs, kLwHi = kLwLo + kR, kLwPtr = kLwHi + kR,
              kLwLen = kLwPtr + 2, kLwRel = kLwLen + 1, kLwExcl = kLwRel + 1;
static_assert(kLwExcl + 1 == kLTaskLaneWords, "task record layout");

uint64_t lftj_ws_task_bytes() { return (uint64_t)kLTaskWords * 4; }

template <int A>
__device__ __forceinline__ void write_ltask(const WsArgs& W, uint64_t slot, int level,
                                            const LLv<A>& l, uint64_t c0, uint64_t c1) {
  uint32_t* rec = W.tasks + slot * (uint64_t)kLTaskWords;
  const int lane = lane_id();
  uint32_t* lw = rec + kLTaskHdr;
  sfor<0, kMaxLftjAttrs>([&](auto a) {
    lw[(kLwB + a) * kWave + lane] = (a < A) ? (uint32_t)l.b[(a < A) ? (int)a : 0] : 0u;
  });
  sfor<0, kR>([&](auto q) {
    lw[(kLwLo + q) * kWave + lane] = l.lo[q];
    lw[(kLwHi + q) * kWave + lane] = l.hi[q];
  });
  const uint64_t ptr = (uint64_t)l.dptr;
  lw[(kLwPtr + 0) * kWave + lane] = (uint32_t)ptr;
  lw[(kLwPtr + 1) * kWave + lane] = (uint32_t)(ptr >> 32);
  lw[kLwLen * kWave + lane] = l.dlen;
  lw[kLwRel * kWave + lane] = (uint32_t)l.drel;
  lw[kLwExcl * kWave + lane] = l.excl;
  if (lane == 0) {
    rec[0] = (uint32_t)level;
    rec[1] = (uint32_t)c0;
    rec[2] = (uint32_t)c1;
    rec[3] = l.T;
  }
  publish_task(W, slot);
}

template <int A>
__device__ __forceinline__ void read_ltask(const WsArgs& W, uint64_t slot, int& level,
                                           uint32_t& c0, uint32_t& c1, LLv<A>& l) {
  const uint32_t* rec = W.tasks + slot * (uint64_t)kLTaskWords;
  const int lane = lane_id();
  level = (int)uniform(rec[0]);
  c0 = uniform(rec[1]);
  c1 = uniform(rec[2]);
  l.T = uniform(rec[3]);
  const uint32_t* lw = rec + kLTaskHdr;
  sfor<0, A>([&](auto a) { l.b[a] = (int32_t)lw[(kLwB + a) * kWave + lane]; });
  sfor<0, kR>([&](auto q) {
    l.lo[q] = lw[(kLwLo + q) * kWave + lane];
    l.hi[q] = lw[(kLwHi + q) * kWave + lane];
  });
  const uint64_t ptr = ((uint64_t)lw[(kLwPtr + 1) * kWave + lane] << 32) |
                       lw[(kLwPtr + 0) * kWave + lane];
  l.dptr = (const int32_t*)ptr;
  l.dlen = lw[kLwLen * kWave + lane];
  l.drel = (int32_t)lw[kLwRel * kWave + lane];
  l.excl = lw[kLwExcl * kWave + lane];
}

// Levels that the unit engine runs cooperatively for a compiled shape.
template <int A, int i, uint64_t SH>
constexpr bool coop_level() {
  return SH != 0 && (i == A - 1 || (SH == kShapeFour && i == 2));
}

// Candidates [c0, c1) of level `par` for attribute i, sharing the tail when it is long.
template <int A, int i, uint64_t SH, typename Sink>
__device__ __forceinline__ void ws_ldescend(const LPlan& P, const LLv<A>& par, uint32_t c0,
                                            uint32_t c1, Sink& sink, const WsArgs& W,
                                            WsWave& q, lds_i32* stage) {
  if constexpr (coop_level<A, i, SH>()) {
    ws_step(W, q);
    LLv<A> r = par;
    restrict_range(r, c0, c1);
    if (r.T) descend<A, i, SH>(P, r, sink, stage);
    return;
  }
  uint64_t end = c1;
  for (uint64_t c = c0; c < end; c += kWave) {
    ws_step(W, q);
    if (!q.qfull && ws_should_split(c, end, W.split)) {
      const uint64_t k = ws_pieces(c, end, W.split);
      uint64_t got = 0;
      bool full = false;
      const uint64_t slot = reserve_tasks(W, k, &got, &full);
      if (full) q.qfull = true;  // slots are never reused within a launch: stop asking
      if (got) {
        for (uint64_t j = 0; j < got; ++j) {
          uint64_t p0 = 0, p1 = 0;
          ws_piece(c, end, W.split, k, got, j, &p0, &p1);
          write_ltask<A>(W, slot + j, i, par, p0, p1);
        }
        if (lane_id() == 0) add_relaxed(&W.ctl->pushed, (unsigned long long)got);
        end = ws_keep_end(c, end, W.split, k, got);
      }
    }
    LLv<A> ch;
    const bool ok = lchild<A, i, SH>(P, par, (uint32_t)c, (uint32_t)end, ch);
    if constexpr (i == A - 1) {
      sink(ok, ch.b);
    } else {
      choose_driver<A, SH>(P, i + 1, ok, ch);
      finish(ch);
      if (ch.T) ws_ldescend<A, i + 1, SH>(P, ch, 0u, ch.T, sink, W, q, stage);
    }
  }
}

// One persistent pass: MODE_COUNT / MODE_HASH count (and hash) the results, MODE_WRITE /
// MODE_WRITE_HASH write them with direct output into cols (W.out_cap rows, reserved through
// ctl->out_rows `chunk` rows at a time, unused chunk tails recorded in tails; chunk 0: one
// reservation per buffer). Units are taken in `order` (nullptr: number order).
template <int A, uint64_t SH, int MODE>
__global__ __launch_bounds__(kBlock) void lftj_ws_kernel(const LPlan P, const LOut cols,
                                                         const WsArgs W,
                                                         const uint64_t* __restrict__ order,
                                                         uint32_t chunk,
                                                         uint64_t* __restrict__ tails) {
  constexpr bool MAT = mode_mat<MODE>();
  constexpr bool HASH = mode_hash<MODE>();
  __shared__ int32_t stage_mem[MAT ? kWavesPerBlock * A * kLftjDirectRows : 1];
  __shared__ int32_t coop_mem[SH != 0 ? kWavesPerBlock * stage_ints<SH>() : 1];
  lds_i32* stage = SH != 0 ? (lds_i32*)&coop_mem[(threadIdx.x / kWave) * stage_ints<SH>()]
                           : nullptr;
  using Sink = std::conditional_t<MAT, DirectSink<A, HASH>, CountHashSink<HASH>>;
  Sink sink{};
  if constexpr (MAT) {
    sink.cols = cols;
    sink.capacity = W.out_cap;
    sink.rows = &W.ctl->out_rows;
    sink.buf = (lds_i32*)&stage_mem[(threadIdx.x / kWave) * A * kLftjDirectRows];
    sink.fill = 0u;
    sink.csum = 0;
    sink.chunk = chunk;
  }
  WsWave q{false, 0u};
  ws_serve(
      W,
      [&](uint64_t u, uint64_t ue) {
        for (uint64_t k = u; k < ue; ++k) {
          LLv<A> l0;
          level0<A, SH>(P, unit_of(order, k), l0);
          if (l0.T) ws_ldescend<A, 1, SH>(P, l0, 0u, l0.T, sink, W, q, stage);
        }
      },
      [&](uint64_t slot) {
        int level = 0;
        uint32_t c0 = 0, c1 = 0;
        LLv<A> l;
        read_ltask<A>(W, slot, level, c0, c1, l);
        sfor<1, A>([&](auto ii) {
          if ((int)ii == level) ws_ldescend<A, ii, SH>(P, l, c0, c1, sink, W, q, stage);
        });
      });
  if constexpr (MAT) {
    sink.flush();
    sink.tail_of(tails);
  }
  ws_accumulate(W, sink.count, sink.csum);
}

template <int A, uint64_t SH, int MODE>
static int ws_occupancy() {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lftj_ws_kernel<A, SH, MODE>, kBlock,
                                                   0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 2;
  return per_cu;
}

template <int V>
using IntC = std::integral_constant<int, V>;
template <uint64_t V>
using ShapeC = std::integral_constant<uint64_t, V>;

// f(IntC<A>, ShapeC<SH>, IntC<MODE>) for the plan's attribute count and compiled shape.
template <typename F>
static bool ws_dispatch(const LPlan& P, int mode, F&& f) {
  const uint64_t sh = lftj_static_shape(P);
  auto with_mode = [&](auto a, auto shc) {
    switch (mode) {
      case MODE_COUNT: f(a, shc, IntC<MODE_COUNT>{}); return true;
      case MODE_HASH: f(a, shc, IntC<MODE_HASH>{}); return true;
      case MODE_WRITE: f(a, shc, IntC<MODE_WRITE>{}); return true;
      case MODE_WRITE_HASH: f(a, shc, IntC<MODE_WRITE_HASH>{}); return true;
      default: return false;
    }
  };
  switch (P.n_attrs) {
    case 2: return with_mode(IntC<2>{}, ShapeC<0>{});
    case 3:
      return sh == kShapeTri ? with_mode(IntC<3>{}, ShapeC<kShapeTri>{})
                             : with_mode(IntC<3>{}, ShapeC<0>{});
    case 4:
      return sh == kShapeFour ? with_mode(IntC<4>{}, ShapeC<kShapeFour>{})
                              : with_mode(IntC<4>{}, ShapeC<0>{});
    default: return false;
  }
}

int lftj_ws_blocks_per_cu(const LPlan& P, int mode) {
  int per_cu = 2;
  ws_dispatch(P, mode, [&](auto a, auto sh, auto m) {
    per_cu = ws_occupancy<decltype(a)::value, decltype(sh)::value, decltype(m)::value>();
  });
  return per_cu;
}

hipError_t launch_lftj_ws(const LPlan& P, const LOut& cols, int mode, const void* args,
                          const uint64_t* order, unsigned grid, hipStream_t s, uint32_t chunk,
                          uint64_t* tails) {
  const WsArgs& W = *(const WsArgs*)args;
  const bool ok = ws_dispatch(P, mode, [&](auto a, auto sh, auto m) {
    hipLaunchKernelGGL(
        (lftj_ws_kernel<decltype(a)::value, decltype(sh)::value, decltype(m)::value>),
        dim3(grid), dim3(kBlock), 0, s, P, cols, W, order, chunk, tails);
  });
  if (!ok) return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace amj
