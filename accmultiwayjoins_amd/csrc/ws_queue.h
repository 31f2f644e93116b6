// Work-sharing substrate of the persistent engines: the AMHJ probe (dynamic.hip) and the
// leapfrog triejoin (lftj_ws.hip) run the same protocol over their own task records.
//
// Reference: TaskBook (MHJ-GPU/TaskBook.cuh:13-109) appends tasks with atomicAdd, GCQueue
// (common-utils/cuda/GCQueue.cuh:30-107) hands them out through CAS on tagged slots, and
// CBarrier (MHJ-GPU/CBarrier.cuh:9-46) ends the kernel when every warp is idle; none of them
// checks capacity. Here:
//   * persistent waves claim units (64 probe rows / 64 roots) from a counter in guided chunks;
//   * task slots are reserved with one fetch-add on the queue tail, clamped to the capacity
//     (requests past the end get nothing, and the wave then keeps its work);
//   * a task record is published by its own ready flag (release store after the record), and
//     the taker acquires it; every control-block counter is otherwise relaxed;
//   * a wave leaves when no unit is left, no published task is unclaimed and no wave holds
//     claimed work; every wait is bounded by a wall-clock limit without progress (error flag,
//     never a hang), and working waves report progress as they go.
#pragma once

#include "amj_internal.h"
#include "ws_split.h"

namespace amj {

// Control block of one launch (zeroed by the host).
struct WsCtl {
  unsigned long long next_unit;  // units claimed
  unsigned long long q_tail;     // task slots reserved (may pass cap: clamp on read)
  unsigned long long q_head;     // task slots claimed
  unsigned long long out_rows;   // output rows reserved (the results, plus unused chunk tails)
  unsigned long long count;      // results
  unsigned long long csum;       // checksum, sum of the waves' partials mod 2^64
  unsigned long long pushed;     // tasks published
  unsigned long long progress;   // units and tasks completed, plus heartbeats
  unsigned int busy;             // waves holding a unit or a task
  unsigned int flags;            // 1: a wait hit the time limit
};

constexpr uint64_t kMaxClaim = 64;  // units claimed at once by a wave
constexpr uint64_t kMinClaim = 4;   // ... and at least, in launches of many units per wave
constexpr uint32_t kFlagSpin = 1u;
// Frontier steps of 64 children between two progress reports of a working wave: idle waves
// then see progress while a wave walks a long level alone, not only when a unit or a task ends.
constexpr uint32_t kBeatSteps = 4096;

struct WsArgs {
  WsCtl* ctl;
  uint32_t* tasks;       // cap records of the engine's record size
  uint32_t* ready;       // cap flags
  uint64_t cap;          // task slots
  uint64_t split;        // children per task (multiple of 64); 0 disables sharing
  uint64_t units;        // units to claim
  uint64_t out_cap;      // output rows available
  uint64_t stall_ticks;  // wall-clock ticks without progress before a wave gives up
  uint64_t max_claim;    // most units claimed at once (0: kMaxClaim)
};

// Per-wave sharing state.
struct WsWave {
  bool qfull;      // the queue has no room left in this launch: stop asking
  uint32_t iters;  // frontier steps taken (uniform), for the heartbeat
};

// Every access to the control block is relaxed: an acquire at agent scope invalidates the
// CU's vector caches, and idle waves polling with acquires (or workers finishing chunks with
// acquire-release updates) evicted the data the working waves on the same CU were reading.
// No data is published through these counters: a task record is published by its own flag,
// and a wave that leaves early on a stale view only gives up parallelism.
__device__ __forceinline__ unsigned long long ld_rlx(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned int ld_rlx(const unsigned int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T add_relaxed(T* p, T v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-uniform result of an atomic done by lane 0.
__device__ __forceinline__ uint64_t lane0_u64(uint64_t v) {
  return uniform(__shfl(v, 0, kWave));
}

// One frontier step of a working wave; every kBeatSteps steps it reports progress. A launch
// without a control block (the size samples) only counts.
__device__ __forceinline__ void ws_step(const WsArgs& A, WsWave& W) {
  W.iters += 1u;
  if (A.ctl && (W.iters % kBeatSteps) == 0u && lane_id() == 0)
    add_relaxed(&A.ctl->progress, 1ull);
}

// Reserve up to k task slots with one fetch-add; returns the first slot, the number granted in
// *got, and sets *full once the queue has no room left.
__device__ __forceinline__ uint64_t reserve_tasks(const WsArgs& A, uint64_t k, uint64_t* got,
                                                  bool* full) {
  uint64_t first = 0, n = 0, f = 0;
  if (lane_id() == 0) {
    const unsigned long long tl = add_relaxed(&A.ctl->q_tail, (unsigned long long)k);
    n = ws_granted(tl, k, A.cap);
    first = n ? tl : 0;
    f = tl + k >= A.cap ? 1 : 0;
  }
  *got = lane0_u64(n);
  *full = lane0_u64(f) != 0;
  return lane0_u64(first);
}

// Task slots published or being published: q_tail clamped to the queue's capacity.
__device__ __forceinline__ unsigned long long queue_end(const WsArgs& A) {
  const unsigned long long tl = ld_rlx(&A.ctl->q_tail);
  return tl < A.cap ? tl : A.cap;
}

// Flag of a written record: the record is visible device-wide before its flag.
__device__ __forceinline__ void publish_task(const WsArgs& A, uint64_t slot) {
  __threadfence();
  if (lane_id() == 0)
    __hip_atomic_store(&A.ready[slot], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Claim a published task; -1: none available, -2: its record never became ready.
__device__ __forceinline__ int64_t try_pop(const WsArgs& A) {
  int64_t r = -1;
  if (lane_id() == 0) {
    unsigned long long h = ld_rlx(&A.ctl->q_head);
    const unsigned long long tl = queue_end(A);
    if (h < tl && __hip_atomic_compare_exchange_strong(&A.ctl->q_head, &h, h + 1,
                                                       __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT)) {
      r = (int64_t)h;
      const uint64_t t0 = wall_clock64();
      while (__hip_atomic_load(&A.ready[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
        if (wall_clock64() - t0 > A.stall_ticks) {
          __hip_atomic_fetch_or(&A.ctl->flags, kFlagSpin, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
          r = -2;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
  }
  r = (int64_t)lane0_u64((uint64_t)r);
  if (r >= 0)  // every lane acquires the record (its flag is set by now)
    (void)__hip_atomic_load(&A.ready[r], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
  return r;
}

// The persistent loop of a wave: claim units in guided chunks and run them
// (run_units(first, end)), then take tasks (run_task(slot)) until the launch has drained.
template <typename RunUnits, typename RunTask>
__device__ __forceinline__ void ws_serve(const WsArgs& A, RunUnits&& run_units,
                                         RunTask&& run_task) {
  uint64_t seen = ~0ull, since = 0;  // progress counter last seen, and when it last moved
  for (;;) {
    // busy counts waves holding claimed work; it only keeps idle waves around to help (a wave
    // that leaves early costs parallelism, never work: a pusher always serves its own tasks).
    // Guided chunks of at most 64 units: with 64 or more units per wave, half the remaining
    // units per wave and at least 4, so cheap units do not pay shared-counter updates of their
    // own (every claim updates one counter, and the updates serialise: a streamed 1.5 G-row
    // probe of 23 slices of 1 M units took 316 ms with a quarter per wave and single-unit
    // claims at the end, 224 ms with half and at least 4, where its PCIe copies alone take
    // about 215 ms); with fewer (the triejoin's units of 64 roots, heaviest first), a quarter
    // and at least 1: 4 units at a time took the 4-cliques of R-MAT 20 from 290 to 860 ms.
    uint64_t u = ~0ull, nu = 0;
    if (lane_id() == 0) {
      const unsigned long long seen_u = ld_rlx(&A.ctl->next_unit);
      if (seen_u < A.units) {
        const uint64_t waves = (uint64_t)gridDim.x * kWavesPerBlock;
        const uint64_t most = A.max_claim ? A.max_claim : kMaxClaim;
        const bool many = A.units >= 64 * waves;
        const uint64_t least = (many && most >= kMinClaim) ? kMinClaim : 1;
        uint64_t k = (A.units - seen_u) / ((many ? 2 : 4) * waves);
        k = k < least ? least : (k > most ? most : k);
        u = add_relaxed(&A.ctl->next_unit, (unsigned long long)k);
        nu = k;
      }
    }
    u = lane0_u64(u);
    nu = lane0_u64(nu);
    if (u < A.units) {
      const uint64_t ue = u + nu < A.units ? u + nu : A.units;
      // busy and progress only serve the waits for tasks: without sharing a wave leaves when
      // the units run out, and the claim's one counter update is all it pays (the counters
      // share a cache line, whose updates serialise: 1 M units of a 64 M-row probe slice took
      // 16.4 ms with three updates per claim, against 6.8 ms for the same rows in one launch)
      if (A.split && lane_id() == 0) add_relaxed(&A.ctl->busy, 1u);
      run_units(u, ue);
      if (A.split && lane_id() == 0) {
        add_relaxed(&A.ctl->progress, 1ull);
        add_relaxed(&A.ctl->busy, ~0u);
      }
      continue;
    }
    if (A.split == 0) break;  // without sharing there are no tasks to wait for
    const int64_t slot = try_pop(A);
    if (slot >= 0) {
      if (lane_id() == 0) add_relaxed(&A.ctl->busy, 1u);
      run_task((uint64_t)slot);
      if (lane_id() == 0) {
        add_relaxed(&A.ctl->progress, 1ull);
        add_relaxed(&A.ctl->busy, ~0u);
      }
      continue;
    }
    if (slot == -2) break;
    uint32_t done = 0;
    if (lane_id() == 0)
      done = (ld_rlx(&A.ctl->busy) == 0u && ld_rlx(&A.ctl->q_head) >= queue_end(A)) ? 1u : 0u;
    if (uniform(__shfl(done, 0, kWave))) break;
    // no work here while other waves hold some: give up only if nothing moves for long
    const uint64_t prog = lane0_u64(lane_id() == 0 ? ld_rlx(&A.ctl->progress) : 0ull);
    const uint64_t now = wall_clock64();
    if (prog != seen) {
      seen = prog;
      since = now;
    } else if (now - since > A.stall_ticks) {
      if (lane_id() == 0)
        __hip_atomic_fetch_or(&A.ctl->flags, kFlagSpin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    __builtin_amdgcn_s_sleep(64);
  }
}

// Count and checksum of a wave added to the control block when it leaves.
__device__ __forceinline__ void ws_accumulate(const WsArgs& A, uint64_t count, uint64_t csum) {
  const uint64_t cs = wave_sum_u64(csum);
  if (lane_id() == 0) {
    add_relaxed(&A.ctl->count, (unsigned long long)count);
    add_relaxed(&A.ctl->csum, (unsigned long long)cs);
  }
}

}  // namespace amj
