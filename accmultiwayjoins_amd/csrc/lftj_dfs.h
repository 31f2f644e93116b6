// Depth-first building blocks of the leapfrog triejoin shared by the unit engine (lftj.hip)
// and the persistent work-sharing engine (lftj_ws.hip): level 0 of a unit, the next 64
// candidates of a level checked against the other lists, and the sinks that receive the
// candidates that pass the last attribute (count, write at scanned positions, direct output).
#pragma once

#include "lftj_common.h"

namespace amj {

// Level 0 from root r of driver0 on every lane where `valid`.
template <int A>
__device__ __forceinline__ void level0_at(const LPlan& P, uint64_t r, bool valid, LLv<A>& l) {
  sfor<0, A>([&](auto a) { l.b[a] = 0; });
  sfor<0, kR>([&](auto q) { l.lo[q] = 0u; l.hi[q] = 0u; });
  int32_t v = 0;
  const LRel& d = P.rel[P.driver0];
  if (valid) v = d.d0[r];
  bool ok = valid;
  if (valid) ok = check_and_open(P, 0, v, -1, true, l);
  l.b[0] = v;
  choose_driver(P, 1, ok, l);
  finish(l);
}

// Level 0 of a unit: 64 consecutive level-0 values of driver0.
template <int A>
__device__ __forceinline__ void level0(const LPlan& P, uint64_t unit, LLv<A>& l) {
  const uint64_t r = unit * kWave + (uint64_t)lane_id();
  level0_at(P, r, r < P.roots, l);
}

// What happens to the candidates that pass the last attribute (`ok` per lane, bindings in b):
// counted, written at their scanned positions, or gathered for direct output.
struct CountSink {
  uint64_t count = 0;
  template <int A>
  __device__ __forceinline__ void operator()(bool ok, const int32_t (&)[A]) {
    count += (uint64_t)__popcll(__ballot(ok));
  }
};

// Count and checksum only (the work-sharing engine's passes without output).
template <bool HASH>
struct CountHashSink {
  uint64_t count = 0;
  uint64_t csum = 0;  // per-lane checksum partial
  template <int A>
  __device__ __forceinline__ void operator()(bool ok, const int32_t (&b)[A]) {
    count += (uint64_t)__popcll(__ballot(ok));
    if (HASH && ok) {
      int32_t w[kMaxAttrs];
      sfor<0, kMaxAttrs>([&](auto a) { w[a] = (a < A) ? b[(a < A) ? (int)a : 0] : 0; });
      csum += hash_tuple(w, A);
    }
  }
};

template <int MODE>
struct WriteSink {
  LOut cols;
  uint64_t out_base;  // absolute row of the next result (uniform)
  uint64_t csum;      // per-lane checksum partial
  template <int A>
  __device__ __forceinline__ void operator()(bool ok, const int32_t (&b)[A]) {
    const uint64_t m = __ballot(ok);
    if (ok) {
      if (mode_hash<MODE>()) {
        int32_t w[kMaxAttrs];
        sfor<0, kMaxAttrs>([&](auto a) { w[a] = (a < A) ? b[(a < A) ? (int)a : 0] : 0; });
        csum += hash_tuple(w, A);
      }
      if (mode_mat<MODE>()) {
        const uint64_t pos = out_base + mbcnt(m);
        sfor<0, A>([&](auto a) { cols.c[a][pos] = b[a]; });
      }
    }
    out_base += (uint64_t)__popcll(m);
  }
};

// Direct output: results gather in a per-wave LDS buffer of kLftjDirectRows rows and leave
// with one atomic reservation per full buffer; rows at or past `capacity` are not written (the
// reserved total then tells the host to run again with an exact allocation).
constexpr uint32_t kLftjDirectRows = 256;
typedef __attribute__((address_space(3))) int32_t lds_i32;

template <bool HASH>
struct DirectSink {
  LOut cols;
  uint64_t capacity;
  unsigned long long* rows;  // reserved rows of the whole launch
  lds_i32* buf;              // attribute a of row i at buf[a * kLftjDirectRows + i]
  uint32_t fill;             // rows held (uniform)
  int n_attrs;
  uint64_t csum;

  __device__ __forceinline__ void flush() {
    if (fill == 0) return;
    unsigned long long r0 = 0;
    if (lane_id() == 0)
      r0 = __hip_atomic_fetch_add(rows, (unsigned long long)fill, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t base = uniform((uint64_t)r0);
    __asm__ __volatile__("" ::: "memory");  // LDS ops of one wave complete in order
    for (uint32_t i = (uint32_t)lane_id(); i < fill; i += kWave) {
      const uint64_t pos = base + i;
      if (pos < capacity)
        sfor<0, kMaxLftjAttrs>([&](auto a) {
          if ((int)a < n_attrs) cols.c[a][pos] = buf[a * kLftjDirectRows + i];
        });
    }
    __asm__ __volatile__("" ::: "memory");
    fill = 0;
  }

  template <int A>
  __device__ __forceinline__ void operator()(bool ok, const int32_t (&b)[A]) {
    const uint64_t m = __ballot(ok);
    const uint32_t n = (uint32_t)__popcll(m);
    if (n == 0) return;
    if (HASH && ok) {
      int32_t w[kMaxAttrs];
      sfor<0, kMaxAttrs>([&](auto a) { w[a] = (a < A) ? b[(a < A) ? (int)a : 0] : 0; });
      csum += hash_tuple(w, A);
    }
    if (fill + n > kLftjDirectRows) flush();
    __asm__ __volatile__("" ::: "memory");
    if (ok) {
      const uint32_t slot = fill + mbcnt(m);
      sfor<0, A>([&](auto a) { buf[a * kLftjDirectRows + slot] = b[a]; });
    }
    __asm__ __volatile__("" ::: "memory");
    fill += n;
  }
};

// Candidates [c, c + 64) of the flattened level `par` (those below `end` are valid) checked
// for attribute i: ch gets the parent's bindings and child ranges, b[i] = the candidate, and
// the ranges of the relations rooted at i opened. Returns whether the candidate passed.
template <int A, int i>
__device__ __forceinline__ bool lchild(const LPlan& P, const LLv<A>& par, uint32_t c,
                                       uint32_t end, LLv<A>& ch) {
  const uint32_t g = c + (uint32_t)lane_id();
  const bool valid = g < end;
  const int p = (int)parent_of_slots(par.excl, par.dlen, c);
  const uint32_t k = g - (uint32_t)__shfl(par.excl, p, kWave);
  const int32_t* src = (const int32_t*)__shfl((uint64_t)par.dptr, p, kWave);
  const int drel = __shfl(par.drel, p, kWave);
  sfor<0, A>([&](auto a) { ch.b[a] = __shfl(par.b[a], p, kWave); });
  sfor<0, kR>([&](auto q) {
    ch.lo[q] = __shfl(par.lo[q], p, kWave);
    ch.hi[q] = __shfl(par.hi[q], p, kWave);
  });
  const int32_t v = valid ? src[k] : 0;
  const bool drel_root = P.child_mask[i] == 0u;
  const bool ok = valid && check_and_open(P, i, v, drel, drel_root, ch);
  ch.b[i] = v;
  return ok;
}

template <int A, int i, typename Sink>
__device__ __forceinline__ void descend(const LPlan& P, const LLv<A>& par, Sink& sink) {
  for (uint32_t c = 0; c < par.T; c += (uint32_t)kWave) {
    LLv<A> ch;
    const bool ok = lchild<A, i>(P, par, c, par.T, ch);
    if constexpr (i == A - 1) {
      sink(ok, ch.b);
    } else {
      choose_driver(P, i + 1, ok, ch);
      finish(ch);
      if (ch.T) descend<A, i + 1>(P, ch, sink);
    }
  }
}

}  // namespace amj
