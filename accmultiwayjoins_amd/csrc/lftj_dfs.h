// Depth-first building blocks of the leapfrog triejoin shared by the unit engine (lftj.hip)
// and the persistent work-sharing engine (lftj_ws.hip): level 0 of a unit, the next 64
// candidates of a level checked against the other lists, and the sinks that receive the
// candidates that pass the last attribute (count, write at scanned positions, direct output).
#pragma once

#include "lftj_common.h"

namespace amj {

// Level 0 from root r of driver0 on every lane where `valid`.
template <int A, uint64_t SH = 0>
__device__ __forceinline__ void level0_at(const LPlan& P, uint64_t r, bool valid, LLv<A>& l) {
  sfor<0, A>([&](auto a) { l.b[a] = 0; });
  sfor<0, kR>([&](auto q) { l.lo[q] = 0u; l.hi[q] = 0u; });
  int32_t v = 0;
  const LRel& d = P.rel[P.driver0];
  if (valid) v = d.d0[r];
  bool ok = valid;
  if (valid) {
    if constexpr (SH != 0) ok = check_and_open_fixed<A, SH, 0>(P, v, -1, l);
    else ok = check_and_open(P, 0, v, -1, true, l);
  }
  l.b[0] = v;
  choose_driver<A, SH>(P, 1, ok, l);
  finish(l);
}

// The entries of `l` cut to candidates [c0, c1) of the flattened level.
template <int A>
__device__ __forceinline__ void restrict_range(LLv<A>& l, uint32_t c0, uint32_t c1) {
  const uint32_t b = l.excl > c0 ? l.excl : c0;
  const uint32_t e = l.excl + l.dlen < c1 ? l.excl + l.dlen : c1;
  const uint32_t n = e > b ? e - b : 0u;
  l.dptr = n ? l.dptr + (b - l.excl) : l.dptr;
  l.dlen = n;
  finish(l);
}

// Level 0 of a unit: P.rpu (0: 64) consecutive level-0 values of driver0, one per lane.
template <int A, uint64_t SH = 0>
__device__ __forceinline__ void level0(const LPlan& P, uint64_t unit, LLv<A>& l) {
  const uint32_t rpu = P.rpu ? P.rpu : (uint32_t)kWave;
  const uint32_t lane = (uint32_t)lane_id();
  const uint64_t r = unit * rpu + lane;
  level0_at<A, SH>(P, r, lane < rpu && r < P.roots, l);
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

// Direct output: results gather in a per-wave LDS buffer of kLftjDirectRows rows; rows at or
// past `capacity` are not written (the reserved total then tells the host to run again with a
// larger allocation).
//   * chunk = 0: one atomic reservation on the launch's row counter per full buffer.
//   * chunk > 0: rows are reserved `chunk` at a time (a multiple of kLftjDirectRows) and a wave
//     fills its chunk buffer by buffer, so the shared counter sees one update per `chunk` rows
//     instead of one per buffer (one counter updated by every wave serialises them: about 14 ns
//     per update on this device, which bounded a whole triangle pass). Only a wave's last chunk
//     is left partly unused; the wave records that tail (`tail_of`) and a fill-up pass moves the
//     rows found past the result count into the tails below it (launch_chunk_fill).
constexpr uint32_t kLftjDirectRows = 256;
typedef __attribute__((address_space(3))) int32_t lds_i32;

// The checksum is taken from the buffer when it is flushed, with every lane busy, rather than
// per result on the lanes whose candidate passed.
template <int A, bool HASH>
struct DirectSink {
  LOut cols;
  uint64_t capacity;
  unsigned long long* rows;  // rows reserved by the whole launch
  lds_i32* buf;              // attribute a of row i at buf[a * kLftjDirectRows + i]
  uint32_t fill;             // rows held (uniform)
  uint64_t csum;
  uint32_t chunk = 0;        // rows per reservation (0: one reservation per buffer)
  uint64_t cur = ~0ull;      // first row of the wave's current chunk (~0: none yet)
  uint32_t used = 0;         // rows of the current chunk taken
  uint64_t count = 0;        // rows passed to the sink (uniform)

  __device__ __forceinline__ void flush() {
    if (fill == 0) return;
    // rows [0, na) of the buffer go to base_a + i, rows [na, fill) to base_b + i - na
    uint64_t base_a = 0, base_b = 0;
    uint32_t na = fill;
    if (chunk) {
      const uint32_t rem = cur == ~0ull ? 0u : chunk - used;
      na = fill < rem ? fill : rem;
      base_a = cur + used;
      if (fill > rem) {
        unsigned long long r0 = 0;
        if (lane_id() == 0)
          r0 = __hip_atomic_fetch_add(rows, (unsigned long long)chunk, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
        base_b = uniform((uint64_t)r0);
        cur = base_b;
        used = fill - na;
      } else {
        used += fill;
      }
    } else {
      unsigned long long r0 = 0;
      if (lane_id() == 0)
        r0 = __hip_atomic_fetch_add(rows, (unsigned long long)fill, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
      base_a = uniform((uint64_t)r0);
    }
    count += fill;
    __asm__ __volatile__("" ::: "memory");  // LDS ops of one wave complete in order
    for (uint32_t i = (uint32_t)lane_id(); i < fill; i += kWave) {
      const uint64_t pos = i < na ? base_a + i : base_b + (i - na);
      int32_t w[kMaxAttrs];
      sfor<0, kMaxAttrs>([&](auto a) {
        if constexpr ((int)a < A) w[a] = buf[(int)a * kLftjDirectRows + i];
        else w[a] = 0;
      });
      if (HASH) csum += hash_tuple(w, A);
      if (pos < capacity) sfor<0, A>([&](auto a) { cols.c[a][pos] = w[a]; });
    }
    __asm__ __volatile__("" ::: "memory");
    fill = 0;
  }

  // Chunked, after the last flush: the unused tail of the wave's last chunk, recorded by lane 0
  // as tail[chunk index] when the chunk lies inside the allocation.
  __device__ __forceinline__ void tail_of(uint64_t* tail) const {
    if (chunk && cur != ~0ull && lane_id() == 0 && cur + chunk <= capacity)
      tail[cur / chunk] = chunk - used;
  }

  __device__ __forceinline__ void operator()(bool ok, const int32_t (&b)[A]) {
    const uint64_t m = __ballot(ok);
    const uint32_t n = (uint32_t)__popcll(m);
    if (n == 0) return;
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
template <int A, int i, uint64_t SH = 0>
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
  const int32_t v = valid ? ld_at<false>(src, k) : 0;
  bool ok = valid;
  if constexpr (SH != 0) {
    ok = ok && check_and_open_fixed<A, SH, i>(P, v, drel, ch);
  } else {
    const bool drel_root = P.child_mask[i] == 0u;
    ok = ok && check_and_open(P, i, v, drel, drel_root, ch);
  }
  ch.b[i] = v;
  return ok;
}

// Last attribute of a compiled shape. An entry whose driver list holds at least
// kCoopMinCands candidates and whose other child lists (every child list of the attribute but
// the driver's) hold at most kStageInts values together is run by the whole wave: the other
// lists are copied to the wave's LDS stage once, and each chunk of 64 candidates is checked
// there by binary searches of uniform length (one global load per candidate instead of a
// chain of dependent ones). The other entries go through the flattened enumeration.
constexpr uint32_t kCoopMinCands = 32;
constexpr uint32_t kStageInts = 1024;  // per wave
// 4-clique pairs (descend_clique_pairs): the other list of an entry (at most kCliqueOther
// values) and the entry's intersection after it, in a stage of 2 * kCliqueOther values.
constexpr uint32_t kCliqueOther = 768;
template <uint64_t SH>
constexpr uint32_t stage_ints() {
  return SH == kShapeFour ? 2 * kCliqueOther : kStageInts;
}

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
  return (uint64_t)readlane_u32((uint32_t)v, lane) |
         ((uint64_t)readlane_u32((uint32_t)(v >> 32), lane) << 32);
}

// Are v[j] in the ascending run st[0, n) of LDS (n the same on every lane)? The K searches
// advance together, so K reads are in flight at every step; ok[j] is cleared for a miss. The
// position is kept as an LDS pointer, so a step is one add, one read, one compare and one
// select per search.
template <int K>
__device__ __forceinline__ void lds_has(const lds_i32* st, uint32_t n, const int32_t (&v)[K],
                                        bool (&ok)[K]) {
  if (n == 0) {
    sfor<0, K>([&](auto j) { ok[j] = false; });
    return;
  }
  const lds_i32* b[K];
  sfor<0, K>([&](auto j) { b[j] = st; });
  while (n > 1) {
    const uint32_t half = n >> 1;
    const lds_i32* c[K];
    int32_t x[K];
    sfor<0, K>([&](auto j) {
      c[j] = b[j] + half;
      x[j] = *c[j];
    });
    sfor<0, K>([&](auto j) { b[j] = x[j] <= v[j] ? c[j] : b[j]; });
    n -= half;
  }
  sfor<0, K>([&](auto j) { ok[j] = ok[j] && *b[j] == v[j]; });
}

// Chunks of 64 candidates a wave checks at once in the cooperative last level.
constexpr int kCoopChunks = 4;

template <int A, int I, uint64_t SH, typename Sink>
__device__ __forceinline__ void descend_last_fixed(const LPlan& P, const LLv<A>& par, Sink& sink,
                                                   lds_i32* stage) {
  constexpr uint32_t cm = (uint32_t)(SH >> (12 * I + 6)) & 63u;
  constexpr int nc = bit_count(cm);
  static_assert(nc >= 2, "the last attribute has a driver and at least one list to check");
  constexpr int no = nc - 1;
  constexpr int K = kCoopChunks;
  // this lane's other lists, as in check_and_open_fixed
  int di = 0;
  sfor<1, nc>([&](auto j) {
    if (par.drel == nth_bit(cm, (int)j)) di = (int)j;
  });
  const int32_t* optr[no];
  uint32_t olen[no];
  uint32_t total = 0;
  sfor<0, no>([&](auto k) {
    constexpr int ra = nth_bit(cm, (int)k), rb = nth_bit(cm, (int)k + 1);
    const bool next = (int)k >= di;
    const uint32_t lo = next ? par.lo[rb] : par.lo[ra];
    const uint32_t hi = next ? par.hi[rb] : par.hi[ra];
    optr[k] = (next ? P.rel[rb].d1 : P.rel[ra].d1) + lo;
    olen[k] = hi - lo;
    total += hi - lo;
  });
  const bool coop = par.dlen >= kCoopMinCands && total <= kStageInts;

  // entries with few candidates (or long other lists): flattened
  {
    LLv<A> sm = par;
    sm.dlen = coop ? 0u : par.dlen;
    finish(sm);
    for (uint32_t c = 0; c < sm.T; c += (uint32_t)kWave) {
      LLv<A> ch;
      const bool ok = lchild<A, I, SH>(P, sm, c, sm.T, ch);
      sink(ok, ch.b);
    }
  }

  // the others, one entry at a time; consecutive entries often check the same list (children
  // of one parent value), which then stays staged
  uint64_t m = __ballot(coop);
  const uint32_t lane = (uint32_t)lane_id();
  uint64_t staged[no];
  uint32_t soff[no], slen[no];
  sfor<0, no>([&](auto k) {
    staged[k] = 0ull;
    soff[k] = 0u;
    slen[k] = 0u;
  });
  while (m) {
    const int p = __ffsll((unsigned long long)m) - 1;
    m &= m - 1;
    const int32_t* dp = (const int32_t*)readlane_u64((uint64_t)par.dptr, p);
    const uint32_t dl = readlane_u32(par.dlen, p);
    const int32_t* src[no];
    uint32_t n[no];
    bool same = true;
    sfor<0, no>([&](auto k) {
      src[k] = (const int32_t*)readlane_u64((uint64_t)optr[k], p);
      n[k] = readlane_u32(olen[k], p);
      same = same && (uint64_t)src[k] == staged[k] && n[k] == slen[k];
    });
    if (!same) {
      __asm__ __volatile__("" ::: "memory");  // earlier reads of the stage are done (in order)
      uint32_t at = 0;
      sfor<0, no>([&](auto k) {
        for (uint32_t j0 = 0; j0 < n[k]; j0 += 4u * kWave) {
          int32_t t[4];
          sfor<0, 4>([&](auto u) {
            const uint32_t j = j0 + (uint32_t)u * kWave + lane;
            t[u] = j < n[k] ? ld_at<false>(src[k], j) : 0;
          });
          sfor<0, 4>([&](auto u) {
            const uint32_t j = j0 + (uint32_t)u * kWave + lane;
            if (j < n[k]) stage[at + j] = t[u];
          });
        }
        staged[k] = (uint64_t)src[k];
        soff[k] = at;
        slen[k] = n[k];
        at += n[k];
      });
      __asm__ __volatile__("" ::: "memory");  // LDS ops of one wave complete in order
    }
    int32_t b[A];
    sfor<0, A>([&](auto a) {
      if constexpr ((int)a < I) b[a] = (int32_t)readlane_u32((uint32_t)par.b[a], p);
      else b[a] = 0;
    });
    for (uint32_t c = 0; c < dl; c += (uint32_t)(K * kWave)) {
      int32_t v[K];
      bool ok[K];
      sfor<0, K>([&](auto j) {
        const uint32_t g = c + (uint32_t)j * kWave + lane;
        ok[j] = g < dl;
        v[j] = ok[j] ? ld_at<false>(dp, g) : 0;
      });
      sfor<0, no>([&](auto k) { lds_has<K>(stage + soff[k], slen[k], v, ok); });
      sfor<0, K>([&](auto j) {
        if (c + (uint32_t)j * kWave < dl) {
          b[I] = v[j];
          sink(ok[j], b);
        }
      });
    }
  }
}

// Attributes 2 and 3 of a 4-clique (compiled shape kShapeFour) whose relations (0,2) and (0,3)
// are one trie, and (1,2) and (1,3) another. For an entry (a0, a1) the third attribute takes
// L = children(a0) ∩ children'(a1), and the fourth takes the same L intersected with the children
// of a2 in relation (2,3). An entry whose list to check holds at most kCliqueOther values is run
// by the whole wave: L is formed once in LDS (the other list staged, the driver's candidates
// checked there and the passing ones appended in order, so L stays sorted), then every a2 of L
// opens its children in (2,3), and each child found in L is a result: one LDS search in the
// short list L per candidate, instead of two searches in long lists per candidate. The other
// entries take the generic path.
template <int A, uint64_t SH, typename Sink>
__device__ __forceinline__ void descend_clique_pairs(const LPlan& P, const LLv<A>& par,
                                                     Sink& sink, lds_i32* stage) {
  static_assert(A == 4 && SH == kShapeFour, "4-clique shape only");
  constexpr int K = kCoopChunks;
  // the relation checked at attribute 2 besides the driver (1: children of a0, 3: of a1)
  const bool drv1 = par.drel == 1;
  const int32_t* optr = drv1 ? P.rel[3].d1 + par.lo[3] : P.rel[1].d1 + par.lo[1];
  const uint32_t olen = drv1 ? par.hi[3] - par.lo[3] : par.hi[1] - par.lo[1];
  const bool clique = par.dlen > 0 && olen <= kCliqueOther;

  // entries with a long list to check: the generic levels
  {
    LLv<A> sm = par;
    sm.dlen = clique ? 0u : par.dlen;
    finish(sm);
    for (uint32_t c = 0; c < sm.T; c += (uint32_t)kWave) {
      LLv<A> ch;
      const bool ok = lchild<A, 2, SH>(P, sm, c, sm.T, ch);
      choose_driver<A, SH>(P, 3, ok, ch);
      finish(ch);
      if (ch.T) descend_last_fixed<A, 3, SH>(P, ch, sink, stage);
    }
  }

  lds_i32* lst = stage + kCliqueOther;  // L
  uint64_t m = __ballot(clique);
  const uint32_t lane = (uint32_t)lane_id();
  uint64_t staged = 0;
  uint32_t slen = 0;
  const LRel& r5 = P.rel[5];
  while (m) {
    const int p = __ffsll((unsigned long long)m) - 1;
    m &= m - 1;
    const int32_t* dp = (const int32_t*)readlane_u64((uint64_t)par.dptr, p);
    const uint32_t dl = readlane_u32(par.dlen, p);
    const int32_t* src = (const int32_t*)readlane_u64((uint64_t)optr, p);
    const uint32_t n = readlane_u32(olen, p);
    if ((uint64_t)src != staged || n != slen) {
      __asm__ __volatile__("" ::: "memory");  // earlier reads of the stage are done (in order)
      for (uint32_t j0 = 0; j0 < n; j0 += 4u * kWave) {
        int32_t t[4];
        sfor<0, 4>([&](auto u) {
          const uint32_t j = j0 + (uint32_t)u * kWave + lane;
          t[u] = j < n ? ld_at<false>(src, j) : 0;
        });
        sfor<0, 4>([&](auto u) {
          const uint32_t j = j0 + (uint32_t)u * kWave + lane;
          if (j < n) stage[j] = t[u];
        });
      }
      staged = (uint64_t)src;
      slen = n;
      __asm__ __volatile__("" ::: "memory");  // LDS ops of one wave complete in order
    }
    // L = the driver's candidates found in the staged list, in order
    uint32_t nl = 0;
    for (uint32_t c = 0; c < dl; c += (uint32_t)(K * kWave)) {
      int32_t v[K];
      bool ok[K];
      sfor<0, K>([&](auto j) {
        const uint32_t g = c + (uint32_t)j * kWave + lane;
        ok[j] = g < dl;
        v[j] = ok[j] ? ld_at<false>(dp, g) : 0;
      });
      lds_has<K>(stage, n, v, ok);
      sfor<0, K>([&](auto j) {
        const uint64_t bm = __ballot(ok[j]);
        if (ok[j]) lst[nl + mbcnt(bm)] = v[j];
        nl += (uint32_t)__popcll(bm);
      });
    }
    __asm__ __volatile__("" ::: "memory");
    if (nl == 0) continue;
    int32_t b[A];
    b[0] = (int32_t)readlane_u32((uint32_t)par.b[0], p);
    b[1] = (int32_t)readlane_u32((uint32_t)par.b[1], p);
    // every a2 of L (64 at a time) with its children in (2,3), flattened over the wave
    for (uint32_t g0 = 0; g0 < nl; g0 += (uint32_t)kWave) {
      const bool has = g0 + lane < nl;
      const int32_t a2 = has ? lst[g0 + lane] : 0;
      uint32_t clo = 0, chi = 0;
      const bool f = has && root_range<true>(r5, a2, clo, chi);
      const uint32_t len = f ? chi - clo : 0u;
      const uint32_t incl = wave_scan_add_dpp(len);
      const uint32_t excl = incl - len;
      const uint32_t T = lane63(incl);
      for (uint32_t c = 0; c < T; c += (uint32_t)(K * kWave)) {
        // K chunks of 64 candidates at once: their loads and searches are in flight together
        int q[K];
        bool ok[K];
        int32_t a3[K];
        sfor<0, K>([&](auto j) {
          const uint32_t cj = c + (uint32_t)j * kWave;
          const uint32_t g = cj + lane;
          q[j] = (int)parent_of_slots(excl, len, cj < T ? cj : 0u);
          const uint32_t k = g - (uint32_t)__shfl(excl, q[j], kWave);
          const uint32_t qlo = (uint32_t)__shfl(clo, q[j], kWave);
          ok[j] = g < T;
          a3[j] = ok[j] ? ld_at<false>(r5.d1, qlo + k) : 0;
        });
        lds_has<K>(lst, nl, a3, ok);
        sfor<0, K>([&](auto j) {
          if (c + (uint32_t)j * kWave < T) {
            b[2] = __shfl(a2, q[j], kWave);
            b[3] = a3[j];
            sink(ok[j], b);
          }
        });
      }
    }
  }
}

template <int A, int i, uint64_t SH = 0, typename Sink>
__device__ __forceinline__ void descend(const LPlan& P, const LLv<A>& par, Sink& sink,
                                        lds_i32* stage = nullptr) {
  if constexpr (SH != 0 && i == A - 1) {
    descend_last_fixed<A, i, SH>(P, par, sink, stage);
  } else if constexpr (SH == kShapeFour && i == 2) {
    // (0,2) and (0,3) one trie, (1,2) and (1,3) another: the clique pairs apply
    const bool same = P.rel[1].d1 == P.rel[2].d1 && P.rel[1].off == P.rel[2].off &&
                      P.rel[1].d0 == P.rel[2].d0 && P.rel[3].d1 == P.rel[4].d1 &&
                      P.rel[3].off == P.rel[4].off && P.rel[3].d0 == P.rel[4].d0;
    if (same) {
      descend_clique_pairs<A, SH>(P, par, sink, stage);
      return;
    }
    for (uint32_t c = 0; c < par.T; c += (uint32_t)kWave) {
      LLv<A> ch;
      const bool ok = lchild<A, i, SH>(P, par, c, par.T, ch);
      choose_driver<A, SH>(P, i + 1, ok, ch);
      finish(ch);
      if (ch.T) descend<A, i + 1, SH>(P, ch, sink, stage);
    }
  } else {
    for (uint32_t c = 0; c < par.T; c += (uint32_t)kWave) {
      LLv<A> ch;
      const bool ok = lchild<A, i, SH>(P, par, c, par.T, ch);
      if constexpr (i == A - 1) {
        sink(ok, ch.b);
      } else {
        choose_driver<A, SH>(P, i + 1, ok, ch);
        finish(ch);
        if (ch.T) descend<A, i + 1, SH>(P, ch, sink, stage);
      }
    }
  }
}

}  // namespace amj
