// Frontier-DFS building blocks of the AMHJ probe shared by the unit engine (probe.hip) and the
// persistent work-sharing / direct-output engine (dynamic.hip): the per-level frontier state,
// level 0 from 64 probe rows, the next 64 children of a level, and the emission of a last
// level's results.
#pragma once

#include "probe_common.h"

namespace amj {

// Frontier state of one level (per lane) and its wave-uniform total; NA bindings (na_of(L)
// for a query of L tables).
template <int NA>
struct LvN {
  int32_t b[NA];         // bindings of this lane's partial tuple
  uint32_t lo;           // run of matches in this level's table
  uint32_t len;
  uint64_t excl;         // exclusive prefix of len over the wave
  uint64_t T;            // total children of the level (uniform)
};
template <int L>
using LvL = LvN<na_of(L)>;

struct Acc {
  uint64_t count;     // results counted (uniform)
  uint64_t out_base;  // absolute output row of the next result (uniform)
  uint64_t csum;      // per-lane checksum partial
};

template <int NA>
__device__ __forceinline__ void finish_level(LvN<NA>& l) {
  if (__ballot(l.len >= (1u << 26)) == 0ull) {  // 64 runs sum below 2^32
    const uint32_t incl = wave_scan_add_dpp(l.len);
    l.excl = incl - l.len;
    l.T = (uint64_t)lane63(incl);
  } else {
    const uint64_t incl = wave_incl_scan_u64((uint64_t)l.len);
    l.excl = incl - l.len;
    l.T = uniform(__shfl(incl, kWave - 1, kWave));
  }
}

// Parent lane of flattened child g (g in [c, c+64)).
template <int NA>
__device__ __forceinline__ int parent_lane(const LvN<NA>& l, uint64_t c, uint64_t gg) {
  if ((l.T >> 32) == 0)
    return (int)parent_of_slots((uint32_t)l.excl, l.len, (uint32_t)c);
  return find_parent(l.excl, gg, l.T);
}

template <int NA>
__device__ __forceinline__ void lookup_level(const Plan& P, int t, LvN<NA>& l, bool valid) {
  const int32_t key = get_attr(l.b, P.key_attr[t]);
  const bool check = P.val_mode[t] == VAL_CHECK;
  const int32_t cval = check ? get_attr(l.b, P.val_attr[t]) : 0;
  lookup_run(P.tab[t], key, check, cval, valid, l.lo, l.len);
}

// Level 0 from probe row r on every lane where `valid`.
template <int NA>
__device__ __forceinline__ void make_level0_at(const Plan& P, uint64_t r, bool valid,
                                               LvN<NA>& l) {
  sfor<0, NA>([&](auto a) { l.b[a] = 0; });
  for (int c = 0; c < P.probe_ncols; ++c)
    set_attr(l.b, P.probe_attr[c], valid ? P.probe_cols[c][r] : 0);
  lookup_level(P, 0, l, valid);
  finish_level(l);
}

// Level 0 of probe rows [row0, row0 + 64).
template <int NA>
__device__ __forceinline__ void make_level0(const Plan& P, uint64_t row0, LvN<NA>& l) {
  const uint64_t r = row0 + (uint64_t)lane_id();
  make_level0_at(P, r, r < P.probe_n, l);
}

// Children [c, min(c+64, end)) of level t become the level-(t+1) frontier (end <= par.T).
template <int t, int NA>
__device__ __forceinline__ void make_child(const Plan& P, const LvN<NA>& par, uint64_t c,
                                           uint64_t end, LvN<NA>& ch) {
  const uint64_t g = c + (uint64_t)lane_id();
  const bool valid = g < end;
  const uint64_t gg = valid ? g : par.T - 1;
  const int p = parent_lane(par, c, gg);
  const uint32_t plo = __shfl(par.lo, p, kWave);
  const uint32_t pex = (uint32_t)__shfl((uint32_t)par.excl, p, kWave);
  const uint32_t row = plo + ((uint32_t)gg - pex);
  sfor<0, NA>([&](auto a) { ch.b[a] = (a < P.n_attrs) ? __shfl(par.b[a], p, kWave) : 0; });
  if (P.val_mode[t] == VAL_BIND) set_attr(ch.b, P.val_attr[t], P.tab[t].vals[row]);
  lookup_level(P, t + 1, ch, valid);
  finish_level(ch);
}

// Called once per 64 results by the emitters below (the persistent engine counts its steps).
struct NoStep {
  __device__ __forceinline__ void operator()() const {}
};

// Emit the results [0, T) of the last level at absolute rows out_row0 + e.
template <bool MATERIALIZE, bool HASH, typename Step = NoStep, int NA>
__device__ __forceinline__ void emit(const Plan& P, int t, const LvN<NA>& l,
                                     uint64_t out_row0, const ProbeOut& out, Acc& acc,
                                     const Step& step = Step{}) {
  const bool bind = P.val_mode[t] == VAL_BIND;
  const int vattr = P.val_attr[t];
  const int32_t* vals = P.tab[t].vals;
  const uint64_t e1 = l.T;
  for (uint64_t c = 0; c < e1; c += kWave) {
    step();
    const uint64_t g = c + (uint64_t)lane_id();
    const bool valid = g < e1;
    const uint64_t gg = valid ? g : e1 - 1;
    const int p = parent_lane(l, c, gg);
    const uint32_t plo = __shfl(l.lo, p, kWave);
    const uint32_t pex = (uint32_t)__shfl((uint32_t)l.excl, p, kWave);
    const uint32_t row = plo + ((uint32_t)gg - pex);
    int32_t v[NA];
    sfor<0, NA>([&](auto a) { v[a] = (a < P.n_attrs) ? __shfl(l.b[a], p, kWave) : 0; });
    if (bind) set_attr(v, vattr, vals[row]);
    if (valid) {
      if (HASH) acc.csum += hash_tuple(v, P.n_attrs);
      if (MATERIALIZE) {
        const uint64_t pos = out_row0 + g;
        if (pos < out.capacity) {
          sfor<0, NA>([&](auto a) {
            if (a < P.n_attrs) out.cols[a][pos] = v[a];
          });
        }
      }
    }
  }
}

// Direct output with dynamic placement: a wave gathers its results in an LDS buffer of
// kDirectRows rows and reserves room in the output with one atomic per full buffer (tiny
// last-level frontiers, as in TRI and FOUR, would otherwise make one update per few results),
// or, when out.chunk != 0, one atomic per out.chunk rows: the wave fills its chunk buffer by
// buffer, since a row counter updated by every wave serialises the updates (about 14 ns each on
// this device), and records the unused tail of its last chunk for the fill-up pass
// (launch_chunk_fill) that makes the output dense. Rows past the output's capacity are not
// written; the reserved total then tells the host that the pass must run again.
constexpr uint32_t kDirectRows = 256;
typedef __attribute__((address_space(3))) int32_t lds_i32;

struct DirectStage {
  lds_i32* buf;          // this wave's rows: attribute a of row i at buf[a * kDirectRows + i]
  uint32_t fill;         // rows held (uniform)
  uint64_t cur = ~0ull;  // chunked: first row of the current chunk (~0: none yet)
  uint32_t used = 0;     // chunked: rows of the current chunk taken
};

// LDS of a block's staging buffers for a query of L tables.
template <int L>
constexpr int direct_stage_words() { return kWavesPerBlock * na_of(L) * kDirectRows; }

template <int NA>
__device__ __forceinline__ void direct_flush(const ProbeOut& out, int na, DirectStage& st,
                                             unsigned long long* rows) {
  if (st.fill == 0) return;
  // rows [0, nb) of the buffer go to base_a + i, rows [nb, fill) to base_b + i - nb
  uint64_t base_a = 0, base_b = 0;
  uint32_t nb = st.fill;
  if (out.chunk) {
    const uint32_t rem = st.cur == ~0ull ? 0u : out.chunk - st.used;
    nb = st.fill < rem ? st.fill : rem;
    base_a = st.cur + st.used;
    if (st.fill > rem) {
      unsigned long long r0 = 0;
      if (lane_id() == 0)
        r0 = __hip_atomic_fetch_add(rows, (unsigned long long)out.chunk, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
      base_b = uniform((uint64_t)r0);
      st.cur = base_b;
      st.used = st.fill - nb;
    } else {
      st.used += st.fill;
    }
  } else {
    unsigned long long r0 = 0;
    if (lane_id() == 0)
      r0 = __hip_atomic_fetch_add(rows, (unsigned long long)st.fill, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
    base_a = uniform((uint64_t)r0);
  }
  __asm__ __volatile__("" ::: "memory");  // LDS ops of one wave complete in order
  for (uint32_t i = (uint32_t)lane_id(); i < st.fill; i += kWave) {
    const uint64_t pos = i < nb ? base_a + i : base_b + (i - nb);
    if (pos < out.capacity)
      sfor<0, NA>([&](auto a) {
        if (a < na) out.cols[a][pos] = st.buf[a * kDirectRows + i];
      });
  }
  __asm__ __volatile__("" ::: "memory");
  st.fill = 0;
}

// After a wave's last flush (chunked output): the unused tail of its last chunk, when the chunk
// lies inside the output.
__device__ __forceinline__ void direct_tail(const ProbeOut& out, const DirectStage& st) {
  if (out.chunk && st.cur != ~0ull && lane_id() == 0 && st.cur + out.chunk <= out.capacity)
    out.tails[st.cur / out.chunk] = out.chunk - st.used;
}

// Results [0, T) of a last level into the wave's buffer (MAT) and the checksum (HASH).
template <bool MAT, bool HASH, typename Step = NoStep, int NA>
__device__ __forceinline__ void direct_emit(const Plan& P, int t, const LvN<NA>& l,
                                            const ProbeOut& out, Acc& acc, DirectStage& st,
                                            unsigned long long* rows,
                                            const Step& step = Step{}) {
  const bool bind = P.val_mode[t] == VAL_BIND;
  const int vattr = P.val_attr[t];
  const int32_t* vals = P.tab[t].vals;
  const uint64_t e1 = l.T;
  for (uint64_t c = 0; c < e1; c += kWave) {
    step();
    const uint64_t g = c + (uint64_t)lane_id();
    const bool valid = g < e1;
    const uint64_t gg = valid ? g : e1 - 1;
    const int p = parent_lane(l, c, gg);
    const uint32_t plo = __shfl(l.lo, p, kWave);
    const uint32_t pex = (uint32_t)__shfl((uint32_t)l.excl, p, kWave);
    const uint32_t row = plo + ((uint32_t)gg - pex);
    int32_t v[NA];
    sfor<0, NA>([&](auto a) { v[a] = (a < P.n_attrs) ? __shfl(l.b[a], p, kWave) : 0; });
    if (bind) set_attr(v, vattr, vals[row]);
    if (HASH && valid) acc.csum += hash_tuple(v, P.n_attrs);
    if (MAT) {
      const uint32_t m = e1 - c < (uint64_t)kWave ? (uint32_t)(e1 - c) : (uint32_t)kWave;
      if (st.fill + m > kDirectRows) direct_flush<NA>(out, P.n_attrs, st, rows);
      __asm__ __volatile__("" ::: "memory");
      if (valid)
        sfor<0, NA>([&](auto a) {
          if (a < P.n_attrs) st.buf[a * kDirectRows + st.fill + lane_id()] = v[a];
        });
      __asm__ __volatile__("" ::: "memory");
      st.fill += m;
    }
  }
}

// Depth-first over the frontier levels of one unit (count, or write at acc.out_base).
template <int L, int t, int MODE>
__device__ __forceinline__ void descend(const Plan& P, const LvL<L>& cur,
                                        const ProbeOut& out, Acc& acc) {
  if constexpr (t == L - 1) {
    if constexpr (MODE == MODE_COUNT) {
      acc.count += cur.T;
    } else {
      emit<mode_mat<MODE>(), mode_hash<MODE>()>(P, t, cur, acc.out_base, out, acc);
      acc.out_base += cur.T;
      acc.count += cur.T;
    }
  } else {
    for (uint64_t c = 0; c < cur.T; c += kWave) {
      LvL<L> ch;
      make_child<t>(P, cur, c, cur.T, ch);
      if (ch.T) descend<L, t + 1, MODE>(P, ch, out, acc);
    }
  }
}

}  // namespace amj
