/*
 * oracle.c -- CPU oracle for the join results. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * mhj_join restates the reference's CPU multi-way hash join, MHJ-CPU/MHJ.cpp:
 *   - buckets B_i = floorPowerOf2(|R_i|)                      (MHJ.cpp:573-575, helper.h:25-34)
 *   - pass 1: parallel radix partition on the high half of the bucket bits
 *     (parallel_radix_partition, MHJ.cpp:122-173; bit split MHJ.cpp:416-419)
 *   - pass 2: serial radix partition of every pass-1 partition on the low half
 *     (serial_radix_partition, MHJ.cpp:91-119; task creation MHJ.cpp:324-349)
 *   - bucket of key k: HASH_BIT_MODULO(k, B-1, 0) = k & (B-1)  (mem_helper.h:26)
 *   - per-thread depth-first probe with backtracking over the bucket chains
 *     (probe_write / probe_tbj, MHJ.cpp:64-88, :176-279)
 *   - used_for_compare / attr_idxes_in_iRes by first occurrence   (MHJ.cpp:555-571)
 * The reference materialises into 64K-tuple chunks (types.h:70-84) and reports the count; this
 * restatement additionally folds every tuple into the order-independent checksum.
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define MAXT 8  /* MHJ-CPU/types.h:11 MAX_NUM_BUILD_TABLES = 7 build tables */
#define MAXA 8  /* MHJ-CPU/types.h:12 MAX_NUM_RES_ATTRS */

static inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

uint64_t oracle_tuple_hash(const int32_t* v, int n_attrs) {
  uint64_t h = (uint64_t)n_attrs;
  for (int i = 0; i < n_attrs; i += 2) {
    const uint64_t hi = (i + 1 < n_attrs) ? (uint64_t)(uint32_t)v[i + 1] : 0ull;
    h = mix64(h ^ ((uint64_t)(uint32_t)v[i] | (hi << 32)));
  }
  return h;
}

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* helper.h:25-34 floorPowerOf2 over the power-of-two table 1..2^31 */
static uint64_t floor_pow2(uint64_t v) {
  uint64_t p = 1;
  if (v == 0) return 0;
  while (p <= v / 2 && p < (1ull << 31)) p <<= 1;
  return p;
}
static int log2_exact(uint64_t b) {
  int r = 0;
  while ((1ull << r) < b) ++r;
  return r;
}

typedef struct {
  int32_t key;
  uint64_t idx; /* idx_in_origin */
} htuple_t;    /* MHJ-CPU/types.h:25-29 hash_tuple_t */

typedef struct {
  htuple_t* tuples;
  uint64_t* buc_ptrs; /* B+1 */
  uint64_t buckets;
  const orel_t* origin;
  uint32_t hash_attr;
} htable_t;

/* Pass 1 + pass 2 of the MHJ-CPU radix partition for one build table. */
static int build_table(const orel_t* r, htable_t* ht, int T) {
  const uint64_t n = r->rows;
  const uint64_t B = floor_pow2(n) ? floor_pow2(n) : 1;
  const int bits = log2_exact(B);
  const int R = bits / 2, D = bits - bits / 2; /* MHJ.cpp:418-419 */
  const uint64_t fan1 = 1ull << R, fan2 = 1ull << D;
  const uint64_t mask1 = (fan1 - 1) << D;
  const int32_t* keys = r->cols[0];
  htuple_t* tmp = (htuple_t*)malloc(sizeof(htuple_t) * (n ? n : 1));
  ht->tuples = (htuple_t*)malloc(sizeof(htuple_t) * (n ? n : 1));
  ht->buc_ptrs = (uint64_t*)calloc(B + 1, sizeof(uint64_t));
  ht->buckets = B;
  ht->origin = r;
  ht->hash_attr = r->attr[0]; /* hashed on the first column, MHJ.cpp:424 */
  uint64_t* hist = (uint64_t*)calloc((size_t)T * fan1, sizeof(uint64_t));
  uint64_t* out1 = (uint64_t*)calloc(fan1 + 1, sizeof(uint64_t));
  if (!tmp || !ht->tuples || !ht->buc_ptrs || !hist || !out1) return -1;
  /* pass 1: per-thread histogram, then per-thread cluster starts (MHJ.cpp:138-163) */
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    const uint64_t per = n / T;
    const uint64_t lo = per * t, hi = (t == T - 1) ? n : per * (t + 1);
    uint64_t* h = hist + (size_t)t * fan1;
    for (uint64_t i = lo; i < hi; ++i) h[((uint64_t)(uint32_t)keys[i] & mask1) >> D]++;
#pragma omp barrier
    uint64_t* dst = (uint64_t*)malloc(sizeof(uint64_t) * fan1);
    uint64_t run = 0;
    for (uint64_t j = 0; j < fan1; ++j) {
      uint64_t before = 0;
      for (int tt = 0; tt < T; ++tt) {
        if (tt == t) before = run;
        run += hist[(size_t)tt * fan1 + j];
      }
      dst[j] = before;
    }
    if (t == 0) {
      uint64_t acc = 0;
      for (uint64_t j = 0; j < fan1; ++j) {
        out1[j] = acc;
        for (int tt = 0; tt < T; ++tt) acc += hist[(size_t)tt * fan1 + j];
      }
      out1[fan1] = n;
    }
    for (uint64_t i = lo; i < hi; ++i) {
      const uint64_t c = ((uint64_t)(uint32_t)keys[i] & mask1) >> D;
      tmp[dst[c]].key = keys[i];
      tmp[dst[c]].idx = i;
      ++dst[c];
    }
    free(dst);
  }
  /* pass 2: one serial partition task per pass-1 cluster (MHJ.cpp:91-119, :324-349) */
#pragma omp parallel for num_threads(T) schedule(dynamic, 16)
  for (int64_t k = 0; k < (int64_t)fan1; ++k) {
    const uint64_t s = out1[k], len = out1[k + 1] - out1[k];
    uint64_t* h2 = (uint64_t*)calloc(fan2 + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < len; ++i) h2[(uint64_t)(uint32_t)tmp[s + i].key & (fan2 - 1)]++;
    uint64_t off = 0;
    for (uint64_t j = 0; j < fan2; ++j) {
      const uint64_t c = h2[j];
      h2[j] = off;
      ht->buc_ptrs[(uint64_t)k * fan2 + j] = s + off;
      off += c;
    }
    for (uint64_t i = 0; i < len; ++i) {
      const uint64_t j = (uint64_t)(uint32_t)tmp[s + i].key & (fan2 - 1);
      ht->tuples[s + h2[j]++] = tmp[s + i];
    }
    free(h2);
  }
  ht->buc_ptrs[B] = n;
  free(hist);
  free(out1);
  free(tmp);
  return 0;
}

typedef struct {
  int32_t* data[MAXA];
  uint64_t n, cap;
} tbuf_t;

static void tbuf_push(tbuf_t* b, const int32_t* v, int A) {
  if (b->n == b->cap) {
    b->cap = b->cap ? b->cap * 2 : 4096;
    for (int a = 0; a < A; ++a) b->data[a] = (int32_t*)realloc(b->data[a], b->cap * 4);
  }
  for (int a = 0; a < A; ++a) b->data[a][b->n] = v[a];
  b->n++;
}

/* does build row (through idx_in_origin) match the bound tuple? MHJ.cpp:64-88 */
static inline int row_matches(const htable_t* ht, uint64_t i, const int32_t* iRes,
                              const int* ufc, int general) {
  if (!general) return iRes[ht->hash_attr] == ht->tuples[i].key;
  const orel_t* o = ht->origin;
  const uint64_t oi = ht->tuples[i].idx;
  for (uint32_t a = 0; a < o->ncols; ++a)
    if (ufc[a] && iRes[o->attr[a]] != o->cols[a][oi]) return 0;
  return 1;
}

int mhj_join(const orel_t* rels, uint32_t n, int nthreads, int general, uint64_t* count,
             uint64_t* checksum, int32_t** out_cols, uint64_t out_cap, double* partition_ms,
             double* probe_ms) {
  if (n < 2 || n - 1 > MAXT) return -1;
  const int T = nthreads > 0 ? nthreads : omp_get_max_threads();
  const uint32_t L = n - 1;
  /* used_for_compare and the attributes carried in iRes (MHJ.cpp:555-571) */
  int ufc[MAXT][2];
  int referred[64] = {0};
  int A = 0;
  int iRes_attrs[MAXA];
  int n_iRes = 0;
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t a = 0; a < rels[i].ncols; ++a) {
      const uint32_t at = rels[i].attr[a];
      if (at >= 64) return -2;
      if (!referred[at]) {
        referred[at] = 1;
        ++A;
        if (i) ufc[i - 1][a] = 0;
        if (i != n - 1) iRes_attrs[n_iRes++] = (int)at;
      } else if (i) {
        ufc[i - 1][a] = 1;
      }
    }
  }
  if (A > MAXA) return -3;
  for (int a = 0; a < A; ++a)
    if (!referred[a]) return -4; /* attributes must be numbered 0..A-1 */

  const double t0 = now_ms();
  htable_t ht[MAXT];
  for (uint32_t i = 0; i < L; ++i)
    if (build_table(&rels[i + 1], &ht[i], T)) return -5;
  const double t1 = now_ms();

  const orel_t* probe = &rels[0];
  uint64_t total = 0, csum = 0;
  tbuf_t* bufs = out_cols ? (tbuf_t*)calloc(T, sizeof(tbuf_t)) : NULL;
#pragma omp parallel num_threads(T) reduction(+ : total, csum)
  {
    const int t = omp_get_thread_num();
    const uint64_t per = probe->rows / T;
    const uint64_t p_lo = per * t, p_hi = (t == T - 1) ? probe->rows : per * (t + 1);
    int32_t iRes[MAXA] = {0};
    uint64_t buc_start[MAXT] = {0}, buc_end[MAXT] = {0};
    uint64_t it[MAXT];
    uint64_t p = p_lo;
    int cur = 0;
    /* MHJ.cpp:202-278 */
    for (;;) {
      if (cur == 0 && buc_start[0] >= buc_end[0]) {
        if (p >= p_hi) break;
        for (uint32_t a = 0; a < probe->ncols; ++a) iRes[probe->attr[a]] = probe->cols[a][p];
        ++p;
        const uint64_t h = (uint64_t)(uint32_t)iRes[ht[0].hash_attr] & (ht[0].buckets - 1);
        buc_start[0] = ht[0].buc_ptrs[h];
        buc_end[0] = ht[0].buc_ptrs[h + 1];
      }
      if (cur == (int)L - 1) { /* last table: emit every match */
        const htable_t* h = &ht[cur];
        for (uint64_t j = buc_start[cur]; j < buc_end[cur]; ++j) {
          if (!row_matches(h, j, iRes, ufc[cur], general)) continue;
          int32_t v[MAXA];
          memcpy(v, iRes, sizeof(v));
          const uint64_t oi = h->tuples[j].idx;
          for (uint32_t a = 0; a < h->origin->ncols; ++a)
            if (!ufc[cur][a]) v[h->origin->attr[a]] = h->origin->cols[a][oi];
          ++total;
          csum += oracle_tuple_hash(v, A);
          if (bufs) tbuf_push(&bufs[t], v, A);
        }
        buc_start[cur] = buc_end[cur];
        if (cur == 0) continue;
        --cur;
        continue;
      }
      /* probe_tbj: next match in the current bucket range (MHJ.cpp:64-88) */
      int found = 0;
      for (uint64_t i = buc_start[cur]; i < buc_end[cur]; ++i) {
        if (row_matches(&ht[cur], i, iRes, ufc[cur], general)) {
          it[cur] = i;
          buc_start[cur] = i + 1;
          found = 1;
          break;
        }
      }
      if (!found) {
        buc_start[cur] = buc_end[cur];
        if (cur > 0) --cur;
        continue;
      }
      /* bind the new attributes of the matched row, open the next bucket (MHJ.cpp:263-277) */
      const htable_t* h = &ht[cur];
      const uint64_t oi = h->tuples[it[cur]].idx;
      for (uint32_t a = 0; a < h->origin->ncols; ++a)
        if (!ufc[cur][a]) iRes[h->origin->attr[a]] = h->origin->cols[a][oi];
      const htable_t* nx = &ht[cur + 1];
      const uint64_t hv = (uint64_t)(uint32_t)iRes[nx->hash_attr] & (nx->buckets - 1);
      buc_start[cur + 1] = nx->buc_ptrs[hv];
      buc_end[cur + 1] = nx->buc_ptrs[hv + 1];
      ++cur;
    }
  }
  const double t2 = now_ms();
  if (bufs) {
    uint64_t pos = 0;
    for (int t = 0; t < T; ++t) {
      for (uint64_t i = 0; i < bufs[t].n && pos < out_cap; ++i, ++pos)
        for (int a = 0; a < A; ++a) out_cols[a][pos] = bufs[t].data[a][i];
      for (int a = 0; a < A; ++a) free(bufs[t].data[a]);
    }
    free(bufs);
  }
  for (uint32_t i = 0; i < L; ++i) {
    free(ht[i].tuples);
    free(ht[i].buc_ptrs);
  }
  (void)iRes_attrs;
  *count = total;
  *checksum = csum;
  if (partition_ms) *partition_ms = t1 - t0;
  if (probe_ms) *probe_ms = t2 - t1;
  return A;
}

/* ---------------------------------------------------------------------------------------- */
/* Graph patterns on an oriented, (src,dst)-sorted edge list.                                  */
/* ---------------------------------------------------------------------------------------- */
typedef struct {
  uint64_t* off; /* nv + 1 */
  const int32_t* adj;
  uint64_t nv;
} csr_t;

static int make_csr(const int32_t* src, const int32_t* dst, uint64_t m, csr_t* g) {
  int32_t mx = 0;
  for (uint64_t i = 0; i < m; ++i) {
    if (src[i] > mx) mx = src[i];
    if (dst[i] > mx) mx = dst[i];
  }
  g->nv = (uint64_t)mx + 1;
  g->off = (uint64_t*)calloc(g->nv + 1, sizeof(uint64_t));
  if (!g->off) return -1;
  for (uint64_t i = 0; i < m; ++i) g->off[(uint64_t)src[i] + 1]++;
  for (uint64_t v = 0; v < g->nv; ++v) g->off[v + 1] += g->off[v];
  g->adj = dst;
  return 0;
}

/* sorted-list intersection into out; returns its size */
static uint64_t isect(const int32_t* a, uint64_t na, const int32_t* b, uint64_t nb, int32_t* out) {
  uint64_t i = 0, j = 0, k = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (a[i] > b[j]) ++j;
    else {
      out[k++] = a[i];
      ++i;
      ++j;
    }
  }
  return k;
}

uint64_t graph_triangles(const int32_t* src, const int32_t* dst, uint64_t m, int nthreads,
                         uint64_t* checksum) {
  csr_t g;
  if (m == 0) {
    *checksum = 0;
    return 0;
  }
  if (make_csr(src, dst, m, &g)) return 0;
  const int T = nthreads > 0 ? nthreads : omp_get_max_threads();
  uint64_t total = 0, csum = 0;
#pragma omp parallel num_threads(T) reduction(+ : total, csum)
  {
    int32_t* buf = NULL;
    uint64_t cap = 0;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t e = 0; e < (int64_t)m; ++e) {
      const int32_t a = src[e], b = dst[e];
      const uint64_t na = g.off[a + 1] - g.off[a];
      const uint64_t nb = (uint64_t)b < g.nv ? g.off[b + 1] - g.off[b] : 0;
      if (na > cap || nb > cap) {
        cap = na > nb ? na : nb;
        buf = (int32_t*)realloc(buf, cap * 4 + 4);
      }
      const uint64_t k = isect(g.adj + g.off[a], na, g.adj + g.off[b], nb, buf);
      for (uint64_t i = 0; i < k; ++i) {
        const int32_t v[3] = {a, b, buf[i]};
        csum += oracle_tuple_hash(v, 3);
      }
      total += k;
    }
    free(buf);
  }
  free(g.off);
  *checksum = csum;
  return total;
}

uint64_t graph_four_cliques(const int32_t* src, const int32_t* dst, uint64_t m, int nthreads,
                            uint64_t* checksum) {
  csr_t g;
  if (m == 0) {
    *checksum = 0;
    return 0;
  }
  if (make_csr(src, dst, m, &g)) return 0;
  const int T = nthreads > 0 ? nthreads : omp_get_max_threads();
  uint64_t total = 0, csum = 0;
#pragma omp parallel num_threads(T) reduction(+ : total, csum)
  {
    int32_t *c1 = NULL, *c2 = NULL;
    uint64_t cap = 0;
#pragma omp for schedule(dynamic, 256)
    for (int64_t e = 0; e < (int64_t)m; ++e) {
      const int32_t a = src[e], b = dst[e];
      const uint64_t na = g.off[a + 1] - g.off[a];
      const uint64_t nb = (uint64_t)b < g.nv ? g.off[b + 1] - g.off[b] : 0;
      if (na > cap || nb > cap) {
        cap = na > nb ? na : nb;
        c1 = (int32_t*)realloc(c1, cap * 4 + 4);
        c2 = (int32_t*)realloc(c2, cap * 4 + 4);
      }
      const uint64_t k = isect(g.adj + g.off[a], na, g.adj + g.off[b], nb, c1);
      for (uint64_t i = 0; i < k; ++i) {
        const int32_t c = c1[i];
        const uint64_t nc = (uint64_t)c < g.nv ? g.off[c + 1] - g.off[c] : 0;
        const uint64_t k2 = isect(c1, k, g.adj + g.off[c], nc, c2);
        for (uint64_t j = 0; j < k2; ++j) {
          const int32_t v[4] = {a, b, c, c2[j]};
          csum += oracle_tuple_hash(v, 4);
        }
        total += k2;
      }
    }
    free(c1);
    free(c2);
  }
  free(g.off);
  *checksum = csum;
  return total;
}
