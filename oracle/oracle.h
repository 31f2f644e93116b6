/*
 * oracle.h -- CPU oracle of the AMHJ / ALFTJ join results. TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it. The product (libamj) never calls it.
 *
 * Checksum: every result tuple (v0..v{A-1}) in attribute order hashes to
 *   h = A; for i = 0, 2, 4, ...: h = splitmix64(h ^ (u32(v_i) | u32(v_{i+1}) << 32))
 * (v_A = 0 when A is odd), and the checksum of a result is the sum of h mod 2^64.
 */
#ifndef AMJ_ORACLE_H_
#define AMJ_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t ncols;            /* 1 or 2 */
  uint64_t rows;
  const int32_t* cols[2];
  uint32_t attr[2];          /* attribute id of each column */
} orel_t;

/* Multi-way hash join of MHJ-CPU (MHJ-CPU/MHJ.cpp:377-500): rels[0] probes, rels[1..n) are
 * built on their first column. general_compare=0 is the reference's single-join-key probe;
 * general_compare=1 compares every already-bound column (the single_join_key=false branch of
 * MHJ-CPU/MHJ.cpp:64-88 and :217-228). Output column a holds attribute a; attributes must be
 * numbered 0..A-1. out_cols may be NULL (count + checksum only); otherwise each of the A
 * columns must hold out_cap rows. Returns the number of result attributes, or < 0. */
int mhj_join(const orel_t* rels, uint32_t n, int nthreads, int general_compare, uint64_t* count,
             uint64_t* checksum, int32_t** out_cols, uint64_t out_cap, double* partition_ms,
             double* probe_ms);

/* Textbook triangle / 4-clique enumeration on an oriented edge list sorted by (src, dst):
 * result tuples (a,b,c) with a->b, b->c, a->c (and (a,b,c,d) with all six edges). */
uint64_t graph_triangles(const int32_t* src, const int32_t* dst, uint64_t m, int nthreads,
                         uint64_t* checksum);
uint64_t graph_four_cliques(const int32_t* src, const int32_t* dst, uint64_t m, int nthreads,
                            uint64_t* checksum);

uint64_t oracle_tuple_hash(const int32_t* v, int n_attrs);

#ifdef __cplusplus
}
#endif
#endif
