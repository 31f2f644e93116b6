/*
 * amj_gen.h -- deterministic workload generators of libamj (host code, OpenMP).
 *
 * Replacements for the reference's data generators, which are unseeded or time-seeded
 * (MHJ-GPU/uniform_writer_.cu:65-133, MHJ-GPU/skewed_writer_.cu:1-88,
 * MHJ-GPU/data_generator.cuh:24-180, LFTJ-GPU/generator_cuda.cuh:41-530). Every value is a pure
 * function of (seed, stream, row), so results do not depend on the thread count.
 */
#ifndef AMJ_GEN_H_
#define AMJ_GEN_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* out[i] uniform in [lo, hi). */
void amj_gen_uniform(int32_t* out, uint64_t n, int64_t lo, int64_t hi, uint64_t seed,
                     uint64_t stream);
/* out[i] = perm[rank], rank ~ Zipf(s) over 1..k (s = 0: uniform), perm a seeded permutation
 * of [0, k). Returns 0, or -1 for an invalid k. */
int amj_gen_zipf(int32_t* out, uint64_t n, uint64_t k, double s, uint64_t seed, uint64_t stream);
/* m R-MAT edges over 2^scale vertices with quadrant probabilities (a, b, c, 1-a-b-c), vertex
 * ids relabelled by a seeded permutation (Graph500 style). */
void amj_gen_rmat(int32_t* src, int32_t* dst, int scale, uint64_t m, double a, double b,
                  double c, uint64_t seed);
/* Symmetrise, drop self-loops and duplicates, orient u->v iff (deg u, u) < (deg v, v), and
 * sort by (src, dst). out_* must hold m entries; returns the number of oriented edges. */
uint64_t amj_graph_orient(const int32_t* src, const int32_t* dst, uint64_t m, uint64_t nv,
                          int32_t* out_src, int32_t* out_dst);

#ifdef __cplusplus
}
#endif
#endif /* AMJ_GEN_H_ */
