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
/* amj_gen_zipf with the rank -> value map chosen: shuffle != 0 is amj_gen_zipf; shuffle == 0
 * maps rank r to value r-1, so value 0 is the most frequent (the reference's skewed writer
 * with SHUFFLE = 0, MHJ-GPU/skewed_writer_.cu:75-81). Returns 0, or -1 for an invalid k. */
int amj_gen_zipf_mapped(int32_t* out, uint64_t n, uint64_t k, double s, uint64_t seed,
                        uint64_t stream, int shuffle);
/* The values j % range for j in [0, n), in a seeded random order: every value of [0, range)
 * occurs n / range or n / range + 1 times (gen type 1 of MHJ-GPU/uniform_writer_.cu:79-84,
 * alpha 0 of MHJ-GPU/skewed_writer_.cu:26-31). Returns 0, or -1 (range 0 or above 2^31,
 * n >= 2^32). */
int amj_gen_unique_column(int32_t* out, uint64_t n, uint64_t range, uint64_t seed,
                          uint64_t stream);
/* n distinct pairs (key0, key1) drawn uniformly from [lo0, hi0) x [lo1, hi1), in (key0, key1)
 * order, or in a seeded random order when shuffle != 0 (uniformUniquePairGenerator,
 * LFTJ-GPU/generator_cuda.cuh:341-419, with and without UNSORTED_DATA). Returns 0; -1 for
 * invalid ranges, n not below the number of pairs (the reference asserts the same) or
 * n >= 2^32; -2 if the sampling rounds did not finish. */
int amj_gen_unique_pairs(int32_t* key0, int32_t* key1, uint64_t n, int64_t lo0, int64_t hi0,
                         int64_t lo1, int64_t hi1, uint64_t seed, int shuffle);
/* n distinct pairs whose column skew_col (0 or 1) is Zipf(z) over its range (seeded
 * rank -> value permutation) and whose other column is uniform, distinct within every value of
 * the skewed column; ordered by (skewed value, uniform value), or shuffled when shuffle != 0
 * (zipfUniquePairGenerator, LFTJ-GPU/generator_cuda.cuh:459-527). Returns 0; -1 for invalid
 * arguments; -2 if some skewed value occurs more often than the uniform range has values (the
 * reference loops forever then). */
int amj_gen_zipf_pairs(int32_t* key0, int32_t* key1, uint64_t n, int64_t lo0, int64_t hi0,
                       int64_t lo1, int64_t hi1, uint64_t seed, int skew_col, double z,
                       int shuffle);
/* (lo0 + i / p, i % p) for i in [0, n) with p = n / (hi0 - lo0): the sequential pairs of
 * uniformUniqueSeqPairGenerator (LFTJ-GPU/generator_cuda.cuh:421-456), for debugging. Returns
 * 0, or -1 when n < hi0 - lo0. */
int amj_gen_seq_pairs(int32_t* key0, int32_t* key1, uint64_t n, int64_t lo0, int64_t hi0);
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
