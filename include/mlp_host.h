/*
 * mlp_host.h -- host-side inputs of the MLP local search (no device work).
 * Each returns 0 on success, a negative value on bad arguments.
 *
 * Reference interfaces replaced:
 *   MTRandom::seed / rand / rand(max) / shuffle  mtrand.cpp:51-121, mtrand.h:168-195,301-312
 *   TSPLibInstance::load (EUC_2D) + inear        tsplib.cpp:236-240, 458-497, 604-687
 *   MLProblem::load coordShift                   mlproblem.cpp:137-151
 *   MLSolution::random(rng, alpha)               mlsolution.cpp:753-874
 * Synthetic generators (uniform square, Gaussian clusters, log-uniform n) have
 * no counterpart in the reference; they define the benchmark inputs.
 */
#ifndef MLP_HOST_H_
#define MLP_HOST_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* (skip+1)-th raw MT19937 output of init_genrand(seed). */
uint32_t mlp_mt_first(uint32_t seed, int32_t skip);

/* n clients, x then y drawn as rand(maxc) from MT19937(seed); client 0 = depot. */
int mlp_gen_uniform(int32_t n, uint32_t seed, int32_t maxc, int32_t* xs, int32_t* ys);
/* ncent centres rand(span)^2, each client: centre rand(ncent-1) plus
 * round-half-away(sigma * Box-Muller normal), then shifted to >= 0. */
int mlp_gen_clustered(int32_t n, uint32_t seed, int32_t ncent, int32_t span, double sigma,
                      int32_t* xs, int32_t* ys);
/* round(exp(U[ln lo, ln hi])) with U the (index+1)-th draw of MT19937(seed). */
int32_t mlp_gen_loguniform_n(uint32_t seed, int32_t index, int32_t lo, int32_t hi);

/* TSPLIB EUC_2D coordinates (float parse, inear rounding); 0 on success,
 * -4 with *n_out = DIMENSION when cap is too small. */
int mlp_tsplib_load(const char* path, int32_t* xs, int32_t* ys, int32_t cap, int32_t* n_out);
int mlp_shift_nonnegative(int32_t n, int32_t* xs, int32_t* ys);

/* int(sqrtf(float(dx^2 + dy^2)) + dr), dr = 0.5 * round (mlproblem.cpp:159-162). */
int mlp_host_dist(int32_t x1, int32_t y1, int32_t x2, int32_t y2, int32_t round);
/* Latency sum_{p=1}^{N-1} (N - p) d(tour[p-1], tour[p]) as int64. */
int64_t mlp_host_latency(const int32_t* xs, const int32_t* ys, int32_t round, const int32_t* tour,
                         int32_t N);

/* [0, shuffle(1..n-1), 0] -- N = n + 1 entries. */
int mlp_tour_perm(uint32_t seed, int32_t n, int32_t* tour);
/* Randomized greedy start tour (alpha RCL), the (skip+1)-th one drawn from
 * MT19937(seed); N = n + 1 entries. */
int mlp_tour_random(const int32_t* xs, const int32_t* ys, int32_t n, int32_t round, uint32_t seed,
                    int32_t skip, float alpha, int32_t* tour);
/* The same randomized construction with the rank-`pick` candidate found by
 * selection (ties in distance ordered by client id): O(n^2), for large
 * synthetic instances.  Tours differ from mlp_tour_random's where distances
 * tie. */
int mlp_tour_random_fast(const int32_t* xs, const int32_t* ys, int32_t n, int32_t round,
                         uint32_t seed, float alpha, int32_t* tour);

#ifdef __cplusplus
}
#endif

#endif /* MLP_HOST_H_ */
