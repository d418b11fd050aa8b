/*
 * dress_oracle.h -- CPU oracle for the DRESS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline.  libdress_hip.so never links or calls it.
 *
 * It restates libdress v0.8.3's sequential algorithm (sort + Kahan-Babuska-
 * Neumaier sums, FP64) and its OpenMP variant; every function cites the
 * reference lines it follows.  Symbols carry an ora_ prefix so the oracle
 * can share a process with libdress_hip.so and with the compiled reference.
 */
#ifndef DRESS_ORACLE_H
#define DRESS_ORACLE_H

#include "dress/dress.h"
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

p_dress_graph_t ora_init_graph(int N, int E, int *U, int *V, double *W,
                               double *NW, dress_variant_t variant,
                               int precompute_intercepts);
void ora_free_graph(p_dress_graph_t g);

/* Graph whose intersection index is supplied by the caller (for the CPU
 * baseline on graphs whose index was built elsewhere).  Takes ownership of
 * every array; the CSR must follow the reference's construction. */
p_dress_graph_t ora_graph_from_parts(int N, int E, dress_variant_t variant,
                                     int *U, int *V, double *W, double *NW,
                                     int *adj_offset, int *adj_target,
                                     int *adj_edge_idx, double *edge_weight,
                                     int *intercept_offset, int *ux, int *vx);

void ora_fit(p_dress_graph_t g, int max_iterations, double epsilon,
             int *iterations, double *delta);
void ora_fit_omp(p_dress_graph_t g, int max_iterations, double epsilon,
                 int *iterations, double *delta);

double ora_get(const p_dress_graph_t g, int u, int v, int max_iterations,
               double epsilon, double edge_weight);

/* Flat histogram drivers: return [bits, count, ...] sorted by value,
 * *hist_size = 2 * distinct.  use_omp != 0 fits subgraphs in parallel
 * (only when n_samples == 0). */
int64_t *ora_delta_fit_flat(p_dress_graph_t g, int k, int iterations,
                            double epsilon, int n_samples, unsigned int seed,
                            int *hist_size, int keep_multisets,
                            double **multisets, int64_t *num_subgraphs,
                            int offset, int stride, int use_omp);
int64_t *ora_nabla_fit_flat(p_dress_graph_t g, int k, int iterations,
                            double epsilon, int n_samples, unsigned int seed,
                            int *hist_size, int keep_multisets,
                            double **multisets, int64_t *num_tuples,
                            int offset, int stride, int use_omp);

int64_t ora_binom(int n, int k);
int64_t ora_perm_count(int n, int k);

/* Sort + KBN sum of buf[0..n) (buf is reordered). */
double ora_kbn_sorted_sum(double *buf, int n);

/* Picks of the sampled drivers (sorted, deduplicated); returns count. */
int ora_sample_picks(int64_t count_per_stride, int64_t n_samples,
                     unsigned int seed, int offset, int64_t *out);

void ora_free(void *p);

#ifdef __cplusplus
}
#endif

#endif
