/*
 * dress/dress.h -- libdress core C API, as exported by libdress_hip.so.
 *
 * This header restates the ABI of libdress v0.8.3
 * (libdress/include/dress/dress.h:1-252): the dress_graph_t field order,
 * the variant enum values, the histogram pair struct and every prototype are
 * kept identical, because the language bindings read dress_graph_t fields by
 * offset (python/src/dress/_ctypes_helpers.py:5-28, julia/src/DRESS.jl,
 * wasm/dress.js) and bind these symbols by name.
 *
 * In libdress_hip.so every compute entry point below runs on the GPU:
 * dress_init_graph builds the CSR (and, when requested, the intersection
 * index) on the device and mirrors it into the host struct; dress_fit,
 * dress_delta_fit* and dress_nabla_fit* are the HIP implementations (the same
 * functions as the *_hip names in dress/hip/dress_hip.h).  Errors are
 * reported through dress_hip_last_error() (dress/hip/dress_hip.h).
 */
#ifndef DRESS_H
#define DRESS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* the library is built with hidden visibility; these names are its interface */
#pragma GCC visibility push(default)

/* One (value, count) entry of a sparse histogram (dress.h:12-15). */
typedef struct __dress_hist_pair_t {
    double  value;
    int64_t count;
} dress_hist_pair_t;

/* How the adjacency N[u] is derived from the edge list (dress.h:24-29).
 *   UNDIRECTED  N[u] = N(u), w_uv = 2w
 *   DIRECTED    N[u] = out(u) + in(u), reciprocal weights summed
 *   FORWARD     N[u] = out(u)
 *   BACKWARD    N[u] = in(u)                                              */
typedef enum __dress_variant_t {
    DRESS_VARIANT_UNDIRECTED = 0,
    DRESS_VARIANT_DIRECTED   = 1,
    DRESS_VARIANT_FORWARD    = 2,
    DRESS_VARIANT_BACKWARD   = 3
} dress_variant_t;

/* Graph + fit state (dress.h:43-75).  Field order is ABI: do not change. */
typedef struct __dress_graph_t {
    dress_variant_t variant;
    int      N;                 /* vertices                                 */
    int      E;                 /* input edge records                       */

    int     *U;                 /* [E] sources (owned)                      */
    int     *V;                 /* [E] targets (owned)                      */

    int     *adj_offset;        /* [N+1] CSR row offsets                    */
    int     *adj_target;        /* [S]   neighbour ids, sorted per row      */
    int     *adj_edge_idx;      /* [S]   input edge of each half-edge       */
    int      max_degree;

    double  *W;                 /* [E] raw weights or NULL (owned)          */
    double  *edge_weight;       /* [E] variant weight w-bar                 */
    double  *edge_dress;        /* [E] current values                       */
    double  *edge_dress_next;   /* [E] double buffer                        */

    double  *NW;                /* [N] vertex weights (all 1.0 if NULL)     */
    double  *vertex_dress;      /* [N] norms of the last vertex phase       */

    int      precompute_intercepts;
    int     *intercept_offset;  /* [E+1]                                    */
    int     *intercept_edge_ux; /* [T] edge (u,x) per common neighbour x    */
    int     *intercept_edge_vx; /* [T] edge (v,x) per common neighbour x    */
} dress_graph_t, *p_dress_graph_t;

/* dress.h:80-83.  Takes ownership of U, V, W, NW (malloc'd by the caller). */
p_dress_graph_t dress_init_graph(int N, int E, int *U, int *V,
                                 double *W, double *NW,
                                 dress_variant_t variant,
                                 int precompute_intercepts);

/* dress.h:89-90.  Continues from g->edge_dress; strict max|delta| < epsilon. */
void   dress_fit(p_dress_graph_t g, int max_iterations, double epsilon,
                 int *iterations, double *delta);

/* dress.h:93 */
void   dress_free_graph(p_dress_graph_t g);

/* dress.h:105-107 */
double dress_get(const p_dress_graph_t g, int u, int v,
                 int max_iterations, double epsilon,
                 double edge_weight);

/* dress.h:148-153 */
dress_hist_pair_t *dress_delta_fit(p_dress_graph_t g, int k, int iterations,
                                   double epsilon,
                                   int n_samples, unsigned int seed,
                                   int *hist_size,
                                   int keep_multisets, double **multisets,
                                   int64_t *num_subgraphs);

/* dress.h:157-162 */
int64_t *dress_delta_fit_flat(p_dress_graph_t g, int k, int iterations,
                              double epsilon,
                              int n_samples, unsigned int seed,
                              int *hist_size,
                              int keep_multisets, double **multisets,
                              int64_t *num_subgraphs);

/* dress.h:173-179 */
dress_hist_pair_t *dress_delta_fit_strided(p_dress_graph_t g, int k, int iterations,
                                           double epsilon,
                                           int n_samples, unsigned int seed,
                                           int *hist_size,
                                           int keep_multisets, double **multisets,
                                           int64_t *num_subgraphs,
                                           int offset, int stride);

/* dress.h:182-188 */
int64_t *dress_delta_fit_strided_flat(p_dress_graph_t g, int k, int iterations,
                                      double epsilon,
                                      int n_samples, unsigned int seed,
                                      int *hist_size,
                                      int keep_multisets, double **multisets,
                                      int64_t *num_subgraphs,
                                      int offset, int stride);

/* dress.h:224-229 */
dress_hist_pair_t *dress_nabla_fit(p_dress_graph_t g, int k, int iterations,
                                   double epsilon,
                                   int n_samples, unsigned int seed,
                                   int *hist_size,
                                   int keep_multisets, double **multisets,
                                   int64_t *num_tuples);

/* dress.h:232-237 */
int64_t *dress_nabla_fit_flat(p_dress_graph_t g, int k, int iterations,
                              double epsilon,
                              int n_samples, unsigned int seed,
                              int *hist_size,
                              int keep_multisets, double **multisets,
                              int64_t *num_tuples);

/* dress.h:240-246 */
int64_t *dress_nabla_fit_strided_flat(p_dress_graph_t g, int k,
                                      int iterations, double epsilon,
                                      int n_samples, unsigned int seed,
                                      int *hist_size,
                                      int keep_multisets, double **multisets,
                                      int64_t *num_tuples,
                                      int offset, int stride);

/* Counting helpers the reference exports from its drivers
 * (delta_dress_impl.h:60, nabla_dress_impl.h:57). */
int64_t dress_delta_binom(int n, int k);
int64_t dress_nabla_perm_count(int n, int k);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif

#endif /* DRESS_H */
