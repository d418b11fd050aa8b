/*
 * dress/hip/dress_hip.h -- the HIP (MI355X, gfx950) backend of libdress.
 *
 * Drop-in counterpart of libdress/include/dress/cuda/dress_cuda.h: the same
 * entry points with the suffix _hip, the same arguments, results and
 * ownership rules (histograms and multisets are malloc'd; the caller frees
 * them with free()).  Include dress/hip/dress.h instead to redirect the plain
 * names (dress_fit, dress_delta_fit, ...) to these.
 *
 * What differs from the CUDA backend:
 *  - the graph lives on the device: dress_init_graph builds its CSR and the
 *    neighbourhood-intersection index there, and later fits reuse them; the
 *    host struct is filled with the same arrays the CPU library would build;
 *  - delta^k / nabla^k subgraphs are not rebuilt on the host: one persistent
 *    kernel runs every subgraph of the request with the parent graph masked
 *    (deleted vertices) or re-weighted (marked vertices) in LDS;
 *  - errors do not exit(): the call leaves its outputs unset (fits set
 *    *iterations = -1 and *delta = NaN, lifted fits return NULL with
 *    *hist_size = 0) and dress_hip_last_error() describes what failed;
 *  - edge lists are checked: an id outside [0, N) is an error, not a write
 *    out of bounds.
 *
 * Threading: calls on distinct graphs may run concurrently from different
 * host threads (each graph has its own stream).  One dress_graph_t must not
 * be used by two calls at once: delta^k / nabla^k of a graph too large for
 * the batched kernel reuse its device state (values, self terms, weights)
 * and restore it before they return.
 *
 * Numerics: FP64 results equal the CPU library's bit for bit, except in the
 * rare case where an exact sum lies within ~n^2*2^-106 of a rounding midpoint
 * (the CPU sorts the terms and adds them with Kahan-Babuska-Neumaier; the GPU
 * adds them in any order with error-free transformations and rounds once).
 * The optional FP32->FP64 precision keeps every value within 1e-9 of the
 * FP64 result (see dress_hip_set_precision).
 */
#ifndef DRESS_HIP_H
#define DRESS_HIP_H

#include "dress/dress.h"
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* the library is built with hidden visibility; these names are its interface */
#pragma GCC visibility push(default)

/* ---- counterparts of libdress/include/dress/cuda/dress_cuda.h -------- */

/* dress_cuda.h:38-39 */
void dress_fit_hip(p_dress_graph_t g, int max_iterations, double epsilon,
                   int *iterations, double *delta);

/* dress_cuda.h:47-52 */
dress_hist_pair_t *dress_delta_fit_hip(p_dress_graph_t g, int k, int iterations,
                                       double epsilon,
                                       int n_samples, unsigned int seed,
                                       int *hist_size,
                                       int keep_multisets, double **multisets,
                                       int64_t *num_subgraphs);

/* dress_cuda.h:60-66 */
dress_hist_pair_t *dress_delta_fit_hip_strided(p_dress_graph_t g, int k,
                                               int iterations, double epsilon,
                                               int n_samples, unsigned int seed,
                                               int *hist_size,
                                               int keep_multisets, double **multisets,
                                               int64_t *num_subgraphs,
                                               int offset, int stride);

/* dress_cuda.h:72-78 (and the unstrided flat form delta_dress_cuda.c:24-36) */
int64_t *dress_delta_fit_hip_flat(p_dress_graph_t g, int k, int iterations,
                                  double epsilon,
                                  int n_samples, unsigned int seed,
                                  int *hist_size,
                                  int keep_multisets, double **multisets,
                                  int64_t *num_subgraphs);
int64_t *dress_delta_fit_hip_strided_flat(p_dress_graph_t g, int k,
                                          int iterations, double epsilon,
                                          int n_samples, unsigned int seed,
                                          int *hist_size,
                                          int keep_multisets, double **multisets,
                                          int64_t *num_subgraphs,
                                          int offset, int stride);

/* dress_cuda.h:82-87, :89-95, :97-102, :104-110 */
dress_hist_pair_t *dress_nabla_fit_hip(p_dress_graph_t g, int k, int iterations,
                                       double epsilon,
                                       int n_samples, unsigned int seed,
                                       int *hist_size,
                                       int keep_multisets, double **multisets,
                                       int64_t *num_tuples);
dress_hist_pair_t *dress_nabla_fit_hip_strided(p_dress_graph_t g, int k,
                                               int iterations, double epsilon,
                                               int n_samples, unsigned int seed,
                                               int *hist_size,
                                               int keep_multisets, double **multisets,
                                               int64_t *num_tuples,
                                               int offset, int stride);
int64_t *dress_nabla_fit_hip_flat(p_dress_graph_t g, int k, int iterations,
                                  double epsilon,
                                  int n_samples, unsigned int seed,
                                  int *hist_size,
                                  int keep_multisets, double **multisets,
                                  int64_t *num_tuples);
int64_t *dress_nabla_fit_hip_strided_flat(p_dress_graph_t g, int k,
                                          int iterations, double epsilon,
                                          int n_samples, unsigned int seed,
                                          int *hist_size,
                                          int keep_multisets, double **multisets,
                                          int64_t *num_tuples,
                                          int offset, int stride);

/* ---- backend extensions (no counterpart in the reference) ------------ */

/* Last error of the calling thread ("" when the last call succeeded). */
const char *dress_hip_last_error(void);

/* Number of visible HIP devices (0 when no GPU / no runtime). */
int dress_hip_device_count(void);

/* Arithmetic of the *_hip fits.
 *   DRESS_HIP_FP64      FP64 throughout (default; the CPU library's values)
 *   DRESS_HIP_FP32_FP64 early iterations store values in FP32 (half the
 *                       gathered bytes) and hand over to FP64 while enough
 *                       iterations remain to bring every value within 1e-9
 *                       of the FP64 result; a fit whose stopping decision
 *                       is within 0.1% of epsilon is redone in FP64.
 * The initial default can also be set with DRESS_HIP_PRECISION=fp64|mixed. */
typedef enum {
    DRESS_HIP_FP64 = 0,
    DRESS_HIP_FP32_FP64 = 1
} dress_hip_precision_t;
void dress_hip_set_precision(dress_hip_precision_t precision);
dress_hip_precision_t dress_hip_get_precision(void);

/* Statistics of the calling thread's last dress_fit_hip / dress_hip_fit_resident. */
typedef struct {
    int    iterations_executed;   /* iterations actually run                  */
    int    fp32_iterations;       /* of which with FP32 storage               */
    int    fp64_reruns;           /* mixed fits redone in FP64                */
    double device_ms;             /* device time of the iteration loop        */
    int64_t intersections;        /* T: entries of the intersection index     */
    int64_t half_edges;           /* S: CSR entries                           */
} dress_hip_fit_stats_t;
void dress_hip_last_fit_stats(dress_hip_fit_stats_t *out);

/* Fit with explicit precision; returns 0 on success, -1 on error.
 * Same semantics as dress_fit_hip otherwise. */
int dress_fit_hip_ex(p_dress_graph_t g, int max_iterations, double epsilon,
                     dress_hip_precision_t precision,
                     int *iterations, double *delta);

/* Shape of the device graph: info[0..20] = N, E, S (CSR entries), T
 * (intersection entries), max degree, 1 if the index is packed (16-bit row
 * positions), edges per work bin (5), vertices per work bin (5), intersection
 * entries per edge bin (5).  Bins group items whose sums
 * have <= 4, 64, 512, 8192, more terms.  Returns 0, or -1 on error. */
int dress_hip_graph_info(p_dress_graph_t g, int64_t *info, int n);

/* Device-resident fits for callers that keep the graph on the GPU between
 * calls (benchmarks, servers): dress_hip_resident_reset() loads
 * g->edge_dress onto the device; dress_hip_fit_resident() then iterates the
 * device copy without any host transfer; dress_hip_resident_download()
 * copies values and norms back into g.  Return 0 on success, -1 on error. */
int dress_hip_resident_reset(p_dress_graph_t g);
int dress_hip_fit_resident(p_dress_graph_t g, int max_iterations, double epsilon,
                           dress_hip_precision_t precision,
                           int *iterations, double *delta);
int dress_hip_resident_download(p_dress_graph_t g);

/* Sets every resident edge value to `value` on the device (a fresh start is
 * value = 1.0, the state dress_init_graph creates).  Vertex weights stay as
 * the last dress_hip_resident_reset or fit read them from g->NW.  Returns 0
 * or -1. */
int dress_hip_resident_fill(p_dress_graph_t g, double value);

/* Device time of the kernels of one FP64 iteration, averaged over `reps`
 * iterations run on the resident state (ms holds 12 entries): ms[0..4] the
 * vertex phase and ms[5..9] the edge phase of each of the five work bins
 * launched on its own (0 for an empty bin); ms[10] the vertex phase and
 * ms[11] the edge phase as fits launch them (one kernel for the vertex
 * phase, one per edge bin, back to back).  Leaves the
 * resident values as they were.  Returns 0 or -1. */
int dress_hip_time_phases(p_dress_graph_t g, int reps, double *ms);

/* Streams n_bytes (a multiple of 16) from one device buffer to another with
 * a plain copy kernel, `reps` times after one warm-up, and stores the mean
 * device time of one copy in *ms.  Measurement aid: the achievable HBM
 * bandwidth, and the reference reading for hardware byte counters.  No
 * counterpart in libdress.  Returns 0 or -1. */
int dress_hip_copy_probe(int64_t n_bytes, int reps, double *ms);
/* Measurement probes (no counterpart in the reference).  Mean ms of a
 * kernel of n_gathers random 8-byte reads from a src_bytes array (the
 * access pattern of the edge kernels' gathers, for calibrating byte
 * counters); and ms of one launch of an LDS read kernel filling the device,
 * with the LDS bytes it reads in *bytes. */
int dress_hip_gather_probe(int64_t src_bytes, int64_t n_gathers, int reps, double *ms);
int dress_hip_lds_probe(int rounds, double *ms, double *bytes);

/* dress_get for many pairs at once (libdress/src/dress.c:651-739 per
 * query), against g's current edge_dress / vertex_dress.  out[i] receives
 * the value of (u[i], v[i]) with hypothetical weight edge_weight[i].
 * Returns 0, or -1 (see dress_hip_last_error). */
int dress_hip_get_batch(p_dress_graph_t g, int n, const int *u, const int *v,
                        const double *edge_weight, int max_iterations, double epsilon,
                        double *out);

/* dress_hip_get_batch against the state the last completed dress_fit (or
 * dress_hip_fit_resident) left on the device -- the graph's values, vertex
 * norms and vertex weights as that fit ended -- instead of uploading g's
 * host arrays for the call: for link-prediction queries in batches after a
 * fit.  No counterpart in libdress.  Answers equal dress_hip_get_batch's
 * as long as g's host arrays still hold what the fit wrote.  Returns -1
 * (see dress_hip_last_error) when no such state is there: before the first
 * fit, or after host values were uploaded (a fit in progress, a reset) or
 * another fit ran on the graph's device copy (delta / nabla subgraphs run
 * one at a time). */
int dress_hip_get_batch_resident(p_dress_graph_t g, int n, const int *u, const int *v,
                                 const double *edge_weight, int max_iterations, double epsilon,
                                 double *out);

/* dress_hip_get_batch_resident without any device allocation, for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * hosts that forbid hipMalloc).  No counterpart in libdress.  At most
 * 65,536 queries per call (more is an error); they are staged in a fixed
 * 1.5 MiB device array of the library, one per device, held until the
 * process ends, so calls that overlap on one device race.  The graph must
 * already hold a fitted state and its query rows on the device: one call of
 * dress_hip_get_batch_resident after the fit places them.  Answers equal
 * dress_hip_get_batch_resident's.  Returns 0, or -1 (see
 * dress_hip_last_error). */
int dress_hip_get_batch_resident_noalloc(p_dress_graph_t g, int n, const int *u, const int *v,
                                         const double *edge_weight, int max_iterations, double epsilon,
                                         double *out);

/* How dress_{delta,nabla}_fit_hip* run their subgraphs:
 *   0 (default) one batched persistent launch when the graph fits the
 *     kernel's LDS state (up to 65535 edges and ~150 KB of state), otherwise
 *     fits on the graph's own device copy: its CSR and intersection index
 *     serve every subgraph, delta^k marks the removed edges dead, nabla^k
 *     re-weights the tuple's vertices; up to 256 subgraphs share one pass
 *     (k <= 8), one at a time otherwise;
 *   1 always one device fit per rebuilt subgraph (the structure of the CUDA
 *     backend, libdress/src/cuda/delta_dress_cuda.c:24-51), kept for
 *     comparison;
 *   2 always the one-at-a-time fits on the graph's device copy;
 *   3 always the multi-subgraph passes where they apply (2 otherwise).
 * All give the same values bit for bit in FP64; the multi-subgraph passes
 * run in FP64 whatever the precision setting. */
void dress_hip_set_lifted_mode(int mode);

/* Host-only helper: the subgraph indices a lifted call with these arguments
 * processes (kind 0 delta / 1 nabla), in order, with their k vertices each
 * (combinations and ordered tuples in lexicographic order, stride/offset and
 * sampling as in libdress/src/delta_dress_impl.c:295-371).  Writes up to
 * `cap` entries; returns the full count, or -1 on bad arguments. */
int64_t dress_hip_enumerate_subgraphs(int kind, int N, int k, int n_samples, unsigned int seed,
                                      int offset, int stride, int64_t *sel, int *verts, int64_t cap);

/* Counters of the calling thread's last lifted call. */
typedef struct {
    int64_t subgraphs;            /* subgraphs fitted                         */
    int64_t values;               /* live edge values produced                */
    int     fp64_reruns;          /* mixed-precision subgraphs redone in FP64 */
    double  device_ms;            /* time of the batched launch(es)           */
} dress_hip_lifted_stats_t;
void dress_hip_last_lifted_stats(dress_hip_lifted_stats_t *out);

/* delta^k (kind 0) or nabla^k (kind 1) of many graphs in one batched
 * launch.  Graph g has N[g] vertices and the edge records
 * U/V/W[edge_offset[g] .. edge_offset[g+1]) (W may be NULL); all graphs use
 * `variant`.  Every subgraph of every graph is fitted (no sampling or
 * striding).  Per graph the result holds
 *   - the exact histogram: distinct values ascending with their counts;
 *   - on request (qtol >= 0; qtol < 0 leaves them empty): the quantised
 *     histogram, counts of floor(d / epsilon), accumulated on the device in
 *     integer atomics, for values farther than qtol from a bin edge; values
 *     within qtol of an edge are counted separately in qbound (their bin
 *     depends on the last bits of d).
 * Large arrays are pinned host memory kept in a pool for later calls;
 * release every array only with dress_hip_lifted_batch_free.
 * Returns 0, or -1 (see dress_hip_last_error). */
typedef struct {
    int      n_graphs;
    int64_t *hist_offset;         /* [n_graphs+1] into hist_value/hist_count  */
    double  *hist_value;
    int64_t *hist_count;
    int64_t *qhist_offset;        /* [n_graphs+1] into qhist_key/qhist_count  */
    int64_t *qhist_key;
    int64_t *qhist_count;
    int64_t *qbound_offset;       /* [n_graphs+1] into qbound_key/qbound_count */
    int64_t *qbound_key;
    int64_t *qbound_count;
    int64_t  subgraphs;
    int64_t  values;
    int      fp64_reruns;
    double   device_ms;           /* batched launch(es)                       */
    double   total_ms;            /* whole call, wall clock                   */
    int64_t  work[4];             /* with DRESS_HIP_LIFT_WORK=1: sums over the
                                     subgraphs of iterations run x their index
                                     entries, half-edges, edges, vertices;
                                     else zeros                               */
} dress_hip_lifted_batch_t;
int dress_hip_lifted_batch(int kind, int k, int n_graphs, const int *N, const int64_t *edge_offset,
                           const int *U, const int *V, const double *W, dress_variant_t variant,
                           int max_iterations, double epsilon, dress_hip_precision_t precision,
                           double qtol, dress_hip_lifted_batch_t *out);
void dress_hip_lifted_batch_free(dress_hip_lifted_batch_t *r);

/* Whole-graph fits of many small graphs in one persistent launch (no
 * counterpart in libdress): graph g is vertices [0, N[g]) and edge records
 * [edge_offset[g], edge_offset[g+1]) of U/V/W (W may be NULL), with vertex
 * weights 1.0.  Each graph runs from d = 1 in one workgroup with its state
 * in LDS, so it must satisfy dress_hip_fit_batch_eligible().  Per graph the
 * FP64 results equal those of dress_init_graph + dress_fit_hip_ex on it, bit
 * for bit (FP32->FP64: the same stopping step and every value within 1e-9
 * of FP64, deterministic but rounded on its own in the FP32 phase):
 * iterations and delta in the reference's convention, edge values in
 * edge order (concatenated like the input), vertex_dress (the norms of the
 * last vertex phase) concatenated at vertex_offset.  Large arrays are
 * pinned host memory kept in a pool for later calls; release every array
 * only with dress_hip_fit_batch_free.  Returns 0, or -1. */
typedef struct {
    int      n_graphs;
    int     *iterations;          /* [n_graphs]                               */
    double  *delta;               /* [n_graphs]                               */
    double  *edge_dress;          /* [edge_offset[n_graphs]]                  */
    int64_t *vertex_offset;       /* [n_graphs+1] into vertex_dress           */
    double  *vertex_dress;
    int      fp64_reruns;
    double   device_ms;
    double   total_ms;
} dress_hip_fit_batch_t;
int dress_hip_fit_batch_eligible(int N, int E);
int dress_hip_fit_batch(int n_graphs, const int *N, const int64_t *edge_offset, const int *U, const int *V,
                        const double *W, dress_variant_t variant, int max_iterations, double epsilon,
                        dress_hip_precision_t precision, dress_hip_fit_batch_t *out);
void dress_hip_fit_batch_free(dress_hip_fit_batch_t *r);

/* ---- serving (no counterpart in libdress) ---------------------------
 * A request server for streams of fit / delta^k / nabla^k requests of mixed
 * sizes.  Worker threads, each with its own stream, take requests as soon
 * as their estimated device-memory footprint fits in the budget next to
 * the requests already running, so work of many requests overlaps on the
 * device; requests whose graph fits the LDS-resident kernel (fits as whole
 * graphs, delta/nabla as subgraphs) are packed with the others of their
 * kind and parameters waiting at that moment into one batched launch
 * (dress_hip_fit_batch, dress_hip_lifted_batch).  Workers are split into
 * lanes -- batched fits, device-wide fits, batched delta/nabla, and the
 * delta/nabla requests too large for the batched kernel, which no other
 * lane takes -- so short requests keep flowing behind long ones.  That
 * lane has one worker, holds at most half of the budget, fills at most a
 * quarter of the device while other requests wait or run, and pauses
 * between its passes for them; batched delta/nabla launches fill at most
 * half the device while fits wait or run.  Delta/nabla requests run in
 * chunks of subgraphs whose values fit a fixed device share (32M values),
 * each chunk's histogram merged on the host, so their footprint is bounded
 * whatever their size.  A request whose footprint still exceeds the budget
 * of its kind (the whole budget, half of it for the large lane), or whose
 * result would exceed the server's host-memory budget for one result
 * (DRESS_HIP_SERVE_HOST_BYTES, default 32 GiB; also the int hist_size of
 * the API), is answered at once with an error.  After a device error that
 * leaves the device unusable (an illegal address, a failed launch, ...)
 * the server stops: queued requests are answered with that error, submit
 * returns -1, and dress_hip_server_stats reports device_failed = 1.
 * A request's result does not depend on what runs beside it;
 * FP64 results are those of the single-graph calls (dress_fit_hip_ex,
 * dress_delta_fit_hip, dress_nabla_fit_hip) bit for bit, FP32->FP64 ones
 * within 1e-9 of them. */
typedef struct dress_hip_server dress_hip_server_t;

typedef enum {
    DRESS_HIP_REQ_FIT   = 0,
    DRESS_HIP_REQ_DELTA = 1,
    DRESS_HIP_REQ_NABLA = 2
} dress_hip_request_kind_t;

typedef struct {
    int kind;                      /* dress_hip_request_kind_t             */
    int k;                         /* delta/nabla order                    */
    int N, E;
    const int *U, *V;              /* copied by dress_hip_server_submit    */
    const double *W;               /* NULL: unweighted                     */
    dress_variant_t variant;
    int max_iterations;
    double epsilon;
    dress_hip_precision_t precision;
    double arrival;                /* dress_hip_clock() seconds; < 0: now   */
} dress_hip_request_t;

typedef struct {
    int64_t id;
    int status;                    /* 0 done, -1 failed or rejected        */
    int kind, N, E;
    int iterations;                /* fits                                 */
    double delta;
    double *edge_dress;            /* fits: E values, malloc'd             */
    double *vertex_dress;          /* fits: N values, malloc'd             */
    dress_hist_pair_t *hist;       /* delta/nabla: malloc'd                */
    int hist_size;
    int64_t subgraphs;
    int batch_size;                /* requests sharing the launch          */
    int64_t footprint;             /* estimated device bytes               */
    double arrival, admitted, finished;   /* dress_hip_clock() seconds     */
    char error[192];
} dress_hip_response_t;

typedef struct {
    int64_t submitted, completed, failed, rejected;
    int64_t queued, running, max_queued;
    int64_t budget_bytes, reserved_bytes, peak_reserved_bytes;
    int64_t lifted_launches, lifted_requests;  /* batched launches, requests in them */
    int64_t rejected_host;         /* of rejected: result beyond the host budget  */
    int64_t chunked_requests;      /* delta/nabla requests run in several chunks   */
    int64_t device_failed;         /* 1 once a device error stopped the server    */
} dress_hip_server_stats_t;

/* Monotonic clock in seconds, the time base of arrival/admitted/finished. */
double dress_hip_clock(void);

/* workers >= 1 threads on the current device; budget_bytes <= 0 takes half
 * of the device's free memory; max_batch bounds the requests packed into
 * one launch.  NULL on error. */
dress_hip_server_t *dress_hip_server_create(int workers, int64_t budget_bytes, int max_batch);

/* Copies the request and queues it.  Returns its id (>= 0), or -1. */
int64_t dress_hip_server_submit(dress_hip_server_t *s, const dress_hip_request_t *req);

/* Waits up to timeout_s for a finished request (any, in completion order).
 * Returns 1 and fills *out (free with dress_hip_response_free), or 0 on
 * timeout. */
int dress_hip_server_next(dress_hip_server_t *s, double timeout_s, dress_hip_response_t *out);
void dress_hip_response_free(dress_hip_response_t *r);
void dress_hip_server_stats(dress_hip_server_t *s, dress_hip_server_stats_t *out);

/* Failure-path test hook: the next request a worker finishes is failed as if
 * the device had reported the HIP error code `hip_error`, and the server's
 * device-error check runs on that code (a code that leaves the device
 * unusable stops the server).  No device work is involved.  Returns 0. */
int dress_hip_server_inject_error(dress_hip_server_t *s, int hip_error);

/* Drops queued requests, stops running delta/nabla requests at their next
 * subgraph (they finish with status -1), waits for the rest, frees the
 * server. */
void dress_hip_server_destroy(dress_hip_server_t *s);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif

#endif /* DRESS_HIP_H */
