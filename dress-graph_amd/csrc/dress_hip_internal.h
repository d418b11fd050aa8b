// dress_hip_internal.h -- internal interfaces of libdress_hip.so (HIP, gfx950).
//
// Device layout of one graph (all arrays in HBM, built by dress_build.hip):
//
//   row_off[N+1]  int32   CSR row offsets of the variant adjacency
//   adj_x[S]      int32   neighbour id of half-edge s, rows sorted by (x, edge)
//   adj_e[S]      int32   input edge of half-edge s
//
// Edges are renumbered internally: internal ids run bin by bin (bin chosen by
// the intersection size) and, inside a bin, in CSR order of the edge's first
// half-edge, so every bin is a contiguous range read with unit stride and
// consecutive edges share a row.
//   ord[i] / inv[e]       internal -> input / input -> internal
//   uv[i]         int2    (U, V) of internal edge i
//   rs[i]         int2    (row_off[U], row_off[V])
//   pos[i]        int2    the half-edges that carry edge i (y = -1 if one)
//   ew[i], c4[i]  f64     w-bar and the self-term constant 4nw_u + 4nw_v
//                         (4nw_u for FORWARD/BACKWARD)
//   ioff[i]       int64   intersection index offsets
//   ipk[k]        u32     (a - rs.x) | (b - rs.y) << 16 for the half-edges a
//                         in row U and b in row V of each common neighbour,
//                         ascending in the neighbour (max degree < 65536)
//   ipair[k]      int2    (a, b) absolute, for graphs with larger degrees
//
// Fit state: d[2][E] edge values (internal order, double buffer); wh[2][S]
// the value w-bar*d copied onto each of the edge's half-edges (double buffer),
// so the vertex phase streams rows and the common-neighbour gathers of one
// edge fall inside its two rows, close together; vd[N] norms.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>

#include "dress/dress.h"

namespace dh {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
void clear_error();
bool has_error();

#define DH_CHECK(call)                                                          \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            ::dh::set_error(std::string(#call) + ": " + hipGetErrorString(e_) + \
                            " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
            return false;                                                       \
        }                                                                       \
    } while (0)

// device buffer helpers (return false and set the error on failure)
bool dmalloc(void **p, size_t bytes, const char *what);
template <typename T>
inline bool dalloc(T **p, size_t n, const char *what) {
    return dmalloc(reinterpret_cast<void **>(p), n * sizeof(T) + 16, what);
}
void dfree(void *p);

// Execution context of the calling host thread (servers): while active,
// graphs built on this thread run on `stream`, and device blocks freed by
// dfree() are kept in a per-thread cache and handed out again by dmalloc()
// instead of being returned with hipFree (which waits for the whole device
// and would serialise concurrent requests).  Blocks are reused only by the
// thread that freed them, i.e. in the order of its stream.
void ctx_begin(hipStream_t stream);
void *host_alloc(size_t bytes);    // malloc-compatible (free() releases it)
void ctx_end();                    // returns the cached blocks to the device
// pinned host blocks for batch results; pinned_free also takes malloc'd ones
// host_writes: the caller fills the block from host threads (its pages are
// touched once, the first time it is handed out so)
void *pinned_alloc(size_t bytes, bool host_writes = false);
// every page of [p, p + bytes) written once (several threads from 32 MiB)
void touch_pages(void *p, size_t bytes);
void pinned_free(void *p);
// large copies between pageable host memory and the device, host-parallel
bool d2h_parallel(void *dst, const void *src, size_t bytes);
// the flat (value bits, count) histogram of runs values into pageable dst,
// from device values and counts narrowed to width (1, 2, 4) bytes
bool d2h_flat_narrow(int64_t *dst, const double *vals, const void *counts, int width, int64_t runs);
bool h2d_parallel(void *dst, const void *src, size_t bytes);
// pinned blocks of SNAPSHOT_BYTES for the state snapshots of a running fit
// or pass: take at its start, give back at its end (nullptr on failure,
// with the error set)
constexpr size_t SNAPSHOT_BYTES = 256;
void *snapshot_take();
void snapshot_give(void *p);
// n doubles set to a value (NaN), by several host threads when n is large
void fill_parallel(double *p, size_t n, double value);
void fill_nan_parallel(double *p, size_t n);
// every id of the edge list in [0, N), else the error is set (host scan)
bool check_vertex_ids(int N, int64_t E, const int *U, const int *V, const char *who);

void ctx_trim(size_t max_block, size_t keep_bytes);
bool ctx_serving();                // the calling thread is a server worker
// the calling thread's block cache for batch calls (kept between calls);
// enter returns its stream, or nullptr when a context is already active
hipStream_t batch_ctx_enter();
void batch_ctx_leave();
// two more streams of the calling thread's batch context (i = 0, 1;
// nullptr outside one, or inside a server worker), for work that overlaps
// the main stream's.  The block cache reuses a freed block on any stream:
// free a block only once the work that reads it has finished.
hipStream_t batch_side_stream(int i);
// The calling thread's batch block cache for the length of one batch call
// (a no-op inside a server worker, which has its own).  Enter it after the
// graph's own device arrays exist: blocks taken inside come back to the
// cache when freed.
struct BatchScope {
    bool on;
    BatchScope() : on(batch_ctx_enter() != nullptr) {}
    ~BatchScope()
    {
        if (on) batch_ctx_leave();
    }
    BatchScope(const BatchScope &) = delete;
    BatchScope &operator=(const BatchScope &) = delete;
};
hipStream_t ctx_stream();          // nullptr when no context is active
int64_t ctx_cached_bytes();        // bytes held by the cache (free + in use)

// ---------------------------------------------------------------- graph
enum Precision { PREC_FP64 = 0, PREC_FP32_FP64 = 1 };

struct Bins {                      // degree-binned vertex work list
    static constexpr int NB = 5;   // group sizes 1, 4, 16, 64, 256
    int *list = nullptr;           // vertex ids, bin by bin
    int count[NB] = {0, 0, 0, 0, 0};
    int start[NB] = {0, 0, 0, 0, 0};
};

// number of summed terms -> bin (group of 1, 4, 16, 64 lanes or a workgroup);
// the bounds were picked by timing the 4M-vertex power-law fit (5.60 ms per
// iteration with 4 / 32 / 256 / 4096, 5.48 ms with 4 / 64 / 512 / 8192)
__host__ __device__ inline int term_bin(int64_t terms)
{
    return terms <= 4 ? 0 : terms <= 64 ? 1 : terms <= 512 ? 2 : terms <= 8192 ? 3 : 4;
}

struct DevGraph {
    int N = 0, E = 0, variant = 0;
    int64_t S = 0, T = 0;
    int max_degree = 0;
    bool weighted = false;
    hipStream_t stream = nullptr;
    bool own_stream = true;        // false when the thread context lent it

    // Vertices carry internal labels: rank[v] is v's position in order of
    // decreasing degree (ties by id).  Rows list neighbours by rank, so a
    // row starts with its high-degree neighbours, where most common
    // neighbours of an edge lie, and the rows of the hubs are adjacent.
    int *U = nullptr, *V = nullptr;        // input order, internal labels
    int *row_off = nullptr, *adj_x = nullptr, *adj_e = nullptr;   // internal CSR
    double *ew_in = nullptr, *c4_in = nullptr;   // input order
    double *nw = nullptr;                  // internal labels
    int *rank = nullptr;                   // input label -> internal label
    // reference-layout CSR in input labels (libdress's adjacency), kept
    // until its download; rebuilt from the host struct for dress_get
    int *ref_U = nullptr, *ref_V = nullptr;
    int *ref_row_off = nullptr, *ref_adj_x = nullptr, *ref_adj_e = nullptr;
    double *vstage = nullptr;              // per-vertex staging (input labels)
    // dress_get's copy of the host state (input order and labels): edge
    // values, norms, vertex weights, written per call where queries read
    double *q_d = nullptr, *q_vd = nullptr, *q_nw = nullptr;
    int2 *pos_in = nullptr;                // half-edges of input edge e

    // internal edge order (valid once have_index)
    int *ord = nullptr, *inv = nullptr;
    int2 *uv = nullptr, *rs = nullptr, *pos = nullptr;
    double *ew = nullptr, *c4 = nullptr;
    int64_t *ioff = nullptr;
    uint32_t *ipk = nullptr;       // packed index (max degree < 65536)
    int2 *ipair = nullptr;         // absolute index otherwise
    bool packed = false;
    int64_t *cnt_in = nullptr;     // |intersection| per input edge
    int ebin[Bins::NB + 1] = {0, 0, 0, 0, 0, 0};  // bin ranges of internal ids
    bool have_index = false;

    Bins vbins;
    bool have_bins = false;
    // d[cur], vd and nw hold the state a completed dress_fit (or resident
    // fit) left on the device: set at its end, cleared by every other fit
    // run and every upload of host values into d
    bool fit_state_valid = false;

    // fit state
    double *d[2] = {nullptr, nullptr};     // internal order
    double *wh[2] = {nullptr, nullptr};    // per half-edge
    double *vd = nullptr;
    float *d32[2] = {nullptr, nullptr};
    float *wh32[2] = {nullptr, nullptr};
    double *stage = nullptr;       // input-order staging for uploads/downloads
    void *state = nullptr;         // FitState (device)
    hipGraphExec_t iter_graph[2] = {nullptr, nullptr};   // one iteration, [mixed]
    double graph_eps[2] = {0.0, 0.0};                      // ... captured with these
    int graph_maxit[2] = {0, 0};
    int graph_uni[2] = {0, 0};
    // bin 0 (one lane per edge): index entries as absolute half-edge pairs
    // behind 32-bit offsets relative to the bin (built with the fit state)
    int *b0off = nullptr;
    int2 *b0pair = nullptr;
    // Uniform per-edge constants: when every w-bar (ew), every self-term
    // constant (c4) and every vertex weight (nw) is the same -- unweighted
    // graphs without vertex weights, except DIRECTED -- the kernels take the
    // three values as arguments instead of reading the arrays.  uni_state:
    // 0 unknown (every writer of ew / c4 / nw resets it), 1 uniform, 2 not.
    int uni_state = 0;
    // the self terms (c4) were last computed from unit vertex weights and
    // nothing has written nw or c4 since: a refresh without vertex weights
    // has nothing to do
    bool self_terms_ones = false;
    // wh[cur] holds w-bar * d[cur] for the current values (set only by a
    // fill that prepared the half-edge copies, taken by the next fit)
    bool wh_ready = false;
    double ew_c = 0.0, c4_c = 0.0, nw_c = 0.0;
    int cur = 0;                   // buffer holding the current values
};

// dress_build.hip
bool build_graph(DevGraph *G, int N, int E, const int *U, const int *V,
                 const double *W, const double *NW, int variant);
bool build_index(DevGraph *G);
bool build_bins(DevGraph *G);
bool download_csr(DevGraph *G, int *adj_offset, int *adj_target, int *adj_edge_idx,
                  double *edge_weight);
bool download_index(DevGraph *G, int **ioff_out, int **ux_out, int **vx_out);
bool refresh_self_terms(DevGraph *G, const double *NW_host);
void release_reference_csr(DevGraph *G);
bool download_vertex_values(DevGraph *G, const double *internal, double *host);
bool upload_reference_csr(DevGraph *G, const int *adj_offset, const int *adj_target, const int *adj_edge_idx);
void free_graph(DevGraph *G);

// dress_fit.hip
struct FitResult {
    int iterations = 0;            // reference convention: 0-based converging
    double delta = 0.0;            // iteration or max_iterations
    int executed = 0;              // iterations actually run
    int fp32_iterations = 0;       // of which in FP32 storage (mixed mode)
    float ms = 0.f;                // device time of the iteration loop
    int reruns = 0;                // mixed fits redone in FP64 (ambiguous stop)
};
bool fit_alloc_state(DevGraph *G, bool fp32);
bool fit_upload_values(DevGraph *G, const double *edge_dress);
bool fit_run(DevGraph *G, int max_iterations, double epsilon, int precision,
             FitResult *res);
bool fit_download(DevGraph *G, double *edge_dress, double *vertex_dress);
// Sets every current edge value to `value`.  With prepare_wh (graphs whose
// w-bar is uniform) the half-edge copies are filled too, so the next FP64
// fit skips their set-up; any later write of the values clears that.
bool fit_fill_values(DevGraph *G, double value, bool prepare_wh = false);
bool fit_time_phases(DevGraph *G, int reps, double *ms);

// dress_prim.hip (rocPRIM wrappers, stream-ordered, scratch allocated inside)
bool prim_sort_pairs_u64_i32(const uint64_t *kin, uint64_t *kout, const int *vin, int *vout,
                             size_t n, int end_bit, hipStream_t s);
bool prim_sort_u64(const uint64_t *kin, uint64_t *kout, size_t n, hipStream_t s, int end_bit = 64);
// runs of equal keys of a sorted array: the distinct keys, their counts, and
// the number of runs (device int64)
// (with keep, the scratch is handed to the caller, to be freed once the
// encoding has run, instead of being freed at once)
bool prim_run_length_encode_u64(const uint64_t *in, size_t n, uint64_t *unique, int64_t *counts, int64_t *runs,
                                hipStream_t s, void **keep = nullptr);
bool prim_sort_pairs_u64_u64(const uint64_t *kin, uint64_t *kout, const uint64_t *vin, uint64_t *vout,
                             size_t n, hipStream_t s);
bool prim_sort_pairs_i32_u64(const int *kin, int *kout, const uint64_t *vin, uint64_t *vout, size_t n,
                             int end_bit, hipStream_t s);
bool prim_exclusive_scan_i64(const int64_t *in, int64_t *out, size_t n, hipStream_t s);
bool prim_exclusive_scan_i32_to_i64(const int *in, int64_t *out, size_t n, hipStream_t s);

// misc
int current_device();
int num_cus();
int num_xcds();   // XCDs (accelerator complex dies) of the current device

}  // namespace dh
