// dress_hip_internal.h -- internal interfaces of libdress_hip.so (HIP, gfx950).
//
// Device layout of one graph (all arrays in HBM, built by dress_build.hip):
//
//   row_off[N+1]  int32   CSR row offsets of the variant adjacency
//   adj_x[S]      int32   neighbour id of half-edge i            (sorted (x, e))
//   adj_e[S]      int32   input edge of half-edge i
//   U[E], V[E]    int32   edge endpoints
//   ew[E]         f64     variant edge weight  w-bar
//   c4[E]         f64     self term constant 4nw_u + 4nw_v (4nw_u for FWD/BWD)
//   pos[E]        int2    the (at most two) half-edges that carry edge e
//   ioff[E+1]     int64   intersection index offsets
//   ipair[T]      int2    (half-edge in row U[e], half-edge in row V[e]) per
//                         common neighbour, in ascending neighbour order
//
// Fit state: d[2][E] (edge values, double buffer), wh[2][S] (w-bar*d copied
// onto every half-edge, double buffer), vd[N] (vertex norms).  Because every
// term of a row or intersection sum is read from wh[] by half-edge position,
// the vertex phase streams contiguous rows and the edge phase gathers inside
// the two rows of the edge.
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

// ---------------------------------------------------------------- graph
enum Precision { PREC_FP64 = 0, PREC_FP32_FP64 = 1 };

struct Bins {                      // degree-binned work lists
    static constexpr int NB = 5;   // group sizes 1, 4, 16, 64, 256
    int *list = nullptr;           // concatenated item ids, bin by bin
    int count[NB] = {0, 0, 0, 0, 0};
    int start[NB] = {0, 0, 0, 0, 0};
};

struct DevGraph {
    int N = 0, E = 0, variant = 0;
    int64_t S = 0, T = 0;
    int max_degree = 0;
    bool weighted = false;
    hipStream_t stream = nullptr;

    int *U = nullptr, *V = nullptr;
    int *row_off = nullptr, *adj_x = nullptr, *adj_e = nullptr;
    double *ew = nullptr, *nw = nullptr, *c4 = nullptr;
    int2 *pos = nullptr;
    int64_t *ioff = nullptr;
    int2 *ipair = nullptr;
    bool have_index = false;

    Bins vbins, ebins;             // vertex rows by degree, edges by |isect|
    bool have_bins = false;

    // fit state
    double *d[2] = {nullptr, nullptr};
    double *wh[2] = {nullptr, nullptr};
    double *vd = nullptr;
    float *d32[2] = {nullptr, nullptr};
    float *wh32[2] = {nullptr, nullptr};
    void *state = nullptr;         // FitState (device)
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

// dress_prim.hip (rocPRIM wrappers, stream-ordered, scratch allocated inside)
bool prim_sort_pairs_u64_i32(const uint64_t *kin, uint64_t *kout, const int *vin, int *vout,
                             size_t n, int end_bit, hipStream_t s);
bool prim_sort_u64(const uint64_t *kin, uint64_t *kout, size_t n, hipStream_t s);
bool prim_exclusive_scan_i64(const int64_t *in, int64_t *out, size_t n, hipStream_t s);
bool prim_exclusive_scan_i32_to_i64(const int *in, int64_t *out, size_t n, hipStream_t s);

// misc
int current_device();
int num_cus();

}  // namespace dh
