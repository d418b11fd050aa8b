// dress_batch.h -- interfaces of the batched delta^k / nabla^k kernel.
#pragma once

#include "dress_hip_internal.h"

namespace dh {

enum { LIFT_DELTA = 0, LIFT_NABLA = 1 };
constexpr int LIFT_KMAX = 64;      // largest k a work item may carry

// one parent graph, as offsets into the concatenated pools
struct LiftParent {
    int N, E, S, T;
    int64_t row_ofs, adj_ofs, edge_ofs, ioff_ofs, pair_ofs;
    int directed;                  // FORWARD/BACKWARD self-term form
    int uni_w;                     // every variant weight is w_c (bit for bit)
    double w_c;
    int64_t ell_ofs;               // column-major edge ids at ell[ell_ofs ...], or -1
    int ell_stride;                // rows per column there (N rounded up to 64)
    int pad;
};

// one subgraph: parent + deleted vertices (delta) or marked tuple (nabla)
struct LiftItem {
    int parent;
    int kind;
    int k;
    int mixed;                     // start in FP32
    int64_t vofs;                  // item vertices at item_verts[vofs .. vofs+k)
    int64_t val_ofs;               // live values at vals[val_ofs ...]
    int64_t ms_row;                // multiset row
    int64_t vn_ofs;                // vertex norms at vnorm[vn_ofs ...] (whole-graph fits)
    int64_t ew_ofs;                // variant weights at ew[ew_ofs ...], or -1: the parent's
};

struct LiftArgs {
    const LiftParent *parents;
    const LiftItem *items;
    const int *order;              // queue order (item ids), or null
    int n_items;
    const int *item_verts;
    const double *marks;           // nabla marker weights [LIFT_KMAX]
    const int *row_off;            // per parent: N+1 local offsets
    const int *adj_e;              // per parent: S local edge ids
    const int *ell;                // per parent (or none): the same, column by column
    const int2 *uv;                // per parent: E (U, V)
    const double *ew;              // per parent: E variant weights
    const int *ioff;               // per parent: E+1 local offsets
    const uint32_t *pairs;         // per parent: T (edge(u,x) | edge(v,x) << 16)
    int *queue;                    // work counter (zeroed before launch)
    int ecap, ncap;                // LDS capacity in edges / vertices
    int scap;                      // ... and half-edges (structure staged in LDS)
    int max_it;
    double eps;
    double *vals;                  // packed live values (or null)
    int sort_vals;                 // 1: each item's live values ascending (by sort key), else edge order
    double *ms;                    // multiset rows (or null)
    int64_t ms_stride;
    double *vnorm;                 // per item: the norms of the last vertex phase (or null)
    int *iters;                    // per item (or null)
    double *deltas;                // per item (or null)
    int *flags;                    // per item: 1 = redo in FP64 (or null)
    // quantised histogram: open-addressing table keyed by
    // (graph << 41) | (boundary << 40) | floor(d / eps), integer counts
    unsigned long long *qkeys;     // capacity qcap (power of two), or null
    unsigned long long *qcounts;
    int64_t qcap;
    double qtol;                   // values within qtol of a bin edge are boundary values
    const int *parent_graph;       // graph id of each parent (for the key)
    // work counters (or null): sums over items of executed iterations x the
    // subgraph's index entries, half-edges, edges and vertices
    unsigned long long *work;
};

constexpr unsigned long long QEMPTY = ~0ull;

// ascending order of doubles as unsigned keys (sign-folded bit patterns):
// the order of the device histograms' keys
__device__ __forceinline__ unsigned long long sort_key(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// quantised-table key of value v of graph gid: floor(v / eps) (clamped to
// 40 bits), with bit 40 set for values within tol of a bin edge
__device__ __forceinline__ unsigned long long quant_key(double v, unsigned long long gid, double eps, double tol)
{
    const double r = v / eps;
    const double q = floor(r);
    const bool bnd = (r - q) * eps < tol || (q + 1.0 - r) * eps < tol;
    const unsigned long long qq =
        (q < 0.0) ? 0ull : (q >= 1099511627775.0 ? 1099511627775ull : (unsigned long long)q);
    return (gid << 41) | ((unsigned long long)bnd << 40) | qq;
}

// first probe slot of a key in a table of cap (a power of two) slots
__device__ __forceinline__ int64_t quant_slot(unsigned long long key, int64_t cap)
{
    return (int64_t)((key * 0x9E3779B97F4A7C15ull) >> 20) & (cap - 1);
}

size_t lifted_lds_bytes(int ecap, int ncap);
// with the parent's structure staged in LDS as well (scap half-edges)
size_t lifted_lds_bytes_staged(int ecap, int ncap, int scap);
size_t lifted_lds_limit();
size_t lifted_lds_per_cu();
bool lifted_launch(const LiftArgs &a, int block, int grid, bool staged, hipStream_t s);

}  // namespace dh
