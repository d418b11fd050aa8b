// dress_masked.hip -- delta^k / nabla^k subgraphs of graphs too large for
// the LDS-resident kernel, as masked fits of the parent's device graph.
//
// The reference rebuilds every subgraph and fits it from scratch
// (delta_dress_impl.c:95-169, nabla_dress_impl.c:61-91).  Here the parent's
// device graph -- CSR, intersection index, bins -- serves every subgraph:
//   delta^k  the edges touching a deleted vertex are marked dead (self term
//            NaN, value 0); the edge kernel keeps their value and half-edge
//            copies at 0 and leaves them out of the change, so every live
//            sum sees exact zeros where the subgraph has no term.  Vertex
//            weights are 1.0 (the subgraph is built with NW = NULL).  In
//            DIRECTED graphs with repeated reciprocal edges of different
//            weights, the subgraph's own reciprocal lookup can pick another
//            of them: those edges' weights are patched for the item.
//   nabla^k  no edge is removed; the tuple's vertices get the marker
//            weights and the self terms are recomputed.
// Sums are rounded once from their exact value, and adding exact zeros
// does not change that value, so each live value equals the value on the
// rebuilt subgraph bit for bit.  Live values stay on the device for the
// histogram.
#include "dress_batch.h"
#include "dress_capi_internal.h"
#include "dress_hip_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace dh {

namespace {

constexpr int TPB = 256;

inline int grid_for(int64_t threads, int cap)
{
    int64_t b = (threads + TPB - 1) / TPB;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

__global__ void k_fill_f64(double *p, int64_t n, double v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// self terms from the vertex weights, internal edge order (the arithmetic
// of the graph build: 4nw_u + 4nw_v, or 4nw_u for FORWARD / BACKWARD)
__global__ void k_self_from_nw(const int2 *uv, int E, const double *nw, int directed, double *c4)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        int2 p = uv[i];
        double a = 4.0 * nw[p.x];
        c4[i] = directed ? a : a + 4.0 * nw[p.y];
    }
}

// delta: marks (dead) or restores the edges touching the k deleted
// vertices (input labels in verts)
__global__ void k_mask_edges(const int2 *uv, int E, const int *verts, int k, const int *rank, int dead,
                             double live_c4, double *c4, double *d)
{
    __shared__ int del[LIFT_KMAX];
    if ((int)threadIdx.x < k) del[threadIdx.x] = rank[verts[threadIdx.x]];
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        int2 p = uv[i];
        bool hit = false;
        for (int j = 0; j < k; j++) hit = hit || p.x == del[j] || p.y == del[j];
        if (!hit) continue;
        if (dead) {
            c4[i] = __longlong_as_double(0x7ff8000000000000ll);
            d[i] = 0.0;
        } else {
            c4[i] = live_c4;
        }
    }
}

// nabla: the tuple's vertex weights (marks) or 1.0 back
__global__ void k_mark_nw(const int *verts, int k, const int *rank, const double *marks, int set, double *nw)
{
    int j = threadIdx.x;
    if (j < k) nw[rank[verts[j]]] = set ? marks[j] : 1.0;
}

// live values (self term not NaN) appended at out[0..): order does not
// matter to the histogram
__global__ void k_collect_live(const double *d, const double *c4, int E, double *out, unsigned long long cap,
                               unsigned long long *count)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        double c = c4[i];
        if (c == c) {
            unsigned long long at = atomicAdd(count, 1ull);
            if (at < cap) out[at] = d[i];
        }
    }
}

// variant weights of input edges e[0..n) set to w[0..n)
__global__ void k_patch_ew(const int *inv, const int *e, const double *w, int n, double *ew)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ew[inv[e[i]]] = w[i];
}

// multiset row in input edge order, NaN for removed edges
__global__ void k_multiset_row(const int *inv, const double *d, const double *c4, int E, double *row)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        int i = inv[e];
        double c = c4[i];
        row[e] = (c == c) ? d[i] : __longlong_as_double(0x7ff8000000000000ll);
    }
}

}  // namespace

// Fits every item (k vertices each in verts) as a masked fit of G (built
// from g).  Live values are appended to d_vals (device, room for cap
// values: the sum of live); multiset rows, when ms_rows is non-null, are
// written into ms_rows + sel[i] * E on the host.  Item i's variant weights
// patch_w[patch_off[i] ..) replace those of input edges patch_e[..] for its
// fit (DIRECTED reciprocal weights of the subgraph) and patch_parent_w
// puts the parent's back; patch_off is empty when no item has patches.
// Vertex weights are restored from g->NW at the end.  Returns false with
// the error set on failure.
bool masked_items(p_dress_graph_t g, DevGraph *G, int kind, int k, const std::vector<int> &verts,
                  const std::vector<int64_t> &live, const std::vector<int64_t> &sel, const double *marks, int max_it,
                  double eps, int precision, double *d_vals, int64_t cap, int64_t *n_vals, double *ms_rows,
                  const std::vector<int64_t> &patch_off, const std::vector<int> &patch_e,
                  const std::vector<double> &patch_w, const std::vector<double> &patch_parent_w)
{
    *n_vals = 0;
    if (!build_bins(G) || !fit_alloc_state(G, precision == PREC_FP32_FP64)) return false;
    // items mark removed edges and vertex weights per fit: not uniform
    // until the graph's own self terms are back
    G->uni_state = 2;
    G->self_terms_ones = false;
    hipStream_t s = G->stream;
    const int E = G->E, N = G->N;
    const int directed = (G->variant == DRESS_VARIANT_FORWARD || G->variant == DRESS_VARIANT_BACKWARD) ? 1 : 0;
    const double live_c4 = directed ? 4.0 : 8.0;
    const int n_items = (int)live.size();
    int *d_verts = nullptr;
    double *d_marks = nullptr, *d_row = nullptr;
    unsigned long long *d_count = nullptr;
    bool ok = dalloc(&d_verts, verts.size() + 1, "item vertices") && dalloc(&d_marks, LIFT_KMAX, "marker weights") &&
              dalloc(&d_count, 1, "value count") && (!ms_rows || dalloc(&d_row, (size_t)E + 1, "multiset row"));
    if (ok && !verts.empty())
        ok = hipMemcpyAsync(d_verts, verts.data(), verts.size() * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok && marks) ok = hipMemcpyAsync(d_marks, marks, (size_t)k * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok) ok = hipMemsetAsync(d_count, 0, sizeof(unsigned long long), s) == hipSuccess;
    const bool patched = !patch_off.empty();
    int *d_pe = nullptr;
    double *d_pw = nullptr, *d_ppw = nullptr;
    if (ok && patched) {
        const size_t np = patch_e.size();
        ok = dalloc(&d_pe, np + 1, "weight patch edges") && dalloc(&d_pw, np + 1, "weight patches") &&
             dalloc(&d_ppw, np + 1, "parent weights") &&
             hipMemcpyAsync(d_pe, patch_e.data(), np * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(d_pw, patch_w.data(), np * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(d_ppw, patch_parent_w.data(), np * sizeof(double), hipMemcpyHostToDevice, s) ==
                 hipSuccess;
    }
    // vertex weights 1.0, self terms to match (both kinds start from there)
    const int egrid = grid_for(E, num_cus() * 8);
    if (ok && N > 0) hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(N, 4096)), dim3(TPB), 0, s, G->nw, (int64_t)N, 1.0);
    if (ok && E > 0) hipLaunchKernelGGL(k_self_from_nw, dim3(egrid), dim3(TPB), 0, s, G->uv, E, G->nw, directed, G->c4);
    // the item whose weight patch is on the device and not yet undone: a
    // failed fit leaves the loop before the undo, and the parent's weights
    // must come back on every exit (later fits and dress_get read them)
    int64_t undo_p0 = -1;
    int undo_np = 0;
    for (int it = 0; ok && it < n_items; it++) {
        if (thread_cancelled()) {
            set_error("cancelled");
            ok = false;
            break;
        }
        if (kind == LIFT_DELTA && live[(size_t)it] == 0) continue;   // delta_dress_impl.c:124-129
        const int *vi = d_verts + (size_t)it * k;
        ok = fit_fill_values(G, 1.0);
        if (!ok) break;
        if (kind == LIFT_DELTA) {
            if (k > 0)
                hipLaunchKernelGGL(k_mask_edges, dim3(egrid), dim3(TPB), 0, s, G->uv, E, vi, k, G->rank, 1, live_c4,
                                   G->c4, G->d[0]);
        } else {
            hipLaunchKernelGGL(k_mark_nw, dim3(1), dim3(LIFT_KMAX), 0, s, vi, k, G->rank, d_marks, 1, G->nw);
            hipLaunchKernelGGL(k_self_from_nw, dim3(egrid), dim3(TPB), 0, s, G->uv, E, G->nw, directed, G->c4);
        }
        const int64_t p0 = patched ? patch_off[(size_t)it] : 0;
        const int np = patched ? (int)(patch_off[(size_t)it + 1] - p0) : 0;
        if (np > 0) {
            hipLaunchKernelGGL(k_patch_ew, dim3(grid_for(np, 64)), dim3(TPB), 0, s, G->inv, d_pe + p0, d_pw + p0, np,
                               G->ew);
            undo_p0 = p0;
            undo_np = np;
        }
        FitResult r;
        ok = hipGetLastError() == hipSuccess && fit_run(G, max_it, eps, precision, &r);
        if (!ok) break;
        const double *dc = G->d[G->cur];
        hipLaunchKernelGGL(k_collect_live, dim3(egrid), dim3(TPB), 0, s, dc, G->c4, E, d_vals,
                           (unsigned long long)cap, d_count);
        if (ms_rows) {
            hipLaunchKernelGGL(k_multiset_row, dim3(egrid), dim3(TPB), 0, s, G->inv, dc, G->c4, E, d_row);
            ok = hipMemcpyAsync(ms_rows + (size_t)sel[(size_t)it] * (size_t)E, d_row, (size_t)E * sizeof(double),
                                hipMemcpyDeviceToHost, s) == hipSuccess;
        }
        // back to the parent: self terms of the removed edges, or weights
        if (np > 0) {
            hipLaunchKernelGGL(k_patch_ew, dim3(grid_for(np, 64)), dim3(TPB), 0, s, G->inv, d_pe + p0, d_ppw + p0, np,
                               G->ew);
            undo_np = 0;
        }
        if (kind == LIFT_DELTA) {
            if (k > 0)
                hipLaunchKernelGGL(k_mask_edges, dim3(egrid), dim3(TPB), 0, s, G->uv, E, vi, k, G->rank, 0, live_c4,
                                   G->c4, G->d[0]);
        } else {
            hipLaunchKernelGGL(k_mark_nw, dim3(1), dim3(LIFT_KMAX), 0, s, vi, k, G->rank, d_marks, 0, G->nw);
            hipLaunchKernelGGL(k_self_from_nw, dim3(egrid), dim3(TPB), 0, s, G->uv, E, G->nw, directed, G->c4);
        }
        ok = ok && hipGetLastError() == hipSuccess;
    }
    if (undo_np > 0) {
        hipLaunchKernelGGL(k_patch_ew, dim3(grid_for(undo_np, 64)), dim3(TPB), 0, s, G->inv, d_pe + undo_p0,
                           d_ppw + undo_p0, undo_np, G->ew);
        (void)hipGetLastError();
    }
    unsigned long long cnt = 0;
    if (ok)
        ok = hipMemcpyAsync(&cnt, d_count, sizeof(cnt), hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    // the graph's own vertex weights and self terms again
    bool restored = refresh_self_terms(G, g->NW) && hipStreamSynchronize(s) == hipSuccess;
    G->uni_state = 0;
    dfree(d_verts); dfree(d_marks); dfree(d_count); dfree(d_row);
    dfree(d_pe); dfree(d_pw); dfree(d_ppw);
    if (ok && !restored) ok = false;
    if (ok && cnt > (unsigned long long)cap) {
        set_error("masked subgraph fits produced more live values than counted");
        ok = false;
    }
    if (!ok && !has_error()) set_error("masked subgraph fits failed");
    *n_vals = (int64_t)cnt;
    return ok;
}

}  // namespace dh
