// dress_query.hip -- dress_get on the device, a group of lanes per query
// (libdress/src/dress.c:651-739).
//
// A query (u, v) returns the value of an existing edge found by the same
// binary search as the reference, or, for a non-edge, the 1-D fixed point of
// the virtual edge against the frozen state: A = sum over common neighbours of
// w-bar*d(u,x) + w-bar*d(v,x) plus the self term, then
//   d <- (A + cross * cw * d) / (sqrt(|u|^2 + cw*d) * sqrt(|v|^2 + cw*d)).
// The state is the graph's host values (edge_dress, vertex_dress, NW),
// uploaded before the queries, so results match what the reference computes
// on g.
#include "dress_capi_internal.h"
#include "dress_device.h"
#include "dress/hip/dress_hip.h"

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace dh {

namespace {

constexpr int TPB = 256;
constexpr int GQ = 16;                 // lanes per query

// first index in [lo, hi) whose value is >= x
__device__ __forceinline__ int lower_bound_q(const int *a, int lo, int hi, int x)
{
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Rows are the reference's adjacency (input labels, neighbours ascending),
// so the binary search (lane 0 of the query's group) probes the same entries
// as dress.c:656-667 and returns the same copy of a repeated edge.  For a
// virtual edge the group walks the shorter row and looks each neighbour up
// in the longer one; the i-th copy of a neighbour in one row pairs with its
// i-th copy in the other, as the reference's merge does (dress.c:679-693).
// Half-edge values are edge_weight * edge_dress of the edge the half-edge
// belongs to; the group's compensated partial sums merge exactly, so the
// sum is the serial merge's.
__global__ void __launch_bounds__(TPB) k_get(int n, const int *__restrict__ qu, const int *__restrict__ qv,
                                             const double *__restrict__ qw, int max_it, double eps,
                                             const int *__restrict__ row_off, const int *__restrict__ adj_x,
                                             const int *__restrict__ adj_e, const int *__restrict__ inv,
                                             const double *__restrict__ d, const double *__restrict__ ew,
                                             const double *__restrict__ vdn, const double *__restrict__ nw,
                                             const int *__restrict__ rank, int variant, double *out)
{
    // d in input edge order (inv null) or internal order; vdn / nw in input
    // labels (rank null) or internal labels
#define EDGE_D(e) d[inv ? inv[e] : (e)]
#define VTX(a, x) a[rank ? rank[x] : (x)]
    const int lane = threadIdx.x % GQ;
    const int64_t groups = (int64_t)gridDim.x * (TPB / GQ);
    for (int64_t q = (int64_t)blockIdx.x * (TPB / GQ) + threadIdx.x / GQ; q < n; q += groups) {
        const int u = qu[q], v = qv[q];
        // dress.c:656-667: existing edge
        double res = 0.0;
        int found = 0;
        if (lane == 0) {
            int lo = row_off[u], hi = row_off[u + 1] - 1;
            while (lo <= hi) {
                int mid = lo + (hi - lo) / 2;
                int nb = adj_x[mid];
                if (nb == v) {
                    res = EDGE_D(adj_e[mid]);
                    found = 1;
                    break;
                }
                if (nb < v) lo = mid + 1; else hi = mid - 1;
            }
        }
        found = __shfl(found, 0, GQ);
        if (!found) {
            // dress.c:671-737: virtual edge
            const int ub = row_off[u], ue = row_off[u + 1];
            const int vb = row_off[v], ve = row_off[v + 1];
            const bool a_is_u = (ue - ub) <= (ve - vb);
            const int ab = a_is_u ? ub : vb, ae = a_is_u ? ue : ve;
            const int bb = a_is_u ? vb : ub, be = a_is_u ? ve : ue;
            double s = 0.0, c = 0.0;
            for (int p = ab + lane; p < ae; p += GQ) {
                const int x = adj_x[p];
                const int first = (p > ab && adj_x[p - 1] == x) ? lower_bound_q(adj_x, ab, p, x) : p;
                const int r = p - first;
                const int lo = lower_bound_q(adj_x, bb, be, x);
                int have = 0;
                while (lo + have < be && adj_x[lo + have] == x && have <= r) have++;
                if (r < have) {
                    const int ea = adj_e[a_is_u ? p : lo + r], eb = adj_e[a_is_u ? lo + r : p];
                    two_sum_add(s, c, ew[ea] * EDGE_D(ea) + ew[eb] * EDGE_D(eb));
                }
            }
            group_reduce<GQ>(s, c);
            if (lane == 0) {
                const bool dir = (variant == DRESS_VARIANT_FORWARD || variant == DRESS_VARIANT_BACKWARD);
                two_sum_add(s, c, dir ? 4.0 * VTX(nw, u) : 4.0 * VTX(nw, u) + 4.0 * VTX(nw, v));
                const double A = s + c;
                const double cross = dir ? 1.0 : 2.0;
                const double cw = (variant == DRESS_VARIANT_UNDIRECTED) ? 2.0 * qw[q] : qw[q];
                const double du2 = VTX(vdn, u) * VTX(vdn, u);
                const double dv2 = VTX(vdn, v) * VTX(vdn, v);
                if (du2 <= 0.0 || dv2 <= 0.0) {
                    res = 0.0;
                } else {
                    double x = 1.0;
                    for (int it = 0; it < max_it; it++) {
                        double wd = cw * x;
                        double den = sqrt(du2 + wd) * sqrt(dv2 + wd);
                        if (den <= 0.0) { x = 0.0; break; }
                        double nx = (A + cross * wd) / den;
                        if (fabs(nx - x) <= eps) { x = nx; break; }
                        x = nx;
                    }
                    res = x;
                }
            }
        }
        if (lane == 0) out[q] = res;
    }
#undef EDGE_D
#undef VTX
}

// the entries of g's host state the queries read: edge values at edge ids,
// norms and vertex weights at vertices
__global__ void k_get_scatter(int ne, const int *__restrict__ eid, const double *__restrict__ ev, int nv,
                              const int *__restrict__ vid, const double *__restrict__ vv,
                              const double *__restrict__ vw, double *d, double *vd, double *nw)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ne + nv; i += gridDim.x * blockDim.x) {
        if (i < ne) {
            d[eid[i]] = ev[i];
        } else {
            const int j = i - ne;
            vd[vid[j]] = vv[j];
            nw[vid[j]] = vw[j];
        }
    }
}

}  // namespace

// Queries against g's host state (edge_dress, vertex_dress, NW, uploaded
// for the call into buffers of its own), or, with resident, against the
// state the last completed fit left on the device.
bool get_batch(p_dress_graph_t g, int n, const int *u, const int *v, const double *w, int max_it, double eps,
               double *out, bool resident)
{
    for (int i = 0; i < n; i++) {
        if (u[i] < 0 || u[i] >= g->N || v[i] < 0 || v[i] >= g->N) {
            set_error("dress_get: vertex out of range");
            return false;
        }
    }
    DevGraph *G = device_graph(g);
    if (!G) return false;
    if (resident && !G->fit_state_valid) {
        set_error("dress_hip_get_batch_resident: no fitted state on the device (run dress_fit on the graph "
                  "first; host values uploaded since, or other fits on its device copy, replace it)");
        return false;
    }
    if (!upload_reference_csr(G, g->adj_offset, g->adj_target, g->adj_edge_idx)) return false;
    hipStream_t s = G->stream;
    if (!resident) {
        // the graph's copy of the host state (input order and labels), kept
        // with the graph and rewritten per call where the queries read
        const bool have = (G->q_d || dalloc(&G->q_d, (size_t)G->E + 1, "query edge values")) &&
                          (G->q_vd || dalloc(&G->q_vd, (size_t)g->N + 1, "query norms")) &&
                          (G->q_nw || dalloc(&G->q_nw, (size_t)g->N + 1, "query vertex weights"));
        if (!have) return false;
    }
    // this call's buffers from the thread's block cache, kept between calls
    BatchScope scope;
    // host gathers of the state entries the queries read (alive until the
    // stream is synchronised below)
    std::vector<int> eid, vid;
    std::vector<double> ev, vv, vw;
    int *du = nullptr, *dv = nullptr, *deid = nullptr, *dvid = nullptr;
    double *dw = nullptr, *dout = nullptr, *dev = nullptr, *dvv = nullptr, *dvw = nullptr;
    bool ok = dalloc(&du, (size_t)n + 1, "queries") && dalloc(&dv, (size_t)n + 1, "queries") &&
              dalloc(&dw, (size_t)n + 1, "queries") && dalloc(&dout, (size_t)n + 1, "answers");
    if (ok && !resident) {
        // A query reads the values of the edges in its two rows and the norms
        // and weights of its two vertices.  When the queries' rows hold less
        // than a quarter of the half-edges, only those entries go up (a
        // single dress_get on a large graph moves a few KB); otherwise the
        // whole host state does.
        int64_t rows = 0;
        for (int i = 0; i < n && rows * 4 < (int64_t)G->S; i++)
            rows += (int64_t)(g->adj_offset[u[i] + 1] - g->adj_offset[u[i]]) +
                    (g->adj_offset[v[i] + 1] - g->adj_offset[v[i]]);
        if (rows * 4 < (int64_t)G->S) {
            eid.reserve((size_t)rows);
            ev.reserve((size_t)rows);
            for (int i = 0; i < n; i++) {
                for (int x : {u[i], v[i]}) {
                    for (int p = g->adj_offset[x]; p < g->adj_offset[x + 1]; p++) {
                        const int e = g->adj_edge_idx[p];
                        eid.push_back(e);
                        ev.push_back(g->edge_dress[e]);
                    }
                    vid.push_back(x);
                    vv.push_back(g->vertex_dress[x]);
                    vw.push_back(g->NW[x]);
                }
            }
            const int ne = (int)eid.size(), nv = (int)vid.size();
            ok = dalloc(&deid, (size_t)ne + 1, "query edge ids") && dalloc(&dev, (size_t)ne + 1, "query edge values") &&
                 dalloc(&dvid, (size_t)nv + 1, "query vertices") && dalloc(&dvv, (size_t)nv + 1, "query norms") &&
                 dalloc(&dvw, (size_t)nv + 1, "query vertex weights") &&
                 (ne == 0 || (hipMemcpyAsync(deid, eid.data(), (size_t)ne * sizeof(int), hipMemcpyHostToDevice, s) ==
                                  hipSuccess &&
                              hipMemcpyAsync(dev, ev.data(), (size_t)ne * sizeof(double), hipMemcpyHostToDevice,
                                             s) == hipSuccess)) &&
                 (nv == 0 || (hipMemcpyAsync(dvid, vid.data(), (size_t)nv * sizeof(int), hipMemcpyHostToDevice, s) ==
                                  hipSuccess &&
                              hipMemcpyAsync(dvv, vv.data(), (size_t)nv * sizeof(double), hipMemcpyHostToDevice,
                                             s) == hipSuccess &&
                              hipMemcpyAsync(dvw, vw.data(), (size_t)nv * sizeof(double), hipMemcpyHostToDevice,
                                             s) == hipSuccess));
            if (ok && ne + nv > 0)
                hipLaunchKernelGGL(k_get_scatter, dim3((int)std::min<int64_t>(4096, ((int64_t)ne + nv + TPB - 1) / TPB)),
                                   dim3(TPB), 0, s, ne, deid, dev, nv, dvid, dvv, dvw, G->q_d, G->q_vd, G->q_nw);
        } else {
            // the caller's arrays may still be read by earlier work on s
            ok = hipStreamSynchronize(s) == hipSuccess &&
                 h2d_parallel(G->q_d, g->edge_dress, (size_t)G->E * sizeof(double)) &&
                 (g->N == 0 || (h2d_parallel(G->q_vd, g->vertex_dress, (size_t)g->N * sizeof(double)) &&
                                h2d_parallel(G->q_nw, g->NW, (size_t)g->N * sizeof(double))));
        }
    }
    if (ok) {
        ok = hipMemcpyAsync(du, u, (size_t)n * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(dv, v, (size_t)n * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(dw, w, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess;
        if (ok && n > 0) {
            const int grid = (int)std::min<int64_t>(8192, ((int64_t)n * GQ + TPB - 1) / TPB);
            if (resident)
                hipLaunchKernelGGL(k_get, dim3(grid), dim3(TPB), 0, s, n, du, dv, dw, max_it, eps, G->ref_row_off,
                                   G->ref_adj_x, G->ref_adj_e, (const int *)G->inv, (const double *)G->d[G->cur],
                                   (const double *)G->ew_in, (const double *)G->vd, (const double *)G->nw,
                                   (const int *)G->rank, G->variant, dout);
            else
                hipLaunchKernelGGL(k_get, dim3(grid), dim3(TPB), 0, s, n, du, dv, dw, max_it, eps, G->ref_row_off,
                                   G->ref_adj_x, G->ref_adj_e, (const int *)nullptr, (const double *)G->q_d,
                                   (const double *)G->ew_in, (const double *)G->q_vd, (const double *)G->q_nw,
                                   (const int *)nullptr, G->variant, dout);
        }
        ok = ok && hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(out, dout, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
        if (!ok && !has_error()) set_error("dress_get kernel failed");
    }
    (void)hipStreamSynchronize(s);          // the cached blocks are idle before they go back
    dfree(du); dfree(dv); dfree(dw); dfree(dout); dfree(deid); dfree(dev); dfree(dvid); dfree(dvv); dfree(dvw);
    return ok;
}

// ---------------------------------------------------------------------------
// dress_get batches without allocation (dress_hip_get_batch_resident_noalloc):
// the queries and answers are staged in this file-scope device array, whose
// device address is looked up once per device (hipGetSymbolAddress); the
// array's name is a host-side stand-in and is never passed to the device.
// Every call on a device shares that device's copy, so calls that overlap on
// one device race.

constexpr int NOALLOC_MAX_QUERIES = 65536;
// per query: u, v (int32), edge weight and answer (double)
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)NOALLOC_MAX_QUERIES * (2 * sizeof(int) + 2 * sizeof(double));
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {nullptr};

// scratch's device address on the current device (nullptr with the error
// set when it cannot be had)
unsigned char *scratch_on_device()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        (void)hipGetLastError();
        set_error("dress_hip_get_batch_resident_noalloc: no current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            set_error(std::string("dress_hip_get_batch_resident_noalloc: hipGetSymbolAddress: ") +
                      hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

}  // namespace dh

extern "C" {

int dress_hip_get_batch_resident_noalloc(p_dress_graph_t g, int n, const int *u, const int *v,
                                         const double *edge_weight, int max_iterations, double epsilon, double *out)
{
    using namespace dh;
    clear_error();
    if (!g || n < 0 || (n > 0 && (!u || !v || !edge_weight || !out))) {
        set_error("dress_hip_get_batch_resident_noalloc: bad arguments");
        return -1;
    }
    if (n > NOALLOC_MAX_QUERIES) {
        set_error("dress_hip_get_batch_resident_noalloc: " + std::to_string(n) + " queries exceed the maximum of " +
                  std::to_string(NOALLOC_MAX_QUERIES) + " per call");
        return -1;
    }
    for (int i = 0; i < n; i++) {
        if (u[i] < 0 || u[i] >= g->N || v[i] < 0 || v[i] >= g->N) {
            set_error("dress_get: vertex out of range");
            return -1;
        }
    }
    DevGraph *G = lookup(g);
    if (!G || !G->fit_state_valid) {
        set_error("dress_hip_get_batch_resident_noalloc: no fitted state on the device (run dress_fit on the graph "
                  "first)");
        return -1;
    }
    if (!G->ref_row_off || !G->ref_adj_x || !G->ref_adj_e) {
        set_error("dress_hip_get_batch_resident_noalloc: the graph's query rows are not on the device (one call "
                  "of dress_hip_get_batch_resident places them)");
        return -1;
    }
    if (n == 0) return 0;
    unsigned char *base = scratch_on_device();
    if (!base) return -1;
    // scratch layout: answers and weights (8-byte aligned), then u and v
    double *d_out = (double *)base;
    double *d_w = d_out + NOALLOC_MAX_QUERIES;
    int *d_u = (int *)(d_w + NOALLOC_MAX_QUERIES);
    int *d_v = d_u + NOALLOC_MAX_QUERIES;
    hipStream_t s = G->stream;
    bool ok = hipMemcpyAsync(d_u, u, (size_t)n * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(d_v, v, (size_t)n * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(d_w, edge_weight, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok) {
        const int grid = (int)std::min<int64_t>(8192, ((int64_t)n * GQ + TPB - 1) / TPB);
        hipLaunchKernelGGL(k_get, dim3(grid), dim3(TPB), 0, s, n, (const int *)d_u, (const int *)d_v,
                           (const double *)d_w, max_iterations, epsilon, G->ref_row_off, G->ref_adj_x, G->ref_adj_e,
                           (const int *)G->inv, (const double *)G->d[G->cur], (const double *)G->ew_in,
                           (const double *)G->vd, (const double *)G->nw, (const int *)G->rank, G->variant, d_out);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(out, d_out, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    }
    if (!ok) {
        (void)hipStreamSynchronize(s);
        set_error("dress_hip_get_batch_resident_noalloc: dress_get kernel failed");
        return -1;
    }
    return 0;
}

// dress.h:105-107.  Returns NaN on error (dress_hip_last_error()).
double dress_get(const p_dress_graph_t g, int u, int v, int max_iterations, double epsilon, double edge_weight)
{
    dh::clear_error();
    double r = NAN;
    if (!g) {
        dh::set_error("dress_get: NULL graph");
        return r;
    }
    if (!dh::get_batch(g, 1, &u, &v, &edge_weight, max_iterations, epsilon, &r, false)) return NAN;
    return r;
}

int dress_hip_get_batch(p_dress_graph_t g, int n, const int *u, const int *v, const double *edge_weight,
                        int max_iterations, double epsilon, double *out)
{
    dh::clear_error();
    if (!g || n < 0 || (n > 0 && (!u || !v || !edge_weight || !out))) {
        dh::set_error("dress_hip_get_batch: bad arguments");
        return -1;
    }
    return dh::get_batch(g, n, u, v, edge_weight, max_iterations, epsilon, out, false) ? 0 : -1;
}

int dress_hip_get_batch_resident(p_dress_graph_t g, int n, const int *u, const int *v,
                                 const double *edge_weight, int max_iterations, double epsilon, double *out)
{
    dh::clear_error();
    if (!g || n < 0 || (n > 0 && (!u || !v || !edge_weight || !out))) {
        dh::set_error("dress_hip_get_batch_resident: bad arguments");
        return -1;
    }
    return dh::get_batch(g, n, u, v, edge_weight, max_iterations, epsilon, out, true) ? 0 : -1;
}

}  // extern "C"
