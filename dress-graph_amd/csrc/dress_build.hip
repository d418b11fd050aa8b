// dress_build.hip -- device-side graph construction for libdress_hip.so.
//
// Follows the construction of libdress v0.8.3 (libdress/src/dress.c:103-296,
// 307-363) but runs it on the GPU:
//   1. half-edges of the variant adjacency are generated per edge record
//      (UNDIRECTED both directions with w-bar = 2w, DIRECTED out-edges plus
//      in-edges without a reciprocal, reciprocal weights summed; FORWARD
//      out-edges; BACKWARD in-edges), keyed (row, neighbour) and radix sorted.
//      The sort is stable and half-edges are generated in edge order, so
//      ties on the neighbour end up ordered by edge index -- the order the
//      reference's stable qsort leaves them in.
//   2. the intersection index: for every edge, the multiset intersection of
//      its two sorted rows.  A group of lanes walks the shorter row and
//      binary-searches each element in the longer one; the i-th copy of a
//      repeated neighbour pairs with the i-th copy on the other side, which
//      is what the reference's merge walk produces.  Matches are emitted in
//      ascending neighbour order with a ballot + prefix popcount, so the
//      index equals the reference's intercept arrays entry for entry.
#include "dress_hip_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace dh {

namespace {

constexpr int TPB = 256;

inline int grid_for(int64_t threads, int cap_blocks = 1 << 20)
{
    int64_t b = (threads + TPB - 1) / TPB;
    if (b < 1) b = 1;
    if (b > cap_blocks) b = cap_blocks;
    return (int)b;
}

__global__ void k_fill_i32(int *p, int64_t n, int v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_f64(double *p, int64_t n, double v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// raw out-adjacency keys for DIRECTED (reciprocal lookup only)
__global__ void k_raw_keys(const int *U, const int *V, int E, uint64_t *key, int *val, int *deg)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        key[e] = ((uint64_t)(uint32_t)U[e] << 32) | (uint32_t)V[e];
        val[e] = e;
        atomicAdd(&deg[U[e]], 1);
    }
}

// Binary search for neighbour v in raw row u with the probe sequence of
// dress.c:70-87, so repeated edges resolve to the same entry.
__device__ int raw_find(const int64_t *roff, const uint64_t *rkey, int u, int v)
{
    int64_t lo = roff[u], hi = roff[u + 1] - 1;
    while (lo <= hi) {
        int64_t mid = lo + (hi - lo) / 2;
        int x = (int)(uint32_t)(rkey[mid] & 0xffffffffu);
        if (x == v) return (int)mid;
        if (x < v) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

// variant half-edges: slot 2e is the out/primary entry, slot 2e+1 the
// in/secondary one; unused slots get the all-ones key and sort last.
__global__ void k_half_keys(const int *U, const int *V, const double *W, int E, int variant,
                            const int64_t *roff, const uint64_t *rkey, const int *rval,
                            uint64_t *key, int *val, double *ew, int *deg)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        int u = U[e], v = V[e];
        double w = W ? W[e] : 1.0;
        uint64_t k0 = ~0ull, k1 = ~0ull;
        double wb = w;
        if (variant == DRESS_VARIANT_UNDIRECTED) {
            k0 = ((uint64_t)(uint32_t)u << 32) | (uint32_t)v;
            k1 = ((uint64_t)(uint32_t)v << 32) | (uint32_t)u;
            wb = 2.0 * w;
        } else if (variant == DRESS_VARIANT_DIRECTED) {
            k0 = ((uint64_t)(uint32_t)u << 32) | (uint32_t)v;
            int r = raw_find(roff, rkey, v, u);
            if (r >= 0) wb = w + (W ? W[rval[r]] : 1.0);
            else k1 = ((uint64_t)(uint32_t)v << 32) | (uint32_t)u;
        } else if (variant == DRESS_VARIANT_FORWARD) {
            k0 = ((uint64_t)(uint32_t)u << 32) | (uint32_t)v;
        } else {
            k0 = ((uint64_t)(uint32_t)v << 32) | (uint32_t)u;
        }
        key[2 * (int64_t)e] = k0;
        key[2 * (int64_t)e + 1] = k1;
        val[2 * (int64_t)e] = 2 * e;
        val[2 * (int64_t)e + 1] = 2 * e + 1;
        ew[e] = wb;
        if (k0 != ~0ull) atomicAdd(&deg[(int)(k0 >> 32)], 1);
        if (k1 != ~0ull) atomicAdd(&deg[(int)(k1 >> 32)], 1);
    }
}

__global__ void k_unpack_half(const uint64_t *key, const int *val, int64_t S, int *adj_x,
                              int *adj_e)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < S;
         i += (int64_t)gridDim.x * blockDim.x) {
        adj_x[i] = (int)(uint32_t)(key[i] & 0xffffffffu);
        adj_e[i] = val[i] >> 1;
    }
}

__global__ void k_i64_to_i32(const int64_t *a, int *b, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (int)a[i];
}

__global__ void k_edge_consts(const int *U, const int *V, const double *nw, int E, int variant,
                              double *c4)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        double a = 4.0 * nw[U[e]];
        c4[e] = (variant == DRESS_VARIANT_FORWARD || variant == DRESS_VARIANT_BACKWARD)
                    ? a : a + 4.0 * nw[V[e]];
    }
}

__global__ void k_max_deg(const int *row_off, int N, int *out)
{
    int m = 0;
    for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < N; u += gridDim.x * blockDim.x)
        m = max(m, row_off[u + 1] - row_off[u]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// first index in [lo, hi) whose value is >= x
__device__ __forceinline__ int lower_bound_i(const int *a, int lo, int hi, int x)
{
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Multiset intersection of rows U[e], V[e].  GS lanes per edge walk the
// shorter row; group g takes edge ord[g] (g without ord).  FILL=false: count
// into cnt[e].  FILL=true: for internal
// edge i (input edge ord[i]) write the half-edge pair (a in row U, b in row V)
// of each common neighbour at ioff[i], ascending in the neighbour, packed as
// row-relative 16-bit positions (ipk) or absolute (ipair).
// Hub rows as bitmaps: hub h (slot[v] = h, or -1) has bits[h * W + i], bit
// j set when label 64 i + j is in v's row, and pre[h * W + i] = the set bits
// of words 0 .. i-1, so the position of a member x in the row is
// pre[x / 64] + popcount(word below x).  Only rows without repeated
// neighbours get a slot.
struct HubBits {
    const int *slot = nullptr;
    const unsigned long long *bits = nullptr;
    const unsigned *pre = nullptr;
    int64_t W = 0;
};

template <int GS, bool FILL>
__global__ void k_isect(const int *U, const int *V, int E, const int *row_off, const int *adj_x,
                        const int *ord, int64_t *cnt, const int64_t *ioff, uint32_t *ipk, int2 *ipair,
                        HubBits hub)
{
    const int lane = threadIdx.x % GS;
    const int wlane = threadIdx.x & 63;
    const int gbase = wlane - lane;              // first lane of my group in the wave
    const int64_t groups = (int64_t)gridDim.x * blockDim.x / GS;
    for (int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / GS; g < E; g += groups) {
        int e = ord ? ord[g] : (int)g;
        int u = U[e], v = V[e];
        int ub = row_off[u], ue = row_off[u + 1];
        int vb = row_off[v], ve = row_off[v + 1];
        bool a_is_u = (ue - ub) <= (ve - vb);
        int ab = a_is_u ? ub : vb, ae = a_is_u ? ue : ve;
        int bb = a_is_u ? vb : ub, be = a_is_u ? ve : ue;
        // the longer row as a hub bitmap, when it has one
        const int hs = hub.slot ? hub.slot[a_is_u ? v : u] : -1;
        const unsigned long long *hbits = hs >= 0 ? hub.bits + (int64_t)hs * hub.W : nullptr;
        const unsigned *hpre = hs >= 0 ? hub.pre + (int64_t)hs * hub.W : nullptr;
        int64_t out = FILL ? ioff[g] : 0;
        int64_t found = 0;
        for (int base = ab; base < ae; base += GS) {
            int p = base + lane;
            bool hit = false;
            int q = 0;
            if (p < ae) {
                int x = adj_x[p];
                const bool again = p > ab && adj_x[p - 1] == x;     // not x's first copy in row a
                if (hbits) {
                    // row b holds x at most once: only x's first copy pairs
                    const unsigned long long w = hbits[x >> 6], m = 1ull << (x & 63);
                    if (!again && (w & m)) {
                        hit = true;
                        q = bb + (int)hpre[x >> 6] + __popcll(w & (m - 1));
                    }
                } else {
                    int first = again ? lower_bound_i(adj_x, ab, p, x) : p;
                    int r = p - first;
                    int lo = lower_bound_i(adj_x, bb, be, x);
                    int have = 0;
                    while (lo + have < be && adj_x[lo + have] == x && have <= r) have++;
                    if (r < have) { hit = true; q = lo + r; }
                }
            }
            if (FILL) {
                unsigned long long m = __ballot(hit);
                unsigned long long gm = (GS == 64) ? m : ((m >> gbase) & ((1ull << GS) - 1));
                if (hit) {
                    int before = __popcll(gm & ((1ull << lane) - 1));
                    int pu = a_is_u ? p : q, pv = a_is_u ? q : p;
                    if (ipk)
                        ipk[out + before] = (uint32_t)(pu - ub) | ((uint32_t)(pv - vb) << 16);
                    else {
                        int2 pr;
                        pr.x = pu;
                        pr.y = pv;
                        ipair[out + before] = pr;
                    }
                }
                out += __popcll(gm);
            } else {
                found += hit ? 1 : 0;
            }
        }
        if (!FILL) {
            for (int o = GS / 2; o > 0; o >>= 1) found += __shfl_xor(found, o, GS);
            if (lane == 0) cnt[e] = found;
        }
    }
}

// vertices with at least `deg_min` neighbours: a slot each (in no order)
__global__ void k_hub_assign(const int *row_off, int N, int deg_min, int *slot, int *hub_vertex, int *count)
{
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        int h = -1;
        if (row_off[v + 1] - row_off[v] >= deg_min) {
            h = atomicAdd(count, 1);
            hub_vertex[h] = v;
        }
        slot[v] = h;
    }
}

// bits of hub rows (one workgroup per hub; bits start at 0); a row with a
// repeated neighbour is marked in bad[]
__global__ void k_hub_bits(const int *row_off, const int *adj_x, const int *hub_vertex, int nh, int64_t W,
                           unsigned long long *bits, int *bad)
{
    for (int h = blockIdx.x; h < nh; h += gridDim.x) {
        const int v = hub_vertex[h];
        const int b = row_off[v], e = row_off[v + 1];
        for (int p = b + threadIdx.x; p < e; p += blockDim.x) {
            const int x = adj_x[p];
            atomicOr(&bits[(int64_t)h * W + (x >> 6)], 1ull << (x & 63));
            if (p > b && adj_x[p - 1] == x) bad[h] = 1;
        }
    }
}

// pre[h * W + i] = set bits of words 0 .. i-1 of hub h (one workgroup per
// hub: each thread sums a contiguous run of words, then a scan over the
// threads); hubs marked bad lose their slot
__global__ void k_hub_prefix(const unsigned long long *bits, int nh, int64_t W, unsigned *pre, const int *bad,
                             const int *hub_vertex, int *slot)
{
    __shared__ unsigned part[256];
    const int T = blockDim.x;                      // 256
    for (int h = blockIdx.x; h < nh; h += gridDim.x) {
        const unsigned long long *wb = bits + (int64_t)h * W;
        unsigned *pb = pre + (int64_t)h * W;
        const int64_t per = (W + T - 1) / T;
        const int64_t i0 = threadIdx.x * per, i1 = i0 + per < W ? i0 + per : W;
        unsigned sum = 0;
        for (int64_t i = i0; i < i1; i++) sum += (unsigned)__popcll(wb[i]);
        part[threadIdx.x] = sum;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned run = 0;
            for (int t = 0; t < T; t++) {
                const unsigned c = part[t];
                part[t] = run;
                run += c;
            }
        }
        __syncthreads();
        unsigned run = part[threadIdx.x];
        for (int64_t i = i0; i < i1; i++) {
            pb[i] = run;
            run += (unsigned)__popcll(wb[i]);
        }
        if (threadIdx.x == 0 && bad[h]) slot[hub_vertex[h]] = -1;
        __syncthreads();
    }
}

// sort keys of the edges by their first half-edge (CSR order of the rows)
__global__ void k_first_half_keys(const int2 *pos_in, int E, uint64_t *key, int *val)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        key[e] = (uint64_t)(uint32_t)pos_in[e].x;
        val[e] = e;
    }
}

// half-edges of each input edge (at most two), ordered x < y
__global__ void k_edge_positions(const int *adj_e, int64_t S, int2 *pos_in)
{
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < S;
         s += (int64_t)gridDim.x * blockDim.x) {
        int e = adj_e[s];
        if (atomicCAS(&pos_in[e].x, -1, (int)s) != -1) pos_in[e].y = (int)s;
    }
}

__global__ void k_order_positions(int2 *pos_in, int E)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        int2 p = pos_in[e];
        if (p.y >= 0 && p.y < p.x) {
            int t = p.x;
            p.x = p.y;
            p.y = t;
            pos_in[e] = p;
        }
    }
}

// internal order key: (bin, half-edge of the edge in its owner's row).  In
// the bins set in low_bins the owner is the endpoint of smaller degree: the
// edges of one low-degree vertex are adjacent, its norm and row are read and
// written in sequence, and the other endpoints' norms come from the short
// prefix of high-degree labels that stays in L2.  Elsewhere the owner is the
// endpoint of larger degree, so the edges of one hub share its row of values.
// Ties go to the lower id.
__global__ void k_order_keys(const int64_t *cnt_in, const int2 *pos_in, const int *U, const int *V,
                             const int *row_off, int E, int low_bins, int tile_bins, int tile_logs, uint64_t *key,
                             int *val)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        int2 p = pos_in[e];
        int own = p.x;
        const int bin = term_bin(cnt_in[e] + 1);
        // (bin, position) with the bin above bit 56
        uint64_t k = ((uint64_t)bin << 56) | (uint32_t)own;
        if (p.y >= 0) {
            int u = U[e], v = V[e];
            int du = row_off[u + 1] - row_off[u], dv = row_off[v + 1] - row_off[v];
            bool u_high = du > dv || (du == dv && u < v);
            int o = (((low_bins >> bin) & 1) ? !u_high : u_high) ? u : v;
            own = (p.x >= row_off[o] && p.x < row_off[o + 1]) ? p.x : p.y;
            k = ((uint64_t)bin << 56) | (uint32_t)own;
            if ((tile_bins >> bin) & 1) {
                // blocks of 2^tile_log owners, within a block by the other
                // end (labels below 2^28: the host checks)
                const int other = o == u ? v : u;
                const int tile_log = (tile_logs >> (5 * bin)) & 31;
                k = ((uint64_t)bin << 56) | ((uint64_t)((unsigned)o >> tile_log) << 28) | (uint64_t)(unsigned)other;
            }
        }
        key[e] = k;
        val[e] = e;
    }
}

__global__ void k_count_bins(const int *ord, const int64_t *cnt_in, int E, int *counts)
{
    __shared__ int c[Bins::NB];
    if (threadIdx.x < Bins::NB) c[threadIdx.x] = 0;
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x)
        atomicAdd(&c[term_bin(cnt_in[ord[i]] + 1)], 1);
    __syncthreads();
    if (threadIdx.x < Bins::NB) atomicAdd(&counts[threadIdx.x], c[threadIdx.x]);
}

__global__ void k_internal_edges(const int *ord, int E, const int *U, const int *V, const int *row_off,
                                 const double *ew_in, const double *c4_in, const int64_t *cnt_in,
                                 const int2 *pos_in, int *inv, int2 *uv, int2 *rs, int2 *pos,
                                 double *ew, double *c4, int64_t *cnt_i)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        int e = ord[i];
        inv[e] = i;
        int2 p;
        p.x = U[e];
        p.y = V[e];
        uv[i] = p;
        int2 r;
        r.x = row_off[p.x];
        r.y = row_off[p.y];
        rs[i] = r;
        pos[i] = pos_in[e];
        ew[i] = ew_in[e];
        c4[i] = c4_in[e];
        if (cnt_i) cnt_i[i] = cnt_in[e];
    }
}

__global__ void k_self_terms(const int *ord, int E, const double *c4_in, double *c4)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x)
        c4[i] = c4_in[ord[i]];
}

// bin key of vertex rows: (bin << 32) | u
__global__ void k_bin_keys_rows(const int *row_off, int N, uint64_t *key)
{
    for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < N; u += gridDim.x * blockDim.x) {
        int b = term_bin(row_off[u + 1] - row_off[u] + 1);
        key[u] = ((uint64_t)b << 32) | (uint32_t)u;
    }
}

__global__ void k_bin_unpack(const uint64_t *key, int n, int *list, int *counts)
{
    __shared__ int c[Bins::NB];
    if (threadIdx.x < Bins::NB) c[threadIdx.x] = 0;
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        list[i] = (int)(uint32_t)(key[i] & 0xffffffffu);
        atomicAdd(&c[(int)(key[i] >> 32)], 1);
    }
    __syncthreads();
    if (threadIdx.x < Bins::NB) atomicAdd(&counts[threadIdx.x], c[threadIdx.x]);
}

// vertex rows sorted by (bin, id)
bool vertex_bins(hipStream_t s, int n, const int *row_off, int *list, int counts_out[Bins::NB])
{
    uint64_t *k0 = nullptr, *k1 = nullptr;
    int *cnt_d = nullptr;
    if (!dalloc(&k0, (size_t)n + 1, "bin keys") || !dalloc(&k1, (size_t)n + 1, "bin keys") ||
        !dalloc(&cnt_d, Bins::NB, "bin counts")) {
        dfree(k0); dfree(k1); dfree(cnt_d);
        return false;
    }
    bool ok = true;
    if (n > 0) {
        hipLaunchKernelGGL(k_bin_keys_rows, dim3(grid_for(n, 4096)), dim3(TPB), 0, s, row_off, n, k0);
        ok = prim_sort_u64(k0, k1, (size_t)n, s);
    }
    if (ok && hipMemsetAsync(cnt_d, 0, Bins::NB * sizeof(int), s) != hipSuccess) {
        set_error("bin count reset failed");
        ok = false;
    }
    if (ok && n > 0)
        hipLaunchKernelGGL(k_bin_unpack, dim3(grid_for(n, 2048)), dim3(TPB), 0, s, k1, n, list, cnt_d);
    if (ok) {
        hipError_t e2 = hipMemcpyAsync(counts_out, cnt_d, Bins::NB * sizeof(int), hipMemcpyDeviceToHost, s);
        hipError_t e3 = hipStreamSynchronize(s);
        if (e2 != hipSuccess || e3 != hipSuccess) { set_error("bin download failed"); ok = false; }
    }
    dfree(k0); dfree(k1); dfree(cnt_d);
    return ok;
}

// ------------------------------------------------------ internal labels

// rank key of vertex u: decreasing degree, then increasing id
__global__ void k_rank_keys(const int *row_off, int N, uint64_t *key)
{
    for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < N; u += gridDim.x * blockDim.x) {
        uint32_t deg = (uint32_t)(row_off[u + 1] - row_off[u]);
        key[u] = ((uint64_t)(0xffffffffu - deg) << 32) | (uint32_t)u;
    }
}

__global__ void k_rank_scatter(const uint64_t *sorted, int N, int *rank)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        rank[(int)(uint32_t)(sorted[i] & 0xffffffffu)] = i;
}

__global__ void k_relabel_edges(const int *U, const int *V, const int *rank, int E, int *U2, int *V2)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        U2[e] = rank[U[e]];
        V2[e] = rank[V[e]];
    }
}

// to[rank[v]] = from[v]
__global__ void k_to_internal(const double *from, const int *rank, int N, double *to)
{
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        to[rank[v]] = from[v];
}

// to[v] = from[rank[v]]
__global__ void k_to_input(const double *from, const int *rank, int N, double *to)
{
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        to[v] = from[rank[v]];
}

// internal degree of every internal row
__global__ void k_internal_degrees(const int *row_off, const int *rank, int N, int *deg)
{
    for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < N; u += gridDim.x * blockDim.x)
        deg[rank[u]] = row_off[u + 1] - row_off[u];
}

// half-edge p of reference row u -> key (rank[u], rank[x]); the stable sort
// keeps equal neighbours in edge order, as in the reference row
__global__ void k_internal_half_keys(const int *row_off, const int *adj_x, const int *adj_e, const int *rank,
                                     int N, uint64_t *key, int *val)
{
    for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < N; u += gridDim.x * blockDim.x) {
        const uint64_t hi = (uint64_t)(uint32_t)rank[u] << 32;
        for (int p = row_off[u]; p < row_off[u + 1]; p++) {
            key[p] = hi | (uint32_t)rank[adj_x[p]];
            val[p] = adj_e[p];
        }
    }
}

__global__ void k_unpack_internal(const uint64_t *key, const int *val, int64_t S, int *adj_x, int *adj_e)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < S;
         i += (int64_t)gridDim.x * blockDim.x) {
        adj_x[i] = (int)(uint32_t)(key[i] & 0xffffffffu);
        adj_e[i] = val[i];
    }
}

__global__ void k_iota(int *p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

// absolute half-edge pairs -> edge ids through the reference adjacency
__global__ void k_pairs_to_edges(const int2 *pairs, int64_t T, const int *adj_e, int *ux, int *vx)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < T;
         i += (int64_t)gridDim.x * blockDim.x) {
        int2 p = pairs[i];
        ux[i] = adj_e[p.x];
        vx[i] = adj_e[p.y];
    }
}

// Internal labels: moves the reference-layout CSR just built (input
// labels) to G->ref_* and builds the CSR the fit runs on, over vertices
// relabelled by decreasing degree with rows ordered by neighbour rank.
// Every sum of the fit is exactly rounded, so results do not depend on
// this order; it only decides which values share cache lines.
// Graphs below this many half-edges keep their input labels: their values
// fit in one XCD's L2, and the extra sort would only add build latency.
// DRESS_HIP_RELABEL=always|never overrides.
constexpr int64_t RELABEL_MIN_HALF_EDGES = 1 << 18;

// bins whose edges are ordered by their low-degree endpoint (bit b = bin b):
// bin 0 (edges of <= 3 common neighbours, most of them between a
// low-degree vertex and a hub) by default, measured 12 % faster there on a
// 4M-vertex power-law graph and slower in bin 1; DRESS_HIP_LOW_ORDER_BINS
// overrides the mask
// bins whose edges go in blocks of 2^tile_log consecutive owners (by
// label), ordered by their other endpoint within a block, so that the
// other endpoint's row serves the block's owners from L2: bin 3 (edges of
// 513..8192 terms, between two hubs) in blocks of 32 owners by default.  On
// the 4M-vertex power-law graph bin 3 took 1.115 -> 0.998 ms per iteration
// (blocks of 2..64 owners: 1.085, 1.061, 1.023, 1.011, 0.998, 1.057 ms); the
// same order slowed bins 0, 1 and 2 at every block size tried (bin 0 with
// 4..65536 owners: 1.07..1.36 ms against 1.056).
// DRESS_HIP_TILE_BINS=<mask> and DRESS_HIP_TILE_LOG=<log2 owners per block:
// one value for every bin, or "l0,l1,l2,l3,l4"> override.
int tile_bins()
{
    const char *e = getenv("DRESS_HIP_TILE_BINS");
    return (e && *e) ? (int)strtol(e, nullptr, 0) & 0x1f : 0x8;
}
// log2 owners per block of each bin, 5 bits a bin (bin b at bits 5b..5b+4)
int tile_logs()
{
    int l[5] = {0, 0, 0, 5, 0};
    if (const char *e = getenv("DRESS_HIP_TILE_LOG")) {
        int v[5], n = sscanf(e, "%d,%d,%d,%d,%d", &v[0], &v[1], &v[2], &v[3], &v[4]);
        if (n == 1)
            for (int b = 0; b < 5; b++) l[b] = v[0];
        else
            for (int b = 0; b < n; b++) l[b] = v[b];
    }
    int packed = 0;
    for (int b = 0; b < 5; b++) packed |= (l[b] < 0 ? 0 : (l[b] > 22 ? 22 : l[b])) << (5 * b);
    return packed;
}

int low_order_bins()
{
    const char *e = getenv("DRESS_HIP_LOW_ORDER_BINS");
    if (e && *e) return (int)strtol(e, nullptr, 0) & 0x1f;
    return 1;
}

bool want_relabel(int64_t S)
{
    const char *e = getenv("DRESS_HIP_RELABEL");
    if (e && (!strcmp(e, "always") || !strcmp(e, "1"))) return true;
    if (e && (!strcmp(e, "never") || !strcmp(e, "0"))) return false;
    return S >= RELABEL_MIN_HALF_EDGES;
}

bool relabel(DevGraph *G)
{
    hipStream_t s = G->stream;
    const int N = G->N, E = G->E;
    const int64_t S = G->S;
    G->ref_U = G->U; G->ref_V = G->V;
    G->ref_row_off = G->row_off; G->ref_adj_x = G->adj_x; G->ref_adj_e = G->adj_e;
    if (!want_relabel(S)) {
        // identity labels: the reference arrays are the fit's arrays
        if (!dalloc(&G->rank, (size_t)N + 1, "vertex ranks")) return false;
        if (N > 0) hipLaunchKernelGGL(k_iota, dim3(grid_for(N, 4096)), dim3(TPB), 0, s, G->rank, N);
        DH_CHECK(hipGetLastError());
        return true;
    }
    G->U = G->V = G->row_off = G->adj_x = G->adj_e = nullptr;
    double *nw_in = G->nw;
    G->nw = nullptr;
    const size_t nk = (size_t)std::max<int64_t>(N, S) + 1;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    int *v0 = nullptr, *v1 = nullptr, *deg = nullptr;
    int64_t *deg64 = nullptr;
    bool ok = dalloc(&G->rank, (size_t)N + 1, "vertex ranks") && dalloc(&G->U, (size_t)E + 1, "U") &&
              dalloc(&G->V, (size_t)E + 1, "V") && dalloc(&G->row_off, (size_t)N + 1, "row offsets") &&
              dalloc(&G->adj_x, (size_t)S + 1, "adj_target") && dalloc(&G->adj_e, (size_t)S + 1, "adj_edge_idx") &&
              dalloc(&G->nw, (size_t)N + 1, "NW") && dalloc(&k0, nk, "relabel keys") &&
              dalloc(&k1, nk, "relabel keys") && dalloc(&v0, (size_t)S + 1, "relabel vals") &&
              dalloc(&v1, (size_t)S + 1, "relabel vals") && dalloc(&deg, (size_t)N + 1, "degrees") &&
              dalloc(&deg64, (size_t)N + 1, "degrees");
    int bits = 1;
    while (bits < 31 && (1 << bits) < N) bits++;
    if (ok && N > 0) {
        hipLaunchKernelGGL(k_rank_keys, dim3(grid_for(N, 4096)), dim3(TPB), 0, s, G->ref_row_off, N, k0);
        ok = prim_sort_u64(k0, k1, (size_t)N, s);
        if (ok) hipLaunchKernelGGL(k_rank_scatter, dim3(grid_for(N, 4096)), dim3(TPB), 0, s, k1, N, G->rank);
    }
    if (ok) {
        if (E > 0)
            hipLaunchKernelGGL(k_relabel_edges, dim3(grid_for(E, 8192)), dim3(TPB), 0, s, G->ref_U, G->ref_V,
                               G->rank, E, G->U, G->V);
        if (N > 0)
            hipLaunchKernelGGL(k_to_internal, dim3(grid_for(N, 4096)), dim3(TPB), 0, s, nw_in, G->rank, N, G->nw);
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(N + 1, 4096)), dim3(TPB), 0, s, deg, (int64_t)N + 1, 0);
        if (N > 0)
            hipLaunchKernelGGL(k_internal_degrees, dim3(grid_for(N, 4096)), dim3(TPB), 0, s, G->ref_row_off,
                               G->rank, N, deg);
        ok = prim_exclusive_scan_i32_to_i64(deg, deg64, (size_t)N + 1, s);
    }
    if (ok) {
        hipLaunchKernelGGL(k_i64_to_i32, dim3(grid_for(N + 1, 4096)), dim3(TPB), 0, s, deg64, G->row_off,
                           (int64_t)N + 1);
        if (S > 0) {
            hipLaunchKernelGGL(k_internal_half_keys, dim3(grid_for(N, 4096)), dim3(TPB), 0, s, G->ref_row_off,
                               G->ref_adj_x, G->ref_adj_e, G->rank, N, k0, v0);
            ok = prim_sort_pairs_u64_i32(k0, k1, v0, v1, (size_t)S, 32 + bits, s);
            if (ok)
                hipLaunchKernelGGL(k_unpack_internal, dim3(grid_for(S, 8192)), dim3(TPB), 0, s, k1, v1, S,
                                   G->adj_x, G->adj_e);
        }
    }
    if (ok) {
        hipError_t e1 = hipGetLastError();
        hipError_t e2 = hipStreamSynchronize(s);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            set_error(std::string("relabelling failed: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
            ok = false;
        }
    }
    dfree(k0); dfree(k1); dfree(v0); dfree(v1); dfree(deg); dfree(deg64); dfree(nw_in);
    return ok;
}

}  // namespace

bool build_graph(DevGraph *G, int N, int E, const int *U, const int *V, const double *W,
                 const double *NW, int variant)
{
    if (N < 0 || E < 0) { set_error("negative graph size"); return false; }
    if ((int64_t)E * 2 >= (int64_t)INT32_MAX) { set_error("too many edges for int32 CSR"); return false; }
    if (variant < 0 || variant > 3) { set_error("unknown variant"); return false; }
    // ids index the degree counters and rows below: reject any outside [0, N)
    if (!check_vertex_ids(N, E, U, V, "graph build")) return false;
    G->N = N; G->E = E; G->variant = variant; G->weighted = W != nullptr;
    if (!G->stream) {
        if (hipStream_t cs = ctx_stream()) {
            G->stream = cs;
            G->own_stream = false;
        } else {
            DH_CHECK(hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking));
            G->own_stream = true;
        }
    }
    hipStream_t s = G->stream;
    // the uploads below read the caller's arrays asynchronously: whatever
    // path leaves this function, they are complete before it returns (the
    // caller may free the arrays at once)
    struct SyncOnExit {
        hipStream_t s;
        ~SyncOnExit() { (void)hipStreamSynchronize(s); }
    } sync_on_exit{s};

    if (!dalloc(&G->U, E, "U") || !dalloc(&G->V, E, "V") || !dalloc(&G->ew_in, E, "edge_weight") ||
        !dalloc(&G->nw, (size_t)N, "NW") || !dalloc(&G->c4_in, E, "self term") ||
        !dalloc(&G->row_off, (size_t)N + 1, "row offsets"))
        return false;
    if (E > 0) {
        DH_CHECK(hipMemcpyAsync(G->U, U, (size_t)E * sizeof(int), hipMemcpyHostToDevice, s));
        DH_CHECK(hipMemcpyAsync(G->V, V, (size_t)E * sizeof(int), hipMemcpyHostToDevice, s));
    }
    double *dW = nullptr;
    if (W && E > 0) {
        if (!dalloc(&dW, E, "W")) return false;
        DH_CHECK(hipMemcpyAsync(dW, W, (size_t)E * sizeof(double), hipMemcpyHostToDevice, s));
    }
    if (NW && N > 0) DH_CHECK(hipMemcpyAsync(G->nw, NW, (size_t)N * sizeof(double), hipMemcpyHostToDevice, s));
    else if (N > 0) hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(N, 4096)), dim3(TPB), 0, s, G->nw, (int64_t)N, 1.0);

    // raw out-adjacency, only needed to find reciprocal edges (DIRECTED)
    int64_t *roff = nullptr;
    uint64_t *rkey = nullptr;
    int *rval = nullptr;
    int *deg = nullptr;
    int64_t *deg64 = nullptr;
    if (!dalloc(&deg, (size_t)N + 1, "degrees") || !dalloc(&deg64, (size_t)N + 1, "degrees")) {
        dfree(deg); dfree(dW);
        return false;
    }
    bool ok = true;
    if (variant == DRESS_VARIANT_DIRECTED && E > 0) {
        uint64_t *k0 = nullptr;
        int *v0 = nullptr;
        ok = dalloc(&k0, E, "raw keys") && dalloc(&v0, E, "raw vals") && dalloc(&rkey, E, "raw keys") &&
             dalloc(&rval, E, "raw vals") && dalloc(&roff, (size_t)N + 1, "raw offsets");
        if (ok) {
            hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(N + 1, 4096)), dim3(TPB), 0, s, deg, (int64_t)N + 1, 0);
            hipLaunchKernelGGL(k_raw_keys, dim3(grid_for(E, 8192)), dim3(TPB), 0, s, G->U, G->V, E, k0, v0, deg);
            ok = prim_sort_pairs_u64_i32(k0, rkey, v0, rval, (size_t)E, 64, s) &&
                 prim_exclusive_scan_i32_to_i64(deg, roff, (size_t)N + 1, s);
        }
        dfree(k0); dfree(v0);
        if (!ok) { dfree(rkey); dfree(rval); dfree(roff); dfree(deg); dfree(deg64); dfree(dW); return false; }
    }

    // variant half-edges
    uint64_t *hk0 = nullptr, *hk1 = nullptr;
    int *hv0 = nullptr, *hv1 = nullptr;
    int64_t slots = 2 * (int64_t)E;
    ok = dalloc(&hk0, (size_t)slots + 1, "half keys") && dalloc(&hk1, (size_t)slots + 1, "half keys") &&
         dalloc(&hv0, (size_t)slots + 1, "half vals") && dalloc(&hv1, (size_t)slots + 1, "half vals");
    if (ok) {
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(N + 1, 4096)), dim3(TPB), 0, s, deg, (int64_t)N + 1, 0);
        if (E > 0)
            hipLaunchKernelGGL(k_half_keys, dim3(grid_for(E, 8192)), dim3(TPB), 0, s, G->U, G->V,
                               (const double *)dW, E, variant, (const int64_t *)roff,
                               (const uint64_t *)rkey, (const int *)rval, hk0, hv0, G->ew_in, deg);
        ok = prim_sort_pairs_u64_i32(hk0, hk1, hv0, hv1, (size_t)slots, 64, s) &&
             prim_exclusive_scan_i32_to_i64(deg, deg64, (size_t)N + 1, s);
    }
    if (ok) {
        int64_t S = 0;
        hipError_t e1 = hipMemcpyAsync(&S, deg64 + N, sizeof(int64_t), hipMemcpyDeviceToHost, s);
        hipError_t e2 = hipStreamSynchronize(s);
        if (e1 != hipSuccess || e2 != hipSuccess) { set_error("half-edge count download failed"); ok = false; }
        G->S = S;
        if (ok && (!dalloc(&G->adj_x, (size_t)S + 1, "adj_target") || !dalloc(&G->adj_e, (size_t)S + 1, "adj_edge_idx")))
            ok = false;
        if (ok) {
            hipLaunchKernelGGL(k_i64_to_i32, dim3(grid_for(N + 1, 4096)), dim3(TPB), 0, s, deg64, G->row_off, (int64_t)N + 1);
            if (S > 0)
                hipLaunchKernelGGL(k_unpack_half, dim3(grid_for(S, 8192)), dim3(TPB), 0, s, hk1, hv1, S,
                                   G->adj_x, G->adj_e);
            if (E > 0)
                hipLaunchKernelGGL(k_edge_consts, dim3(grid_for(E, 8192)), dim3(TPB), 0, s, G->U, G->V,
                                   G->nw, E, variant, G->c4_in);
            int *dmax = deg;  // reuse: degrees no longer needed
            hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, dmax, (int64_t)1, 0);
            if (N > 0)
                hipLaunchKernelGGL(k_max_deg, dim3(grid_for(N, 2048)), dim3(TPB), 0, s, G->row_off, N, dmax);
            int md = 0;
            hipError_t e3 = hipMemcpyAsync(&md, dmax, sizeof(int), hipMemcpyDeviceToHost, s);
            hipError_t e4 = hipStreamSynchronize(s);
            if (e3 != hipSuccess || e4 != hipSuccess) { set_error("graph build failed: " + std::string(hipGetErrorString(e4))); ok = false; }
            G->max_degree = md;
        }
    }
    dfree(hk0); dfree(hk1); dfree(hv0); dfree(hv1);
    dfree(rkey); dfree(rval); dfree(roff); dfree(deg); dfree(deg64); dfree(dW);
    return ok && relabel(G);
}

// least row length that gets a hub bitmap (DRESS_HIP_HUB_DEGREE overrides,
// read at every index build; 0 switches bitmaps off)
int hub_degree()
{
    const char *e = getenv("DRESS_HIP_HUB_DEGREE");
    return e ? std::max(0, atoi(e)) : 1024;
}

// device buffers of the hub bitmaps of one index build
struct HubBuffers {
    int *slot = nullptr, *vertex = nullptr, *count = nullptr, *bad = nullptr;
    unsigned long long *bits = nullptr;
    unsigned *pre = nullptr;
    void release()
    {
        dfree(slot); dfree(vertex); dfree(count); dfree(bad); dfree(bits); dfree(pre);
        slot = vertex = count = bad = nullptr;
        bits = nullptr;
        pre = nullptr;
    }
    HubBuffers() = default;
    HubBuffers(const HubBuffers &) = delete;
    HubBuffers &operator=(const HubBuffers &) = delete;
    ~HubBuffers() { release(); }
};

// Bitmaps of the rows (row_off / adj_x: the internal CSR, or the
// reference-layout one, both sorted by label) with at least hub_degree()
// neighbours, the least degree doubled until they take at most 2 GiB
// (ceil(N / 64) words of 12 bytes per hub).  Returns an empty view when
// there are none.
HubBits build_hubs(DevGraph *G, HubBuffers &B, const int *row_off, const int *adj_x)
{
    HubBits view;
    const int N = G->N;
    int deg_min = hub_degree();
    if (deg_min <= 0 || N <= 0 || G->max_degree < deg_min) return view;
    hipStream_t s = G->stream;
    const int64_t W = ((int64_t)N + 63) / 64;
    const int64_t cap = (int64_t)2 << 30;
    if (!dalloc(&B.slot, (size_t)N + 1, "hub slots") || !dalloc(&B.vertex, (size_t)N + 1, "hub vertices") ||
        !dalloc(&B.count, 1, "hub count")) {
        hipGetLastError();
        clear_error();
        B.release();
        return view;
    }
    int nh = 0;
    for (;;) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, B.count, (int64_t)1, 0);
        hipLaunchKernelGGL(k_hub_assign, dim3(grid_for(N, 4096)), dim3(TPB), 0, s, row_off, N, deg_min, B.slot,
                           B.vertex, B.count);
        if (hipMemcpyAsync(&nh, B.count, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            nh = 0;
            break;
        }
        if ((int64_t)nh * W * 12 <= cap || deg_min > G->max_degree) break;
        deg_min *= 2;
    }
    if (nh <= 0 || (int64_t)nh * W * 12 > cap) {
        B.release();
        return view;
    }
    if (!dalloc(&B.bits, (size_t)nh * (size_t)W, "hub bitmaps") || !dalloc(&B.pre, (size_t)nh * (size_t)W, "hub counts") ||
        !dalloc(&B.bad, (size_t)nh, "hub flags")) {
        // no room: the index is built without bitmaps
        hipGetLastError();
        clear_error();
        B.release();
        return view;
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for((int64_t)nh * W * 2, 8192)), dim3(TPB), 0, s, (int *)B.bits,
                       (int64_t)nh * W * 2, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(nh, 4096)), dim3(TPB), 0, s, B.bad, (int64_t)nh, 0);
    const int hgrid = std::min(nh, 65535);
    hipLaunchKernelGGL(k_hub_bits, dim3(hgrid), dim3(TPB), 0, s, row_off, adj_x, B.vertex, nh, W, B.bits, B.bad);
    hipLaunchKernelGGL(k_hub_prefix, dim3(hgrid), dim3(256), 0, s, B.bits, nh, W, B.pre, B.bad, B.vertex, B.slot);
    view.slot = B.slot;
    view.bits = B.bits;
    view.pre = B.pre;
    view.W = W;
    return view;
}

bool build_index(DevGraph *G)
{
    if (G->have_index) return true;
    hipStream_t s = G->stream;
    const int E = G->E;
    const int GS = 16;
    int grid = grid_for((int64_t)E * GS, num_cus() * 32);
    int64_t *cnt_i = nullptr;
    if (!dalloc(&G->cnt_in, (size_t)E + 1, "isect counts") || !dalloc(&cnt_i, (size_t)E + 1, "isect counts") ||
        !dalloc(&G->ioff, (size_t)E + 1, "isect offsets") || !dalloc(&G->ord, (size_t)E + 1, "edge order") ||
        !dalloc(&G->inv, (size_t)E + 1, "edge order") || !dalloc(&G->uv, (size_t)E + 1, "edge ends") ||
        !dalloc(&G->rs, (size_t)E + 1, "row starts") || !dalloc(&G->pos, (size_t)E + 1, "edge half-edges") ||
        !dalloc(&G->pos_in, (size_t)E + 1, "edge half-edges") ||
        !dalloc(&G->ew, (size_t)E + 1, "edge weights") || !dalloc(&G->c4, (size_t)E + 1, "self terms")) {
        dfree(cnt_i);
        return false;
    }
    // 1. half-edges of every edge, |N(u) & N(v)| per input edge
    HubBuffers hubs;
    HubBits hub;
    if (E > 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(2 * (int64_t)E, 8192)), dim3(TPB), 0, s, (int *)G->pos_in,
                           2 * (int64_t)E, -1);
        if (G->S > 0)
            hipLaunchKernelGGL(k_edge_positions, dim3(grid_for(G->S, 8192)), dim3(TPB), 0, s, G->adj_e, G->S,
                               G->pos_in);
        hipLaunchKernelGGL(k_order_positions, dim3(grid_for(E, 8192)), dim3(TPB), 0, s, G->pos_in, E);
        hub = build_hubs(G, hubs, G->row_off, G->adj_x);
        // from 1M edges on, the count walks the edges in the order of their
        // first half-edge, so the edges of one row run together (71 -> 33
        // ms on the 4M-vertex graph against input order); smaller graphs,
        // or a build without room for the sort, walk them in input order
        uint64_t *ck0 = nullptr, *ck1 = nullptr;
        int *cv0 = nullptr, *cord = nullptr;
        bool sorted = E >= (1 << 20) && dalloc(&ck0, (size_t)E + 1, "count order keys") &&
                      dalloc(&ck1, (size_t)E + 1, "count order keys") && dalloc(&cv0, (size_t)E + 1, "count order") &&
                      dalloc(&cord, (size_t)E + 1, "count order");
        if (sorted) {
            int bits = 1;
            while (bits < 32 && ((int64_t)1 << bits) <= G->S) bits++;
            hipLaunchKernelGGL(k_first_half_keys, dim3(grid_for(E, 8192)), dim3(TPB), 0, s, G->pos_in, E, ck0, cv0);
            sorted = prim_sort_pairs_u64_i32(ck0, ck1, cv0, cord, (size_t)E, bits, s);
        }
        if (!sorted) {
            (void)hipGetLastError();
            clear_error();
        }
        hipLaunchKernelGGL((k_isect<GS, false>), dim3(grid), dim3(TPB), 0, s, G->U, G->V, E, G->row_off,
                           G->adj_x, (const int *)(sorted ? cord : nullptr), G->cnt_in, (const int64_t *)nullptr,
                           (uint32_t *)nullptr, (int2 *)nullptr, hub);
        dfree(ck0); dfree(ck1); dfree(cv0); dfree(cord);
    }
    // 2. internal order: by bin, then by first half-edge (CSR order)
    int counts[Bins::NB] = {0, 0, 0, 0, 0};
    if (E > 0) {
        uint64_t *k0 = nullptr, *k1 = nullptr;
        int *v0 = nullptr, *cd = nullptr;
        bool ok2 = dalloc(&k0, (size_t)E + 1, "order keys") && dalloc(&k1, (size_t)E + 1, "order keys") &&
                   dalloc(&v0, (size_t)E + 1, "order vals") && dalloc(&cd, Bins::NB, "bin counts");
        if (ok2) {
            hipLaunchKernelGGL(k_order_keys, dim3(grid_for(E, 8192)), dim3(TPB), 0, s, G->cnt_in, G->pos_in, G->U, G->V,
                               G->row_off, E, low_order_bins(), G->N < (1 << 28) ? tile_bins() : 0, tile_logs(), k0,
                               v0);
            ok2 = prim_sort_pairs_u64_i32(k0, k1, v0, G->ord, (size_t)E, 64, s);
        }
        ok2 = ok2 && hipMemsetAsync(cd, 0, Bins::NB * sizeof(int), s) == hipSuccess;
        if (ok2)
            hipLaunchKernelGGL(k_count_bins, dim3(grid_for(E, 2048)), dim3(TPB), 0, s, G->ord, G->cnt_in, E, cd);
        ok2 = ok2 && hipMemcpyAsync(counts, cd, sizeof(counts), hipMemcpyDeviceToHost, s) == hipSuccess &&
              hipStreamSynchronize(s) == hipSuccess;
        dfree(k0); dfree(k1); dfree(v0); dfree(cd);
        if (!ok2) {
            if (!has_error()) set_error("edge ordering failed");
            dfree(cnt_i);
            return false;
        }
        hipLaunchKernelGGL(k_internal_edges, dim3(grid_for(E, 8192)), dim3(TPB), 0, s, G->ord, E, G->U, G->V,
                           G->row_off, G->ew_in, G->c4_in, G->cnt_in, G->pos_in, G->inv, G->uv, G->rs, G->pos,
                           G->ew, G->c4, cnt_i);
    }
    G->ebin[0] = 0;
    for (int b = 0; b < Bins::NB; b++) G->ebin[b + 1] = G->ebin[b] + counts[b];
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, (int *)(cnt_i + E), (int64_t)2, 0);
    bool ok = prim_exclusive_scan_i64(cnt_i, G->ioff, (size_t)E + 1, s);
    dfree(cnt_i);
    if (!ok) return false;
    int64_t T = 0;
    DH_CHECK(hipMemcpyAsync(&T, G->ioff + E, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    DH_CHECK(hipStreamSynchronize(s));
    G->T = T;
    // 3. fill, in internal order
    G->packed = G->max_degree <= 65535;
    if (G->packed) {
        if (!dalloc(&G->ipk, (size_t)T + 1, "intersection index")) return false;
    } else {
        if (!dalloc(&G->ipair, (size_t)T + 1, "intersection index")) return false;
    }
    if (E > 0 && T > 0)
        hipLaunchKernelGGL((k_isect<GS, true>), dim3(grid), dim3(TPB), 0, s, G->U, G->V, E, G->row_off,
                           G->adj_x, G->ord, (int64_t *)nullptr, (const int64_t *)G->ioff, G->ipk, G->ipair, hub);
    DH_CHECK(hipGetLastError());
    DH_CHECK(hipStreamSynchronize(s));
    G->have_index = true;
    return true;
}

bool build_bins(DevGraph *G)
{
    if (G->have_bins) return true;
    if (!build_index(G)) return false;
    if (!dalloc(&G->vbins.list, (size_t)G->N + 1, "vertex bins")) return false;
    int counts[Bins::NB] = {0, 0, 0, 0, 0};
    if (!vertex_bins(G->stream, G->N, G->row_off, G->vbins.list, counts)) return false;
    int st = 0;
    for (int b = 0; b < Bins::NB; b++) { G->vbins.start[b] = st; G->vbins.count[b] = counts[b]; st += counts[b]; }
    G->have_bins = true;
    return true;
}

bool refresh_self_terms(DevGraph *G, const double *NW_host)
{
    G->fit_state_valid = false;
    if (!NW_host && G->self_terms_ones) return true;
    G->self_terms_ones = false;
    G->uni_state = 0;
    hipStream_t s = G->stream;
    // no vertex weights: every vertex weighs 1.0 (the device copy may still
    // hold an earlier call's weights)
    if (!NW_host && G->N > 0)
        hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(G->N, 4096)), dim3(TPB), 0, s, G->nw, (int64_t)G->N, 1.0);
    if (NW_host && G->N > 0) {
        if (!G->vstage && !dalloc(&G->vstage, (size_t)G->N + 1, "vertex staging")) return false;
        DH_CHECK(hipMemcpyAsync(G->vstage, NW_host, (size_t)G->N * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_to_internal, dim3(grid_for(G->N, 4096)), dim3(TPB), 0, s, (const double *)G->vstage,
                           (const int *)G->rank, G->N, G->nw);
    }
    if (G->E > 0) {
        hipLaunchKernelGGL(k_edge_consts, dim3(grid_for(G->E, 8192)), dim3(TPB), 0, s, G->U, G->V, G->nw,
                           G->E, G->variant, G->c4_in);
        if (G->have_index)
            hipLaunchKernelGGL(k_self_terms, dim3(grid_for(G->E, 8192)), dim3(TPB), 0, s, G->ord, G->E, G->c4_in,
                               G->c4);
    }
    DH_CHECK(hipGetLastError());
    // the internal self terms exist only once the index is built
    G->self_terms_ones = !NW_host && (G->have_index || G->E == 0);
    return true;
}

bool download_csr(DevGraph *G, int *adj_offset, int *adj_target, int *adj_edge_idx,
                  double *edge_weight)
{
    hipStream_t s = G->stream;
    if (!G->ref_row_off) { set_error("reference adjacency is no longer on the device"); return false; }
    // the caller's pageable arrays: large ones through the host-parallel
    // staged copy (its source must be complete)
    DH_CHECK(hipStreamSynchronize(s));
    bool ok = d2h_parallel(adj_offset, G->ref_row_off, ((size_t)G->N + 1) * sizeof(int)) &&
              (G->S == 0 || (d2h_parallel(adj_target, G->ref_adj_x, (size_t)G->S * sizeof(int)) &&
                             d2h_parallel(adj_edge_idx, G->ref_adj_e, (size_t)G->S * sizeof(int)))) &&
              (G->E == 0 || d2h_parallel(edge_weight, G->ew_in, (size_t)G->E * sizeof(double)));
    if (!ok) set_error("download of the host adjacency failed");
    return ok;
}

// Intercept arrays in the reference's form (input edge ids, int offsets in
// input order), malloc'd: the multiset intersections are recomputed on the
// reference-layout rows, so every edge lists its common neighbours in
// increasing input label, as dress.c:307-363 does.
bool download_index(DevGraph *G, int **ioff_out, int **ux_out, int **vx_out)
{
    if (!build_index(G)) return false;
    if (!G->ref_row_off) { set_error("reference adjacency is no longer on the device"); return false; }
    if (G->T >= (int64_t)INT32_MAX) { set_error("intersection index exceeds int32 offsets of dress_graph_t"); return false; }
    hipStream_t s = G->stream;
    const int E = G->E;
    const int64_t T = G->T;
    const int GS = 16;
    // the caller frees these with free(): large ones on huge pages
    int *io = (int *)host_alloc(((size_t)E + 1) * sizeof(int));
    int *ux = (int *)host_alloc(((size_t)T + 1) * sizeof(int));
    int *vx = (int *)host_alloc(((size_t)T + 1) * sizeof(int));
    int *dux = nullptr, *dvx = nullptr, *ident = nullptr, *dio32 = nullptr;
    int64_t *dio = nullptr, *dcnt = nullptr;
    int2 *pairs = nullptr;
    if (!io || !ux || !vx) set_error("host allocation failed for intercept arrays");
    bool ok = io && ux && vx && dalloc(&dux, (size_t)T + 1, "ux") && dalloc(&dvx, (size_t)T + 1, "vx") &&
              dalloc(&dio, (size_t)E + 1, "input offsets") && dalloc(&dio32, (size_t)E + 1, "input offsets") &&
              dalloc(&dcnt, (size_t)E + 1, "input counts") &&
              dalloc(&ident, (size_t)E + 1, "identity order") && dalloc(&pairs, (size_t)T + 1, "half-edge pairs");
    if (ok) {
        hipError_t e0 = hipMemcpyAsync(dcnt, G->cnt_in, (size_t)E * sizeof(int64_t), hipMemcpyDeviceToDevice, s);
        if (e0 != hipSuccess) { set_error("intercept copy failed"); ok = false; }
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, (int *)(dcnt + E), (int64_t)2, 0);
        ok = ok && prim_exclusive_scan_i64(dcnt, dio, (size_t)E + 1, s);
    }
    HubBuffers ref_hubs;
    if (ok) {
        if (E > 0 && T > 0) {
            const HubBits ref_hub = build_hubs(G, ref_hubs, G->ref_row_off, G->ref_adj_x);
            hipLaunchKernelGGL(k_iota, dim3(grid_for(E, 8192)), dim3(TPB), 0, s, ident, E);
            hipLaunchKernelGGL((k_isect<GS, true>), dim3(grid_for((int64_t)E * GS, num_cus() * 32)), dim3(TPB), 0,
                               s, (const int *)G->ref_U, (const int *)G->ref_V, E, (const int *)G->ref_row_off,
                               (const int *)G->ref_adj_x, (const int *)ident, (int64_t *)nullptr,
                               (const int64_t *)dio, (uint32_t *)nullptr, pairs, ref_hub);
            hipLaunchKernelGGL(k_pairs_to_edges, dim3(grid_for(T, 8192)), dim3(TPB), 0, s, (const int2 *)pairs, T,
                               (const int *)G->ref_adj_e, dux, dvx);
        }
        // offsets to int on the device (T < 2^31 was checked), then the
        // host-parallel staged copies (their sources complete)
        hipLaunchKernelGGL(k_i64_to_i32, dim3(grid_for((int64_t)E + 1, 8192)), dim3(TPB), 0, s, dio, dio32,
                           (int64_t)E + 1);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess &&
             d2h_parallel(io, dio32, ((size_t)E + 1) * sizeof(int)) &&
             (T == 0 || (d2h_parallel(ux, dux, (size_t)T * sizeof(int)) &&
                         d2h_parallel(vx, dvx, (size_t)T * sizeof(int))));
        if (!ok) set_error("intercept download failed");
    }
    dfree(dux); dfree(dvx); dfree(dio); dfree(dio32); dfree(dcnt); dfree(ident); dfree(pairs);
    if (!ok) { free(io); free(ux); free(vx); return false; }
    *ioff_out = io; *ux_out = ux; *vx_out = vx;
    return true;
}

// frees the reference-layout arrays that are not also the fit's arrays
void free_reference_arrays(DevGraph *G)
{
    if (G->ref_U != G->U) dfree(G->ref_U);
    if (G->ref_V != G->V) dfree(G->ref_V);
    if (G->ref_row_off != G->row_off) dfree(G->ref_row_off);
    if (G->ref_adj_x != G->adj_x) dfree(G->ref_adj_x);
    if (G->ref_adj_e != G->adj_e) dfree(G->ref_adj_e);
    G->ref_U = G->ref_V = G->ref_row_off = G->ref_adj_x = G->ref_adj_e = nullptr;
}

// per-vertex device values in internal labels -> host array in input labels
bool download_vertex_values(DevGraph *G, const double *internal, double *host)
{
    if (G->N <= 0) return true;
    hipStream_t s = G->stream;
    if (!G->vstage && !dalloc(&G->vstage, (size_t)G->N + 1, "vertex staging")) return false;
    hipLaunchKernelGGL(k_to_input, dim3(grid_for(G->N, 4096)), dim3(TPB), 0, s, internal, (const int *)G->rank, G->N,
                       G->vstage);
    DH_CHECK(hipMemcpyAsync(host, G->vstage, (size_t)G->N * sizeof(double), hipMemcpyDeviceToHost, s));
    DH_CHECK(hipStreamSynchronize(s));
    return true;
}

void release_reference_csr(DevGraph *G)
{
    if (G->stream) hipStreamSynchronize(G->stream);
    free_reference_arrays(G);
}

// the reference-layout adjacency of a dress_graph_t (input labels), for
// queries that must probe rows in the reference's order
bool upload_reference_csr(DevGraph *G, const int *adj_offset, const int *adj_target, const int *adj_edge_idx)
{
    if (G->ref_row_off && G->ref_adj_x && G->ref_adj_e) return true;
    if (!adj_offset || (G->S > 0 && (!adj_target || !adj_edge_idx))) {
        set_error("graph has no adjacency arrays");
        return false;
    }
    hipStream_t s = G->stream;
    if (!G->ref_row_off && !dalloc(&G->ref_row_off, (size_t)G->N + 1, "reference row offsets")) return false;
    if (!G->ref_adj_x && !dalloc(&G->ref_adj_x, (size_t)G->S + 1, "reference adjacency")) return false;
    if (!G->ref_adj_e && !dalloc(&G->ref_adj_e, (size_t)G->S + 1, "reference adjacency")) return false;
    // the caller's pageable arrays, large ones through the host-parallel
    // staged copy (the blocks may still be read by earlier work on s)
    DH_CHECK(hipStreamSynchronize(s));
    const bool ok = h2d_parallel(G->ref_row_off, adj_offset, ((size_t)G->N + 1) * sizeof(int)) &&
                    (G->S == 0 || (h2d_parallel(G->ref_adj_x, adj_target, (size_t)G->S * sizeof(int)) &&
                                   h2d_parallel(G->ref_adj_e, adj_edge_idx, (size_t)G->S * sizeof(int))));
    if (!ok) set_error("upload of the reference adjacency failed");
    return ok;
}

void free_graph(DevGraph *G)
{
    if (!G) return;
    if (G->stream) hipStreamSynchronize(G->stream);
    dfree(G->U); dfree(G->V); dfree(G->row_off); dfree(G->adj_x); dfree(G->adj_e);
    dfree(G->ew_in); dfree(G->c4_in); dfree(G->nw); dfree(G->pos_in);
    dfree(G->rank); dfree(G->vstage); dfree(G->q_d); dfree(G->q_vd); dfree(G->q_nw);
    free_reference_arrays(G);
    dfree(G->ord); dfree(G->inv); dfree(G->uv); dfree(G->rs); dfree(G->pos); dfree(G->ew); dfree(G->c4);
    dfree(G->ioff); dfree(G->ipk); dfree(G->ipair); dfree(G->cnt_in); dfree(G->b0off); dfree(G->b0pair);
    dfree(G->vbins.list);
    for (int i = 0; i < 2; i++) { dfree(G->d[i]); dfree(G->wh[i]); dfree(G->d32[i]); dfree(G->wh32[i]); }
    dfree(G->vd); dfree(G->stage); dfree(G->state);
    for (int i = 0; i < 2; i++)
        if (G->iter_graph[i]) hipGraphExecDestroy(G->iter_graph[i]);
    if (G->stream && G->own_stream) hipStreamDestroy(G->stream);
    delete G;
}

}  // namespace dh
