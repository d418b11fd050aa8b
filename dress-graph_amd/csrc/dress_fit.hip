// dress_fit.hip -- the DRESS fixed-point iteration on one device-resident graph.
//
// One iteration (libdress/src/dress.c:546-608) is three kernels:
//   vertex phase   vd[u] = sqrt( 4nw[u] + sum_{s in row u} wh[s] )
//   edge phase     d'[e] = ( sum_k (wh[a_k] + wh[b_k]) + t0_e ) / (vd[U]*vd[V])
//                  t0_e = c4[e] + 2*w[e]*d[e]   (c4[e] + w[e]*d[e] FWD/BWD)
//                  with max|d - d'| folded into one device word, and the new
//                  w[e]*d'[e] written onto the edge's half-edges
//   end            convergence bookkeeping on the device (strict < epsilon)
// so the host only reads the state every few iterations.  wh[] holds one
// copy of w-bar*d per CSR entry: the vertex phase streams rows, and the
// common-neighbour terms of an edge are read at increasing positions of its
// two rows, so neighbouring lanes share cache lines.
//
// What bounds these kernels is the number of L1 -> L2 requests: a CU keeps
// about 100 of them in flight, so time ~ requests x latency.  The index and
// the per-edge data of bins 1-4 and the vertex rows are therefore read
// through the caches: the G lanes of an edge read its index in slices of G
// entries, four slices a round, and one cached line serves all four.
// Non-temporal loads kept no line in L1 and made each slice a request of
// its own (4M-vertex fit 134.0 -> 125.4 ms, same values).  Bin 0 streams
// each edge's data once and keeps non-temporal loads, which measured
// slightly faster there.
//
// Sums are exact-rounding compensated sums (TwoSum per term, pairwise
// TwoSum merges across lanes): the result is the correctly rounded sum of
// the terms, whatever order the lanes add them in.  The reference's sort +
// Kahan-Babuska-Neumaier sum (dress.c:51-66) produces the same correctly
// rounded value, so values match it bit for bit (except where an exact sum
// lies within ~n^2 u^2 of a rounding midpoint), and they are invariant
// under relabelling like the reference's.  All FP64 code is compiled with
// -ffp-contract=off: the reference computes w*d + w*d with two roundings.
//
// Work is degree binned: rows with <= 4, 64, 512, 8192 terms use groups of
// 1, 4, 16, 64 lanes, edges groups of 1, 4, 8, 32 lanes; larger ones a
// 256-lane workgroup.  Edges of one bin are a contiguous range of internal
// ids.  The vertex bins run as one launch.
#include "dress_device.h"
#include "dress_hip_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdlib>

namespace dh {

struct FitState {
    int iter;                // iterations finished
    int done;
    int iterations;          // reference convention result
    int cur;                 // buffer index holding the current values
    double delta;            // max change of the last finished iteration
    double prev_delta;       // ... and of the one before
    unsigned long long maxd; // running max |change| (bit pattern, >= 0)
    int fp32;                // 1 while values live in the FP32 buffers,
                             // 2 during the iteration that converts them
    int switched_at;         // iteration at which FP32 -> FP64 happened
    double switch_thr;       // hand over once the FP32 change is below this
    int switch_by;           // ... or at this iteration at the latest
    int mixed;               // fit started in FP32
    int ambiguous;           // an FP64-phase change fell within 0.1% of
                             // epsilon: the stop step may differ from a
                             // pure FP64 run, so the host reruns in FP64
};

static_assert(2 * sizeof(FitState) <= SNAPSHOT_BYTES, "two fit state snapshots per block");

namespace {

constexpr int TPB = 256;

// workgroup-wide merge of (s, c); result valid in thread 0
__device__ __forceinline__ void block_reduce(double &s, double &c)
{
    __shared__ double ss[TPB / 64], sc[TPB / 64];
    group_reduce<64>(s, c);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { ss[w] = s; sc[w] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < TPB / 64; i++) two_sum_merge(s, c, ss[i], sc[i]);
    }
    __syncthreads();
}

// max over the wave into *out (bit pattern of a value >= 0); the atomic is
// skipped where the wave's max cannot raise the running one
__device__ __forceinline__ void wave_max_atomic(double m, unsigned long long *out)
{
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0 && m > 0.0) {
        const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
        if (bits > __atomic_load_n(out, __ATOMIC_RELAXED)) atomicMax(out, bits);
    }
}

struct EdgeArgs {
    const int2 *uv;
    const int2 *rs;
    const int2 *pos;
    const double *ew;
    const double *c4;
    const int64_t *ioff;
    const uint32_t *ipk;
    const int2 *ipair;
    const double *vd;
    const int *b0off;              // bin 0: offsets relative to the bin
    const int2 *b0pair;            // bin 0: absolute half-edge pairs
    double ew_c, c4_c;             // the uniform w-bar and self-term constant
    int directed;
};

// half-edge pair of index entry k of an edge whose rows start at r
template <bool PK>
__device__ __forceinline__ void entry(const EdgeArgs &a, int64_t k, int2 r, int &x, int &y)
{
    if (PK) {
        uint32_t q = a.ipk[k];
        x = r.x + (int)(q & 0xffffu);
        y = r.y + (int)(q >> 16);
    } else {
        int2 p = a.ipair[k];
        x = p.x;
        y = p.y;
    }
}

// Vertex phase: vd[u] = sqrt(4nw[u] + sum of w-bar*d over row u).
// VT = double (FP64 phase) or float (FP32 phase: values stored in FP32,
// summed in FP64, exact for FP32 terms within 2^29 of each other).
template <int G, typename VT, bool UNI>
__device__ __forceinline__ void vertex_rows(const int *__restrict__ list, int count, int blk, int nblk,
                                            const int *__restrict__ row_off, const VT *__restrict__ wh,
                                            const double *__restrict__ nw, double nw_c, double *__restrict__ vd)
{
    const int lane = (G == TPB) ? threadIdx.x : (threadIdx.x % G);
    const int64_t groups = (int64_t)nblk * (TPB / G);
    for (int64_t g = (int64_t)blk * (TPB / G) + threadIdx.x / G; g < count; g += groups) {
        int u = list[g];
        int b = row_off[u], e = row_off[u + 1];
        double s = 0.0, c = 0.0;
        if (lane == 0) s = 4.0 * (UNI ? nw_c : nw[u]);
        int i = b + lane;
        for (; i + G < e; i += 2 * G) {
            double x0 = (double)wh[i], x1 = (double)wh[i + G];
            two_sum_add(s, c, x0);
            two_sum_add(s, c, x1);
        }
        for (; i < e; i += G) two_sum_add(s, c, (double)wh[i]);
        if (G == TPB) block_reduce(s, c);
        else group_reduce<G>(s, c);
        if (lane == 0) vd[u] = sqrt(s + c);
    }
}

template <int G, typename VT, bool UNI>
__global__ void __launch_bounds__(TPB) k_vertex(const int *__restrict__ list, int count,
                                                const int *__restrict__ row_off, const VT *w0, const VT *w1,
                                                const double *__restrict__ nw, double nw_c, double *__restrict__ vd,
                                                const FitState *__restrict__ st)
{
    const bool f32 = sizeof(VT) == 4;
    if (st->done || (st->fp32 != 0) != f32) return;
    vertex_rows<G, VT, UNI>(list, count, blockIdx.x, gridDim.x, row_off, st->cur ? w1 : w0, nw, nw_c, vd);
}

// the five vertex bins of an iteration in one launch: blocks [first[b],
// first[b+1]) serve bin b (rows of up to 4, 64, 512, 8192 and more entries
// with 1, 4, 16, 64 and 256 lanes), the widest rows first
struct VertexBins {
    const int *list[Bins::NB];
    int count[Bins::NB];
    int first[Bins::NB + 1];       // in launch order: bins 4, 3, 2, 1, 0
};

template <typename VT, bool UNI>
__global__ void __launch_bounds__(TPB) k_vertex_all(VertexBins vb, const int *__restrict__ row_off, const VT *w0,
                                                    const VT *w1, const double *__restrict__ nw, double nw_c,
                                                    double *__restrict__ vd, const FitState *__restrict__ st)
{
    const bool f32 = sizeof(VT) == 4;
    if (st->done || (st->fp32 != 0) != f32) return;
    const VT *wh = st->cur ? w1 : w0;
    const int b = blockIdx.x;
    if (b < vb.first[1]) {
        vertex_rows<TPB, VT, UNI>(vb.list[4], vb.count[4], b, vb.first[1], row_off, wh, nw, nw_c, vd);
    } else if (b < vb.first[2]) {
        vertex_rows<64, VT, UNI>(vb.list[3], vb.count[3], b - vb.first[1], vb.first[2] - vb.first[1], row_off, wh, nw,
                                 nw_c, vd);
    } else if (b < vb.first[3]) {
        vertex_rows<16, VT, UNI>(vb.list[2], vb.count[2], b - vb.first[2], vb.first[3] - vb.first[2], row_off, wh, nw,
                                 nw_c, vd);
    } else if (b < vb.first[4]) {
        vertex_rows<4, VT, UNI>(vb.list[1], vb.count[1], b - vb.first[3], vb.first[4] - vb.first[3], row_off, wh, nw,
                                nw_c, vd);
    } else {
        vertex_rows<1, VT, UNI>(vb.list[0], vb.count[0], b - vb.first[4], vb.first[5] - vb.first[4], row_off, wh, nw,
                                nw_c, vd);
    }
}

// first edge of slice `j` of `nx` over [lo, hi): slices hold about the same
// cost, counted as index entries + SLICE_EDGE_COST per edge
constexpr int64_t SLICE_EDGE_COST = 8;
__device__ __forceinline__ int slice_start(const int64_t *ioff, int lo, int hi, int j, int nx)
{
    if (j <= 0) return lo;
    if (j >= nx) return hi;
    const int64_t base = ioff[lo];
    const int64_t total = (ioff[hi] - base) + SLICE_EDGE_COST * (hi - lo);
    const int64_t target = total / nx * j + (total % nx) * j / nx;
    int a = lo, b = hi;                  // first e with cost(e) >= target
    while (a < b) {
        int m = a + (b - a) / 2;
        int64_t c = (ioff[m] - base) + SLICE_EDGE_COST * (m - lo);
        if (c < target) a = m + 1;
        else b = m;
    }
    return a;
}

// Edge phase over internal edges [lo, hi).  The range is cut into `nx`
// slices of equal cost and block b works in slice b % nx, which is the XCD
// the hardware dispatches it to (workgroups go round-robin over the XCDs):
// edges of one owner vertex are adjacent in a bin, so the owner's row stays
// in that XCD's L2 while its edges are processed.  The grid is what fits on
// the device at once, so the blocks of a slice sweep it together.
// Edges [lo, hi) of one bin, G lanes per edge, by block j of the K blocks
// that work in slice xs of nx; the largest change into mymax.
template <int G, bool PK, typename VT, bool UNI>
__device__ __forceinline__ void edge_range(int lo, int hi, int nx, int xs, int j, int K, const EdgeArgs &a,
                                           const VT *__restrict__ dc, const VT *__restrict__ wc,
                                           VT *__restrict__ dn, VT *__restrict__ wn, double &mymax)
{
    const int lane = (G == TPB) ? threadIdx.x : (threadIdx.x % G);
    const int s_lo = slice_start(a.ioff, lo, hi, xs, nx);
    const int s_hi = slice_start(a.ioff, lo, hi, xs + 1, nx);
    const int64_t groups = (int64_t)K * (TPB / G);
    for (int64_t g = s_lo + (int64_t)j * (TPB / G) + threadIdx.x / G; g < s_hi; g += groups) {
        const int e = (int)g;
        int64_t kb = a.ioff[e], ke = a.ioff[e + 1];
        int2 r = PK ? a.rs[e] : make_int2(0, 0);
        double s = 0.0, c = 0.0;
        int64_t k = kb + lane;
        // four index entries and eight gathers in flight per lane
        for (; k + 3 * G < ke; k += 4 * G) {
            int x0, y0, x1, y1, x2, y2, x3, y3;
            entry<PK>(a, k, r, x0, y0);
            entry<PK>(a, k + G, r, x1, y1);
            entry<PK>(a, k + 2 * G, r, x2, y2);
            entry<PK>(a, k + 3 * G, r, x3, y3);
            double t0 = (double)wc[x0] + (double)wc[y0];
            double t1 = (double)wc[x1] + (double)wc[y1];
            double t2 = (double)wc[x2] + (double)wc[y2];
            double t3 = (double)wc[x3] + (double)wc[y3];
            two_sum_add(s, c, t0);
            two_sum_add(s, c, t1);
            two_sum_add(s, c, t2);
            two_sum_add(s, c, t3);
        }
        for (; k < ke; k += G) {
            int x, y;
            entry<PK>(a, k, r, x, y);
            two_sum_add(s, c, (double)wc[x] + (double)wc[y]);
        }
        double de = 0.0, w = 0.0, c4 = 0.0;
        if (lane == 0) {
            de = (double)dc[e];
            w = UNI ? a.ew_c : a.ew[e];
            double own = w * de;
            c4 = UNI ? a.c4_c : a.c4[e];
            double t0 = a.directed ? (c4 + own) : (c4 + 2.0 * own);
            two_sum_add(s, c, t0);
        }
        if (G == TPB) block_reduce(s, c);
        else group_reduce<G>(s, c);
        if (lane == 0) {
            // a NaN self term marks an edge removed from a masked subgraph:
            // its value and half-edge copies stay 0 (exact zeros in every
            // other sum) and it takes no part in the change
            const bool dead = !(c4 == c4);
            double num = s + c;
            int2 p = a.uv[e];
            double den = a.vd[p.x] * a.vd[p.y];
            double x = dead ? 0.0 : (den > 0.0 ? num / den : 0.0);
            if (!dead) mymax = fmax(mymax, fabs(de - x));
            dn[e] = (VT)x;
            VT wx = (VT)(w * x);
            int2 q = a.pos[e];
            wn[q.x] = wx;
            if (q.y >= 0) wn[q.y] = wx;
        }
    }
}

template <int G, bool PK, typename VT, bool UNI>
__global__ void __launch_bounds__(TPB) k_edge(int lo, int hi, int nx, EdgeArgs a, VT *d0, VT *d1, VT *w0,
                                              VT *w1, FitState *st)
{
    const bool f32 = sizeof(VT) == 4;
    if (st->done || (st->fp32 != 0) != f32) return;
    const int cur = st->cur;
    double mymax = 0.0;
    edge_range<G, PK, VT, UNI>(lo, hi, nx, blockIdx.x % nx, blockIdx.x / nx, gridDim.x / nx, a, cur ? d1 : d0,
                               cur ? w1 : w0, cur ? d0 : d1, cur ? w0 : w1, mymax);
    wave_max_atomic(mymax, &st->maxd);
}

// first edge (relative to the bin) of slice j of nx over the bin's n edges,
// cost as in slice_start, from the bin's 32-bit offsets
__device__ __forceinline__ int slice_start32(const int *off, int n, int j, int nx)
{
    if (j <= 0) return 0;
    if (j >= nx) return n;
    const int64_t total = (int64_t)off[n] + SLICE_EDGE_COST * n;
    const int64_t target = total / nx * j + (total % nx) * j / nx;
    int a = 0, b = n;
    while (a < b) {
        int m = a + (b - a) / 2;
        int64_t c = (int64_t)off[m] + SLICE_EDGE_COST * m;
        if (c < target) a = m + 1;
        else b = m;
    }
    return a;
}

// Edge phase of bin 0 (at most 3 common neighbours, one lane per edge):
// the index entries are absolute half-edge pairs behind a 32-bit offset, so
// an edge streams its ends, its half-edge slots, one offset and its value
// (and, for graphs whose w-bar or self terms vary, those two), with no row
// starts and no 64-bit offsets.  Same arithmetic, in the same order, as
// k_edge<1>.
template <typename VT, bool UNI>
__device__ __forceinline__ void edge0_range(int lo, int hi, int nx, int xs, int j, int K, const EdgeArgs &a,
                                            const VT *__restrict__ dc, const VT *__restrict__ wc,
                                            VT *__restrict__ dn, VT *__restrict__ wn, double &mymax)
{
    const int n = hi - lo;
    const int s_lo = slice_start32(a.b0off, n, xs, nx);
    const int s_hi = slice_start32(a.b0off, n, xs + 1, nx);
    const int groups = K * TPB;
    for (int g = s_lo + j * TPB + threadIdx.x; g < s_hi; g += groups) {
        const int e = lo + g;
        const int kb = ntl(a.b0off + g), ke = ntl(a.b0off + g + 1);
        double s = 0.0, c = 0.0;
        for (int k = kb; k < ke; k++) {
            const int2 p = ntl2(a.b0pair + k);
            two_sum_add(s, c, (double)wc[p.x] + (double)wc[p.y]);
        }
        const double de = (double)dc[e];
        const double w = UNI ? a.ew_c : ntl(a.ew + e);
        const double c4 = UNI ? a.c4_c : ntl(a.c4 + e);
        const double own = w * de;
        two_sum_add(s, c, a.directed ? (c4 + own) : (c4 + 2.0 * own));
        // a NaN self term marks an edge removed from a masked subgraph
        const bool dead = !(c4 == c4);
        const double num = s + c;
        const int2 p = ntl2(a.uv + e);
        const double den = a.vd[p.x] * a.vd[p.y];
        const double x = dead ? 0.0 : (den > 0.0 ? num / den : 0.0);
        if (!dead) mymax = fmax(mymax, fabs(de - x));
        dn[e] = (VT)x;
        const VT wx = (VT)(w * x);
        const int2 q = ntl2(a.pos + e);
        wn[q.x] = wx;
        if (q.y >= 0) wn[q.y] = wx;
    }
}

template <typename VT, bool UNI>
__global__ void __launch_bounds__(TPB) k_edge0(int lo, int hi, int nx, EdgeArgs a, VT *d0, VT *d1, VT *w0,
                                               VT *w1, FitState *st)
{
    const bool f32 = sizeof(VT) == 4;
    if (st->done || (st->fp32 != 0) != f32) return;
    const int cur = st->cur;
    double mymax = 0.0;
    edge0_range<VT, UNI>(lo, hi, nx, blockIdx.x % nx, blockIdx.x / nx, gridDim.x / nx, a, cur ? d1 : d0,
                         cur ? w1 : w0, cur ? d0 : d1, cur ? w0 : w1, mymax);
    wave_max_atomic(mymax, &st->maxd);
}

__global__ void k_fill_int(int *p, int n, int v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// bin 0's offsets relative to the bin and its entries as absolute pairs
__global__ void k_pack_bin0(int lo, int hi, const int64_t *ioff, const int2 *rs, const uint32_t *ipk,
                            const int2 *ipair, int *b0off, int2 *b0pair)
{
    const int64_t base = ioff[lo];
    for (int e = lo + blockIdx.x * blockDim.x + threadIdx.x; e < hi; e += gridDim.x * blockDim.x) {
        b0off[e - lo] = (int)(ioff[e] - base);
        if (e == hi - 1) b0off[hi - lo] = (int)(ioff[hi] - base);
        for (int64_t k = ioff[e]; k < ioff[e + 1]; k++) {
            int2 p;
            if (ipk) {
                const uint32_t q = ipk[k];
                const int2 r = rs[e];
                p = make_int2(r.x + (int)(q & 0xffffu), r.y + (int)(q >> 16));
            } else {
                p = ipair[k];
            }
            b0pair[k - base] = p;
        }
    }
}

// sets *flag when ew, c4 or nw hold more than one bit pattern
__global__ void k_check_uniform(const double *ew, const double *c4, int E, const double *nw, int N, int *flag)
{
    const long long e0 = E > 0 ? __double_as_longlong(ew[0]) : 0, c0 = E > 0 ? __double_as_longlong(c4[0]) : 0;
    const long long n0 = N > 0 ? __double_as_longlong(nw[0]) : 0;
    int bad = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < E || i < N; i += gridDim.x * blockDim.x) {
        if (i < E && (__double_as_longlong(ew[i]) != e0 || __double_as_longlong(c4[i]) != c0)) bad = 1;
        if (i < N && __double_as_longlong(nw[i]) != n0) bad = 1;
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// ------------------------------------------------------------- glue kernels

// current buffer from input-order staging: d[i] = stage[ord[i]], and the
// half-edge copies wh = w-bar*d
__global__ void k_gather_in(int E, const int *ord, const double *stage, double *d0, double *d1,
                            const FitState *st)
{
    double *d = st->cur ? d1 : d0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x)
        d[i] = stage[ord[i]];
}

// stage[ord[i]] = d[cur][i] (input order out)
__global__ void k_scatter_out(int E, const int *ord, const double *d, double *stage)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x)
        stage[ord[i]] = d[i];
}

// half-edge copies of the current values
template <typename VT>
__global__ void k_init_wh(int E, const double *ew, const int2 *pos, const double *d0, const double *d1,
                          VT *f0, VT *f1, VT *w0, VT *w1, const FitState *st)
{
    const double *d = st->cur ? d1 : d0;
    VT *f = st->cur ? f1 : f0;
    VT *w = st->cur ? w1 : w0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        double x = d[e];
        if (f) f[e] = (VT)x;
        VT wx = (VT)(ew[e] * x);
        int2 q = pos[e];
        w[q.x] = wx;
        if (q.y >= 0) w[q.y] = wx;
    }
}

// FP32 values -> FP64 buffers once the FP32 phase reached its hand-over
// point; runs only in the iteration where the end kernel flagged it.
__global__ void k_to64(int E, const double *ew, const int2 *pos, double *d0, double *d1, double *w0,
                       double *w1, const float *f0, const float *f1, FitState *st)
{
    if (st->done || st->fp32 != 2) return;
    const float *f = st->cur ? f1 : f0;
    double *d = st->cur ? d1 : d0;
    double *w = st->cur ? w1 : w0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        double x = (double)f[e];
        d[e] = x;
        double wx = ew[e] * x;
        int2 q = pos[e];
        w[q.x] = wx;
        if (q.y >= 0) w[q.y] = wx;
    }
}

__global__ void k_switch_done(FitState *st)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && st->fp32 == 2) st->fp32 = 0;
}

// convergence bookkeeping for the iteration just finished
__global__ void k_end(FitState *st, double eps, int max_it)
{
    if (threadIdx.x != 0 || blockIdx.x != 0 || st->done) return;
    double md = __longlong_as_double((long long)st->maxd);
    st->maxd = 0ull;
    st->prev_delta = st->delta;
    st->delta = md;
    st->cur ^= 1;
    int it = st->iter;
    st->iter = it + 1;
    if (st->fp32) {
        // FP32 phase: never decide convergence on FP32 values; hand over to
        // FP64 while enough iterations remain to contract the FP32 error
        if (md < st->switch_thr || it + 1 >= st->switch_by) {
            st->fp32 = 2;
            st->switched_at = it + 1;
        }
        if (it + 1 >= max_it) { st->done = 1; st->iterations = max_it; st->ambiguous = 1; }
        return;
    }
    if (st->mixed && fabs(md - eps) <= 1e-3 * md) st->ambiguous = 1;
    if (md < eps) { st->done = 1; st->iterations = it; }
    else if (it + 1 >= max_it) { st->done = 1; st->iterations = max_it; }
}

inline int grid_for(int64_t threads, int cap)
{
    int64_t b = (threads + TPB - 1) / TPB;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

template <int G, typename VT, bool UNI>
void launch_vertex_t(DevGraph *Gr, int bin, int cap, VT *w0, VT *w1)
{
    const Bins &B = Gr->vbins;
    int n = B.count[bin];
    if (!n) return;
    int grid = grid_for((int64_t)n * G, cap);
    hipLaunchKernelGGL((k_vertex<G, VT, UNI>), dim3(grid), dim3(TPB), 0, Gr->stream, B.list + B.start[bin], n,
                       Gr->row_off, w0, w1, Gr->nw, Gr->nw_c, Gr->vd, (const FitState *)Gr->state);
}

// all vertex bins in one launch (the block counts per bin the separate
// launches would use)
template <typename VT, bool UNI>
void launch_vertex_all_t(DevGraph *Gr, VT *w0, VT *w1)
{
    const Bins &B = Gr->vbins;
    const int cap = num_cus() * 32;
    static const int lanes[Bins::NB] = {1, 4, 16, 64, TPB};
    VertexBins vb;
    vb.first[0] = 0;
    for (int j = 0; j < Bins::NB; j++) {
        const int bin = Bins::NB - 1 - j;      // launch order: widest rows first
        vb.list[bin] = B.list + B.start[bin];
        vb.count[bin] = B.count[bin];
        vb.first[j + 1] = vb.first[j] + (B.count[bin] ? grid_for((int64_t)B.count[bin] * lanes[bin], cap) : 0);
    }
    if (vb.first[Bins::NB] == 0) return;
    hipLaunchKernelGGL((k_vertex_all<VT, UNI>), dim3(vb.first[Bins::NB]), dim3(TPB), 0, Gr->stream, vb, Gr->row_off,
                       w0, w1, Gr->nw, Gr->nw_c, Gr->vd, (const FitState *)Gr->state);
}

// workgroups of a kernel resident on the whole device at once
template <typename K>
int resident_blocks(K kernel)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, TPB, 0) != hipSuccess || nb < 1) {
        (void)hipGetLastError();
        nb = 4;
    }
    return nb * num_cus();
}

template <int G, bool PK, typename VT, bool UNI>
int edge_capacity()
{
    static const int cap = resident_blocks(k_edge<G, PK, VT, UNI>);
    return cap;
}

template <typename VT, bool UNI>
int edge0_capacity()
{
    static const int cap = resident_blocks(k_edge0<VT, UNI>);
    return cap;
}

EdgeArgs edge_args(DevGraph *Gr)
{
    EdgeArgs a;
    a.uv = Gr->uv; a.rs = Gr->rs; a.pos = Gr->pos; a.ew = Gr->ew; a.c4 = Gr->c4; a.ioff = Gr->ioff;
    a.ipk = Gr->ipk; a.ipair = Gr->ipair; a.vd = Gr->vd;
    a.b0off = Gr->b0off; a.b0pair = Gr->b0pair;
    a.ew_c = Gr->ew_c; a.c4_c = Gr->c4_c;
    a.directed = (Gr->variant == DRESS_VARIANT_FORWARD || Gr->variant == DRESS_VARIANT_BACKWARD);
    return a;
}

// Workgroups per CU of each edge bin's launch.  Fewer than fit (8) keep
// fewer rows competing for each CU's L1 and each XCD's L2: on the 4M-vertex
// graph 4 per CU took bins 0, 1 and 3 from 1.053, 1.015, 0.958 to 1.024,
// 0.969, 0.943 ms and 3 took bin 2 from 1.723 to 1.546 ms (2 and 1 were
// slower everywhere), the same values throughout.
// The FP32 kernels of a mixed fit have caps of their own.
// DRESS_HIP_EDGE_BLOCKS="b0,b1,b2,b3,b4" and DRESS_HIP_EDGE_BLOCKS32 (FP32)
// override them (1..8 each; a tuning knob: results do not depend on it).
const int *edge_blocks_per_cu(bool fp32)
{
    static int blocks[2][Bins::NB] = {{4, 4, 3, 4, 4}, {4, 4, 4, 4, 4}};
    static bool once = [] {
        const char *names[2] = {"DRESS_HIP_EDGE_BLOCKS", "DRESS_HIP_EDGE_BLOCKS32"};
        for (int p = 0; p < 2; p++) {
            if (const char *e = getenv(names[p])) {
                int v[Bins::NB];
                if (sscanf(e, "%d,%d,%d,%d,%d", &v[0], &v[1], &v[2], &v[3], &v[4]) == Bins::NB)
                    for (int i = 0; i < Bins::NB; i++)
                        if (v[i] >= 1 && v[i] <= 8) blocks[p][i] = v[i];
            }
        }
        return true;
    }();
    (void)once;
    return blocks[fp32 ? 1 : 0];
}

// grid of an edge launch of bin `bin` over n edges with g lanes each (at most
// `capacity` workgroups, what fits on the device at once), and its XCD slices
inline void edge_grid(int bin, bool fp32, int64_t n, int g, int capacity, int *grid, int *nx)
{
    capacity = std::min(capacity, edge_blocks_per_cu(fp32)[bin] * num_cus());
    *grid = grid_for(n * g, capacity);
    *nx = num_xcds();
    if (*grid >= 4 * *nx) *grid -= *grid % *nx;
    else *nx = 1;
}

template <int G, bool PK, typename VT, bool UNI>
void launch_edge_t(DevGraph *Gr, int bin, VT *d0, VT *d1, VT *w0, VT *w1)
{
    int lo = Gr->ebin[bin], hi = Gr->ebin[bin + 1];
    if (hi <= lo) return;
    int grid, nx;
    if (G == 1 && Gr->b0off) {
        edge_grid(bin, sizeof(VT) == 4, hi - lo, 1, edge0_capacity<VT, UNI>(), &grid, &nx);
        hipLaunchKernelGGL((k_edge0<VT, UNI>), dim3(grid), dim3(TPB), 0, Gr->stream, lo, hi, nx, edge_args(Gr), d0, d1,
                           w0, w1, (FitState *)Gr->state);
        return;
    }
    edge_grid(bin, sizeof(VT) == 4, hi - lo, G, edge_capacity<G, PK, VT, UNI>(), &grid, &nx);
    hipLaunchKernelGGL((k_edge<G, PK, VT, UNI>), dim3(grid), dim3(TPB), 0, Gr->stream, lo, hi, nx, edge_args(Gr), d0,
                       d1, w0, w1, (FitState *)Gr->state);
}

// DRESS_HIP_PHASE_SPLIT=1 launches the vertex bins of an iteration one by
// one instead of in one kernel (0.19 against 0.14 ms on the 4M-vertex graph)
bool phases_split()
{
    static const bool split = [] {
        const char *e = getenv("DRESS_HIP_PHASE_SPLIT");
        return e && !strcmp(e, "1");
    }();
    return split;
}

// lanes per edge of edge bins 1-3 (4, 8, 32 by default: on the 4M-vertex
// graph 139.8 ms per fit against 141.1 with 4, 16, 64; bin 0 takes one lane
// per edge, bin 4 a workgroup).  DRESS_HIP_EDGE_LANES="l1,l2,l3"
// overrides them with powers of two from 2 to 64 (a tuning knob: results
// do not depend on it).
const int *edge_lanes()
{
    static int lanes[3] = {4, 8, 32};
    static bool once = [] {
        if (const char *e = getenv("DRESS_HIP_EDGE_LANES")) {
            int v[3];
            if (sscanf(e, "%d,%d,%d", &v[0], &v[1], &v[2]) == 3)
                for (int i = 0; i < 3; i++)
                    if (v[i] >= 2 && v[i] <= 64 && (v[i] & (v[i] - 1)) == 0) lanes[i] = v[i];
        }
        return true;
    }();
    (void)once;
    return lanes;
}

template <bool PK, typename VT, bool UNI>
void launch_edge_lanes(DevGraph *Gr, int bin, int lanes, VT *d0, VT *d1, VT *w0, VT *w1)
{
    switch (lanes) {
    case 2: launch_edge_t<2, PK, VT, UNI>(Gr, bin, d0, d1, w0, w1); break;
    case 4: launch_edge_t<4, PK, VT, UNI>(Gr, bin, d0, d1, w0, w1); break;
    case 8: launch_edge_t<8, PK, VT, UNI>(Gr, bin, d0, d1, w0, w1); break;
    case 16: launch_edge_t<16, PK, VT, UNI>(Gr, bin, d0, d1, w0, w1); break;
    case 32: launch_edge_t<32, PK, VT, UNI>(Gr, bin, d0, d1, w0, w1); break;
    default: launch_edge_t<64, PK, VT, UNI>(Gr, bin, d0, d1, w0, w1); break;
    }
}

// Kernel b of an iteration: b < NB vertex bin b, b < 2 NB edge bin b - NB.
// Unless `split`, the vertex phase is one launch over every bin (b = 0) and
// b = 1..4 are empty.  Edge bins are always launched one by one: the edge
// phase as one launch (each block through its share of every bin in turn)
// took 5.87 ms per iteration on the 4M-vertex graph against 5.10 ms for the
// five launches, whose boundaries keep each bin's rows together in L2.
template <bool PK, typename VT, bool UNI>
void launch_one_t(DevGraph *Gr, int b, bool split, VT *d0, VT *d1, VT *w0, VT *w1)
{
    const int cap = num_cus() * 32;
    switch (b) {
    case 0:
        if (split) launch_vertex_t<1, VT, UNI>(Gr, 0, cap, w0, w1);
        else launch_vertex_all_t<VT, UNI>(Gr, w0, w1);
        break;
    case 1: if (split) launch_vertex_t<4, VT, UNI>(Gr, 1, cap, w0, w1); break;
    case 2: if (split) launch_vertex_t<16, VT, UNI>(Gr, 2, cap, w0, w1); break;
    case 3: if (split) launch_vertex_t<64, VT, UNI>(Gr, 3, cap, w0, w1); break;
    case 4: if (split) launch_vertex_t<TPB, VT, UNI>(Gr, 4, cap, w0, w1); break;
    case 5: launch_edge_t<1, PK, VT, UNI>(Gr, 0, d0, d1, w0, w1); break;
    case 6: case 7: case 8: launch_edge_lanes<PK, VT, UNI>(Gr, b - 5, edge_lanes()[b - 6], d0, d1, w0, w1); break;
    default: launch_edge_t<TPB, PK, VT, UNI>(Gr, 4, d0, d1, w0, w1); break;
    }
}

// kernel b of an iteration in the graph's precision form
void launch_one(DevGraph *Gr, int b, bool fp32, bool split)
{
    const bool uni = Gr->uni_state == 1;
    if (fp32) {
        float *d0 = Gr->d32[0], *d1 = Gr->d32[1], *w0 = Gr->wh32[0], *w1 = Gr->wh32[1];
        if (Gr->packed) {
            if (uni) launch_one_t<true, float, true>(Gr, b, split, d0, d1, w0, w1);
            else launch_one_t<true, float, false>(Gr, b, split, d0, d1, w0, w1);
        } else {
            if (uni) launch_one_t<false, float, true>(Gr, b, split, d0, d1, w0, w1);
            else launch_one_t<false, float, false>(Gr, b, split, d0, d1, w0, w1);
        }
    } else {
        double *d0 = Gr->d[0], *d1 = Gr->d[1], *w0 = Gr->wh[0], *w1 = Gr->wh[1];
        if (Gr->packed) {
            if (uni) launch_one_t<true, double, true>(Gr, b, split, d0, d1, w0, w1);
            else launch_one_t<true, double, false>(Gr, b, split, d0, d1, w0, w1);
        } else {
            if (uni) launch_one_t<false, double, true>(Gr, b, split, d0, d1, w0, w1);
            else launch_one_t<false, double, false>(Gr, b, split, d0, d1, w0, w1);
        }
    }
}

template <bool PK, typename VT, bool UNI>
void prepare_t()
{
    (void)edge_capacity<1, PK, VT, UNI>(); (void)edge_capacity<2, PK, VT, UNI>();
    (void)edge_capacity<4, PK, VT, UNI>(); (void)edge_capacity<8, PK, VT, UNI>();
    (void)edge_capacity<16, PK, VT, UNI>(); (void)edge_capacity<32, PK, VT, UNI>();
    (void)edge_capacity<64, PK, VT, UNI>();
    (void)edge_capacity<TPB, PK, VT, UNI>(); (void)edge0_capacity<VT, UNI>();
}

// resolves the occupancy of every edge kernel the iteration launches (a
// one-time query that must not happen inside a stream capture)
void launch_phase_prepare(DevGraph *Gr)
{
    (void)num_xcds();
    (void)num_cus();
    if (Gr->packed) {
        prepare_t<true, double, true>(); prepare_t<true, double, false>();
        prepare_t<true, float, true>(); prepare_t<true, float, false>();
    } else {
        prepare_t<false, double, true>(); prepare_t<false, double, false>();
        prepare_t<false, float, true>(); prepare_t<false, float, false>();
    }
}

void launch_phase(DevGraph *Gr, bool fp32)
{
    const bool split = phases_split();
    for (int b = 0; b < 2 * Bins::NB; b++) launch_one(Gr, b, fp32, split);
}

}  // namespace

// Are ew, c4 and nw uniform (G->uni_state 1) or not (2)?  Decided on the
// device when unknown; the kernels then take the three constants as
// arguments.
bool settle_uniform(DevGraph *G)
{
    if (G->uni_state != 0) return true;
    hipStream_t s = G->stream;
    int *flag = nullptr;
    if (!dalloc(&flag, 1, "uniformity flag")) return false;
    int bad = 1;
    double c[3] = {0.0, 0.0, 0.0};
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, flag, 1, 0);
    const int64_t n = std::max<int64_t>(G->E, G->N);
    if (n > 0)
        hipLaunchKernelGGL(k_check_uniform, dim3(grid_for(n, num_cus() * 8)), dim3(TPB), 0, s, (const double *)G->ew,
                           (const double *)G->c4, G->E, (const double *)G->nw, G->N, flag);
    bool ok = hipGetLastError() == hipSuccess &&
              hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, s) == hipSuccess &&
              (G->E == 0 || (hipMemcpyAsync(&c[0], G->ew, sizeof(double), hipMemcpyDeviceToHost, s) == hipSuccess &&
                             hipMemcpyAsync(&c[1], G->c4, sizeof(double), hipMemcpyDeviceToHost, s) == hipSuccess)) &&
              (G->N == 0 || hipMemcpyAsync(&c[2], G->nw, sizeof(double), hipMemcpyDeviceToHost, s) == hipSuccess) &&
              hipStreamSynchronize(s) == hipSuccess;
    dfree(flag);
    if (!ok) {
        set_error("uniformity check failed");
        return false;
    }
    G->uni_state = bad ? 2 : 1;
    G->ew_c = c[0];
    G->c4_c = c[1];
    G->nw_c = c[2];
    return true;
}

bool fit_alloc_state(DevGraph *G, bool fp32)
{
    if (!build_bins(G)) return false;
    if (!G->state && !dmalloc(&G->state, sizeof(FitState), "fit state")) return false;
    if (!G->stage && !dalloc(&G->stage, (size_t)G->E + 1, "staging")) return false;
    for (int i = 0; i < 2; i++) {
        if (!G->d[i] && !dalloc(&G->d[i], (size_t)G->E + 1, "edge values")) return false;
        if (!G->wh[i] && !dalloc(&G->wh[i], (size_t)G->S + 1, "half-edge values")) return false;
    }
    if (!G->vd && !dalloc(&G->vd, (size_t)G->N + 1, "vertex norms")) return false;
    if (!G->b0off && G->ebin[1] > G->ebin[0]) {
        // bin 0's index as absolute pairs behind 32-bit offsets (its
        // entries are at most 3 per edge, so the bin holds far fewer than
        // 2^31 of them)
        const int lo = G->ebin[0], hi = G->ebin[1];
        int64_t t[2] = {0, 0};
        DH_CHECK(hipMemcpyAsync(&t[0], G->ioff + lo, sizeof(int64_t), hipMemcpyDeviceToHost, G->stream));
        DH_CHECK(hipMemcpyAsync(&t[1], G->ioff + hi, sizeof(int64_t), hipMemcpyDeviceToHost, G->stream));
        DH_CHECK(hipStreamSynchronize(G->stream));
        const int64_t t0 = t[1] - t[0];
        if (t0 < (int64_t)INT32_MAX) {
            if (!dalloc(&G->b0off, (size_t)(hi - lo) + 1, "bin-0 offsets") ||
                !dalloc(&G->b0pair, (size_t)t0 + 1, "bin-0 entries")) {
                dfree(G->b0off);
                G->b0off = nullptr;
                return false;
            }
            hipLaunchKernelGGL(k_pack_bin0, dim3(grid_for(hi - lo, num_cus() * 8)), dim3(TPB), 0, G->stream, lo, hi,
                               (const int64_t *)G->ioff, (const int2 *)G->rs, (const uint32_t *)G->ipk,
                               (const int2 *)G->ipair, G->b0off, G->b0pair);
            DH_CHECK(hipGetLastError());
        }
    }
    if (fp32) {
        for (int i = 0; i < 2; i++) {
            if (!G->d32[i] && !dalloc(&G->d32[i], (size_t)G->E + 1, "edge values fp32")) return false;
            if (!G->wh32[i] && !dalloc(&G->wh32[i], (size_t)G->S + 1, "half-edge values fp32")) return false;
        }
    }
    return true;
}

// host edge values (input order) -> current device buffer (internal order)
bool fit_upload_values(DevGraph *G, const double *edge_dress)
{
    G->fit_state_valid = false;
    G->wh_ready = false;
    if (!fit_alloc_state(G, false)) return false;
    G->cur = 0;
    if (G->E == 0) return true;
    hipStream_t s = G->stream;
    // the stage may still be read by earlier work on s
    DH_CHECK(hipStreamSynchronize(s));
    if (!h2d_parallel(G->stage, edge_dress, (size_t)G->E * sizeof(double))) {
        set_error("edge value upload failed");
        return false;
    }
    FitState h;
    memset(&h, 0, sizeof(h));
    h.cur = 0;
    DH_CHECK(hipMemcpyAsync(G->state, &h, sizeof(h), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_gather_in, dim3(grid_for(G->E, num_cus() * 8)), dim3(TPB), 0, s, G->E, G->ord, G->stage,
                       G->d[0], G->d[1], (const FitState *)G->state);
    DH_CHECK(hipGetLastError());
    DH_CHECK(hipStreamSynchronize(s));
    return true;
}

namespace {

// the kernels of one iteration on stream s (state-driven: the same launches
// serve every iteration)
void launch_iteration(DevGraph *G, bool mixed, double epsilon, int max_iterations, int egrid, hipStream_t s)
{
    FitState *st = (FitState *)G->state;
    if (mixed) launch_phase(G, true);
    launch_phase(G, false);
    hipLaunchKernelGGL(k_end, dim3(1), dim3(64), 0, s, st, epsilon, max_iterations);
    if (mixed) {
        hipLaunchKernelGGL(k_to64, dim3(egrid), dim3(TPB), 0, s, G->E, G->ew, G->pos, G->d[0], G->d[1], G->wh[0],
                           G->wh[1], G->d32[0], G->d32[1], st);
        hipLaunchKernelGGL(k_switch_done, dim3(1), dim3(64), 0, s, st);
    }
}

// With DRESS_HIP_GRAPH=1, one iteration is captured as a HIP graph (kernel
// nodes only) and kept with the device graph, so each iteration is one
// graph launch instead of one launch per kernel.  The captured arguments are
// the graph's buffers, epsilon and max_iterations; a fit with other values
// recaptures.  Off by default: on MI355X the replay saves nothing
// measurable for fits of 10^4..10^5 edges (the time is the kernels' device
// dispatch, not the host launch), and the capture and instantiation cost
// of a one-shot fit is a loss under many concurrent requests.
hipGraphExec_t iteration_graph(DevGraph *G, bool mixed, double epsilon, int max_iterations, int egrid)
{
    const char *env = getenv("DRESS_HIP_GRAPH");
    if (!(env && !strcmp(env, "1"))) return nullptr;
    const int m = mixed ? 1 : 0;
    if (G->iter_graph[m] && G->graph_eps[m] == epsilon && G->graph_maxit[m] == max_iterations &&
        G->graph_uni[m] == G->uni_state)
        return G->iter_graph[m];
    if (G->iter_graph[m]) {
        hipGraphExecDestroy(G->iter_graph[m]);
        G->iter_graph[m] = nullptr;
    }
    // occupancy queries and device attributes are settled before capturing
    launch_phase_prepare(G);
    hipStream_t s = G->stream;
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        hipGetLastError();
        return nullptr;
    }
    launch_iteration(G, mixed, epsilon, max_iterations, egrid, s);
    hipError_t e = hipStreamEndCapture(s, &graph);
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess && graph) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) hipGraphDestroy(graph);
    if (e != hipSuccess || !exec) {
        hipGetLastError();
        return nullptr;                     // fall back to eager launches
    }
    G->iter_graph[m] = exec;
    G->graph_eps[m] = epsilon;
    G->graph_maxit[m] = max_iterations;
    G->graph_uni[m] = G->uni_state;
    return exec;
}

bool fit_run_once(DevGraph *G, int max_iterations, double epsilon, bool mixed, FitResult *res,
                  bool *ambiguous)
{
    hipStream_t s = G->stream;
    if (!fit_alloc_state(G, mixed)) return false;
    FitState h;
    memset(&h, 0, sizeof(h));
    h.cur = G->cur;
    h.fp32 = mixed ? 1 : 0;
    h.mixed = mixed ? 1 : 0;
    // The FP32 phase carries a relative error of ~1.5e-7 (u32 / (1 - rho)).
    // The FP64 phase contracts that error and the change at the same rate,
    // so handing over at change ~ 300 * epsilon leaves the error below
    // 1.5e-7 * epsilon / (300 epsilon) = 5e-10 when the change reaches
    // epsilon, within the 1e-9 tolerance.  On the 4M-vertex graph the
    // measured max edge error against FP64 was 2.2e-13, 5.9e-13, 4.4e-12,
    // 3.0e-11 and 1.9e-10 for hand-over at 3000, 1000, 300, 100 and 30
    // epsilon, with 13, 14, 16, 18 and 20 of 26 iterations in FP32.  At
    // least 16 FP64 iterations are kept before max_iterations.
    h.switch_thr = 300.0 * epsilon;
    h.switch_by = max_iterations - 16;
    FitState *st = (FitState *)G->state;
    DH_CHECK(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, s));
    hipEvent_t ev0, ev1;
    DH_CHECK(hipEventCreate(&ev0));
    DH_CHECK(hipEventCreate(&ev1));
    int egrid = grid_for(G->E, num_cus() * 8);
    DH_CHECK(hipEventRecord(ev0, s));
    // half-edge copies prepared by a fill of the current values (FP64 only:
    // a mixed fit starts from its FP32 copies)
    const bool wh_ready = G->wh_ready && !mixed && G->cur == 0 && G->uni_state == 1;
    G->wh_ready = false;
    if (G->E > 0 && !wh_ready) {
        if (mixed)
            hipLaunchKernelGGL((k_init_wh<float>), dim3(egrid), dim3(TPB), 0, s, G->E, G->ew, G->pos, G->d[0],
                               G->d[1], G->d32[0], G->d32[1], G->wh32[0], G->wh32[1], st);
        else
            hipLaunchKernelGGL((k_init_wh<double>), dim3(egrid), dim3(TPB), 0, s, G->E, G->ew, G->pos, G->d[0],
                               G->d[1], (double *)nullptr, (double *)nullptr, G->wh[0], G->wh[1], st);
    }
    hipGraphExec_t ig = iteration_graph(G, mixed, epsilon, max_iterations, egrid);
    // Iterations go in chunks, each followed by a copy of the state into a
    // pinned snapshot slot.  The next chunk is launched before the last
    // snapshot is read, so the device does not wait for the host between
    // chunks; iterations launched past convergence return at once.
    FitState *snap = (FitState *)snapshot_take();
    if (!snap) return false;
    // back to the pool once no copy into it can be in flight
    struct Give {
        void *p;
        hipStream_t s;
        ~Give()
        {
            if (p) (void)hipStreamSynchronize(s);
            snapshot_give(p);
        }
    } give{snap, s};
    hipEvent_t sev[2];
    DH_CHECK(hipEventCreateWithFlags(&sev[0], hipEventDisableTiming));
    DH_CHECK(hipEventCreateWithFlags(&sev[1], hipEventDisableTiming));
    int launched = 0;
    auto issue = [&](int n, int slot) {
        for (int j = 0; j < n; j++) {
            if (ig) (void)hipGraphLaunch(ig, s);
            else launch_iteration(G, mixed, epsilon, max_iterations, egrid, s);
        }
        launched += n;
        return hipGetLastError() == hipSuccess &&
               hipMemcpyAsync(&snap[slot], st, sizeof(FitState), hipMemcpyDeviceToHost, s) == hipSuccess &&
               hipEventRecord(sev[slot], s) == hipSuccess;
    };
    bool ok = max_iterations <= 0 || issue(std::min(4, max_iterations), 0);
    int pend = 0, chunk = 4;                 // the second chunk goes in blind
    while (ok && launched > 0) {
        int nxt = -1, n = 0;
        if (launched < max_iterations) {
            n = std::min(chunk, max_iterations - launched);
            ok = issue(n, pend ^ 1);
            nxt = pend ^ 1;
        }
        ok = ok && hipEventSynchronize(sev[pend]) == hipSuccess;
        if (!ok) break;
        h = snap[pend];
        if (h.done || nxt < 0) break;
        // the chunk after the one in flight: what a geometric decay of the
        // change still needs to reach epsilon, so launches past convergence
        // stay few
        chunk = 16;
        if (h.prev_delta > 0.0 && h.delta > 0.0 && h.delta < h.prev_delta && epsilon > 0.0) {
            double r = h.delta / h.prev_delta;
            double need = std::ceil(std::log(epsilon / h.delta) / std::log(r)) - n;
            if (need >= 1.0 && need < 16.0) chunk = (int)need;
            else if (need < 1.0) chunk = 1;
        }
        pend = nxt;
    }
    DH_CHECK(hipEventRecord(ev1, s));
    DH_CHECK(hipEventSynchronize(ev1));
    hipEventDestroy(sev[0]);
    hipEventDestroy(sev[1]);
    float ms = 0.f;
    hipEventElapsedTime(&ms, ev0, ev1);
    hipEventDestroy(ev0);
    hipEventDestroy(ev1);
    if (!ok) { set_error("fit iteration launch failed"); return false; }
    // the state after everything launched (the chunk in flight at the stop
    // changed nothing)
    DH_CHECK(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
    DH_CHECK(hipStreamSynchronize(s));
    if (!h.done) { set_error("fit did not finish within the launched iterations"); return false; }
    G->cur = h.cur;
    res->iterations = h.iterations;
    res->delta = h.delta;
    res->executed = h.iter;
    res->fp32_iterations = mixed ? h.switched_at : 0;
    res->ms += ms;
    *ambiguous = h.ambiguous != 0;
    return true;
}

__global__ void k_copy_f64(const double *a, double *b, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

}  // namespace

bool fit_run(DevGraph *G, int max_iterations, double epsilon, int precision, FitResult *res)
{
    G->fit_state_valid = false;
    bool mixed = (precision == PREC_FP32_FP64);
    res->ms = 0.f;
    res->reruns = 0;
    if (!build_bins(G) || !settle_uniform(G)) return false;
    if (max_iterations <= 0) {
        res->iterations = max_iterations;
        res->executed = 0;
        return true;
    }
    // mixed mode keeps the start values so an ambiguous stop can be redone
    double *start = nullptr;
    if (mixed) {
        if (!fit_alloc_state(G, true) || !dalloc(&start, (size_t)G->E + 1, "start values")) return false;
        if (G->E > 0)
            hipLaunchKernelGGL(k_copy_f64, dim3(grid_for(G->E, num_cus() * 8)), dim3(TPB), 0, G->stream,
                               G->d[G->cur], start, (int64_t)G->E);
    }
    bool amb = false;
    int start_cur = G->cur;
    bool ok = fit_run_once(G, max_iterations, epsilon, mixed, res, &amb);
    if (ok && mixed && amb) {
        // redo in FP64 from the same start: bitwise the reference trajectory
        G->cur = start_cur;
        if (G->E > 0)
            hipLaunchKernelGGL(k_copy_f64, dim3(grid_for(G->E, num_cus() * 8)), dim3(TPB), 0, G->stream,
                               start, G->d[start_cur], (int64_t)G->E);
        res->reruns = 1;
        ok = fit_run_once(G, max_iterations, epsilon, false, res, &amb);
        res->fp32_iterations = 0;
    }
    dfree(start);
    return ok;
}

namespace {
__global__ void k_fill_vals(double *d, int64_t n, double v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = v;
}
}  // namespace

bool fit_fill_values(DevGraph *G, double value, bool prepare_wh)
{
    G->wh_ready = false;
    if (!fit_alloc_state(G, false)) return false;
    G->cur = 0;
    if (G->E > 0)
        hipLaunchKernelGGL(k_fill_vals, dim3(grid_for(G->E, num_cus() * 8)), dim3(TPB), 0, G->stream, G->d[0],
                           (int64_t)G->E, value);
    if (prepare_wh && G->uni_state == 1 && G->E > 0) {
        // every half-edge copy is w-bar * value, the product k_init_wh forms
        hipLaunchKernelGGL(k_fill_vals, dim3(grid_for(G->S, num_cus() * 8)), dim3(TPB), 0, G->stream, G->wh[0],
                           (int64_t)G->S, G->ew_c * value);
        G->wh_ready = true;
    }
    DH_CHECK(hipGetLastError());
    DH_CHECK(hipStreamSynchronize(G->stream));
    return true;
}

// Average device time of each phase kernel of one FP64 iteration over reps
// iterations, bracketed by events on the graph's stream.  ms[0..4] vertex
// bins, ms[5..9] edge bins, each launched alone (0 where the bin is empty);
// ms[10] the vertex phase and ms[11] the edge phase as the fit launches
// them (one kernel for the vertex phase, one per edge bin).
bool fit_time_phases(DevGraph *G, int reps, double *ms)
{
    if (!fit_alloc_state(G, false) || !settle_uniform(G)) return false;
    hipStream_t s = G->stream;
    FitState h;
    memset(&h, 0, sizeof(h));
    h.cur = G->cur;
    DH_CHECK(hipMemcpyAsync(G->state, &h, sizeof(h), hipMemcpyHostToDevice, s));
    if (G->E > 0)
        hipLaunchKernelGGL((k_init_wh<double>), dim3(grid_for(G->E, num_cus() * 8)), dim3(TPB), 0, s, G->E, G->ew,
                           G->pos, G->d[0], G->d[1], (double *)nullptr, (double *)nullptr, G->wh[0], G->wh[1],
                           (const FitState *)G->state);
    std::vector<hipEvent_t> ev(24);
    for (auto &e : ev) DH_CHECK(hipEventCreate(&e));
    for (int i = 0; i < 12; i++) ms[i] = 0.0;
    for (int r = 0; r < reps; r++) {
        for (int b = 0; b < 2 * Bins::NB; b++) {
            DH_CHECK(hipEventRecord(ev[2 * b], s));
            launch_one(G, b, false, true);
            DH_CHECK(hipEventRecord(ev[2 * b + 1], s));
        }
        // the vertex phase in one launch, then the edge phase as fits launch it
        DH_CHECK(hipEventRecord(ev[20], s));
        launch_one(G, 0, false, false);
        DH_CHECK(hipEventRecord(ev[21], s));
        for (int b = Bins::NB; b < 2 * Bins::NB; b++) launch_one(G, b, false, false);
        DH_CHECK(hipEventRecord(ev[22], s));
        DH_CHECK(hipEventSynchronize(ev[22]));
        for (int i = 0; i < 12; i++) {
            if (i >= 10) {
                float t = 0.f;
                DH_CHECK(hipEventElapsedTime(&t, ev[20 + i - 10], ev[21 + i - 10]));
                ms[i] += t;
                continue;
            }
            float t = 0.f;
            DH_CHECK(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
            ms[i] += t;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    for (int i = 0; i < 12; i++) ms[i] /= std::max(reps, 1);
    // the timing iterations only wrote the other buffer: the current values
    // (buffer G->cur) are unchanged
    return true;
}

bool fit_download(DevGraph *G, double *edge_dress, double *vertex_dress)
{
    hipStream_t s = G->stream;
    if (edge_dress && G->E > 0) {
        hipLaunchKernelGGL(k_scatter_out, dim3(grid_for(G->E, num_cus() * 8)), dim3(TPB), 0, s, G->E, G->ord,
                           G->d[G->cur], G->stage);
        DH_CHECK(hipStreamSynchronize(s));
        if (!d2h_parallel(edge_dress, G->stage, (size_t)G->E * sizeof(double))) {
            set_error("edge value download failed");
            return false;
        }
    }
    DH_CHECK(hipStreamSynchronize(s));
    if (vertex_dress && G->N > 0 && !download_vertex_values(G, G->vd, vertex_dress)) return false;
    return true;
}

}  // namespace dh
