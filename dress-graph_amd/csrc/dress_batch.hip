// dress_batch.hip -- delta^k / nabla^k subgraph fits in one persistent kernel.
//
// The reference fits every subgraph separately: it rebuilds the subgraph on
// the host (libdress/src/delta_dress_impl.c:95-169, nabla_dress_impl.c:61-91)
// and calls the fit function on it (delta_dress_impl.c:349,
// nabla_dress_impl.c:280).  Here a subgraph is its parent graph plus an item
// descriptor:
//   delta^k  the k deleted vertices: an edge is live when neither end is
//            deleted; vertex weights are all 1.0 (delta_dress_impl.c:166-168
//            passes NW = NULL), edge weights are the parent's;
//   nabla^k  the ordered k-tuple: vertex weights are 1.0 except tuple[i],
//            which gets marker[i] (nabla_dress_impl.c:82-86).
// Because relabelling the surviving vertices keeps every row's order and the
// intersection of two rows in the subgraph is the parent's intersection
// without the deleted neighbours, the live terms of every sum are exactly the
// subgraph's terms; sums are rounded once from their exact value, so each
// subgraph value equals the one the reference computes on the rebuilt graph.
//
// One workgroup runs one item at a time, taking items from a queue sorted by
// decreasing size.  The item's state (edge values, double buffered, and vertex
// norms) lives in LDS; the parent's structure is read through the cache.
// Mixed precision: FP32 edge values, terms summed exactly in FP64, until the
// change is below 3000*epsilon, then FP64 with compensated sums; an item
// whose stopping decision falls within 0.1% of epsilon is flagged and redone
// in FP64 by the host.
#include "dress_batch.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace dh {

namespace {

__device__ __forceinline__ void two_sum_add(double &s, double &c, double x)
{
    double t = s + x;
    double z = t - s;
    double e = (s - (t - z)) + (x - z);
    s = t;
    c += e;
}

// vertex weight of u for the current item
__device__ __forceinline__ double item_nw(int kind, int u, const int *tup, const double *mark, int k)
{
    if (kind == LIFT_NABLA) {
        for (int i = 0; i < k; i++)
            if (tup[i] == u) return mark[i];
    }
    return 1.0;
}

template <int B>
__device__ __forceinline__ double block_max(double m, double *red)
{
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if (B > 64) {
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
        __syncthreads();
        m = red[0];
        for (int i = 1; i < B / 64; i++) m = fmax(m, red[i]);
    }
    return m;
}

// One persistent kernel over all items.  LDS layout (dynamic):
//   double dA[ecap], dB[ecap], vd[ncap]; int alive[ncap] (bytes as int)
// FP32 phase uses the first halves of dA / dB as float arrays.  With ST the
// parent's structure follows, copied in once per parent a workgroup meets:
//   int row[ncap + 2], ioff[ecap + 2]; ushort2 uv[ecap]; ushort adj[scap]
// (edge and vertex ids below 2^16: the batched kernel takes E <= 65535,
// and N is bounded by the LDS far below that), so each iteration reads the
// structure from LDS instead of through L1, whose accesses bounded the
// kernel.  Edge weights that differ per edge and the index pairs stay in
// global memory.
template <int B, bool ST>
__global__ void __launch_bounds__(B) k_lifted(LiftArgs a)
{
    extern __shared__ double smem[];
    double *dA = smem;
    double *dB = dA + a.ecap;
    double *vd = dB + a.ecap;
    int *alive = (int *)(vd + a.ncap);
    int *s_row = alive + a.ncap;
    int *s_ioff = s_row + a.ncap + 2;
    ushort2 *s_uv = (ushort2 *)(s_ioff + a.ecap + 2);
    unsigned short *s_adj = (unsigned short *)(s_uv + a.ecap);
    __shared__ int s_staged;               // parent whose structure is in LDS
    if (ST && threadIdx.x == 0) s_staged = -1;
    __shared__ int s_item;
    __shared__ double s_red[B / 64];
    __shared__ int s_tup[LIFT_KMAX];
    __shared__ double s_mark[LIFT_KMAX];

    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(a.queue, 1);
        __syncthreads();
        const int qi = s_item;
        __syncthreads();
        if (qi >= a.n_items) break;
        const int it = a.order ? a.order[qi] : qi;
        const LiftItem I = a.items[it];
        const LiftParent P = a.parents[I.parent];
        const int N = P.N, E = P.E;
        const int k = I.k;
        const int *__restrict__ row_off = a.row_off + P.row_ofs;
        const int *__restrict__ adj_e = a.adj_e + P.adj_ofs;
        const int2 *__restrict__ uv = a.uv + P.edge_ofs;
        const double *__restrict__ ew = a.ew + (I.ew_ofs >= 0 ? I.ew_ofs : P.edge_ofs);
        // the parent's own weights all equal w_c: no weight read per term
        // (an item with weight patches reads its own array)
        const bool uw = P.uni_w && I.ew_ofs < 0;
        const double wc = P.w_c;
        auto W = [&](int e) -> double { return uw ? wc : ew[e]; };
        const int *__restrict__ ioff = a.ioff + P.ioff_ofs;
        const uint32_t *__restrict__ pairs = a.pairs + P.pair_ofs;
        const int directed = P.directed;
        if (ST && s_staged != I.parent) {
            // every thread has read s_staged above; the previous item is done
            for (int u = threadIdx.x; u <= N; u += B) s_row[u] = row_off[u];
            for (int e = threadIdx.x; e <= E; e += B) s_ioff[e] = ioff[e];
            for (int e = threadIdx.x; e < E; e += B) {
                const int2 p = uv[e];
                s_uv[e] = make_ushort2((unsigned short)p.x, (unsigned short)p.y);
            }
            for (int i = threadIdx.x; i < P.S; i += B) s_adj[i] = (unsigned short)adj_e[i];
            __syncthreads();
            if (threadIdx.x == 0) s_staged = I.parent;
        }
        // the structure from LDS (ST) or through the cache
        auto ROW = [&](int u) -> int {
            if constexpr (ST) return s_row[u];
            else return row_off[u];
        };
        auto ADJ = [&](int i) -> int {
            if constexpr (ST) return (int)s_adj[i];
            else return adj_e[i];
        };
        auto UV = [&](int e) -> int2 {
            if constexpr (ST) {
                const ushort2 p = s_uv[e];
                return make_int2((int)p.x, (int)p.y);
            } else {
                return uv[e];
            }
        };
        auto IOFF = [&](int e) -> int {
            if constexpr (ST) return s_ioff[e];
            else return ioff[e];
        };
        // entry j of row u (starting at rb): from LDS (ST), column by
        // column (lanes on consecutive rows read consecutive words), or
        // from the row itself
        const int *__restrict__ ell = P.ell_ofs >= 0 ? a.ell + P.ell_ofs : nullptr;
        const int ell_stride = P.ell_stride;
        auto ROWE = [&](int u, int rb, int j) -> int {
            if constexpr (ST) return (int)s_adj[rb + j];
            else return ell ? ell[j * ell_stride + u] : adj_e[rb + j];
        };

        if (threadIdx.x < k) {
            s_tup[threadIdx.x] = a.item_verts[I.vofs + threadIdx.x];
            s_mark[threadIdx.x] = a.marks[threadIdx.x];
        }
        for (int u = threadIdx.x; u < N; u += B) {
            alive[u] = 1;
            vd[u] = 0.0;        // norms before any vertex phase (calloc'd in dress_init_graph)
        }
        __syncthreads();
        if (I.kind == LIFT_DELTA && threadIdx.x < k) alive[s_tup[threadIdx.x]] = 0;
        __syncthreads();
        // d = 1.0 on live edges (delta_dress_impl.c:166 -> dress.c:418-421);
        // a dead edge keeps value 0 so it adds nothing to a row sum
        // (with no iteration to run, the values stay at their FP64 start)
        bool fp32 = I.mixed != 0 && a.max_it > 0;
        float *fA = (float *)dA, *fB = (float *)dB;
        for (int e = threadIdx.x; e < E; e += B) {
            int2 p = UV(e);
            bool live = alive[p.x] && alive[p.y];
            if (fp32) { fA[e] = live ? 1.0f : 0.0f; fB[e] = 0.0f; }
            else { dA[e] = live ? 1.0 : 0.0; dB[e] = 0.0; }
        }
        __syncthreads();
        // the subgraph's own size (live entries, half-edges, edges,
        // vertices) for the work counters; live edges hold 1 here, dead 0
        unsigned long long w_t = 0, w_s = 0, w_e = 0, w_n = 0;
        if (a.work) {
            for (int u = threadIdx.x; u < N; u += B) {
                if (!alive[u]) continue;
                w_n++;
                for (int i = ROW(u); i < ROW(u + 1); i++) {
                    const int e = ADJ(i);
                    w_s += (fp32 ? fA[e] != 0.0f : dA[e] != 0.0) ? 1 : 0;
                }
            }
            for (int e = threadIdx.x; e < E; e += B) {
                if (!(fp32 ? fA[e] != 0.0f : dA[e] != 0.0)) continue;
                w_e++;
                for (int q = IOFF(e); q < IOFF(e + 1); q++) {
                    const uint32_t pr = pairs[q];
                    const int ea = (int)(pr & 0xffffu), eb = (int)(pr >> 16);
                    const bool la = fp32 ? fA[ea] != 0.0f : dA[ea] != 0.0;
                    const bool lb = fp32 ? fA[eb] != 0.0f : dA[eb] != 0.0;
                    w_t += (la && lb) ? 1 : 0;
                }
            }
        }
        int executed = 0;       // iterations run
        const double thr = 3000.0 * a.eps;
        int cur = 0;            // 0: A holds the current values, 1: B
        int iters = a.max_it;   // reference convention result
        double last = 0.0;
        int ambiguous = 0;
        for (int iter = 0; iter < a.max_it; iter++) {
            executed = iter + 1;
            // ---- vertex phase
            for (int u = threadIdx.x; u < N; u += B) {
                if (!alive[u]) continue;
                int rb = ROW(u), re = ROW(u + 1);
                const double self = 4.0 * item_nw(I.kind, u, s_tup, s_mark, k);
                double s = 0.0, c = 0.0;
                if (fp32) {
                    // float products, summed exactly in FP64 (independent of
                    // order), then the self term
                    const float *dc = cur ? fB : fA;
                    for (int j = 0; j < re - rb; j++) {
                        int e = ROWE(u, rb, j);
                        s += (double)((float)W(e) * dc[e]);
                    }
                    vd[u] = sqrt(s + self);
                } else {
                    const double *dc = cur ? dB : dA;
                    s = self;
                    for (int j = 0; j < re - rb; j++) {
                        int e = ROWE(u, rb, j);
                        two_sum_add(s, c, W(e) * dc[e]);
                    }
                    vd[u] = sqrt(s + c);
                }
            }
            __syncthreads();
            // ---- edge phase
            double mymax = 0.0;
            for (int e = threadIdx.x; e < E; e += B) {
                int2 p = UV(e);
                if (!alive[p.x] || !alive[p.y]) continue;
                double cu = 4.0 * item_nw(I.kind, p.x, s_tup, s_mark, k);
                double c4 = directed ? cu : cu + 4.0 * item_nw(I.kind, p.y, s_tup, s_mark, k);
                int kb = IOFF(e), ke = IOFF(e + 1);
                double w = W(e);
                double x, de;
                if (fp32) {
                    const float *dc = cur ? fB : fA;
                    float *dn = cur ? fA : fB;
                    double s = 0.0;
                    for (int q = kb; q < ke; q++) {
                        uint32_t pr = pairs[q];
                        int ea = (int)(pr & 0xffffu), eb = (int)(pr >> 16);
                        // a dead neighbour kills both of its edges: adds 0
                        s += (double)((float)W(ea) * dc[ea]) + (double)((float)W(eb) * dc[eb]);
                    }
                    de = (double)dc[e];
                    double own = (double)((float)w * dc[e]);
                    s += directed ? (c4 + own) : (c4 + 2.0 * own);
                    double den = vd[p.x] * vd[p.y];
                    x = den > 0.0 ? s / den : 0.0;
                    dn[e] = (float)x;
                } else {
                    const double *dc = cur ? dB : dA;
                    double *dn = cur ? dA : dB;
                    double s = 0.0, c = 0.0;
                    for (int q = kb; q < ke; q++) {
                        uint32_t pr = pairs[q];
                        int ea = (int)(pr & 0xffffu), eb = (int)(pr >> 16);
                        double ta = dc[ea], tb = dc[eb];
                        if (ta == 0.0 && tb == 0.0) continue;   // deleted common neighbour
                        two_sum_add(s, c, W(ea) * ta + W(eb) * tb);
                    }
                    de = dc[e];
                    double own = w * de;
                    two_sum_add(s, c, directed ? (c4 + own) : (c4 + 2.0 * own));
                    double den = vd[p.x] * vd[p.y];
                    x = den > 0.0 ? (s + c) / den : 0.0;
                    dn[e] = x;
                }
                mymax = fmax(mymax, fabs(de - x));
            }
            double md = block_max<B>(mymax, s_red);
            __syncthreads();
            cur ^= 1;
            last = md;
            if (fp32) {
                // hand over to FP64 (never decide convergence on FP32 values)
                if (md < thr || iter + 1 >= a.max_it - 16) {
                    // float values of buffer cur -> doubles in the other
                    // buffer, then copied back so both buffers hold doubles
                    const float *fc = cur ? fB : fA;
                    double *dn = cur ? dA : dB;
                    double *dm = cur ? dB : dA;
                    for (int e = threadIdx.x; e < E; e += B) dn[e] = (double)fc[e];
                    __syncthreads();
                    for (int e = threadIdx.x; e < E; e += B) dm[e] = dn[e];
                    __syncthreads();
                    cur ^= 1;
                    fp32 = false;
                }
                if (iter + 1 >= a.max_it) { ambiguous = 1; iters = a.max_it; break; }
                continue;
            }
            if (I.mixed && fabs(md - a.eps) <= 1e-3 * md) ambiguous = 1;
            if (md < a.eps) { iters = iter; break; }
        }
        // ---- outputs
        const double *dc = cur ? dB : dA;
        if (a.ms) {
            double *row = a.ms + (int64_t)I.ms_row * a.ms_stride;
            for (int e = threadIdx.x; e < E; e += B) {
                int2 p = UV(e);
                row[e] = (alive[p.x] && alive[p.y]) ? dc[e] : __longlong_as_double(0x7ff8000000000000ll);
            }
        }
        if (a.vals) {
            // live values packed in edge order at vals[I.val_ofs ...], or
            // first packed into the free buffer and sorted there
            double *sv = cur ? dA : dB;            // the previous values: free now
            __shared__ int s_base;
            if (threadIdx.x == 0) s_base = 0;
            __syncthreads();
            for (int e0 = 0; e0 < E; e0 += B) {
                int e = e0 + threadIdx.x;
                bool live = false;
                if (e < E) {
                    int2 p = UV(e);
                    live = alive[p.x] && alive[p.y];
                }
                // block-wide exclusive prefix of live flags, wave ballots
                unsigned long long m = __ballot(live);
                int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
                int before = __popcll(m & ((1ull << lane) - 1));
                __shared__ int s_wave[B / 64];
                if (lane == 0) s_wave[wv] = __popcll(m);
                __syncthreads();
                int off = s_base;
                for (int i = 0; i < wv; i++) off += s_wave[i];
                if (live) {
                    if (a.sort_vals) sv[off + before] = dc[e];
                    else a.vals[I.val_ofs + off + before] = dc[e];
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    int tot = 0;
                    for (int i = 0; i < B / 64; i++) tot += s_wave[i];
                    s_base += tot;
                }
                __syncthreads();
            }
            if (a.sort_vals) {
                // bitonic network over the next power of two P >= L with
                // the slots from L on as +infinity: every compare-exchange
                // puts the smaller at the lower index, so a pair reaching
                // past L never exchanges
                const int L = s_base;
                int lgP = 0;
                while ((1 << lgP) < L) lgP++;
                const int half = (1 << lgP) >> 1;
                for (int lk = 1; lk <= lgP; lk++) {
                    // flip within blocks of 2^lk: i = block start + off,
                    // j = block end - off
                    for (int t = threadIdx.x; t < half; t += B) {
                        const int off = t & ((1 << (lk - 1)) - 1);
                        const int i = ((t >> (lk - 1)) << lk) + off, j = i + (1 << lk) - 1 - 2 * off;
                        if (j < L && sort_key(sv[j]) < sort_key(sv[i])) {
                            const double x = sv[i];
                            sv[i] = sv[j];
                            sv[j] = x;
                        }
                    }
                    __syncthreads();
                    // half-cleaners at distance 2^lh
                    for (int lh = lk - 2; lh >= 0; lh--) {
                        for (int t = threadIdx.x; t < half; t += B) {
                            const int i = ((t >> lh) << (lh + 1)) + (t & ((1 << lh) - 1)), j = i + (1 << lh);
                            if (j < L && sort_key(sv[j]) < sort_key(sv[i])) {
                                const double x = sv[i];
                                sv[i] = sv[j];
                                sv[j] = x;
                            }
                        }
                        __syncthreads();
                    }
                }
                for (int i = threadIdx.x; i < L; i += B) a.vals[I.val_ofs + i] = sv[i];
            }
        }
        if (a.qkeys) {
            // quantised histogram in integer atomics: key floor(d / eps) of
            // every live value; values within qtol of a bin edge counted apart
            const unsigned long long gid = (unsigned long long)a.parent_graph[I.parent];
            for (int e = threadIdx.x; e < E; e += B) {
                int2 p = UV(e);
                if (!(alive[p.x] && alive[p.y])) continue;
                const unsigned long long key = quant_key((double)dc[e], gid, a.eps, a.qtol);
                int64_t slot = quant_slot(key, a.qcap);
                for (int64_t probe = 0; probe < a.qcap; probe++) {
                    unsigned long long prev = atomicCAS(&a.qkeys[slot], QEMPTY, key);
                    if (prev == QEMPTY || prev == key) {
                        atomicAdd(&a.qcounts[slot], 1ull);
                        break;
                    }
                    slot = (slot + 1) & (a.qcap - 1);
                }
            }
        }
        if (a.vnorm) {
            // vertex_dress of dress.c:555-574: norms of the last vertex phase
            for (int u = threadIdx.x; u < N; u += B)
                a.vnorm[I.vn_ofs + u] = alive[u] ? vd[u] : __longlong_as_double(0x7ff8000000000000ll);
        }
        if (a.work) {
            // wave sums, one atomic per wave and counter
            for (int o = 32; o > 0; o >>= 1) {
                w_t += __shfl_xor(w_t, o);
                w_s += __shfl_xor(w_s, o);
                w_e += __shfl_xor(w_e, o);
                w_n += __shfl_xor(w_n, o);
            }
            if ((threadIdx.x & 63) == 0) {
                const unsigned long long x = (unsigned long long)executed;
                if (w_t) atomicAdd(&a.work[0], w_t * x);
                if (w_s) atomicAdd(&a.work[1], w_s * x);
                if (w_e) atomicAdd(&a.work[2], w_e * x);
                if (w_n) atomicAdd(&a.work[3], w_n * x);
            }
        }
        if (threadIdx.x == 0) {
            if (a.iters) a.iters[it] = iters;
            if (a.deltas) a.deltas[it] = last;
            if (a.flags) a.flags[it] = ambiguous;
        }
        __syncthreads();
    }
}

}  // namespace

size_t lifted_lds_bytes(int ecap, int ncap)
{
    return (size_t)(2 * ecap + ncap) * sizeof(double) + (size_t)ncap * sizeof(int) + 16;
}

size_t lifted_lds_bytes_staged(int ecap, int ncap, int scap)
{
    return lifted_lds_bytes(ecap, ncap) + (size_t)(ncap + 2 + ecap + 2) * sizeof(int) +
           (size_t)ecap * sizeof(ushort2) + (size_t)scap * sizeof(unsigned short) + 16;
}

// Largest dynamic LDS one k_lifted launch may take on the current device:
// the per-workgroup maximum less the kernel's own static LDS (and a margin),
// at most 150 KiB.  Launches above it are refused: the workgroups would
// address LDS beyond their allocation.
size_t lifted_lds_limit()
{
    static size_t cache[64] = {0};
    int d = current_device();
    if (d < 0 || d >= 64) d = 0;
    if (!cache[d]) {
        int maxblk = 0;
        if (hipDeviceGetAttribute(&maxblk, hipDeviceAttributeMaxSharedMemoryPerBlock, d) != hipSuccess || maxblk <= 0) {
            hipGetLastError();
            maxblk = 64 * 1024;
        }
        size_t stat = 0;
        const void *fns[8] = {reinterpret_cast<const void *>(&k_lifted<64, false>),
                              reinterpret_cast<const void *>(&k_lifted<128, false>),
                              reinterpret_cast<const void *>(&k_lifted<256, false>),
                              reinterpret_cast<const void *>(&k_lifted<512, false>),
                              reinterpret_cast<const void *>(&k_lifted<64, true>),
                              reinterpret_cast<const void *>(&k_lifted<128, true>),
                              reinterpret_cast<const void *>(&k_lifted<256, true>),
                              reinterpret_cast<const void *>(&k_lifted<512, true>)};
        for (const void *f : fns) {
            hipFuncAttributes fa;
            if (hipFuncGetAttributes(&fa, f) == hipSuccess) stat = std::max(stat, (size_t)fa.sharedSizeBytes);
            else {
                hipGetLastError();
                stat = std::max<size_t>(stat, 4096);
            }
        }
        size_t lim = (size_t)maxblk > stat + 512 ? (size_t)maxblk - stat - 512 : 0;
        cache[d] = std::min<size_t>(lim, 150 * 1024);
    }
    return cache[d];
}

// LDS of one compute unit (bounds how many launch groups share a CU)
size_t lifted_lds_per_cu()
{
    static size_t cache[64] = {0};
    int d = current_device();
    if (d < 0 || d >= 64) d = 0;
    if (!cache[d]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, d) != hipSuccess || v <= 0) {
            hipGetLastError();
            v = 64 * 1024;
        }
        cache[d] = (size_t)v;
    }
    return cache[d];
}

bool lifted_launch(const LiftArgs &a, int block, int grid, bool staged, hipStream_t s)
{
    size_t lds = staged ? lifted_lds_bytes_staged(a.ecap, a.ncap, a.scap) : lifted_lds_bytes(a.ecap, a.ncap);
    if (lds > lifted_lds_limit()) {
        set_error("batched launch needs " + std::to_string(lds) + " B of LDS, above the device limit of " +
                  std::to_string(lifted_lds_limit()) + " B");
        return false;
    }
    if (staged) {
        if (block == 64)
            hipLaunchKernelGGL((k_lifted<64, true>), dim3(grid), dim3(64), lds, s, a);
        else if (block == 128)
            hipLaunchKernelGGL((k_lifted<128, true>), dim3(grid), dim3(128), lds, s, a);
        else if (block == 256)
            hipLaunchKernelGGL((k_lifted<256, true>), dim3(grid), dim3(256), lds, s, a);
        else
            hipLaunchKernelGGL((k_lifted<512, true>), dim3(grid), dim3(512), lds, s, a);
    } else {
        if (block == 64)
            hipLaunchKernelGGL((k_lifted<64, false>), dim3(grid), dim3(64), lds, s, a);
        else if (block == 128)
            hipLaunchKernelGGL((k_lifted<128, false>), dim3(grid), dim3(128), lds, s, a);
        else if (block == 256)
            hipLaunchKernelGGL((k_lifted<256, false>), dim3(grid), dim3(256), lds, s, a);
        else
            hipLaunchKernelGGL((k_lifted<512, false>), dim3(grid), dim3(512), lds, s, a);
    }
    DH_CHECK(hipGetLastError());
    return true;
}

}  // namespace dh
