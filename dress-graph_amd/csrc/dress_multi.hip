// dress_multi.hip -- delta^k / nabla^k subgraphs of a graph too large for the
// LDS-resident kernel, MB subgraphs at a time on the parent's device graph.
//
// The reference fits every subgraph on its own rebuilt graph
// (delta_dress_impl.c:95-169, nabla_dress_impl.c:61-91).  Here one pass
// fits MB subgraphs ("slots") together against the parent's CSR,
// intersection index and degree bins: values, half-edge copies and norms
// are stored item-minor (x[e * MB + slot]), so a lane group of MB slots
// reads one index entry once and gathers one contiguous run of MB values
// for it -- the index traffic of a fit is shared by MB subgraphs and every
// gathered cache line is used whole.
//   delta^k  an edge touching one of the slot's deleted vertices is dead:
//            its value and copies are 0 from the start and are never
//            written, so every live sum sees exact zeros where the
//            subgraph has no term; vertex weights are 1.0 (the subgraph is
//            built with NW = NULL, delta_dress_impl.c:104-106).
//   nabla^k  vertex weights are 1.0 except the slot's tuple, which carries
//            the marker weights (nabla_dress_impl.c:82-86).
// Sums are the exactly rounded compensated sums of the single-graph fit, so
// each live value equals the value on the rebuilt subgraph bit for bit.
// Each slot keeps its own convergence state (strict max|change| < epsilon,
// dress.c:590-600) and stops on its own; a pass ends when every slot has.
// The pass runs in FP64.
#include "dress_batch.h"
#include "dress_capi_internal.h"
#include "dress_device.h"
#include "dress_hip_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace dh {

namespace {

constexpr int TPB = 256;
constexpr int MKMAX = 8;           // largest k of a pass (tuples staged in LDS)

struct SlotState {
    int active;                    // 0 once stopped (or an unused slot)
    int iter;                      // iterations finished
    int cur;                       // buffer holding the slot's values
    int pad;
    unsigned long long maxd;       // running max |change| (bit pattern)
    unsigned long long count;      // live values collected
};

struct MultiArgs {
    const int *row_off;
    const int2 *uv, *rs, *pos;
    const double *ew;
    const int64_t *ioff;
    const uint32_t *ipk;
    const int2 *ipair;
    int directed;                  // FORWARD / BACKWARD self-term form
    int kind, k, n;                // n slots in use
    const int *tup;                // [n * k] internal labels, slot by slot
    const double *marks;           // [k] nabla marker weights
    double *d0, *d1, *w0, *w1;     // item-minor values and half-edge copies
    double *vd;                    // item-minor norms
    SlotState *st;
    double eps;                    // stopping rule of every slot
    int max_it;
    int *remaining;                // slots still active after the iteration
};

inline int grid_for(int64_t threads, int cap)
{
    int64_t b = (threads + TPB - 1) / TPB;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

// the slots' tuples, activity and current buffers in LDS
template <int MB>
struct SlotCache {
    int tup[MB * MKMAX];
    int act[MB];
    int cur[MB];
};

template <int MB>
__device__ __forceinline__ void load_slots(const MultiArgs &a, SlotCache<MB> &sc)
{
    for (int j = threadIdx.x; j < a.n * a.k; j += blockDim.x) sc.tup[j] = a.tup[j];
    for (int i = threadIdx.x; i < MB; i += blockDim.x) {
        sc.act[i] = i < a.n ? a.st[i].active : 0;
        sc.cur[i] = i < a.n ? a.st[i].cur : 0;
    }
    __syncthreads();
}

template <int MB>
__device__ __forceinline__ bool slot_has(const MultiArgs &a, const SlotCache<MB> &sc, int i, int x)
{
    bool h = false;
    for (int j = 0; j < a.k; j++) h = h || sc.tup[i * a.k + j] == x;
    return h;
}

// vertex weight of x in slot i: nabla marker weight or 1.0
template <int MB>
__device__ __forceinline__ double slot_nw(const MultiArgs &a, const SlotCache<MB> &sc, int i, int x)
{
    double w = 1.0;
    if (a.kind == LIFT_NABLA)
        for (int j = 0; j < a.k; j++)
            if (sc.tup[i * a.k + j] == x) w = a.marks[j];
    return w;
}

// merge the (s, c) of the GS lanes of one slot inside a group of L = MB * GS
// lanes (slot = lane index % MB, sub = lane index / MB).  Subs of one slot
// in the same wave merge by shuffles; with L = TPB the group is the
// workgroup and subs in different waves merge through LDS, the result valid
// in the waves of sub 0 (the first MB threads).  MB = TPB has one lane per
// slot and nothing to merge.
template <int MB, int GS>
__device__ __forceinline__ void slot_reduce(double &s, double &c, double (*ls)[64], double (*lc)[64])
{
    if (GS == 1) return;
    constexpr int L = MB * GS;
    constexpr int HI = L < 64 ? L : 64;
#pragma unroll
    for (int o = MB; o < HI; o <<= 1) {
        double s2 = __shfl_xor(s, o);
        double c2 = __shfl_xor(c, o);
        two_sum_merge(s, c, s2, c2);
    }
    if (L == TPB && MB < TPB) {
        // waves w and w + k * (MB / 64) hold the same slots (MB >= 64), or
        // every wave holds all MB slots (MB < 64)
        constexpr int WS = MB >= 64 ? MB / 64 : 1;     // waves of one sub
        const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
        if (l < (MB < 64 ? MB : 64)) { ls[w][l] = s; lc[w][l] = c; }
        __syncthreads();
        if (w < WS && l < (MB < 64 ? MB : 64))
            for (int v = w + WS; v < TPB / 64; v += WS) two_sum_merge(s, c, ls[v][l], lc[v][l]);
        __syncthreads();
    }
}

// vertex phase of every active slot: vd = sqrt(4 nw + row sum of copies)
template <int MB, int GS>
__global__ void __launch_bounds__(TPB) k_mvertex(const int *__restrict__ list, int count, MultiArgs a)
{
    constexpr int L = MB * GS;
    if (*a.remaining == 0) return;             // every slot has stopped
    __shared__ SlotCache<MB> sc;
    __shared__ double ls[TPB / 64][64], lc[TPB / 64][64];
    load_slots<MB>(a, sc);
    const int i = threadIdx.x % MB;
    const int sub = (threadIdx.x / MB) % GS;
    const bool act = sc.act[i] != 0;
    const double *__restrict__ wh = sc.cur[i] ? a.w1 : a.w0;
    const int64_t units = (int64_t)gridDim.x * (TPB / L);
    for (int64_t g = (int64_t)blockIdx.x * (TPB / L) + threadIdx.x / L; g < count; g += units) {
        const int u = list[g];
        const int b = a.row_off[u], e = a.row_off[u + 1];
        double s = 0.0, c = 0.0;
        if (act) {
            if (sub == 0) s = 4.0 * slot_nw<MB>(a, sc, i, u);
            for (int p = b + sub; p < e; p += GS) two_sum_add(s, c, wh[(int64_t)p * MB + i]);
        }
        slot_reduce<MB, GS>(s, c, ls, lc);
        if (act && sub == 0) a.vd[(int64_t)u * MB + i] = sqrt(s + c);
    }
}

template <bool PK>
__device__ __forceinline__ void mentry(const MultiArgs &a, int64_t k, int2 r, int &x, int &y)
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

// convergence bookkeeping of every slot for the iteration just finished
// (dress.c:590-600: stop once the max change is < epsilon, or after
// max_iterations); *remaining = slots still active.  One workgroup.  (A
// last-workgroup-done form inside the edge launch measured 3.4x slower per
// pass: the device-scope fence each workgroup needs writes back and
// invalidates its XCD's L2, which holds the slots' state.)
__global__ void k_mend(MultiArgs a)
{
    if (*a.remaining == 0) return;
    SlotState *st = a.st;
    const int n = a.n;
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        if (!st[i].active) continue;
        const double md = __longlong_as_double((long long)st[i].maxd);
        st[i].maxd = 0ull;
        st[i].cur ^= 1;
        const int it = st[i].iter;
        st[i].iter = it + 1;
        if (md < a.eps || it + 1 >= a.max_it) st[i].active = 0;
        if (st[i].active) atomicAdd(&cnt, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) *a.remaining = cnt;
}

// edge phase of every active slot over internal edges [lo, hi)
template <int MB, int GS, bool PK>
__global__ void __launch_bounds__(TPB) k_medge(int lo, int hi, MultiArgs a)
{
    constexpr int L = MB * GS;
    if (*a.remaining == 0) return;             // every slot has stopped
    __shared__ SlotCache<MB> sc;
    __shared__ double ls[TPB / 64][64], lc[TPB / 64][64];
    load_slots<MB>(a, sc);
    const int i = threadIdx.x % MB;
    const int sub = (threadIdx.x / MB) % GS;
    const bool act = sc.act[i] != 0;
    const int cur = sc.cur[i];
    const double *__restrict__ dc = cur ? a.d1 : a.d0;
    const double *__restrict__ wc = cur ? a.w1 : a.w0;
    double *__restrict__ dn = cur ? a.d0 : a.d1;
    double *__restrict__ wn = cur ? a.w0 : a.w1;
    const int64_t units = (int64_t)gridDim.x * (TPB / L);
    double mymax = 0.0;
    for (int64_t g = lo + (int64_t)blockIdx.x * (TPB / L) + threadIdx.x / L; g < hi; g += units) {
        const int e = (int)g;
        const int2 p = a.uv[e];
        const bool dead = a.kind == LIFT_DELTA && (slot_has<MB>(a, sc, i, p.x) || slot_has<MB>(a, sc, i, p.y));
        const bool live = act && !dead;
        double s = 0.0, c = 0.0, de = 0.0, w = 0.0;
        if (live) {
            const int64_t kb = a.ioff[e], ke = a.ioff[e + 1];
            const int2 r = PK ? a.rs[e] : make_int2(0, 0);
            for (int64_t k = kb + sub; k < ke; k += GS) {
                int x, y;
                mentry<PK>(a, k, r, x, y);
                two_sum_add(s, c, wc[(int64_t)x * MB + i] + wc[(int64_t)y * MB + i]);
            }
            if (sub == 0) {
                de = dc[(int64_t)e * MB + i];
                w = a.ew[e];
                const double own = w * de;
                const double nu = slot_nw<MB>(a, sc, i, p.x);
                const double c4 = a.directed ? 4.0 * nu : 4.0 * nu + 4.0 * slot_nw<MB>(a, sc, i, p.y);
                two_sum_add(s, c, a.directed ? (c4 + own) : (c4 + 2.0 * own));
            }
        }
        slot_reduce<MB, GS>(s, c, ls, lc);
        if (live && sub == 0) {
            const double den = a.vd[(int64_t)p.x * MB + i] * a.vd[(int64_t)p.y * MB + i];
            const double x = den > 0.0 ? (s + c) / den : 0.0;
            mymax = fmax(mymax, fabs(de - x));
            dn[(int64_t)e * MB + i] = x;
            const double wx = w * x;
            const int2 q = a.pos[e];
            wn[(int64_t)q.x * MB + i] = wx;
            if (q.y >= 0) wn[(int64_t)q.y * MB + i] = wx;
        }
    }
    // per-slot max over the wave, then over the workgroup's waves (LDS),
    // then one atomic per slot and workgroup, only where it can raise the
    // slot's running max
    if (MB < 64) {
#pragma unroll
        for (int o = MB; o < 64; o <<= 1) mymax = fmax(mymax, __shfl_xor(mymax, o));
    }
    __shared__ double smax[TPB];
    smax[threadIdx.x] = mymax;
    __syncthreads();
    if ((int)threadIdx.x < MB) {
        // slot t's lanes: t, t + MB, t + 2 MB, ... (and, for MB < 64, the
        // wave-reduced copies at t + 64 k)
        constexpr int STEP = MB < 64 ? 64 : MB;
        double m = smax[threadIdx.x];
        for (int v = threadIdx.x + STEP; v < TPB; v += STEP) m = fmax(m, smax[v]);
        unsigned long long *dst = &a.st[threadIdx.x].maxd;
        const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
        if (m > 0.0 && bits > __atomic_load_n(dst, __ATOMIC_RELAXED)) atomicMax(dst, bits);
    }
}

// fresh slot states for a pass of n items, all of them active
__global__ void k_mslots(SlotState *st, int n, int mb, int *remaining)
{
    const int i = threadIdx.x;
    if (i == 0) *remaining = n;
    if (i < mb) {
        SlotState z;
        z.active = i < n ? 1 : 0;
        z.iter = 0;
        z.cur = 0;
        z.pad = 0;
        z.maxd = 0ull;
        z.count = 0ull;
        st[i] = z;
    }
}

// labels of the items' vertices: input -> internal
__global__ void k_mrank(const int *in, int n, const int *rank, int *out)
{
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) out[j] = rank[in[j]];
}

// start values of a pass in both buffers: 1.0 (0 for dead edges) and the
// half-edge copies w-bar * d
template <int MB>
__global__ void k_minit(int E, MultiArgs a)
{
    __shared__ SlotCache<MB> sc;
    load_slots<MB>(a, sc);
    const int64_t total = (int64_t)E * MB;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int e = (int)(t / MB), i = (int)(t % MB);
        const int2 p = a.uv[e];
        const bool dead = i >= a.n ||
                          (a.kind == LIFT_DELTA && (slot_has<MB>(a, sc, i, p.x) || slot_has<MB>(a, sc, i, p.y)));
        const double x = dead ? 0.0 : 1.0;
        a.d0[t] = x;
        a.d1[t] = x;
        const double wx = dead ? 0.0 : a.ew[e] * x;
        const int2 q = a.pos[e];
        a.w0[(int64_t)q.x * MB + i] = wx;
        a.w1[(int64_t)q.x * MB + i] = wx;
        if (q.y >= 0) {
            a.w0[(int64_t)q.y * MB + i] = wx;
            a.w1[(int64_t)q.y * MB + i] = wx;
        }
    }
}

// live values of every slot appended at vals[ofs[i] .. lim[i]) (order does
// not matter to the histogram); with rows, slot i's multiset row in input
// edge order (NaN for removed edges) at rows[i * E ..).  Thread t takes
// slot t / Ep and input edge t % Ep (Ep = E rounded up to a wave), so a
// wave holds one slot: it compacts its live values with a ballot and takes
// their places with one atomic.
template <int MB>
__global__ void k_mcollect(int E, MultiArgs a, const int *inv, const int64_t *ofs, const int64_t *lim, double *vals,
                           double *rows)
{
    __shared__ SlotCache<MB> sc;
    load_slots<MB>(a, sc);
    const int64_t Ep = ((int64_t)E + 63) & ~(int64_t)63;
    const int64_t total = Ep * a.n;
    const int lane = threadIdx.x & 63;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t - lane < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / Ep);
        const int ein = (int)(t % Ep);
        bool live = false;
        double v = 0.0;
        if (ein < E) {
            const int e = inv[ein];
            const int2 p = a.uv[e];
            const bool dead =
                a.kind == LIFT_DELTA && (slot_has<MB>(a, sc, i, p.x) || slot_has<MB>(a, sc, i, p.y));
            v = (sc.cur[i] ? a.d1 : a.d0)[(int64_t)e * MB + i];
            live = !dead;
            if (rows) rows[(int64_t)i * E + ein] = dead ? __longlong_as_double(0x7ff8000000000000ll) : v;
        }
        const unsigned long long b = __ballot(live);
        unsigned long long base = 0;
        if (lane == 0 && b) base = atomicAdd(&a.st[i].count, (unsigned long long)__popcll(b));
        base = __shfl(base, 0);
        if (live) {
            const int64_t at = ofs[i] + (int64_t)base + __popcll(b & ((1ull << lane) - 1ull));
            if (at < lim[i]) vals[at] = v;          // a miscount is reported by the host
        }
    }
}

template <int MB>
struct Pass {
    // lanes per slot of vertex / edge bins 0..4 (<= 4, 64, 512, 8192, more
    // terms): a slot's sum over one row or edge is split over GS lanes
    static constexpr int G1 = MB >= 32 ? 1 : 2, G2 = MB >= 64 ? 1 : 64 / MB, G3 = MB >= TPB ? 1 : TPB / MB;
    static constexpr int gs(int bin) { return bin == 0 ? 1 : bin == 1 ? G1 : bin == 2 ? G2 : G3; }

    template <int GS>
    static void vertex(DevGraph *G, int b0, int b1, const MultiArgs &a, int cap)
    {
        const Bins &B = G->vbins;
        const int lo = B.start[b0], n = B.start[b1 - 1] + B.count[b1 - 1] - lo;
        if (n <= 0) return;
        hipLaunchKernelGGL((k_mvertex<MB, GS>), dim3(grid_for((int64_t)n * MB * GS, cap)), dim3(TPB), 0, G->stream,
                           B.list + lo, n, a);
    }
    template <int GS, bool PK>
    static void edge(DevGraph *G, int lo, int hi, const MultiArgs &a, int cap)
    {
        hipLaunchKernelGGL((k_medge<MB, GS, PK>), dim3(grid_for((int64_t)(hi - lo) * MB * GS, cap)), dim3(TPB), 0,
                           G->stream, lo, hi, a);
    }
    template <int GS>
    static void vertex_gs(int gsv, DevGraph *G, int b0, int b1, const MultiArgs &a, int cap)
    {
        if constexpr (MB * GS <= TPB) {
            if (gsv == GS) vertex<GS>(G, b0, b1, a, cap);
        }
    }
    template <int GS, bool PK>
    static void edge_gs(int gsv, DevGraph *G, int lo, int hi, const MultiArgs &a, int cap)
    {
        if constexpr (MB * GS <= TPB) {
            if (gsv == GS) edge<GS, PK>(G, lo, hi, a, cap);
        }
    }
    // one iteration: adjacent bins with the same lanes per slot share a
    // launch; k_mend closes it
    template <bool PK>
    static void iteration(DevGraph *G, const MultiArgs &a, int cap)
    {
        int runs[Bins::NB][2], nr = 0;
        for (int b = 0; b < Bins::NB; b++) {
            if (nr > 0 && gs(runs[nr - 1][0]) == gs(b)) runs[nr - 1][1] = b + 1;
            else { runs[nr][0] = b; runs[nr][1] = b + 1; nr++; }
        }
        for (int r = 0; r < nr; r++) {
            const int g = gs(runs[r][0]);
            vertex_gs<1>(g, G, runs[r][0], runs[r][1], a, cap);
            vertex_gs<2>(g, G, runs[r][0], runs[r][1], a, cap);
            vertex_gs<4>(g, G, runs[r][0], runs[r][1], a, cap);
            vertex_gs<16>(g, G, runs[r][0], runs[r][1], a, cap);
        }
        for (int r = 0; r < nr; r++) {
            const int lo = G->ebin[runs[r][0]], hi = G->ebin[runs[r][1]];
            if (hi <= lo) continue;
            const int g = gs(runs[r][0]);
            edge_gs<1, PK>(g, G, lo, hi, a, cap);
            edge_gs<2, PK>(g, G, lo, hi, a, cap);
            edge_gs<4, PK>(g, G, lo, hi, a, cap);
            edge_gs<16, PK>(g, G, lo, hi, a, cap);
        }
        hipLaunchKernelGGL(k_mend, dim3(1), dim3(TPB), 0, G->stream, a);
    }
};

// bytes of one pass's state with mb slots
size_t pass_bytes(const DevGraph *G, int mb)
{
    return (size_t)mb * (4 * (size_t)G->E + 2 * (size_t)G->S + (size_t)G->N + 8) * sizeof(double);
}

template <int MB>
bool run_passes(DevGraph *G, int kind, int k, const std::vector<int> &verts, const std::vector<int64_t> &live,
                const std::vector<int64_t> &sel, const double *marks, int max_it, double eps, double *d_vals,
                double *ms_rows)
{
    hipStream_t s = G->stream;
    const int E = G->E, N = G->N;
    const int64_t S = G->S;
    const int n_items = (int)live.size();
    // value offsets of the items
    std::vector<int64_t> vofs((size_t)n_items + 1, 0);
    for (int i = 0; i < n_items; i++) vofs[(size_t)i + 1] = vofs[(size_t)i] + live[(size_t)i];
    MultiArgs a;
    memset(&a, 0, sizeof(a));
    a.row_off = G->row_off; a.uv = G->uv; a.rs = G->rs; a.pos = G->pos; a.ew = G->ew; a.ioff = G->ioff;
    a.ipk = G->ipk; a.ipair = G->ipair;
    a.directed = (G->variant == DRESS_VARIANT_FORWARD || G->variant == DRESS_VARIANT_BACKWARD) ? 1 : 0;
    a.kind = kind;
    a.k = k;
    int *d_in = nullptr, *d_tup = nullptr, *d_rem = nullptr;
    double *d_marks = nullptr, *d_rows = nullptr;
    int64_t *d_ofs = nullptr, *d_lim = nullptr;
    SlotState *d_st = nullptr;
    bool ok = dalloc(&a.d0, (size_t)E * MB + 1, "slot values") && dalloc(&a.d1, (size_t)E * MB + 1, "slot values") &&
              dalloc(&a.w0, (size_t)S * MB + 1, "slot copies") && dalloc(&a.w1, (size_t)S * MB + 1, "slot copies") &&
              dalloc(&a.vd, (size_t)N * MB + 1, "slot norms") && dalloc(&d_st, MB, "slot states") &&
              dalloc(&d_in, (size_t)MB * k + 1, "slot vertices") && dalloc(&d_tup, (size_t)MB * k + 1, "slot tuples") &&
              dalloc(&d_marks, LIFT_KMAX, "marker weights") && dalloc(&d_rem, 1, "active slots") &&
              dalloc(&d_ofs, MB, "slot offsets") && dalloc(&d_lim, MB, "slot limits") && (!ms_rows || dalloc(&d_rows, (size_t)MB * E + 1, "slot rows"));
    if (ok && marks) ok = hipMemcpyAsync(d_marks, marks, (size_t)k * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess;
    a.tup = d_tup;
    a.marks = d_marks;
    a.st = d_st;
    a.eps = eps;
    a.max_it = max_it;
    a.remaining = d_rem;
    int cap = std::max(1, num_cus() * 8 * thread_grid_share() / 100);
    int *snap = (int *)snapshot_take();
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
    hipEvent_t sev[2] = {nullptr, nullptr};
    ok = ok && snap && hipEventCreateWithFlags(&sev[0], hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&sev[1], hipEventDisableTiming) == hipSuccess;
    std::vector<int> slot_item, tin;
    std::vector<int64_t> sofs, slim;
    std::vector<SlotState> hst(MB);
    int next = 0;
    while (ok && next < n_items) {
        if (thread_cancelled()) {
            set_error("cancelled");
            ok = false;
            break;
        }
        // requests of higher priority first (a server's short requests):
        // wait up to 50 ms while any are waiting or running
        yield_to_pressure(0.05);
        // the next MB items with live edges (delta items without any are
        // skipped, delta_dress_impl.c:124-129)
        slot_item.clear();
        tin.clear();
        sofs.clear();
        slim.clear();
        while (next < n_items && (int)slot_item.size() < MB) {
            if (!(kind == LIFT_DELTA && live[(size_t)next] == 0)) {
                slot_item.push_back(next);
                tin.insert(tin.end(), verts.begin() + (size_t)next * k, verts.begin() + (size_t)(next + 1) * k);
                sofs.push_back(vofs[(size_t)next]);
                slim.push_back(vofs[(size_t)next + 1]);
            }
            next++;
        }
        const int n = (int)slot_item.size();
        if (n == 0) break;
        a.n = n;
        ok = hipMemcpyAsync(d_in, tin.data(), tin.size() * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(d_ofs, sofs.data(), sofs.size() * sizeof(int64_t), hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(d_lim, slim.data(), slim.size() * sizeof(int64_t), hipMemcpyHostToDevice, s) == hipSuccess;
        if (!ok) break;
        if (!tin.empty())
            hipLaunchKernelGGL(k_mrank, dim3(grid_for((int64_t)tin.size(), 64)), dim3(TPB), 0, s, d_in,
                               (int)tin.size(), G->rank, d_tup);
        hipLaunchKernelGGL(k_mslots, dim3(1), dim3(TPB), 0, s, d_st, n, MB, d_rem);
        if (E > 0) hipLaunchKernelGGL(k_minit<MB>, dim3(grid_for((int64_t)E * MB, cap)), dim3(TPB), 0, s, E, a);
        // chunks of 4 iterations, each followed by a copy of the active-slot
        // count into a pinned snapshot slot; the next chunk is launched
        // before the last snapshot is read, and iterations launched after
        // every slot stopped return at once
        int launched = 0, remaining = n;
        auto issue = [&](int slot) {
            const int m = std::min(4, max_it - launched);
            for (int j = 0; j < m; j++) {
                if (G->packed) Pass<MB>::template iteration<true>(G, a, cap);
                else Pass<MB>::template iteration<false>(G, a, cap);
            }
            launched += m;
            return hipGetLastError() == hipSuccess &&
                   hipMemcpyAsync(&snap[slot], d_rem, sizeof(int), hipMemcpyDeviceToHost, s) == hipSuccess &&
                   hipEventRecord(sev[slot], s) == hipSuccess;
        };
        int pend = 0;
        if (max_it > 0) ok = issue(0);
        while (ok && max_it > 0) {
            int nxt = -1;
            if (launched < max_it) {
                yield_to_pressure(0.05);
                cap = std::max(1, num_cus() * 8 * thread_grid_share() / 100);
                ok = issue(pend ^ 1);
                nxt = pend ^ 1;
            }
            ok = ok && hipEventSynchronize(sev[pend]) == hipSuccess;
            if (!ok) break;
            remaining = snap[pend];
            if (remaining == 0 || nxt < 0) break;
            pend = nxt;
        }
        // the count after everything launched
        ok = ok && hipMemcpyAsync(&remaining, d_rem, sizeof(int), hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
        if (!ok) break;
        if (max_it > 0 && remaining > 0) {
            set_error("subgraph fits did not stop within max_iterations");
            ok = false;
            break;
        }
        if (E > 0)
            hipLaunchKernelGGL(k_mcollect<MB>, dim3(grid_for((((int64_t)E + 63) & ~(int64_t)63) * n, cap)),
                               dim3(TPB), 0, s, E, a, G->inv, d_ofs, d_lim, d_vals, d_rows);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(hst.data(), d_st, (size_t)MB * sizeof(SlotState), hipMemcpyDeviceToHost, s) == hipSuccess;
        if (ok && ms_rows)
            for (int i = 0; ok && i < n; i++)
                ok = hipMemcpyAsync(ms_rows + (size_t)sel[(size_t)slot_item[(size_t)i]] * (size_t)E,
                                    d_rows + (size_t)i * E, (size_t)E * sizeof(double), hipMemcpyDeviceToHost,
                                    s) == hipSuccess;
        ok = ok && hipStreamSynchronize(s) == hipSuccess;
        for (int i = 0; ok && i < n; i++)
            if ((int64_t)hst[(size_t)i].count != live[(size_t)slot_item[(size_t)i]]) {
                set_error("subgraph fits: live value count mismatch");
                ok = false;
            }
    }
    for (hipEvent_t x : sev)
        if (x) (void)hipEventDestroy(x);
    dfree(a.d0); dfree(a.d1); dfree(a.w0); dfree(a.w1); dfree(a.vd); dfree(d_st);
    dfree(d_in); dfree(d_tup); dfree(d_marks); dfree(d_rem); dfree(d_ofs); dfree(d_lim); dfree(d_rows);
    if (!ok && !has_error()) set_error("multi-subgraph fits failed");
    return ok;
}

}  // namespace

int multi_max_k() { return MKMAX; }

// slots per pass for G: 256 for graphs without edges of more than 512 terms
// (one lane per slot then carries a whole sum) whose pass state stays within
// 2 GiB; else 64 within 2 GiB, else 16 within 4 GiB; 0 beyond.  A thread's
// slot cap (set_thread_max_slots) keeps its passes narrower.
int multi_slots(const DevGraph *G)
{
    const bool light = G->ebin[5] == G->ebin[3];          // edge bins 3 and 4 empty
    const int cap = thread_max_slots();
    if (cap >= 256 && light && pass_bytes(G, 256) <= ((size_t)2 << 30)) return 256;
    if (cap >= 64 && pass_bytes(G, 64) <= ((size_t)2 << 30)) return 64;
    if (pass_bytes(G, 16) <= ((size_t)4 << 30)) return 16;
    return 0;
}

size_t multi_pass_bytes(const DevGraph *G)
{
    const int mb = multi_slots(G);
    return mb ? pass_bytes(G, mb) : 0;
}

// Fits the items (k vertices each in verts, input labels) of G in passes of
// multi_slots(G) subgraphs.  Item i's live values go to d_vals at the sum
// of live[0..i) (device, room for the sum of live); multiset rows, when
// ms_rows is non-null, into ms_rows + sel[i] * E on the host.  The parent's
// own device state (values, norms, self terms) is not touched.  Returns
// false with the error set on failure.
bool multi_items(DevGraph *G, int kind, int k, const std::vector<int> &verts, const std::vector<int64_t> &live,
                 const std::vector<int64_t> &sel, const double *marks, int max_it, double eps, double *d_vals,
                 double *ms_rows)
{
    if (!build_bins(G)) return false;
    if (k < 1 || k > MKMAX) {
        set_error("multi-subgraph fits need 1 <= k <= 8");
        return false;
    }
    switch (multi_slots(G)) {
    case 256: return run_passes<256>(G, kind, k, verts, live, sel, marks, max_it, eps, d_vals, ms_rows);
    case 64: return run_passes<64>(G, kind, k, verts, live, sel, marks, max_it, eps, d_vals, ms_rows);
    case 16: return run_passes<16>(G, kind, k, verts, live, sel, marks, max_it, eps, d_vals, ms_rows);
    default: set_error("graph too large for multi-subgraph passes"); return false;
    }
}

}  // namespace dh
