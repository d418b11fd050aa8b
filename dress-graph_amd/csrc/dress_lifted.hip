// dress_lifted.hip -- host driver of delta^k / nabla^k on the GPU and the
// C ABI entry points dress_{delta,nabla}_fit_hip*, plus the multi-graph batch
// API dress_hip_lifted_batch.
//
// Enumeration, striding and sampling follow the reference drivers
// (libdress/src/delta_dress_impl.c:233-380, nabla_dress_impl.c:142-311):
// subgraph s is the s-th k-combination (delta) or ordered k-tuple (nabla) in
// lexicographic order; s is processed when s % stride == offset and, when
// sampling, when its rank within the stride is one of the sorted, deduplicated
// picks drawn with the reference's LCG.  The histogram is exact in the value:
// all live values are radix sorted on the device and run-length encoded, and
// returned ascending as [value bits, count] like dress_hist_flatten
// (dress_histogram.c:124-155).
#include "dress_batch.h"
#include "dress_capi_internal.h"
#include "dress/hip/dress_hip.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <string>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>
#include <atomic>
#include <thread>

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

// ------------------------------------------------------------ host parents

// Structure of one small parent graph, built on the host: the variant CSR
// with rows ordered by (neighbour, edge), variant weights, and per edge the
// multiset intersection of its two rows as pairs of edge ids.  Same rules as
// the device build (dress_build.hip) and the reference (dress.c:103-363).
// Phase timings printed to stderr when DRESS_HIP_TIMING is set.
struct PhaseTimer {
    const char *what;
    bool on;
    std::chrono::steady_clock::time_point last;
    explicit PhaseTimer(const char *w) : what(w), on(getenv("DRESS_HIP_TIMING") != nullptr),
                                         last(std::chrono::steady_clock::now()) {}
    void mark(const char *phase)
    {
        if (!on) return;
        auto now = std::chrono::steady_clock::now();
        fprintf(stderr, "%s: %-18s %8.3f ms\n", what, phase,
                std::chrono::duration<double, std::milli>(now - last).count());
        last = now;
    }
};

struct HostParent {
    int N = 0, E = 0, S = 0, T = 0;
    int directed = 0;
    std::vector<int> row_off, adj_x, adj_e, ioff;
    // the rows' edge ids column by column (entry i of row u at
    // ell[i * ell_stride + u], ell_stride = N rounded up to 64), for
    // parents of more than 1,024 edges whose padded copy is at most twice
    // the rows (empty otherwise)
    std::vector<int> ell;
    int ell_stride = 0;
    std::vector<int2> uv;
    std::vector<double> ew;
    std::vector<uint32_t> pairs;
    // DIRECTED: edges whose reciprocal lookup meets repeated reciprocal
    // edges of different weights (amb, with their own weights amb_w), and
    // the raw out-rows the lookup searches (kept only when amb is not empty)
    std::vector<int> amb;
    std::vector<double> amb_w;
    std::vector<int> raw_off, raw_x;
    std::vector<double> raw_w;
};

struct HalfEdge {
    int x, e;
    double w;
};

bool host_parent_build(int N, int E, const int *U, const int *V, const double *W, int variant,
                       HostParent &P)
{
    P.N = N;
    P.E = E;
    P.directed = (variant == DRESS_VARIANT_FORWARD || variant == DRESS_VARIANT_BACKWARD);
    auto by_xe = [](const HalfEdge &a, const HalfEdge &b) { return a.x != b.x ? a.x < b.x : a.e < b.e; };
    // raw out-adjacency for reciprocal lookups (DIRECTED)
    std::vector<int> roff;
    std::vector<HalfEdge> raw;
    if (variant == DRESS_VARIANT_DIRECTED) {
        roff.assign(N + 1, 0);
        for (int e = 0; e < E; e++) roff[U[e] + 1]++;
        for (int u = 0; u < N; u++) roff[u + 1] += roff[u];
        raw.resize(E);
        std::vector<int> fill(roff.begin(), roff.end() - 1);
        for (int e = 0; e < E; e++) raw[fill[U[e]]++] = HalfEdge{V[e], e, W ? W[e] : 1.0};
        for (int u = 0; u < N; u++) std::sort(raw.begin() + roff[u], raw.begin() + roff[u + 1], by_xe);
    }
    auto raw_find = [&](int u, int v) {  // probe sequence of dress.c:70-87
        int lo = roff[u], hi = roff[u + 1] - 1;
        while (lo <= hi) {
            int mid = lo + (hi - lo) / 2;
            int x = raw[mid].x;
            if (x == v) return mid;
            if (x < v) lo = mid + 1; else hi = mid - 1;
        }
        return -1;
    };
    std::vector<std::pair<int, HalfEdge>> half;
    half.reserve(2 * (size_t)E);
    P.ew.assign(E, 0.0);
    P.uv.resize(E);
    for (int e = 0; e < E; e++) {
        int u = U[e], v = V[e];
        double w = W ? W[e] : 1.0;
        P.uv[e] = make_int2(u, v);
        switch (variant) {
        case DRESS_VARIANT_UNDIRECTED:
            half.push_back({u, HalfEdge{v, e, 2.0 * w}});
            half.push_back({v, HalfEdge{u, e, 2.0 * w}});
            P.ew[e] = 2.0 * w;
            break;
        case DRESS_VARIANT_DIRECTED: {
            int r = raw_find(v, u);
            double wb = r >= 0 ? w + raw[r].w : w;
            half.push_back({u, HalfEdge{v, e, wb}});
            if (r < 0) half.push_back({v, HalfEdge{u, e, w}});
            P.ew[e] = wb;
            if (r >= 0) {
                // the run of v -> u edges around r: a subgraph's lookup can
                // stop at another of them (dress.c:70-87 on a shorter row)
                bool differ = false;
                for (int j = r - 1; j >= roff[v] && raw[j].x == u; j--) differ = differ || raw[j].w != raw[r].w;
                for (int j = r + 1; j < roff[v + 1] && raw[j].x == u; j++) differ = differ || raw[j].w != raw[r].w;
                if (differ) {
                    P.amb.push_back(e);
                    P.amb_w.push_back(w);
                }
            }
            break;
        }
        case DRESS_VARIANT_FORWARD:
            half.push_back({u, HalfEdge{v, e, w}});
            P.ew[e] = w;
            break;
        default:
            half.push_back({v, HalfEdge{u, e, w}});
            P.ew[e] = w;
            break;
        }
    }
    if (!P.amb.empty()) {
        P.raw_off = roff;
        P.raw_x.resize(raw.size());
        P.raw_w.resize(raw.size());
        for (size_t i = 0; i < raw.size(); i++) { P.raw_x[i] = raw[i].x; P.raw_w[i] = raw[i].w; }
    }
    P.S = (int)half.size();
    P.row_off.assign(N + 1, 0);
    for (auto &h : half) P.row_off[h.first + 1]++;
    for (int u = 0; u < N; u++) P.row_off[u + 1] += P.row_off[u];
    std::vector<HalfEdge> rows(P.S);
    {
        std::vector<int> fill(P.row_off.begin(), P.row_off.end() - 1);
        for (auto &h : half) rows[fill[h.first]++] = h.second;
    }
    for (int u = 0; u < N; u++) std::sort(rows.begin() + P.row_off[u], rows.begin() + P.row_off[u + 1], by_xe);
    P.adj_x.resize(P.S);
    P.adj_e.resize(P.S);
    for (int i = 0; i < P.S; i++) { P.adj_x[i] = rows[i].x; P.adj_e[i] = rows[i].e; }
    {
        // a wave's lanes walk 64 consecutive rows together: column by
        // column their edge ids are consecutive words
        int maxdeg = 0;
        for (int u = 0; u < N; u++) maxdeg = std::max(maxdeg, P.row_off[u + 1] - P.row_off[u]);
        const int stride = (N + 63) & ~63;
        P.ell.clear();
        P.ell_stride = 0;
        // (parents of up to 1,024 edges are read from the LDS copy of their
        // rows, which the batched kernel makes when the launch still holds
        // 16 waves per CU: no column copy for them)
        if (E > 1024 && maxdeg > 0 && (int64_t)stride * maxdeg <= 2 * (int64_t)P.S) {
            P.ell.assign((size_t)stride * maxdeg, -1);
            for (int u = 0; u < N; u++)
                for (int i = P.row_off[u]; i < P.row_off[u + 1]; i++)
                    P.ell[(size_t)(i - P.row_off[u]) * stride + u] = P.adj_e[i];
            P.ell_stride = stride;
        }
    }
    // intersections (dress.c:307-363 pairing of repeated neighbours)
    P.ioff.assign(E + 1, 0);
    P.pairs.clear();
    for (int e = 0; e < E; e++) {
        int a = P.row_off[U[e]], ae = P.row_off[U[e] + 1];
        int b = P.row_off[V[e]], be = P.row_off[V[e] + 1];
        while (a < ae && b < be) {
            int x = P.adj_x[a], y = P.adj_x[b];
            if (x == y) {
                P.pairs.push_back((uint32_t)P.adj_e[a] | ((uint32_t)P.adj_e[b] << 16));
                a++;
                b++;
            } else if (x < y) a++;
            else b++;
        }
        P.ioff[e + 1] = (int)P.pairs.size();
    }
    P.T = (int)P.pairs.size();
    return true;
}

// DIRECTED delta^k: the variant weights that differ between the subgraph
// without the k vertices in del and the parent.  The subgraph's weight of
// u -> v is w + the weight of the v -> u edge its own lookup finds
// (dress.c:70-87, dress.c:239-246) in v's raw row without the deleted
// vertices; only edges in P.amb can change.  Appends (edge, weight).
void directed_patches(const HostParent &P, const int *del, int k, std::vector<int> &pe, std::vector<double> &pw)
{
    auto gone = [&](int x) {
        for (int j = 0; j < k; j++)
            if (del[j] == x) return true;
        return false;
    };
    std::vector<int> row;
    for (size_t i = 0; i < P.amb.size(); i++) {
        const int e = P.amb[i], u = P.uv[(size_t)e].x, v = P.uv[(size_t)e].y;
        if (gone(u) || gone(v)) continue;
        const int b = P.raw_off[(size_t)v], en = P.raw_off[(size_t)v + 1];
        row.clear();
        bool shorter = false;
        for (int j = b; j < en; j++) {
            if (gone(P.raw_x[(size_t)j])) shorter = true;
            else row.push_back(j);
        }
        if (!shorter) continue;
        int lo = 0, hi = (int)row.size() - 1, r = -1;
        while (lo <= hi) {
            int mid = lo + (hi - lo) / 2;
            int x = P.raw_x[(size_t)row[(size_t)mid]];
            if (x == u) { r = row[(size_t)mid]; break; }
            if (x < u) lo = mid + 1; else hi = mid - 1;
        }
        if (r < 0) continue;                 // not reached: u's edges stay
        const double wb = P.amb_w[i] + P.raw_w[(size_t)r];
        if (wb != P.ew[(size_t)e]) {
            pe.push_back(e);
            pw.push_back(wb);
        }
    }
}

// per-item weight patches (CSR over items); empty when no item has any
struct EwPatches {
    std::vector<int64_t> off;
    std::vector<int> e;
    std::vector<double> w;
    bool any() const { return !e.empty(); }
};

EwPatches delta_patches(const HostParent &P, const std::vector<int> &verts, int k, size_t n)
{
    EwPatches R;
    if (P.amb.empty() || k <= 0) return R;
    R.off.assign(n + 1, 0);
    for (size_t i = 0; i < n; i++) {
        directed_patches(P, &verts[i * (size_t)k], k, R.e, R.w);
        R.off[i + 1] = (int64_t)R.e.size();
    }
    if (R.e.empty()) R.off.clear();
    return R;
}

// ------------------------------------------------------------ enumeration

int host_rand(unsigned int *s)  // delta_dress_impl.h:21-25
{
    *s = *s * 1103515245u + 12345u;
    return (int)((*s >> 16) & 0x7fff);
}

// delta_dress_impl.c:215-219, 296-316: sorted, deduplicated sample picks
std::vector<int64_t> sample_picks(int64_t per, int64_t n, unsigned int seed, int offset)
{
    std::vector<int64_t> p((size_t)n);
    unsigned int s = seed + (unsigned int)offset;
    for (int64_t i = 0; i < n; i++) {
        int64_t hi = (int64_t)host_rand(&s);
        int64_t lo = (int64_t)host_rand(&s);
        p[(size_t)i] = (((hi << 31) | lo) & 0x7FFFFFFFFFFFFFFFLL) % per;
    }
    std::sort(p.begin(), p.end());
    p.erase(std::unique(p.begin(), p.end()), p.end());
    return p;
}

bool next_combo(std::vector<int> &c, int N)
{
    int k = (int)c.size(), i = k - 1;
    while (i >= 0 && c[i] == N - k + i) i--;
    if (i < 0) return false;
    c[i]++;
    for (int j = i + 1; j < k; j++) c[j] = c[j - 1] + 1;
    return true;
}

bool next_tuple(std::vector<int> &t, int N, std::vector<char> &used)
{
    int k = (int)t.size();
    std::fill(used.begin(), used.end(), 0);
    for (int v : t) used[v] = 1;
    for (int i = k - 1; i >= 0; i--) {
        used[t[i]] = 0;
        int x = t[i] + 1;
        while (x < N && used[x]) x++;
        if (x < N) {
            t[i] = x;
            used[x] = 1;
            int y = 0;
            for (int j = i + 1; j < k; j++) {
                while (used[y]) y++;
                t[j] = y;
                used[y] = 1;
            }
            return true;
        }
    }
    return false;
}

// selected subgraph indices and their vertex tuples
void enumerate_items(int kind, int N, int k, int n_samples, unsigned int seed, int offset, int stride,
                     int64_t total, std::vector<int64_t> &sel, std::vector<int> &verts)
{
    int64_t per = (total + stride - 1 - offset) / stride;
    if (per < 0) per = 0;
    int64_t limit = (n_samples == 0) ? per : std::min<int64_t>(n_samples, per);
    bool sampling = n_samples > 0 && limit < per;
    std::vector<int64_t> picks;
    if (sampling) picks = sample_picks(per, limit, seed, offset);
    std::vector<int> cur(k);
    for (int i = 0; i < k; i++) cur[i] = i;
    std::vector<char> used(N + 1, 0);
    int64_t s = 0, my = 0;
    size_t pi = 0;
    if (total <= 0) return;
    do {
        if (s % stride == offset) {
            bool take = !sampling || (pi < picks.size() && my == picks[pi]);
            if (take && sampling) pi++;
            if (take) {
                sel.push_back(s);
                verts.insert(verts.end(), cur.begin(), cur.end());
            }
            my++;
            if (sampling && pi >= picks.size()) break;
        }
        s++;
    } while (kind == LIFT_NABLA ? next_tuple(cur, N, used) : next_combo(cur, N));
}

// ------------------------------------------------------------ device pools

// Host arrays bound for the device in one copy: packed at 256-byte offsets
// into one staging block (pinned when the pool gives one) and one device
// block.  The arrays must stay alive until run(); the device pointers are
// set by run().
struct PackedUpload {
    struct Part {
        std::vector<std::pair<const void *, size_t>> pieces;     // copied one after another
        size_t at;
        std::function<void(char *)> set;
    };
    std::vector<Part> parts;
    size_t total = 0;
    template <typename T>
    void add(T **dst, const std::vector<T> &v)
    {
        add_concat(dst, std::vector<const std::vector<T> *>{&v});
    }
    // one device array holding the vectors vs one after another
    template <typename T>
    void add_concat(T **dst, const std::vector<const std::vector<T> *> &vs)
    {
        const size_t at = total;
        Part p;
        p.at = at;
        size_t bytes = 0;
        for (const std::vector<T> *v : vs) {
            p.pieces.push_back({v->data(), v->size() * sizeof(T)});
            bytes += v->size() * sizeof(T);
        }
        p.set = [dst, at](char *b) { *dst = reinterpret_cast<T *>(b + at); };
        parts.push_back(std::move(p));
        total += (bytes + 256) & ~(size_t)255;        // at least one byte past the end, as upload() gives
    }
    char *stage = nullptr;
    // host side: the arrays copied into the staging block (after this the
    // source arrays may go)
    bool pack()
    {
        stage = (char *)pinned_alloc(total);
        if (!stage) {
            set_error("host staging of the batch inputs failed");
            return false;
        }
        for (Part &p : parts) {
            size_t off = p.at;
            for (auto &pc : p.pieces) {
                if (pc.second) memcpy(stage + off, pc.first, pc.second);
                off += pc.second;
            }
        }
        return true;
    }
    // device side: the block allocated, the pointers set and the copy
    // queued on s; the staging block is released by release() once s has
    // passed the copy
    bool upload(char **blk, hipStream_t s)
    {
        if (!dalloc(blk, total + 256, "batch inputs")) return false;
        for (Part &p : parts) p.set(*blk);
        if (hipMemcpyAsync(*blk, stage, total, hipMemcpyHostToDevice, s) != hipSuccess) {
            set_error("upload of the batch inputs failed");
            return false;
        }
        return true;
    }
    void release()
    {
        pinned_free(stage);
        stage = nullptr;
    }
    ~PackedUpload() { release(); }
    bool run(char **blk, hipStream_t s)
    {
        bool ok = pack() && upload(blk, s);
        if (ok && hipStreamSynchronize(s) != hipSuccess) {
            set_error("upload of the batch inputs failed");
            ok = false;
        }
        release();
        return ok;
    }
};

// the parents' arrays on the device: views into one block
struct DevPools {
    char *blk = nullptr;
    LiftParent *parents = nullptr;
    int *parent_graph = nullptr;
    int *row_off = nullptr, *adj_e = nullptr, *ioff = nullptr, *ell = nullptr;
    int2 *uv = nullptr;
    double *ew = nullptr;
    uint32_t *pairs = nullptr;
    void release()
    {
        dfree(blk);
        *this = DevPools();
    }
};

template <typename T>
bool upload(T **dst, const std::vector<T> &src, hipStream_t s, const char *what)
{
    if (!dalloc(dst, src.size() + 1, what)) return false;
    if (!src.empty()) DH_CHECK(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return true;
}

// registers the parents' arrays with pk, each kind concatenated over the
// parents (DIRECTED patched weight copies after the parents' own weights),
// with the per-parent offsets (meta) and graph_of, for D
void pack_parents(const std::vector<const HostParent *> &ps, const std::vector<int> &graph_of,
                  std::vector<LiftParent> &meta, DevPools &D, PackedUpload &pk,
                  const std::vector<double> *extra_ew = nullptr)
{
    meta.clear();
    std::vector<const std::vector<int> *> row_off, adj_e, ioff, ell;
    std::vector<const std::vector<int2> *> uv;
    std::vector<const std::vector<double> *> ew;
    std::vector<const std::vector<uint32_t> *> pairs;
    int64_t nr = 0, na = 0, ne = 0, ni = 0, np = 0, nl = 0;
    for (const HostParent *P : ps) {
        LiftParent m;
        m.N = P->N; m.E = P->E; m.S = P->S; m.T = P->T;
        m.row_ofs = nr;
        m.adj_ofs = na;
        m.edge_ofs = ne;
        m.ioff_ofs = ni;
        m.pair_ofs = np;
        m.directed = P->directed;
        m.ell_ofs = P->ell.empty() ? -1 : nl;
        m.ell_stride = P->ell_stride;
        nl += (int64_t)P->ell.size();
        ell.push_back(&P->ell);
        // uniform variant weights (unweighted UNDIRECTED, FORWARD, BACKWARD):
        // the kernel takes the constant instead of reading a weight per term
        m.uni_w = 1;
        m.w_c = P->ew.empty() ? 0.0 : P->ew[0];
        for (double w : P->ew)
            if (std::memcmp(&w, &m.w_c, sizeof(double)) != 0) { m.uni_w = 0; break; }
        meta.push_back(m);
        nr += (int64_t)P->row_off.size(); na += (int64_t)P->adj_e.size(); ne += (int64_t)P->uv.size();
        ni += (int64_t)P->ioff.size(); np += (int64_t)P->pairs.size();
        row_off.push_back(&P->row_off);
        adj_e.push_back(&P->adj_e);
        uv.push_back(&P->uv);
        ew.push_back(&P->ew);
        ioff.push_back(&P->ioff);
        pairs.push_back(&P->pairs);
    }
    if (extra_ew) ew.push_back(extra_ew);
    pk.add(&D.parents, meta);
    pk.add(&D.parent_graph, graph_of);
    pk.add_concat(&D.row_off, row_off);
    pk.add_concat(&D.adj_e, adj_e);
    pk.add_concat(&D.ell, ell);
    pk.add_concat(&D.uv, uv);
    pk.add_concat(&D.ew, ew);
    pk.add_concat(&D.ioff, ioff);
    pk.add_concat(&D.pairs, pairs);
}

// ------------------------------------------------------------ histogram

__global__ void k_zero_int(int *p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}
__global__ void k_zero_u64(unsigned long long *p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// ascending sort keys of doubles (the bit pattern, sign-folded), and each
// block's smallest and largest key into part[2b], part[2b + 1]; zero[0] and
// zero[1] (when given) are cleared
constexpr int RANGE_BLOCKS = 1024;
__global__ void __launch_bounds__(TPB) k_sort_keys(const double *v, int64_t n, uint64_t *key,
                                                   unsigned long long *part, int64_t *zero)
{
    if (zero && blockIdx.x == 0 && threadIdx.x < 2) zero[threadIdx.x] = 0;
    unsigned long long lo = ~0ull, hi = 0ull;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t b = (uint64_t)__double_as_longlong(v[i]);
        const uint64_t k = (b >> 63) ? ~b : (b | 0x8000000000000000ull);
        key[i] = k;
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(lo, o), y = __shfl_xor(hi, o);
        lo = x < lo ? x : lo;
        hi = y > hi ? y : hi;
    }
    __shared__ unsigned long long slo[TPB / 64], shi[TPB / 64];
    if ((threadIdx.x & 63) == 0) {
        slo[threadIdx.x >> 6] = lo;
        shi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < TPB / 64; w++) {
            lo = slo[w] < lo ? slo[w] : lo;
            hi = shi[w] > hi ? shi[w] : hi;
        }
        part[2 * blockIdx.x] = lo;
        part[2 * blockIdx.x + 1] = hi;
    }
}

// range[0], range[1] = the smallest and largest of nb block results (one
// workgroup)
__global__ void __launch_bounds__(TPB) k_key_range(const unsigned long long *part, int nb, unsigned long long *range)
{
    unsigned long long lo = ~0ull, hi = 0ull;
    for (int i = threadIdx.x; i < nb; i += TPB) {
        lo = part[2 * i] < lo ? part[2 * i] : lo;
        hi = part[2 * i + 1] > hi ? part[2 * i + 1] : hi;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(lo, o), y = __shfl_xor(hi, o);
        lo = x < lo ? x : lo;
        hi = y > hi ? y : hi;
    }
    __shared__ unsigned long long slo[TPB / 64], shi[TPB / 64];
    if ((threadIdx.x & 63) == 0) {
        slo[threadIdx.x >> 6] = lo;
        shi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < TPB / 64; w++) {
            lo = slo[w] < lo ? slo[w] : lo;
            hi = shi[w] > hi ? shi[w] : hi;
        }
        range[0] = lo;
        range[1] = hi;
    }
}

// run starts: flag where the key or the segment changes
__global__ void k_run_flags(const uint64_t *key, const int *seg, int64_t n, int *flag)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = (i == 0 || key[i] != key[i - 1] || (seg && seg[i] != seg[i - 1])) ? 1 : 0;
}

__global__ void k_run_emit(const uint64_t *key, const int *seg, const int *flag, const int64_t *pos, int64_t n,
                           double *uval, int64_t *ustart, int *useg)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[i]) continue;
        uint64_t k = key[i];
        uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
        int64_t j = pos[i];
        uval[j] = __longlong_as_double((long long)b);
        ustart[j] = i;
        if (useg) useg[j] = seg[i];
    }
}

// the segment holding value i: the last g in [lo, hi] with seg_ofs[g] <= i
__device__ __forceinline__ int segment_in(const int64_t *seg_ofs, int lo, int hi, int64_t i)
{
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (seg_ofs[m] <= i) lo = m;
        else hi = m - 1;
    }
    return lo;
}
__device__ __forceinline__ int segment_of(const int64_t *seg_ofs, int nseg, int64_t i)
{
    return segment_in(seg_ofs, 0, nseg - 1, i);
}

// one 64-bit key per value ordered as (segment, value): the segment above
// bit vbits, the value's key minus the smallest below it.  Up to
// PACK_LDS_SEGS segment offsets are searched in LDS.
constexpr int PACK_LDS_SEGS = 4096;
// vbits < 0: the bits of the key span, read from range on the device; when
// they and the sbits of the segment ids exceed 64, *bad is set and nothing
// is written (the caller takes another layout)
__device__ __forceinline__ int span_bits(const unsigned long long *range)
{
    const unsigned long long span = range[1] - range[0];
    return span ? 64 - __clzll((long long)span) : 0;
}
__global__ void __launch_bounds__(TPB) k_pack_keys(const uint64_t *key, const int64_t *seg_ofs, int nseg, int64_t n,
                                                   const unsigned long long *range, int vbits, uint64_t *out,
                                                   int sbits, int *bad)
{
    if (vbits < 0) {
        vbits = span_bits(range);
        if (vbits + sbits > 64) {
            if (blockIdx.x == 0 && threadIdx.x == 0) *bad = 1;
            return;
        }
    }
    extern __shared__ int64_t s_ofs[];
    const bool lds = nseg < PACK_LDS_SEGS;
    if (lds) {
        for (int i = threadIdx.x; i <= nseg; i += TPB) s_ofs[i] = seg_ofs[i];
        __syncthreads();
    }
    const int64_t *so = lds ? s_ofs : seg_ofs;
    const uint64_t lo = range[0];
    // a contiguous chunk per block: its values lie in the segments between
    // those of its first and last value
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t c0 = blockIdx.x * chunk, c1 = c0 + chunk < n ? c0 + chunk : n;
    if (c0 >= c1) return;
    const int g0 = nseg > 1 ? segment_of(so, nseg, c0) : 0, g1 = nseg > 1 ? segment_of(so, nseg, c1 - 1) : 0;
    for (int64_t i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
        const uint64_t seg = (uint64_t)segment_in(so, g0, g1, i);
        out[i] = (vbits < 64 ? seg << vbits : 0) | (key[i] - lo);
    }
}

// the runs' packed keys back to values, in place, with their segments
__global__ void k_unpack_runs(uint64_t *ukey, int64_t runs, const unsigned long long *range, int vbits, int *useg)
{
    const uint64_t lo = range[0];
    const uint64_t mask = vbits < 64 ? (((uint64_t)1 << vbits) - 1) : ~(uint64_t)0;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < runs; j += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = ukey[j];
        const uint64_t k = (key & mask) + lo;
        ukey[j] = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
        useg[j] = vbits < 64 ? (int)(key >> vbits) : 0;
    }
}

// The runs' packed keys (key layout from the key range, run count from
// the device) to values, into val, each segment's first run into ofs
// (nseg + 1 entries, as k_seg_run_starts), and the largest count into
// *maxc (cleared before); nothing after a layout failure
__global__ void k_runs_out(const uint64_t *ukey, const int64_t *cnt, const unsigned long long *range,
                           const int64_t *runs_dev, const int *bad, int nseg, double *val, int64_t *ofs,
                           unsigned long long *maxc)
{
    if (*bad) return;
    const int64_t runs = *runs_dev;
    const int vbits = span_bits(range);
    const uint64_t lo = range[0];
    const uint64_t mask = vbits < 64 ? (((uint64_t)1 << vbits) - 1) : ~(uint64_t)0;
    unsigned long long most = 0;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < runs; j += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = ukey[j];
        const uint64_t k = (key & mask) + lo;
        val[j] = __longlong_as_double((long long)((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k));
        const int g = vbits < 64 ? (int)(key >> vbits) : 0;
        const int prev = j ? (vbits < 64 ? (int)(ukey[j - 1] >> vbits) : 0) : -1;
        for (int x = prev + 1; x <= g; x++) ofs[x] = j;
        if (j == runs - 1)
            for (int x = g + 1; x <= nseg; x++) ofs[x] = runs;
        const unsigned long long c = (unsigned long long)cnt[j];
        most = c > most ? c : most;
    }
    // one atomic per workgroup: one per wave on the one address cost
    // 100 us a group of 1.2M runs
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(most, o);
        most = x > most ? x : most;
    }
    __shared__ unsigned long long s_most[TPB / 64];
    if ((threadIdx.x & 63) == 0) s_most[threadIdx.x >> 6] = most;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) most = s_most[w] > most ? s_most[w] : most;
        if (most) atomicMax(maxc, most);
    }
}

// bytes per count that hold every count up to most
__host__ __device__ inline int count_width(unsigned long long most)
{
    return most <= 0xffull ? 1 : most <= 0xffffull ? 2 : most <= 0xffffffffull ? 4 : 8;
}

// the counts narrowed to count_width(*maxc) bytes into out (nothing to do
// at 8 bytes, or after a layout failure)
__global__ void k_pack_counts(const int64_t *cnt, const int64_t *runs_dev, const int *bad,
                              const unsigned long long *maxc, void *out)
{
    if (*bad) return;
    const int w = count_width(*maxc);
    if (w == 8) return;
    const int64_t runs = *runs_dev;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < runs; j += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t c = (uint64_t)cnt[j];
        if (w == 1) static_cast<uint8_t *>(out)[j] = (uint8_t)c;
        else if (w == 2) static_cast<uint16_t *>(out)[j] = (uint16_t)c;
        else static_cast<uint32_t *>(out)[j] = (uint32_t)c;
    }
}

// One tile of a round of pairwise merges of sorted runs: outputs [o0, o1)
// (relative to a0) of the merge of A = in[a0, b0) and B = in[b0, b1), written
// to out[a0 + o0, a0 + o1); an empty B copies A.  Ties take A first.
struct MergeTile {
    int64_t a0, b0, b1;
    int o0, o1;
};
constexpr int MERGE_TPB = 128, MERGE_PER_THREAD = 8, MERGE_TILE = MERGE_TPB * MERGE_PER_THREAD;

// number of A elements among the first d outputs of the merge of a[0, la)
// and b[0, lb)
template <typename P>
__device__ __forceinline__ int64_t merge_split(P a, int64_t la, P b, int64_t lb, int64_t d)
{
    int64_t lo = d > lb ? d - lb : 0, hi = d < la ? d : la;
    while (lo < hi) {
        const int64_t m = (lo + hi) >> 1;
        if (a[m] <= b[d - 1 - m]) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__global__ void __launch_bounds__(MERGE_TPB) k_merge_tiles(const uint64_t *in, uint64_t *out, const MergeTile *tiles)
{
    __shared__ uint64_t s[MERGE_TILE];
    __shared__ int64_t s_ia[2];
    const MergeTile T = tiles[blockIdx.x];
    const int64_t la = T.b0 - T.a0, lb = T.b1 - T.b0;
    if (threadIdx.x < 2)
        s_ia[threadIdx.x] = merge_split(in + T.a0, la, in + T.b0, lb, threadIdx.x == 0 ? T.o0 : T.o1);
    __syncthreads();
    // (clamped: on input that is not sorted runs the splits need not be
    // monotone, and the tile must still read and write in bounds)
    const int64_t ia0 = s_ia[0], ia1 = s_ia[1];
    const int na = (int)min(max(ia1 - ia0, (int64_t)0), (int64_t)(T.o1 - T.o0)), nb = (T.o1 - T.o0) - na;
    const int64_t ib0 = T.o0 - ia0;
    for (int i = threadIdx.x; i < na; i += MERGE_TPB) s[i] = in[T.a0 + ia0 + i];
    for (int i = threadIdx.x; i < nb; i += MERGE_TPB) s[na + i] = in[T.b0 + ib0 + i];
    __syncthreads();
    const int n = na + nb;
    const int d0 = min(n, (int)threadIdx.x * MERGE_PER_THREAD), d1 = min(n, d0 + MERGE_PER_THREAD);
    int ia = (int)merge_split(s, na, s + na, nb, d0), ib = d0 - ia;
    uint64_t *o = out + T.a0 + T.o0;
    for (int d = d0; d < d1; d++) {
        const bool take_a = ib >= nb || (ia < na && s[ia] <= s[na + ib]);
        o[d] = take_a ? s[ia++] : s[na + ib++];
    }
}

// Tiles of the rounds that merge the sorted runs of every segment into one:
// runs end at run_end (ascending, the last = the values), segments at
// seg_ofs (each segment boundary is a run boundary).  Each round merges
// neighbouring runs of a segment in pairs (a lone last run is copied), so a
// segment of r runs is one after ceil(log2 r) rounds; segments already done
// are copied, keeping every round's output whole.  emit(tile) receives the
// tiles in order; round_ofs[i] is the number emitted before round i (and
// the last entry the total).
template <typename Emit>
void merge_rounds(const std::vector<int64_t> &run_end, const std::vector<int64_t> &seg_ofs, Emit &&emit,
                  std::vector<size_t> &round_ofs)
{
    round_ofs.assign(1, 0);
    const size_t nseg = seg_ofs.size() - 1;
    // each segment's run boundaries at bnd[first[g], first[g] + cnt[g])
    std::vector<int64_t> bnd;
    bnd.reserve(run_end.size() + 2 * nseg);
    std::vector<size_t> first(nseg);
    std::vector<int> cnt(nseg);
    size_t r = 0;
    int most = 0;
    for (size_t g = 0; g < nseg; g++) {
        first[g] = bnd.size();
        bnd.push_back(seg_ofs[g]);
        while (r < run_end.size() && run_end[r] <= seg_ofs[g + 1]) {
            if (run_end[r] > bnd.back()) bnd.push_back(run_end[r]);
            r++;
        }
        if (bnd.back() < seg_ofs[g + 1]) bnd.push_back(seg_ofs[g + 1]);
        cnt[g] = (int)(bnd.size() - first[g]);
        most = std::max(most, cnt[g]);
    }
    size_t emitted = 0;
    for (; most > 2; most = (most + 2) / 2) {
        for (size_t g = 0; g < nseg; g++) {
            int64_t *b = bnd.data() + first[g];
            const int c = cnt[g];
            int w = 1;                 // boundaries of the next round, written behind the reads
            for (int j = 0; j + 1 < c; j += 2) {
                const int64_t a0 = b[j], b0 = b[j + 1], b1 = j + 2 < c ? b[j + 2] : b[j + 1];
                for (int64_t o = 0; o < b1 - a0; o += MERGE_TILE) {
                    emit(MergeTile{a0, b0, b1, (int)o, (int)std::min<int64_t>(b1 - a0, o + MERGE_TILE)});
                    emitted++;
                }
                b[w++] = b1;
            }
            cnt[g] = w;
        }
        round_ofs.push_back(emitted);
    }
}

// seg[i] = the segment holding value i (the last g with seg_ofs[g] <= i),
// one thread per value
__global__ void k_fill_seg(const int64_t *seg_ofs, int nseg, int64_t n, int *seg)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        seg[i] = segment_of(seg_ofs, nseg, i);
}

// run lengths from run starts (ustart ascending, n values in all)
__global__ void k_run_counts(const int64_t *ustart, int64_t runs, int64_t n, int64_t *cnt)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < runs; j += (int64_t)gridDim.x * blockDim.x)
        cnt[j] = (j + 1 < runs ? ustart[j + 1] : n) - ustart[j];
}

// first run of every segment: ofs[g] for g in [0, nseg], runs ordered by
// segment (segments without runs get the next run's index)
__global__ void k_seg_run_starts(const int *useg, int64_t runs, int nseg, int64_t *ofs)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < runs; j += (int64_t)gridDim.x * blockDim.x) {
        int prev = j ? useg[j - 1] : -1;
        for (int g = prev + 1; g <= useg[j]; g++) ofs[g] = j;
        if (j == runs - 1)
            for (int g = useg[j] + 1; g <= nseg; g++) ofs[g] = runs;
    }
}

// the largest of cnt[0, runs) into *maxc (cleared before): one atomic per
// workgroup
__global__ void __launch_bounds__(TPB) k_count_max(const int64_t *cnt, int64_t runs, unsigned long long *maxc)
{
    unsigned long long most = 0;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < runs; j += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long c = (unsigned long long)cnt[j];
        most = c > most ? c : most;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(most, o);
        most = x > most ? x : most;
    }
    __shared__ unsigned long long s_most[TPB / 64];
    if ((threadIdx.x & 63) == 0) s_most[threadIdx.x >> 6] = most;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < TPB / 64; w++) most = s_most[w] > most ? s_most[w] : most;
        if (most) atomicMax(maxc, most);
    }
}

// cnt[0, runs) narrowed to width (1, 2 or 4) bytes each into out
__global__ void k_narrow_counts(const int64_t *cnt, int64_t runs, int width, void *out)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < runs; j += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t c = (uint64_t)cnt[j];
        if (width == 1) static_cast<uint8_t *>(out)[j] = (uint8_t)c;
        else if (width == 2) static_cast<uint16_t *>(out)[j] = (uint16_t)c;
        else static_cast<uint32_t *>(out)[j] = (uint32_t)c;
    }
}

// (value bits, count) pairs, the layout of dress_hist_pair_t and of the
// flat histogram arrays
__global__ void k_interleave(const double *v, const int64_t *c, int64_t runs, int64_t *flat)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < runs; j += (int64_t)gridDim.x * blockDim.x) {
        flat[2 * j] = __double_as_longlong(v[j]);
        flat[2 * j + 1] = c[j];
    }
}

// Exact (value, count) histogram of vals[0..n) per segment (segments are
// contiguous ranges, seg_ofs has nseg+1 entries), on the device.  When the
// values' sort keys span vbits bits and the segment ids fit in the bits
// above, one 64-bit key per value holds (segment, value): one key sort and
// a run-length encoding give the runs.  Otherwise the values are sorted
// with their segment ids along, then stably by segment, and the runs are
// found by flags and a scan.  Leaves d_val / d_cnt (runs entries, ascending
// by (segment, value)) on the device for the caller to download and dfree;
// run_ofs[g] is the first run of segment g (nseg+1 entries).
bool exact_histogram_dev(const double *vals, int64_t n, const std::vector<int64_t> &seg_ofs, hipStream_t s,
                         double **d_val, int64_t **d_cnt, int64_t *runs_out, std::vector<int64_t> &run_ofs,
                         const std::vector<int64_t> *sorted_runs = nullptr)
{
    PhaseTimer pt("  exact_histogram_dev");
    *d_val = nullptr;
    *d_cnt = nullptr;
    *runs_out = 0;
    const int nseg = (int)seg_ofs.size() - 1;
    run_ofs.assign((size_t)nseg + 1, 0);
    if (n <= 0) return true;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    MergeTile *dtiles = nullptr, *tile_host = nullptr;   // merge tiles, uploaded from a pinned block
    unsigned long long *part = nullptr, *range = nullptr;
    int *g0 = nullptr, *g1 = nullptr, *flag = nullptr, *usg = nullptr;
    int64_t *pos = nullptr, *ustart = nullptr, *dofs = nullptr, *cnt = nullptr, *rofs = nullptr, *druns = nullptr;
    double *uval = nullptr;
    unsigned long long *host = (unsigned long long *)snapshot_take();   // pinned read-backs
    bool ok = host && dalloc(&k0, (size_t)n, "histogram keys") && dalloc(&k1, (size_t)n, "histogram keys") &&
              dalloc(&flag, (size_t)n + 1, "run flags") && dalloc(&dofs, (size_t)nseg + 1, "segment offsets") &&
              dalloc(&rofs, (size_t)nseg + 1, "run offsets") && dalloc(&part, 2 * RANGE_BLOCKS, "key range") &&
              dalloc(&range, 2, "key range") && dalloc(&druns, 1, "run count");
    pt.mark("alloc");
    const int grid = grid_for(n, num_cus() * 8);
    const int rgrid = grid_for(n, RANGE_BLOCKS);
    int64_t runs = 0;
    int vbits = -1, sbits = 0;
    while ((1 << sbits) < nseg) sbits++;
    if (ok) {
        hipLaunchKernelGGL(k_sort_keys, dim3(rgrid), dim3(TPB), 0, s, vals, n, k0, part, (int64_t *)nullptr);
        hipLaunchKernelGGL(k_key_range, dim3(1), dim3(TPB), 0, s, part, rgrid, range);
        ok = hipMemcpyAsync(dofs, seg_ofs.data(), seg_ofs.size() * sizeof(int64_t), hipMemcpyHostToDevice, s) ==
                 hipSuccess &&
             hipMemcpyAsync(host, range, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
        const char *pe = getenv("DRESS_HIP_HIST_PACKED");
        const unsigned long long span = host[1] - host[0];
        const int vb = span ? 64 - __builtin_clzll(span) : 0;
        if (ok && vb + sbits <= 64 && n <= (int64_t)0xffffffffll && !(pe && !strcmp(pe, "0"))) vbits = vb;
    }
    pt.mark("key range");
    if (ok && vbits >= 0) {
        const size_t shm = nseg < PACK_LDS_SEGS ? ((size_t)nseg + 1) * sizeof(int64_t) : 0;
        hipLaunchKernelGGL(k_pack_keys, dim3(grid), dim3(TPB), shm, s, k0, dofs, nseg, n, range, vbits, k1, 0,
                           (int *)nullptr);
        if (pt.on) { hipStreamSynchronize(s); pt.mark("pack"); }
        // values that arrive as sorted runs (each segment's runs merged in
        // ceil(log2 runs) rounds) or in any order (one radix sort)
        uint64_t *sorted = k1, *other = k0;
        std::vector<size_t> round_ofs;
        if (sorted_runs) merge_rounds(*sorted_runs, seg_ofs, [](const MergeTile &) {}, round_ofs);
        if (sorted_runs && round_ofs.size() <= 9) {
            // the tiles written straight into a pinned block (pooled, so
            // already mapped), then one copy
            const size_t nt = round_ofs.back();
            if (nt) {
                tile_host = (MergeTile *)pinned_alloc(nt * sizeof(MergeTile));
                ok = tile_host && dalloc(&dtiles, nt, "merge tiles");
                if (!tile_host) set_error("host staging of the merge tiles failed");
                size_t at = 0;
                if (ok) merge_rounds(*sorted_runs, seg_ofs, [&](const MergeTile &t) { tile_host[at++] = t; }, round_ofs);
                ok = ok && hipMemcpyAsync(dtiles, tile_host, nt * sizeof(MergeTile), hipMemcpyHostToDevice, s) ==
                               hipSuccess;
            }
            if (pt.on) pt.mark("merge plan");
            for (size_t r = 0; ok && r + 1 < round_ofs.size(); r++) {
                hipLaunchKernelGGL(k_merge_tiles, dim3((unsigned)(round_ofs[r + 1] - round_ofs[r])), dim3(MERGE_TPB),
                                   0, s, sorted, other, dtiles + round_ofs[r]);
                std::swap(sorted, other);
            }
            if (pt.on) { hipStreamSynchronize(s); pt.mark("merge rounds"); }
        } else {
            ok = prim_sort_u64(k1, k0, (size_t)n, s, std::max(1, vbits + sbits));
            sorted = k0;
            other = k1;
        }
        ok = ok && dalloc(&cnt, (size_t)n + 1, "histogram counts") &&
             prim_run_length_encode_u64(sorted, (size_t)n, other, cnt, druns, s) &&
             hipMemcpyAsync(host, druns, sizeof(int64_t), hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
        if (pt.on) pt.mark("sort+runs");
        if (ok) {
            runs = (int64_t)host[0];
            const int ugrid = grid_for(runs, num_cus() * 8);
            usg = flag;                                          // (flags unused on this path)
            hipLaunchKernelGGL(k_unpack_runs, dim3(ugrid), dim3(TPB), 0, s, other, runs, range, vbits, usg);
            hipLaunchKernelGGL(k_seg_run_starts, dim3(ugrid), dim3(TPB), 0, s, usg, runs, nseg, rofs);
            uval = reinterpret_cast<double *>(other);
            if (other == k0) k0 = nullptr;
            else k1 = nullptr;
            usg = nullptr;
        }
    } else if (ok) {
        ok = dalloc(&g0, (size_t)n, "segments") && dalloc(&g1, (size_t)n, "segments") &&
             dalloc(&pos, (size_t)n + 1, "run positions");
        if (ok) hipLaunchKernelGGL(k_fill_seg, dim3(grid), dim3(TPB), 0, s, dofs, nseg, n, g0);
        if (nseg > 1) {
            // by value (segment ids travel along) ...
            ok = ok && prim_sort_pairs_u64_i32(k0, k1, g0, g1, (size_t)n, 64, s);
            // ... then stably by segment
            int bits = 1;
            while ((1 << bits) <= nseg) bits++;
            ok = ok && prim_sort_pairs_i32_u64(g1, g0, k1, k0, (size_t)n, bits, s);
        } else {
            // one segment: the keys alone (g0 holds its id, 0, throughout)
            ok = ok && prim_sort_u64(k0, k1, (size_t)n, s);
            if (ok) std::swap(k0, k1);
        }
        // now k0 / g0 hold the sorted (segment, value) keys
        if (ok) {
            hipLaunchKernelGGL(k_run_flags, dim3(grid), dim3(TPB), 0, s, k0, g0, n, flag);
            hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, flag + n, 1);
            ok = prim_exclusive_scan_i32_to_i64(flag, pos, (size_t)n + 1, s) &&
                 hipMemcpyAsync(host, pos + n, sizeof(int64_t), hipMemcpyDeviceToHost, s) == hipSuccess &&
                 hipStreamSynchronize(s) == hipSuccess;
        }
        if (pt.on) pt.mark("sorts+flags+scan");
        if (ok) {
            runs = (int64_t)host[0];
            ok = dalloc(&uval, (size_t)runs + 1, "histogram values") &&
                 dalloc(&ustart, (size_t)runs + 1, "histogram runs") &&
                 dalloc(&usg, (size_t)runs + 1, "histogram segs") && dalloc(&cnt, (size_t)runs + 1, "histogram counts");
        }
        if (ok) {
            const int ugrid = grid_for(runs, num_cus() * 8);
            hipLaunchKernelGGL(k_run_emit, dim3(grid), dim3(TPB), 0, s, k0, g0, flag, pos, n, uval, ustart, usg);
            hipLaunchKernelGGL(k_run_counts, dim3(ugrid), dim3(TPB), 0, s, ustart, runs, n, cnt);
            hipLaunchKernelGGL(k_seg_run_starts, dim3(ugrid), dim3(TPB), 0, s, usg, runs, nseg, rofs);
        }
    }
    if (ok)
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(run_ofs.data(), rofs, ((size_t)nseg + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s) ==
                 hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    if (!ok && !has_error()) set_error("histogram assembly failed");
    if (!ok) (void)hipStreamSynchronize(s);     // queued kernels may still read the blocks
    pt.mark("runs");
    dfree(k0); dfree(k1); dfree(g0); dfree(g1); dfree(flag); dfree(pos); dfree(ustart); dfree(dofs);
    dfree(usg); dfree(rofs); dfree(part); dfree(range); dfree(druns); dfree(dtiles);
    pinned_free(tile_host);
    if (host) snapshot_give(host);
    pt.mark("free");
    if (!ok) {
        dfree(uval);
        dfree(cnt);
        return false;
    }
    *d_val = uval;
    *d_cnt = cnt;
    *runs_out = runs;
    return true;
}

// The exact histogram of one group of a pipelined batch, queued behind the
// group's launches without the host waiting on the device.  The values come
// as sorted runs per item, so the rounds of merges are planned on the host
// from sizes known before the launch (hist_plan, on a planner thread); the
// key layout is read from the key range on the device (k_pack_keys), and
// the run count, a layout failure and each segment's first run come back
// into the plan's pinned block.
struct HistPlan {
    int nseg = 0;
    std::vector<size_t> round_ofs;        // tiles before each round of merges
    // pinned: [0] runs, [1] failure, [2] the largest count, [3, nseg + 4)
    // each segment's first run (read back in one copy), [nseg + 4,
    // 2 nseg + 5) the segment offsets, then the tiles (uploaded in one)
    int64_t *host = nullptr;
    static constexpr int HEAD = 3;
    HistPlan() = default;
    HistPlan(const HistPlan &) = delete;
    HistPlan &operator=(const HistPlan &) = delete;
    ~HistPlan() { pinned_free(host); }
    size_t words() const { return 2 * ((size_t)nseg + 1) + HEAD; }
    MergeTile *tiles() const { return reinterpret_cast<MergeTile *>(host + words()); }
};

// 1 planned, 0 this form does not apply (more than 8 rounds of merges, more
// than 2^32 values, or packed keys switched off), -1 an error
int hist_plan(HistPlan &P, const std::vector<int64_t> &seg_ofs, const std::vector<int64_t> &item_end)
{
    const int nseg = (int)seg_ofs.size() - 1;
    const int64_t n = seg_ofs.back();
    const char *pe = getenv("DRESS_HIP_HIST_PACKED");
    if (n <= 0 || n > (int64_t)0xffffffffll || nseg < 1 || (pe && !strcmp(pe, "0"))) return 0;
    merge_rounds(item_end, seg_ofs, [](const MergeTile &) {}, P.round_ofs);
    if (P.round_ofs.size() > 9) return 0;
    P.nseg = nseg;
    // at least the pool's smallest pinned size, so that the copies are
    // asynchronous
    const size_t nt = P.round_ofs.back();
    P.host = (int64_t *)pinned_alloc(std::max<size_t>(P.words() * sizeof(int64_t) + nt * sizeof(MergeTile), 1 << 20));
    if (!P.host) {
        set_error("host staging of the histogram failed");
        return -1;
    }
    for (int i = 0; i < HistPlan::HEAD; i++) P.host[i] = 0;
    for (int g = 0; g <= nseg; g++) P.host[HistPlan::HEAD + (nseg + 1) + g] = seg_ofs[(size_t)g];
    MergeTile *t = P.tiles();
    size_t at = 0;
    merge_rounds(item_end, seg_ofs, [&](const MergeTile &x) { t[at++] = x; }, P.round_ofs);
    return 1;
}

// Every block the queued work reads is held here until the caller has
// waited for that work (hist_collect, then hist_release).
struct AsyncHist {
    int nseg = 0;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    double *val = nullptr;                // k0 or k1: the runs' values
    void *packed = nullptr;               // the other: the counts at their narrowest width
    int64_t *cnt = nullptr;
    // [0] runs, [1] layout failure, [2] the largest count, [3, nseg + 4)
    // each segment's first run: read back in one copy into the same words
    // of the plan's block
    int64_t *stat = nullptr;
    int64_t *segt = nullptr;              // segment offsets, then the tiles: uploaded in one copy
    unsigned long long *part = nullptr, *range = nullptr;
    void *scratch = nullptr;              // run-length encoding scratch
    int64_t *host = nullptr;              // the plan's pinned block
};

// keep_outputs: the values, counts and packed counts stay for the caller
// to download and free
void hist_release(AsyncHist &H, bool keep_outputs)
{
    if (!keep_outputs) {
        dfree(H.k0);
        dfree(H.k1);
        dfree(H.cnt);
    }
    dfree(H.stat); dfree(H.segt); dfree(H.part); dfree(H.range); dfree(H.scratch);
    pinned_free(H.host);
    H = AsyncHist();
}

// queues the planned histogram of vals[0, n) on s, taking the plan's
// pinned block; false on an error (then nothing is left queued)
bool hist_enqueue(AsyncHist &H, HistPlan &P, const double *vals, int64_t n, hipStream_t s)
{
    H = AsyncHist();
    const int nseg = P.nseg;
    H.nseg = nseg;
    H.host = P.host;
    P.host = nullptr;
    int sbits = 0;
    while ((1 << sbits) < nseg) sbits++;
    const size_t nt = P.round_ofs.back();
    const size_t seg_words = (size_t)nseg + 1, tile_words = nt * (sizeof(MergeTile) / sizeof(int64_t));
    static_assert(sizeof(MergeTile) % sizeof(int64_t) == 0, "tiles are whole words");
    const int64_t *h_seg = H.host + HistPlan::HEAD + seg_words;           // the tiles follow
    bool ok = h_seg[nseg] == n && dalloc(&H.k0, (size_t)n, "histogram keys") &&
              dalloc(&H.k1, (size_t)n, "histogram keys") &&
              dalloc(&H.stat, seg_words + HistPlan::HEAD, "run offsets") &&
              dalloc(&H.segt, seg_words + tile_words, "segment offsets and merge tiles") &&
              dalloc(&H.part, 2 * RANGE_BLOCKS, "key range") && dalloc(&H.range, 2, "key range") &&
              dalloc(&H.cnt, (size_t)n + 1, "histogram counts");
    const int grid = grid_for(n, num_cus() * 8);
    const int rgrid = grid_for(n, RANGE_BLOCKS);
    int *bad = reinterpret_cast<int *>(H.stat + 1);
    const MergeTile *dtiles = reinterpret_cast<const MergeTile *>(H.segt + seg_words);
    ok = ok && hipMemcpyAsync(H.segt, h_seg, (seg_words + tile_words) * sizeof(int64_t), hipMemcpyHostToDevice, s) ==
                   hipSuccess;
    uint64_t *sorted = H.k1, *other = H.k0;
    if (ok) {
        hipLaunchKernelGGL(k_sort_keys, dim3(rgrid), dim3(TPB), 0, s, vals, n, H.k0, H.part, H.stat + 1);
        hipLaunchKernelGGL(k_key_range, dim3(1), dim3(TPB), 0, s, H.part, rgrid, H.range);
        const size_t shm = nseg < PACK_LDS_SEGS ? seg_words * sizeof(int64_t) : 0;
        hipLaunchKernelGGL(k_pack_keys, dim3(grid), dim3(TPB), shm, s, H.k0, (const int64_t *)H.segt, nseg, n,
                           H.range, -1, H.k1, sbits, bad);
        for (size_t r = 0; r + 1 < P.round_ofs.size(); r++) {
            hipLaunchKernelGGL(k_merge_tiles, dim3((unsigned)(P.round_ofs[r + 1] - P.round_ofs[r])), dim3(MERGE_TPB),
                               0, s, sorted, other, dtiles + P.round_ofs[r]);
            std::swap(sorted, other);
        }
        ok = prim_run_length_encode_u64(sorted, (size_t)n, other, H.cnt, H.stat, s, &H.scratch);
    }
    if (ok) {
        // the values go into the merged keys' block, free after the
        // encoding, and the narrowed counts into the unique keys' block
        H.val = reinterpret_cast<double *>(sorted);
        H.packed = other;
        unsigned long long *maxc = reinterpret_cast<unsigned long long *>(H.stat + 2);
        // (2 workgroups per CU: each ends with one atomic on the same word)
        hipLaunchKernelGGL(k_runs_out, dim3(grid_for(n, num_cus() * 2)), dim3(TPB), 0, s, (const uint64_t *)other,
                           (const int64_t *)H.cnt,
                           (const unsigned long long *)H.range, (const int64_t *)H.stat, (const int *)bad, nseg,
                           H.val, H.stat + HistPlan::HEAD, maxc);
        hipLaunchKernelGGL(k_pack_counts, dim3(grid), dim3(TPB), 0, s, (const int64_t *)H.cnt,
                           (const int64_t *)H.stat, (const int *)bad, (const unsigned long long *)maxc, H.packed);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(H.host, H.stat, (seg_words + HistPlan::HEAD) * sizeof(int64_t), hipMemcpyDeviceToHost,
                            s) == hipSuccess;
    }
    if (!ok) {
        if (!has_error()) set_error("histogram assembly failed");
        (void)hipStreamSynchronize(s);              // queued kernels may read the blocks
        hist_release(H, false);
    }
    return ok;
}

// after the queued work has finished: the run count, each segment's first
// run, the bytes per count of H.packed (8: use H.cnt), and whether the key
// layout failed (then none is valid)
void hist_collect(const AsyncHist &H, int64_t *runs, std::vector<int64_t> &run_ofs, int *width, bool *failed)
{
    int fail = 0;                                   // (the slot's first four bytes)
    memcpy(&fail, &H.host[1], sizeof(int));
    *failed = fail != 0;
    *runs = H.host[0];
    *width = count_width((unsigned long long)H.host[2]);
    run_ofs.assign(H.host + HistPlan::HEAD, H.host + HistPlan::HEAD + H.nseg + 1);
}

// the same, downloaded into vectors (runs of all segments; useg unused here)
bool exact_histogram(const double *vals, int64_t n, const std::vector<int64_t> &seg_ofs, hipStream_t s,
                     std::vector<double> &uvals, std::vector<int64_t> &counts, std::vector<int> &useg)
{
    uvals.clear();
    counts.clear();
    useg.clear();
    double *dv = nullptr;
    int64_t *dc = nullptr, runs = 0;
    std::vector<int64_t> rofs;
    if (!exact_histogram_dev(vals, n, seg_ofs, s, &dv, &dc, &runs, rofs)) return false;
    uvals.resize((size_t)runs);
    counts.resize((size_t)runs);
    bool ok = runs == 0 ||
              (hipMemcpyAsync(uvals.data(), dv, (size_t)runs * sizeof(double), hipMemcpyDeviceToHost, s) ==
                   hipSuccess &&
               hipMemcpyAsync(counts.data(), dc, (size_t)runs * sizeof(int64_t), hipMemcpyDeviceToHost, s) ==
                   hipSuccess &&
               hipStreamSynchronize(s) == hipSuccess);
    dfree(dv);
    dfree(dc);
    if (!ok) set_error("histogram download failed");
    return ok;
}

// The pairs of a histogram of `runs` entries are counted by an int in the
// reference API (hist_size: 2 x pairs in the flat form)
bool hist_size_fits(int64_t runs)
{
    if (runs <= (int64_t)(INT32_MAX / 2)) return true;
    set_error("the histogram has " + std::to_string(runs) +
              " distinct values, more than the int hist_size of the API can count");
    return false;
}

// Flat histogram (dress_histogram.c:184-200 layout: value bits, count, ...)
// of vals[0..n), malloc'd; NULL with *hist_size = 0 when empty.  (Faulting
// the result block in from host threads while the device worked made calls
// slower, 790 MB results 32-36 -> 71-92 ms, most likely because the faults
// contend with the runtime's own mappings for the address space.)
// Pages of a host block touched by a host thread (several from 32 MiB)
// while the caller goes on; join() waits, as does the destructor.
class BackgroundTouch {
public:
    BackgroundTouch() = default;
    BackgroundTouch(const BackgroundTouch &) = delete;
    BackgroundTouch &operator=(const BackgroundTouch &) = delete;
    ~BackgroundTouch() { join(); }
    void start(void *p, size_t bytes) { th_ = std::thread([p, bytes] { touch_pages(p, bytes); }); }
    void join()
    {
        if (th_.joinable()) th_.join();
    }

private:
    std::thread th_;
};

// A host block for a result the caller frees with free(), allocated and
// page-touched by host threads while the device computes: writing 0.79 GB
// into fresh memory otherwise takes its page faults inside the download.
// take(bytes) hands it over when it holds bytes and no more than twice
// that (else nullptr, and the block goes).
class PretouchedBlock {
public:
    PretouchedBlock() = default;
    PretouchedBlock(const PretouchedBlock &) = delete;
    PretouchedBlock &operator=(const PretouchedBlock &) = delete;
    ~PretouchedBlock()
    {
        if (th_.joinable()) th_.join();
        free(p_);
    }
    void start(size_t bytes)
    {
        p_ = host_alloc(bytes);
        bytes_ = p_ ? bytes : 0;
        if (p_) th_ = std::thread([this] { touch_pages(p_, bytes_); });
    }
    void *take(size_t bytes)
    {
        if (th_.joinable()) th_.join();
        void *q = nullptr;
        if (p_ && bytes <= bytes_ && 2 * bytes >= bytes_) std::swap(q, p_);
        free(p_);
        p_ = nullptr;
        return q;
    }

private:
    void *p_ = nullptr;
    size_t bytes_ = 0;
    std::thread th_;
};

bool histogram_flat(const double *vals, int64_t n, hipStream_t s, int64_t **flat_out, int *hist_size,
                    PretouchedBlock *pre = nullptr)
{
    *flat_out = nullptr;
    if (hist_size) *hist_size = 0;
    double *dv = nullptr;
    int64_t *dc = nullptr, runs = 0;
    std::vector<int64_t> rofs;
    if (!exact_histogram_dev(vals, n, {0, n}, s, &dv, &dc, &runs, rofs)) return false;
    bool ok = hist_size_fits(runs);
    if (ok && runs > 0) {
        int64_t *dflat = nullptr;
        void *q = pre ? pre->take((size_t)runs * 2 * sizeof(int64_t)) : nullptr;
        int64_t *flat = (int64_t *)(q ? q : host_alloc((size_t)runs * 2 * sizeof(int64_t)));
        if (!flat) set_error("host allocation of the histogram failed");
        // Large histograms whose counts fit 4 bytes come down as values and
        // narrowed counts, interleaved by host threads (delta^2 of the
        // 12-regular graph on 256 vertices, 49M distinct values: 17.7 ms
        // for the 0.79 GB of pairs through d2h_parallel).
        int width = 8;
        unsigned long long *dmax = nullptr;
        uint8_t *dpacked = nullptr;
        if (flat && runs >= ((int64_t)1 << 20) && !ctx_serving() && dalloc(&dmax, 1, "largest count")) {
            unsigned long long most = 0;
            hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, dmax, 1);
            hipLaunchKernelGGL(k_count_max, dim3(grid_for(runs, num_cus() * 2)), dim3(TPB), 0, s, dc, runs, dmax);
            if (hipMemcpyAsync(&most, dmax, sizeof(most), hipMemcpyDeviceToHost, s) == hipSuccess &&
                hipStreamSynchronize(s) == hipSuccess)
                width = count_width(most);
            if (width < 8 && dalloc(&dpacked, (size_t)runs * (size_t)width, "narrowed counts"))
                hipLaunchKernelGGL(k_narrow_counts, dim3(grid_for(runs, num_cus() * 8)), dim3(TPB), 0, s, dc, runs,
                                   width, (void *)dpacked);
            else
                width = 8;
        }
        PhaseTimer pt("  histogram_flat");
        if (flat && width < 8) {
            ok = hipStreamSynchronize(s) == hipSuccess;
            pt.mark("narrow");
            ok = ok && d2h_flat_narrow(flat, dv, dpacked, width, runs);
            pt.mark("download");
            if (!ok) set_error("histogram download failed");
        } else {
            ok = flat && dalloc(&dflat, (size_t)runs * 2, "flat histogram");
            if (ok) {
                hipLaunchKernelGGL(k_interleave, dim3(grid_for(runs, num_cus() * 8)), dim3(TPB), 0, s, dv, dc, runs,
                                   dflat);
                ok = hipStreamSynchronize(s) == hipSuccess;
                pt.mark("interleave");
                ok = ok && d2h_parallel(flat, dflat, (size_t)runs * 2 * sizeof(int64_t));
                pt.mark("download");
                if (!ok) set_error("histogram download failed");
            }
        }
        if (!ok) (void)hipStreamSynchronize(s);
        dfree(dflat);
        dfree(dmax);
        dfree(dpacked);
        if (ok) {
            *flat_out = flat;
            if (hist_size) *hist_size = (int)(2 * runs);
        } else {
            free(flat);
        }
    }
    dfree(dv);
    dfree(dc);
    return ok;
}

// items [b, e) of a chunk and their live values
struct ItemChunk {
    size_t b, e;
    int64_t values;
};

// consecutive items in chunks of at most cap live values (an item larger
// than cap makes a chunk of its own); no chunk for no items
std::vector<ItemChunk> item_chunks(const std::vector<int64_t> &live, int64_t cap)
{
    std::vector<ItemChunk> out;
    size_t b = 0;
    int64_t acc = 0;
    for (size_t i = 0; i < live.size(); i++) {
        if (i > b && acc + live[i] > cap) {
            out.push_back({b, i, acc});
            b = i;
            acc = 0;
        }
        acc += live[i];
    }
    if (b < live.size()) out.push_back({b, live.size(), acc});
    return out;
}

// order of the device histogram's sort keys (k_sort_keys): ascending value
inline uint64_t hist_order(double x)
{
    uint64_t b;
    memcpy(&b, &x, sizeof(b));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// (av, ac) += (v, c): both sorted ascending by value, exact keys (equal
// bit patterns add their counts, as dress_hist_add compares with ==)
void merge_histograms(std::vector<double> &av, std::vector<int64_t> &ac, const std::vector<double> &v,
                      const std::vector<int64_t> &c)
{
    if (v.empty()) return;
    if (av.empty()) {
        av = v;
        ac = c;
        return;
    }
    std::vector<double> mv;
    std::vector<int64_t> mc;
    mv.reserve(av.size() + v.size());
    mc.reserve(av.size() + v.size());
    size_t i = 0, j = 0;
    while (i < av.size() || j < v.size()) {
        if (j == v.size() || (i < av.size() && hist_order(av[i]) < hist_order(v[j]))) {
            mv.push_back(av[i]);
            mc.push_back(ac[i++]);
        } else if (i == av.size() || hist_order(v[j]) < hist_order(av[i])) {
            mv.push_back(v[j]);
            mc.push_back(c[j++]);
        } else {
            mv.push_back(av[i]);
            mc.push_back(ac[i++] + c[j++]);
        }
    }
    av.swap(mv);
    ac.swap(mc);
}

// the flat (value bits, count, ...) form of a host histogram, malloc'd
bool flat_from_vectors(const std::vector<double> &v, const std::vector<int64_t> &c, int64_t **flat_out,
                       int *hist_size)
{
    *flat_out = nullptr;
    if (hist_size) *hist_size = 0;
    const int64_t runs = (int64_t)v.size();
    if (runs == 0) return true;
    if (!hist_size_fits(runs)) return false;
    int64_t *flat = (int64_t *)host_alloc((size_t)runs * 2 * sizeof(int64_t));
    if (!flat) {
        set_error("host allocation of the histogram failed");
        return false;
    }
    for (int64_t j = 0; j < runs; j++) {
        memcpy(&flat[2 * j], &v[(size_t)j], sizeof(double));
        flat[2 * j + 1] = c[(size_t)j];
    }
    *flat_out = flat;
    if (hist_size) *hist_size = (int)(2 * runs);
    return true;
}

// ------------------------------------------------------------ quantised table

__global__ void k_qtable_clear(unsigned long long *keys, unsigned long long *counts, int64_t cap)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cap; i += (int64_t)gridDim.x * blockDim.x) {
        keys[i] = QEMPTY;
        counts[i] = 0ull;
    }
}

// takes the keys of values vals[rng[2i] .. rng[2i+1]) of graph gid[i] back
// out of the table (one count each; a key stays with its count, which may
// reach 0): the mixed-precision values of items about to be fitted again in
// FP64, whose new values the fit counts itself.  One block per item.
__global__ void k_unquant(const double *vals, const int64_t *rng, const int *gid, int n, double eps, double tol,
                          const unsigned long long *keys, unsigned long long *counts, int64_t cap)
{
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const unsigned long long g = (unsigned long long)gid[it];
        for (int64_t i = rng[2 * it] + threadIdx.x; i < rng[2 * it + 1]; i += blockDim.x) {
            const unsigned long long key = quant_key(vals[i], g, eps, tol);
            int64_t slot = quant_slot(key, cap);
            for (int64_t probe = 0; probe < cap; probe++) {
                const unsigned long long k = keys[slot];
                if (k == key) {
                    atomicAdd(&counts[slot], ~0ull);     // minus one
                    break;
                }
                if (k == QEMPTY) break;
                slot = (slot + 1) & (cap - 1);
            }
        }
    }
}

bool unquantise(const double *d_vals, const std::vector<int64_t> &val_ofs, const std::vector<int> &redo,
                const std::vector<int> &item_parent, const std::vector<int> &graph_of, double eps, double tol,
                const unsigned long long *keys, unsigned long long *counts, int64_t cap, hipStream_t s)
{
    const int n = (int)redo.size();
    std::vector<int64_t> rng(2 * (size_t)n);
    std::vector<int> gid((size_t)n);
    for (int j = 0; j < n; j++) {
        const int i = redo[(size_t)j];
        rng[2 * (size_t)j] = val_ofs[(size_t)i];
        rng[2 * (size_t)j + 1] = val_ofs[(size_t)i + 1];
        gid[(size_t)j] = graph_of[(size_t)item_parent[(size_t)i]];
    }
    int64_t *d_rng = nullptr;
    int *d_gid = nullptr;
    bool ok = upload(&d_rng, rng, s, "redo ranges") && upload(&d_gid, gid, s, "redo graphs");
    if (ok && n > 0) {
        hipLaunchKernelGGL(k_unquant, dim3(std::min(n, 65536)), dim3(TPB), 0, s, d_vals, d_rng, d_gid, n, eps, tol,
                           keys, counts, cap);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
    }
    if (!ok && !has_error()) set_error("quantised histogram update failed");
    dfree(d_rng);
    dfree(d_gid);
    return ok;
}

// occupied slots with a count (a key whose values were all taken back out
// keeps its slot with count 0)
__global__ void k_q_flags(const unsigned long long *keys, const unsigned long long *counts, int64_t cap, int *flag)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cap; i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = (keys[i] != QEMPTY && counts[i] != 0ull) ? 1 : 0;
}

__global__ void k_q_compact(const unsigned long long *keys, const unsigned long long *counts, const int *flag,
                            const int64_t *pos, int64_t cap, uint64_t *ok, uint64_t *oc)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cap; i += (int64_t)gridDim.x * blockDim.x)
        if (flag[i]) {
            ok[pos[i]] = keys[i];
            oc[pos[i]] = counts[i];
        }
}

__global__ void k_bnd_flags(const uint64_t *key, int64_t n, int *flag)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = (int)((key[i] >> 40) & 1);
}

// sorted (graph << 41 | boundary << 40 | bin) keys -> the per-graph bin
// table and the per-graph boundary table, with their graph offsets
__global__ void k_qsplit(const uint64_t *key, const uint64_t *cnt, const int64_t *bpos, int64_t n, int ngraph,
                         int64_t *qk, int64_t *qc, int64_t *bk, int64_t *bc, int64_t *qofs, int64_t *bofs)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = key[i];
        const int g = (int)(k >> 41);
        const int64_t b = bpos[i];
        const int64_t q = (int64_t)(k & 0xffffffffffull);
        if ((k >> 40) & 1) { bk[b] = q; bc[b] = (int64_t)cnt[i]; }
        else { qk[i - b] = q; qc[i - b] = (int64_t)cnt[i]; }
        const int prev = i ? (int)(key[i - 1] >> 41) : -1;
        for (int h = prev + 1; h <= g && h <= ngraph; h++) { qofs[h] = i - b; bofs[h] = b; }
        if (i == n - 1)
            for (int h = g + 1; h <= ngraph; h++) { qofs[h] = n - bpos[n]; bofs[h] = bpos[n]; }
    }
}

// Table -> per-graph (bin, count) tables ascending by bin, values near a
// bin edge in a separate per-graph table; all left on the device (the caller
// downloads and dfrees them).  qofs / bofs hold n_graphs + 1 offsets.
bool qtable_split(const unsigned long long *keys, const unsigned long long *counts, int64_t cap, int n_graphs,
                  hipStream_t s, int64_t **qk, int64_t **qc, int64_t **bk, int64_t **bc, int64_t **qofs,
                  int64_t **bofs, int64_t *nq_out, int64_t *nb_out)
{
    *qk = *qc = *bk = *bc = *qofs = *bofs = nullptr;
    *nq_out = *nb_out = 0;
    int *flag = nullptr, *bflag = nullptr;
    int64_t *pos = nullptr, *bpos = nullptr;
    uint64_t *k0 = nullptr, *c0 = nullptr, *k1 = nullptr, *c1 = nullptr;
    int64_t n = 0, nb = 0;
    bool ok = dalloc(&flag, (size_t)cap + 1, "table flags") && dalloc(&pos, (size_t)cap + 1, "table positions");
    if (ok) {
        hipLaunchKernelGGL(k_q_flags, dim3(grid_for(cap, 4096)), dim3(TPB), 0, s, keys, counts, cap, flag);
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, flag + cap, 1);
        ok = prim_exclusive_scan_i32_to_i64(flag, pos, (size_t)cap + 1, s) &&
             hipMemcpyAsync(&n, pos + cap, sizeof(int64_t), hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    }
    ok = ok && dalloc(&k0, (size_t)n + 1, "table keys") && dalloc(&c0, (size_t)n + 1, "table counts") &&
         dalloc(&k1, (size_t)n + 1, "table keys") && dalloc(&c1, (size_t)n + 1, "table counts") &&
         dalloc(&bflag, (size_t)n + 1, "boundary flags") && dalloc(&bpos, (size_t)n + 1, "boundary positions");
    if (ok) {
        hipLaunchKernelGGL(k_q_compact, dim3(grid_for(cap, 4096)), dim3(TPB), 0, s, keys, counts, flag, pos, cap, k0, c0);
        ok = prim_sort_pairs_u64_u64(k0, k1, c0, c1, (size_t)n, s);
        if (ok) {
            if (n > 0)
                hipLaunchKernelGGL(k_bnd_flags, dim3(grid_for(n, 4096)), dim3(TPB), 0, s, (const uint64_t *)k1, n, bflag);
            hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, bflag + n, 1);
            ok = prim_exclusive_scan_i32_to_i64(bflag, bpos, (size_t)n + 1, s) &&
                 hipMemcpyAsync(&nb, bpos + n, sizeof(int64_t), hipMemcpyDeviceToHost, s) == hipSuccess &&
                 hipStreamSynchronize(s) == hipSuccess;
        }
    }
    const int64_t nq = n - nb;
    ok = ok && dalloc(qk, (size_t)nq + 1, "bin keys") && dalloc(qc, (size_t)nq + 1, "bin counts") &&
         dalloc(bk, (size_t)nb + 1, "boundary keys") && dalloc(bc, (size_t)nb + 1, "boundary counts") &&
         dalloc(qofs, (size_t)n_graphs + 1, "bin offsets") && dalloc(bofs, (size_t)n_graphs + 1, "boundary offsets");
    if (ok) {
        if (n > 0) {
            hipLaunchKernelGGL(k_qsplit, dim3(grid_for(n, 4096)), dim3(TPB), 0, s, (const uint64_t *)k1,
                               (const uint64_t *)c1, (const int64_t *)bpos, n, n_graphs, *qk, *qc, *bk, *bc, *qofs,
                               *bofs);
        } else {
            ok = hipMemsetAsync(*qofs, 0, ((size_t)n_graphs + 1) * sizeof(int64_t), s) == hipSuccess &&
                 hipMemsetAsync(*bofs, 0, ((size_t)n_graphs + 1) * sizeof(int64_t), s) == hipSuccess;
        }
        ok = ok && hipGetLastError() == hipSuccess;
    }
    if (!ok && !has_error()) set_error("quantised histogram assembly failed");
    dfree(flag); dfree(pos); dfree(k0); dfree(c0); dfree(k1); dfree(c1); dfree(bflag); dfree(bpos);
    if (!ok) {
        dfree(*qk); dfree(*qc); dfree(*bk); dfree(*bc); dfree(*qofs); dfree(*bofs);
        *qk = *qc = *bk = *bc = *qofs = *bofs = nullptr;
        return false;
    }
    *nq_out = nq;
    *nb_out = nb;
    return true;
}

// ------------------------------------------------------------ runner

struct LiftJob {
    int kind = LIFT_DELTA;
    int k = 1;
    int max_it = 100;
    double eps = 1e-6;
    bool mixed = false;
    double marks[LIFT_KMAX];
};

struct LiftOut {
    std::vector<int64_t> val_ofs;        // per item + 1: packed live values
    double *d_vals = nullptr;            // device, val_ofs.back() entries
    int reruns = 0;
    double device_ms = 0.0;
    int block = 0;
    // with DRESS_HIP_LIFT_WORK=1: executed iterations x the subgraphs' index
    // entries, half-edges, edges and vertices, summed over items (the work
    // counted by k_lifted; a measurement aid, off by default)
    unsigned long long work[4] = {0, 0, 0, 0};
};

bool lift_work_counting()
{
    const char *e = getenv("DRESS_HIP_LIFT_WORK");
    return e && !strcmp(e, "1");
}

// per-item outputs of whole-graph fits run as k = 0 items
struct FitOutputs {
    double *d_vnorm;                       // vertex norms, items at vn_ofs
    std::vector<int64_t> vn_ofs;
    std::vector<int> *iters;
    std::vector<double> *deltas;
};

int forced_block()
{
    const char *b = getenv("DRESS_HIP_LIFT_BLOCK");
    if (b) {
        int v = atoi(b);
        if (v == 64 || v == 128 || v == 256 || v == 512) return v;
    }
    return 0;
}

// Whether the batched kernel copies each parent's structure into LDS
// (DRESS_HIP_LIFT_STAGE=0 turns it off; results do not depend on it).
int lift_stage_mode()
{
    static const int mode = [] {
        const char *e = getenv("DRESS_HIP_LIFT_STAGE");
        return (e && !strcmp(e, "0")) ? 0 : 1;
    }();
    return mode;
}

// static block size by the launch group's largest edge count, where the
// autotuner does not calibrate (delta^1 of 6-regular graphs with E <= 384:
// 2.80 ms of device time with 128 threads, 3.06 ms with 64, 3.11 ms with
// 256; delta^2 of a 12-regular graph, E = 1,524: 256 calibrates fastest)
int choose_block(int E)
{
    if (int f = forced_block()) return f;
    return E <= 128 ? 64 : (E <= 1024 ? 128 : 256);
}

// Block-size autotuner of the batched kernel.  A size class is (bit length
// of the largest item's edge count, precision, kind).  The first launch
// group of an untuned class splits its leading items over the candidate
// block sizes -- every candidate computes real items, results do not
// depend on the block size -- and the size with the least device time per
// item is kept for the rest of the process.  DRESS_HIP_LIFT_BLOCK forces a
// size; DRESS_HIP_LIFT_TUNE=0 keeps the static choice.
struct BlockTuner {
    std::mutex mu;
    std::map<int, int> best;           // class -> block
    static int size_class(int E, bool mixed, int kind)
    {
        int bits = 0;
        while ((1 << bits) < std::max(E, 1)) bits++;
        return bits * 4 + (mixed ? 2 : 0) + (kind ? 1 : 0);
    }
    int lookup(int cls)
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = best.find(cls);
        return it == best.end() ? 0 : it->second;
    }
    void store(int cls, int block)
    {
        std::lock_guard<std::mutex> lk(mu);
        best[cls] = block;
    }
};
BlockTuner g_tuner;

bool tuning_enabled()
{
    static const bool on = [] {
        const char *e = getenv("DRESS_HIP_LIFT_TUNE");
        return !(e && !strcmp(e, "0"));
    }();
    return on;
}

// Runs every item (parent index, vertex tuple) in one persistent launch (plus
// an FP64 relaunch of items whose mixed-precision stop was ambiguous).
// One run of items through k_lifted in three steps: run_prepare (host
// only: items, queue order, DIRECTED weight patches, the parents packed
// into a pinned staging block), run_launch (upload and launches queued on
// the stream, no wait) and run_finish (waits; FP64 redo of ambiguous
// mixed-precision items; per-item outputs; frees).  A RunState is filled
// in place and not moved: the staging registrations point into it.
struct RunState {
    int n = 0;
    std::vector<LiftItem> items;
    std::vector<int> order;
    std::vector<double> extra_ew, marks;
    std::vector<LiftParent> meta;
    PackedUpload pk;
    DevPools D;
    LiftItem *d_items = nullptr;
    int *d_order = nullptr, *d_verts = nullptr, *d_queue = nullptr, *d_flags = nullptr, *d_iters = nullptr;
    double *d_deltas = nullptr, *d_marks = nullptr;
    unsigned long long *d_work = nullptr;
    std::vector<hipEvent_t> ev;          // start / end pairs of the queued launches
    LiftArgs a;
    int block = 0;
    RunState() = default;
    RunState(const RunState &) = delete;
    RunState &operator=(const RunState &) = delete;
    ~RunState() { release(); }
    void release()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        ev.clear();
        dfree(d_queue); dfree(d_flags); dfree(d_iters); dfree(d_deltas); dfree(d_work);
        d_queue = d_flags = d_iters = nullptr;
        d_deltas = nullptr;
        d_work = nullptr;
        D.release();               // items, order, vertices and marks share its block
        pk.release();
    }
};

bool run_prepare(const std::vector<const HostParent *> &ps, const std::vector<int> &graph_of,
                 const std::vector<int> &item_parent, const std::vector<int> &item_verts,
                 const std::vector<int64_t> &item_live, const LiftJob &J, const std::vector<int64_t> &ms_rows,
                 const FitOutputs *fo, RunState &R, LiftOut &out)
{
    const int n = (int)item_parent.size();
    R.n = n;
    out.val_ofs.assign((size_t)n + 1, 0);
    for (int i = 0; i < n; i++) out.val_ofs[(size_t)i + 1] = out.val_ofs[(size_t)i] + item_live[(size_t)i];
    if (n == 0) return true;
    R.items.resize((size_t)n);
    for (int i = 0; i < n; i++) {
        LiftItem &I = R.items[(size_t)i];
        I.parent = item_parent[(size_t)i];
        I.kind = J.kind;
        I.k = J.k;
        I.mixed = J.mixed ? 1 : 0;
        I.vofs = (int64_t)i * J.k;
        I.val_ofs = out.val_ofs[(size_t)i];
        I.ms_row = ms_rows.empty() ? 0 : ms_rows[(size_t)i];
        I.vn_ofs = fo ? fo->vn_ofs[(size_t)i] : 0;
        I.ew_ofs = -1;
    }
    // DIRECTED delta items whose own reciprocal lookups give other weights
    // than the parent's get a patched copy of the weights (one per distinct
    // patch set), placed after the parents' weights
    if (J.kind == LIFT_DELTA && J.k > 0) {
        int64_t base = 0;
        for (const HostParent *P : ps) base += P->E;
        std::map<std::pair<int, std::vector<std::pair<int, double>>>, int64_t> seen;
        std::vector<int> pe;
        std::vector<double> pw;
        for (int i = 0; i < n; i++) {
            const int pi = item_parent[(size_t)i];
            const HostParent &P = *ps[(size_t)pi];
            if (P.amb.empty()) continue;
            pe.clear();
            pw.clear();
            directed_patches(P, &item_verts[(size_t)i * J.k], J.k, pe, pw);
            if (pe.empty()) continue;
            std::vector<std::pair<int, double>> key;
            for (size_t j = 0; j < pe.size(); j++) key.push_back({pe[j], pw[j]});
            auto it = seen.find({pi, key});
            if (it != seen.end()) {
                R.items[(size_t)i].ew_ofs = it->second;
                continue;
            }
            int64_t at = base + (int64_t)R.extra_ew.size();
            R.extra_ew.insert(R.extra_ew.end(), P.ew.begin(), P.ew.end());
            for (size_t j = 0; j < pe.size(); j++) R.extra_ew[(size_t)(at - base) + (size_t)pe[j]] = pw[j];
            seen.emplace(std::make_pair(pi, std::move(key)), at);
            R.items[(size_t)i].ew_ofs = at;
        }
    }
    // queue order: larger parents first, items of equal size in index order
    // (a counting sort over the parents' distinct edge counts)
    R.order.resize((size_t)n);
    {
        std::vector<int> sizes;
        sizes.reserve(ps.size());
        for (const HostParent *P : ps) sizes.push_back(P->E);
        std::sort(sizes.begin(), sizes.end(), std::greater<int>());
        sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
        std::vector<int> prank(ps.size());
        for (size_t p = 0; p < ps.size(); p++)
            prank[p] = (int)(std::lower_bound(sizes.begin(), sizes.end(), ps[p]->E, std::greater<int>()) -
                             sizes.begin());
        std::vector<int> start(sizes.size() + 1, 0);
        for (int i = 0; i < n; i++) start[(size_t)prank[(size_t)item_parent[(size_t)i]] + 1]++;
        for (size_t r = 0; r < sizes.size(); r++) start[r + 1] += start[r];
        for (int i = 0; i < n; i++) R.order[(size_t)start[(size_t)prank[(size_t)item_parent[(size_t)i]]]++] = i;
    }
    R.marks.assign(J.marks, J.marks + LIFT_KMAX);
    PhaseTimer pt("    run_prepare");
    pack_parents(ps, graph_of, R.meta, R.D, R.pk, &R.extra_ew);
    pt.mark("concatenate");
    R.pk.add(&R.d_items, R.items);
    R.pk.add(&R.d_order, R.order);
    R.pk.add(&R.d_verts, item_verts);
    R.pk.add(&R.d_marks, R.marks);
    // the arrays are in the staging block from here on
    const bool ok = R.pk.pack();
    pt.mark("staging");
    return ok;
}

// Launches of the items ids[] (positions in the queue order), grouped so
// that each launch's largest parent fits the LDS limit: the ids are in
// order of decreasing edge count, and a group ends where the next item
// would push the larger of the group's vertex counts past it.  Launches are
// queued with start / end events; only a block-size calibration waits.
bool launch_ids(RunState &R, const std::vector<const HostParent *> &ps, const std::vector<int> &item_parent,
                const LiftJob &J, const int *d_ids, const std::vector<int> &ids, hipStream_t s, LiftOut &out)
{
    LiftArgs &a = R.a;
    const size_t limit = lifted_lds_limit();
    const size_t per_cu_bytes = lifted_lds_per_cu();
    size_t g0 = 0;
    while (g0 < ids.size()) {
        int gE = 0, gN = 0, gS = 0;
        size_t g1 = g0;
        while (g1 < ids.size()) {
            const HostParent *P = ps[(size_t)item_parent[(size_t)ids[g1]]];
            int nE = std::max(gE, P->E), nN = std::max(gN, P->N);
            gS = std::max(gS, P->S);
            if (lifted_lds_bytes((nE + 1) & ~1, (nN + 1) & ~1) > limit) {
                if (g1 == g0) {
                    set_error("a graph of " + std::to_string(P->N) + " vertices and " + std::to_string(P->E) +
                              " edges exceeds the LDS of the batched kernel");
                    return false;
                }
                break;
            }
            gE = nE;
            gN = nN;
            g1++;
        }
        a.ecap = (gE + 1) & ~1;
        a.ncap = (gN + 1) & ~1;
        a.scap = (gS + 1) & ~1;
        // The parent's structure goes into LDS when it fits and the launch
        // still holds 16 waves per CU (or as many as without it) at the
        // static block size: delta^1 of 6-regular graphs of <= 128 vertices
        // (13 instead of 16 workgroups of 128 per CU) took 1.30 instead of
        // 1.50 ms; delta^2 of the 12-regular graph on 256 vertices (3
        // instead of 5 workgroups of 256) would take 7.4 instead of 7.1 ms.
        const size_t lds_plain = lifted_lds_bytes(a.ecap, a.ncap);
        const size_t lds_staged = lifted_lds_bytes_staged(a.ecap, a.ncap, a.scap);
        auto waves_per_cu = [&](size_t lds_b, int block) {
            const int per_cu_lds = (int)std::max<size_t>(1, per_cu_bytes / (lds_b + 1024));
            const int per_cu_waves = std::max(1, 32 / (block / 64));
            return std::max(1, std::min(per_cu_lds, per_cu_waves)) * (block / 64);
        };
        const int sb = choose_block(gE);
        const bool staged = lift_stage_mode() != 0 && lds_staged <= limit &&
                            waves_per_cu(lds_staged, sb) >= std::min(16, waves_per_cu(lds_plain, sb));
        const size_t lds = staged ? lds_staged : lds_plain;
        // workgroups of `block` threads one CU holds at once
        auto per_cu_of = [&](int block) {
            int per_cu_lds = (int)std::max<size_t>(1, per_cu_bytes / (lds + 1024));
            int per_cu_waves = std::max(1, 32 / (block / 64));
            return std::max(1, std::min(per_cu_lds, per_cu_waves));
        };
        // one launch of items [b0, b1) of the list with `block` threads;
        // with ms_out it waits and returns the launch's time
        auto launch_range = [&](size_t b0, size_t b1, int block, float *ms_out) -> bool {
            a.order = d_ids + b0;
            a.n_items = (int)(b1 - b0);
            const int resident = std::max(1, num_cus() * per_cu_of(block) * thread_grid_share() / 100);
            int grid = std::max(1, std::min(a.n_items, resident));
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (hipEventCreate(&e0) != hipSuccess) return false;
            if (hipEventCreate(&e1) != hipSuccess) {
                (void)hipEventDestroy(e0);
                return false;
            }
            R.ev.push_back(e0);
            R.ev.push_back(e1);
            if (!(hipMemsetAsync(R.d_queue, 0, sizeof(int), s) == hipSuccess && hipEventRecord(e0, s) == hipSuccess &&
                  lifted_launch(a, block, grid, staged, s) && hipEventRecord(e1, s) == hipSuccess))
                return false;
            if (ms_out) {
                if (hipEventSynchronize(e1) != hipSuccess) return false;
                (void)hipEventElapsedTime(ms_out, e0, e1);
            }
            return true;
        };
        int block = forced_block();
        const int cls = BlockTuner::size_class(gE, J.mixed, J.kind);
        if (!block && tuning_enabled()) block = g_tuner.lookup(cls);
        size_t b0 = g0;
        if (!block && tuning_enabled()) {
            // calibrate on the group's leading items, each candidate on two
            // full waves of its own resident workgroups, when the group holds
            // at least twice what that takes
            static const int cand[4] = {64, 128, 256, 512};
            size_t chunk[4], need = 0;
            for (int i = 0; i < 4; i++) {
                chunk[i] = 2 * (size_t)num_cus() * (size_t)per_cu_of(cand[i]);
                need += chunk[i];
            }
            if (g1 - g0 >= 2 * need) {
                double best_per_item = 0.0;
                int best = 0;
                for (int i = 0; i < 4; i++) {
                    float ms = 0.f;
                    if (!launch_range(b0, b0 + chunk[i], cand[i], &ms)) return false;
                    double per_item = ms / (double)chunk[i];
                    if (!best || per_item < best_per_item) { best = cand[i]; best_per_item = per_item; }
                    b0 += chunk[i];
                }
                g_tuner.store(cls, best);
                block = best;
            }
        }
        if (!block) block = choose_block(gE);
        out.block = std::max(out.block, block);
        if (b0 < g1 && !launch_range(b0, g1, block, nullptr)) return false;
        g0 = g1;
    }
    return true;
}

bool run_launch(RunState &R, const std::vector<const HostParent *> &ps, const std::vector<int> &item_parent,
                const LiftJob &J, double *d_ms, int64_t ms_stride, hipStream_t s, LiftOut &out,
                unsigned long long *qkeys, unsigned long long *qcounts, int64_t qcap, double qtol,
                const FitOutputs *fo, bool sort_vals = false)
{
    const int n = R.n;
    if (n == 0) return true;
    bool ok = R.pk.upload(&R.D.blk, s) && dalloc(&R.d_queue, 1, "work queue") &&
              dalloc(&R.d_flags, (size_t)n, "item flags") &&
              dalloc(&out.d_vals, (size_t)out.val_ofs.back() + 1, "subgraph values");
    if (ok && fo) ok = dalloc(&R.d_iters, (size_t)n, "item iterations") && dalloc(&R.d_deltas, (size_t)n, "item deltas");
    if (ok && lift_work_counting())
        ok = dalloc(&R.d_work, 4, "work counters") &&
             hipMemsetAsync(R.d_work, 0, 4 * sizeof(unsigned long long), s) == hipSuccess;
    if (!ok) return false;
    LiftArgs &a = R.a;
    memset(&a, 0, sizeof(a));
    a.parents = R.D.parents;
    a.items = R.d_items;
    a.order = R.d_order;
    a.n_items = n;
    a.item_verts = R.d_verts;
    a.marks = R.d_marks;
    a.row_off = R.D.row_off;
    a.adj_e = R.D.adj_e;
    a.ell = R.D.ell;
    a.uv = R.D.uv;
    a.ew = R.D.ew;
    a.ioff = R.D.ioff;
    a.pairs = R.D.pairs;
    a.queue = R.d_queue;
    a.max_it = J.max_it;
    a.eps = J.eps;
    a.vals = out.d_vals;
    a.ms = d_ms;
    a.ms_stride = ms_stride;
    a.flags = R.d_flags;
    a.qkeys = qkeys;
    a.qcounts = qcounts;
    a.qcap = qcap;
    a.qtol = qtol;
    a.parent_graph = R.D.parent_graph;
    a.work = R.d_work;
    a.sort_vals = sort_vals ? 1 : 0;
    if (fo) {
        a.vnorm = fo->d_vnorm;
        a.iters = R.d_iters;
        a.deltas = R.d_deltas;
    }
    // (no clearing of the item flags: the launches run every item, and
    // k_lifted writes each item's flag when it ends the item)
    return launch_ids(R, ps, item_parent, J, R.d_order, R.order, s, out);
}

// rs: the stream for the FP64 redo of ambiguous mixed-precision items
// (default s), so that the redo need not wait for later work queued on s
bool run_finish(RunState &R, const std::vector<const HostParent *> &ps, const std::vector<int> &graph_of,
                const std::vector<int> &item_parent, const LiftJob &J, hipStream_t s, LiftOut &out,
                unsigned long long *qkeys, unsigned long long *qcounts, int64_t qcap, double qtol,
                const FitOutputs *fo, bool ok, hipStream_t rs = nullptr)
{
    hipStream_t xs = rs ? rs : s;
    const int n = R.n;
    PhaseTimer pt("  run_items");
    // (the end of the last launch: later work may already be queued on s)
    if (ok && n > 0)
        ok = (R.ev.empty() ? hipStreamSynchronize(s) : hipEventSynchronize(R.ev.back())) == hipSuccess;
    // launch times from the queued events
    auto add_times = [&] {
        for (size_t i = 0; i + 1 < R.ev.size(); i += 2) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, R.ev[i], R.ev[i + 1]) == hipSuccess) out.device_ms += ms;
        }
        for (hipEvent_t e : R.ev) (void)hipEventDestroy(e);
        R.ev.clear();
    };
    if (ok) add_times();
    pt.mark("launches");
    // FP64 redo of ambiguous mixed-precision items
    if (ok && n > 0 && J.mixed) {
        std::vector<int> flags((size_t)n);
        ok = hipMemcpyAsync(flags.data(), R.d_flags, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, xs) ==
                 hipSuccess && hipStreamSynchronize(xs) == hipSuccess;
        std::vector<int> redo;
        for (int qi = 0; qi < n && ok; qi++)        // keeps the decreasing-size order
            if (flags[(size_t)R.order[(size_t)qi]]) redo.push_back(R.order[(size_t)qi]);
        if (ok && !redo.empty()) {
            out.reruns = (int)redo.size();
            std::vector<LiftItem> items2 = R.items;
            for (int i : redo) items2[(size_t)i].mixed = 0;
            int *d_redo = nullptr;
            // the quantised counts of the redone items' mixed values come
            // back out first; the FP64 redo counts their final values
            ok = (!qkeys || unquantise(out.d_vals, out.val_ofs, redo, item_parent, graph_of, J.eps, qtol, qkeys,
                                       qcounts, qcap, xs)) &&
                 hipMemcpyAsync(R.d_items, items2.data(), items2.size() * sizeof(LiftItem), hipMemcpyHostToDevice,
                                xs) == hipSuccess &&
                 upload(&d_redo, redo, xs, "redo list");
            if (ok) ok = launch_ids(R, ps, item_parent, J, d_redo, redo, xs, out) && hipStreamSynchronize(xs) == hipSuccess;
            if (ok) add_times();
            dfree(d_redo);
        }
    }
    if (ok && n > 0 && fo) {
        fo->iters->resize((size_t)n);
        fo->deltas->resize((size_t)n);
        ok = hipMemcpyAsync(fo->iters->data(), R.d_iters, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s) ==
                 hipSuccess &&
             hipMemcpyAsync(fo->deltas->data(), R.d_deltas, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s) ==
                 hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    }
    if (ok && R.d_work)
        ok = hipMemcpyAsync(out.work, R.d_work, sizeof(out.work), hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    pt.mark("reruns");
    // (on a failure the streams may still hold work that reads these blocks)
    if (!ok) {
        (void)hipStreamSynchronize(s);
        if (rs) (void)hipStreamSynchronize(rs);
    }
    R.release();
    pt.mark("free");
    if (!ok) {
        if (!has_error()) set_error("lifted launch failed");
        dfree(out.d_vals);
        out.d_vals = nullptr;
    }
    return ok;
}

bool run_items(const std::vector<const HostParent *> &ps, const std::vector<int> &graph_of,
               const std::vector<int> &item_parent, const std::vector<int> &item_verts,
               const std::vector<int64_t> &item_live, const LiftJob &J, double *d_ms, int64_t ms_stride,
               const std::vector<int64_t> &ms_rows, hipStream_t s, LiftOut &out, unsigned long long *qkeys,
               unsigned long long *qcounts, int64_t qcap, double qtol, const FitOutputs *fo = nullptr)
{
    RunState R;
    PhaseTimer pt("  run_items");
    bool ok = run_prepare(ps, graph_of, item_parent, item_verts, item_live, J, ms_rows, fo, R, out);
    pt.mark("prepare");
    ok = ok && run_launch(R, ps, item_parent, J, d_ms, ms_stride, s, out, qkeys, qcounts, qcap, qtol, fo);
    return run_finish(R, ps, graph_of, item_parent, J, s, out, qkeys, qcounts, qcap, qtol, fo, ok);
}

}  // namespace

}  // namespace dh

namespace dh {

namespace {

std::mutex g_lift_mu;
int g_lift_mode = 0;   // 0 batched when the parent fits LDS else multi/masked, 1 per-subgraph fits,
                       // 2 masked one at a time, 3 multi-subgraph passes

bool lds_eligible(const HostParent &P) { return lds_eligible_size(P.N, P.E); }

// live edge count of each delta item by inclusion-exclusion: E minus the
// edge records touching a deleted vertex, plus those joining two deleted
// vertices (counted twice above)
std::vector<int64_t> delta_live_counts(const p_dress_graph_t g, const std::vector<int> &verts, int k, size_t n)
{
    std::vector<int64_t> inc(g->N + 1, 0);
    std::unordered_map<uint64_t, int64_t> mult;
    for (int e = 0; e < g->E; e++) {
        int u = g->U[e], v = g->V[e];
        inc[u]++;
        if (v != u) inc[v]++;
        if (k >= 2 && u != v) {
            uint64_t key = ((uint64_t)std::min(u, v) << 32) | (uint32_t)std::max(u, v);
            mult[key]++;
        }
    }
    std::vector<int64_t> live(n);
    for (size_t i = 0; i < n; i++) {
        const int *d = &verts[i * (size_t)k];
        int64_t x = g->E;
        for (int a = 0; a < k; a++) x -= inc[d[a]];
        for (int a = 0; a < k; a++)
            for (int b = a + 1; b < k; b++) {
                uint64_t key = ((uint64_t)std::min(d[a], d[b]) << 32) | (uint32_t)std::max(d[a], d[b]);
                auto it = mult.find(key);
                if (it != mult.end()) x += it->second;
            }
        live[i] = x;
    }
    return live;
}

__global__ void k_fill_nan(double *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = __longlong_as_double(0x7ff8000000000000ll);
}

// nabla_dress_impl.c:33-46
void marker_weights(double *out, int k)
{
    static const int primes[16] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53};
    for (int i = 0; i < k && i < LIFT_KMAX; i++)
        out[i] = (i < 16) ? std::sqrt((double)primes[i]) : std::sqrt((double)(2 * i + 3));
}

// One subgraph by a separate device fit, as the reference's GPU drivers do
// (delta_dress_impl.c:95-169 rebuild, then fit_fn): used for parents too
// large for the batched kernel, and on request for comparison.
bool plain_item(p_dress_graph_t g, int kind, const int *items, int k, const double *marks, int max_it,
                double eps, int precision, std::vector<double> &vals, double *ms_row)
{
    const int N = g->N, E = g->E;
    std::vector<int> sU, sV, emap(E, -1);
    std::vector<double> sW, nw;
    int sN = N;
    if (kind == LIFT_DELTA) {
        std::vector<int> map(N, 0);
        for (int i = 0; i < k; i++) map[items[i]] = -1;
        int id = 0;
        for (int v = 0; v < N; v++)
            if (map[v] >= 0) map[v] = id++;
        sN = N - k;
        for (int e = 0; e < E; e++) {
            int a = map[g->U[e]], b = map[g->V[e]];
            if (a >= 0 && b >= 0) {
                emap[e] = (int)sU.size();
                sU.push_back(a);
                sV.push_back(b);
                if (g->W) sW.push_back(g->W[e]);
            }
        }
        if (sU.empty()) return true;   // delta_dress_impl.c:124-129: no edge, no values
    } else {
        sU.assign(g->U, g->U + E);
        sV.assign(g->V, g->V + E);
        if (g->W) sW.assign(g->W, g->W + E);
        nw.assign(N, 1.0);
        for (int i = 0; i < k; i++) nw[items[i]] = marks[i];
        for (int e = 0; e < E; e++) emap[e] = e;
    }
    DevGraph *G = new DevGraph();
    std::vector<double> d(sU.size(), 1.0);
    FitResult r;
    bool ok = build_graph(G, sN, (int)sU.size(), sU.data(), sV.data(), g->W ? sW.data() : nullptr,
                          kind == LIFT_NABLA ? nw.data() : nullptr, (int)g->variant) &&
              fit_upload_values(G, d.data()) && fit_run(G, max_it, eps, precision, &r) &&
              fit_download(G, d.data(), nullptr);
    free_graph(G);
    if (!ok) return false;
    vals.insert(vals.end(), d.begin(), d.end());
    if (ms_row)
        for (int e = 0; e < E; e++) ms_row[e] = emap[e] >= 0 ? d[(size_t)emap[e]] : NAN;
    return true;
}

// flat histogram of host values (uploaded, then histogram_flat); NULL on
// error or when empty
int64_t *host_values_flat(const std::vector<double> &vals, hipStream_t s, int *hist_size)
{
    double *dv = nullptr;
    int64_t *flat = nullptr;
    if (upload(&dv, vals, s, "values")) histogram_flat(dv, (int64_t)vals.size(), s, &flat, hist_size);
    dfree(dv);
    return flat;
}

// delta_dress_impl.c:233-380 / nabla_dress_impl.c:142-311 on the GPU
int64_t *lifted_flat(int kind, p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                     unsigned int seed, int *hist_size, int keep_multisets, double **multisets,
                     int64_t *num_items, int offset, int stride)
{
    clear_error();
    if (hist_size) *hist_size = 0;
    if (!g) { set_error("NULL graph"); return nullptr; }
    if (stride <= 0 || offset < 0) { set_error("stride must be positive and offset non-negative"); return nullptr; }
    if (k > LIFT_KMAX && k < g->N) { set_error("k larger than supported"); return nullptr; }
    const int N = g->N, E = g->E;
    const bool do_hist = hist_size != nullptr;
    const int64_t total = (k == 0) ? 1 : (kind == LIFT_DELTA ? dress_delta_binom(N, k) : dress_nabla_perm_count(N, k));
    if (num_items) *num_items = total;
    int mode;
    {
        std::lock_guard<std::mutex> lk(g_lift_mu);
        mode = g_lift_mode;
    }
    // the host parent (CSR, intersection pairs, DIRECTED reciprocal
    // lookups) serves the LDS-resident kernel and the DIRECTED delta weight
    // patches; the paths on the device graph need nothing else of it
    const bool lds = mode == 0 && lds_eligible_size(N, E) && iterations > 0;
    double *host_ms = nullptr;
    BackgroundTouch ms_touch;                    // (joined before the download, and on every return)
    if (keep_multisets && multisets) {
        size_t len = (size_t)std::max<int64_t>(total, 0) * (size_t)E;
        host_ms = (double *)host_alloc((len ? len : 1) * sizeof(double));
        if (host_ms) {
            // The LDS kernel's path fills the device copy with NaN and
            // downloads all of it, so the host block is only page-touched,
            // beside the items (0.39 GB for delta^2 of a 256-vertex graph);
            // the other paths write rows into a NaN-filled block.
            if (lds && k > 0 && k < N && total > 0) ms_touch.start(host_ms, len * sizeof(double));
            else fill_nan_parallel(host_ms, len);
            *multisets = host_ms;
        }
    }
    int precision = default_precision();
    hipStream_t s = nullptr;
    DevGraph *G = device_graph(g);
    if (!G) return nullptr;
    s = G->stream;
    // work buffers of this call from the thread's batch block cache (the
    // device graph itself was built above, outside it)
    BatchScope scope;
    if (k == 0) {
        // delta_dress_impl.c:268-286: fit g itself in place
        if (offset != 0) return nullptr;
        int it = 0;
        double dl = 0;
        if (dress_fit_hip_ex(g, iterations, epsilon, (dress_hip_precision_t)precision, &it, &dl) != 0) return nullptr;
        std::vector<double> v(g->edge_dress, g->edge_dress + E);
        if (host_ms) memcpy(host_ms, g->edge_dress, (size_t)E * sizeof(double));
        if (!do_hist) return nullptr;
        return host_values_flat(v, s, hist_size);
    }
    if (k >= N) return nullptr;     // delta_dress_impl.c:289-293, nabla_dress_impl.c:197-201

    PhaseTimer pt("lifted_flat");
    std::vector<int64_t> sel;
    std::vector<int> verts;
    enumerate_items(kind, N, k, n_samples, seed, offset, stride, total, sel, verts);
    pt.mark("enumerate");
    LiftJob J;
    J.kind = kind;
    J.k = k;
    J.max_it = iterations;
    J.eps = epsilon;
    J.mixed = precision == PREC_FP32_FP64;
    memset(J.marks, 0, sizeof(J.marks));
    if (kind == LIFT_NABLA) marker_weights(J.marks, k);

    HostParent P;
    if (lds || (kind == LIFT_DELTA && g->variant == DRESS_VARIANT_DIRECTED && (mode == 0 || mode == 2 || mode == 3)))
        host_parent_build(N, E, g->U, g->V, g->W, (int)g->variant, P);
    pt.mark("parent");
    if (lds || mode == 0 || mode == 2 || mode == 3) {
        std::vector<int64_t> live = kind == LIFT_DELTA ? delta_live_counts(g, verts, k, sel.size())
                                                       : std::vector<int64_t>(sel.size(), (int64_t)E);
        pt.mark("live counts");
        // Chunks of items whose live values (and their histogram assembly)
        // fit the call's device budget; one chunk unless the request is
        // larger than that.  Each chunk's exact histogram is merged on the
        // host, so the result does not depend on the chunking.
        const std::vector<ItemChunk> chunks = item_chunks(live, lifted_chunk_values());
        double *d_ms = nullptr;
        bool ok = true;
        if (lds && host_ms && total > 0) {
            ok = dalloc(&d_ms, (size_t)total * (size_t)E + 1, "multisets");
            if (ok)
                hipLaunchKernelGGL(k_fill_nan, dim3(grid_for((int64_t)total * E, 8192)), dim3(TPB), 0, s, d_ms,
                                   (int64_t)total * E);
        }
        EwPatches ep;
        bool multi = false;
        if (!lds) {
            // parent beyond LDS (or forced): the parent's resident device
            // graph serves every subgraph -- many subgraphs per pass, or one
            // masked fit at a time for tuples longer than a pass takes,
            // DIRECTED reciprocal weight patches, or graphs whose pass state
            // is too large
            if (kind == LIFT_DELTA) ep = delta_patches(P, verts, k, sel.size());
            multi = mode != 2 && !ep.any() && k <= multi_max_k() && build_bins(G) && multi_slots(G) > 0;
            pt.mark("patches, bins");
        }
        int reruns = 0;
        double dev_ms = 0.0;
        int64_t n_values = 0;
        // fits items [c.b, c.e) into *d_vals (device, dfree'd by the caller)
        auto run_chunk = [&](const ItemChunk &c, double **d_vals, int64_t *nv) -> bool {
            *d_vals = nullptr;
            *nv = 0;
            const std::vector<int> cv(verts.begin() + (ptrdiff_t)(c.b * (size_t)k),
                                      verts.begin() + (ptrdiff_t)(c.e * (size_t)k));
            const std::vector<int64_t> cl(live.begin() + (ptrdiff_t)c.b, live.begin() + (ptrdiff_t)c.e);
            const std::vector<int64_t> cs(sel.begin() + (ptrdiff_t)c.b, sel.begin() + (ptrdiff_t)c.e);
            if (lds) {
                LiftOut out;
                std::vector<const HostParent *> ps{&P};
                std::vector<int> gof{0};
                std::vector<int> item_parent(cl.size(), 0);
                if (!run_items(ps, gof, item_parent, cv, cl, J, d_ms, E, cs, s, out, nullptr, nullptr, 0, 0.0)) {
                    dfree(out.d_vals);
                    return false;
                }
                reruns += out.reruns;
                dev_ms += out.device_ms;
                *d_vals = out.d_vals;
                *nv = out.val_ofs.back();
                return true;
            }
            if (!dalloc(d_vals, (size_t)c.values + 1, "subgraph values")) return false;
            if (multi) {
                if (!multi_items(G, kind, k, cv, cl, cs, kind == LIFT_NABLA ? J.marks : nullptr, iterations, epsilon,
                                 *d_vals, host_ms))
                    return false;
                *nv = c.values;
                return true;
            }
            std::vector<int64_t> po;
            std::vector<int> pe;
            std::vector<double> pw, parent_w;
            if (ep.any()) {
                const int64_t p0 = ep.off[c.b], p1 = ep.off[c.e];
                for (size_t i = c.b; i <= c.e; i++) po.push_back(ep.off[i] - p0);
                pe.assign(ep.e.begin() + p0, ep.e.begin() + p1);
                pw.assign(ep.w.begin() + p0, ep.w.begin() + p1);
                for (int e : pe) parent_w.push_back(P.ew[(size_t)e]);
            }
            if (!masked_items(g, G, kind, k, cv, cl, cs, kind == LIFT_NABLA ? J.marks : nullptr, iterations, epsilon,
                              precision, *d_vals, c.values, nv, host_ms, po, pe, pw, parent_w))
                return false;
            if (*nv != c.values) {
                set_error("masked subgraph fits: live value count mismatch");
                return false;
            }
            return true;
        };
        int64_t *flat = nullptr;
        if (ok && chunks.size() <= 1) {
            // the histogram's host block (every value distinct: 16 bytes a
            // value) made ready while the items run, for large results
            PretouchedBlock pre;
            if (do_hist && !chunks.empty() && chunks[0].values >= ((int64_t)4 << 20) && !ctx_serving())
                pre.start((size_t)chunks[0].values * 2 * sizeof(int64_t));
            double *d_vals = nullptr;
            ok = chunks.empty() || run_chunk(chunks[0], &d_vals, &n_values);
            pt.mark("items");
            if (ok && lds && host_ms && total > 0) {
                ms_touch.join();
                ok = hipStreamSynchronize(s) == hipSuccess &&
                     d2h_parallel(host_ms, d_ms, (size_t)total * (size_t)E * sizeof(double));
            }
            if (ok && do_hist) ok = histogram_flat(d_vals, n_values, s, &flat, hist_size, &pre);
            pt.mark("histogram");
            dfree(d_vals);
        } else if (ok) {
            std::vector<double> acc_v, v;
            std::vector<int64_t> acc_c, c;
            std::vector<int> useg;
            for (const ItemChunk &ch : chunks) {
                if (thread_cancelled()) {
                    set_error("cancelled");
                    ok = false;
                    break;
                }
                double *d_vals = nullptr;
                int64_t nv = 0;
                ok = run_chunk(ch, &d_vals, &nv);
                if (ok && do_hist) ok = exact_histogram(d_vals, nv, {0, nv}, s, v, c, useg);
                dfree(d_vals);
                if (!ok) break;
                n_values += nv;
                if (do_hist) merge_histograms(acc_v, acc_c, v, c);
            }
            if (ok && lds && host_ms && total > 0) {
                ms_touch.join();
                ok = hipStreamSynchronize(s) == hipSuccess &&
                     d2h_parallel(host_ms, d_ms, (size_t)total * (size_t)E * sizeof(double));
            }
            if (ok && do_hist) ok = flat_from_vectors(acc_v, acc_c, &flat, hist_size);
            pt.mark("chunks");
        }
        dfree(d_ms);
        if (!ok) {
            free(flat);
            if (hist_size) *hist_size = 0;
            return nullptr;
        }
        record_lifted(reruns, dev_ms, (int64_t)sel.size(), n_values);
        return flat;
    } else {
        std::vector<double> vals;
        for (size_t i = 0; i < sel.size(); i++) {
            if (thread_cancelled()) {
                set_error("cancelled");
                return nullptr;
            }
            double *row = host_ms ? host_ms + (size_t)sel[i] * (size_t)E : nullptr;
            if (!plain_item(g, kind, &verts[i * (size_t)k], k, J.marks, iterations, epsilon, precision, vals, row))
                return nullptr;
        }
        record_lifted(0, 0.0, (int64_t)sel.size(), (int64_t)vals.size());
        if (!do_hist) return nullptr;
        return host_values_flat(vals, s, hist_size);
    }
}

dress_hist_pair_t *to_pairs(int64_t *flat, int flat_len, int *hist_size)
{
    // dress_histogram.c:184-200: the flat array (value bits, count, ...) has
    // the byte layout of dress_hist_pair_t {double; int64_t}, so the same
    // malloc'd block is handed over (the caller frees it with free())
    static_assert(sizeof(dress_hist_pair_t) == 2 * sizeof(int64_t), "pair layout");
    if (!flat || flat_len <= 0) {
        free(flat);
        if (hist_size) *hist_size = 0;
        return nullptr;
    }
    if (hist_size) *hist_size = flat_len / 2;
    return reinterpret_cast<dress_hist_pair_t *>(flat);
}

}  // namespace

// a graph of N vertices and E edge records runs in the LDS-resident kernel
bool lds_eligible_size(int N, int E)
{
    return E <= 65535 && lifted_lds_bytes(E + 2, N + 2) <= lifted_lds_limit();
}

}  // namespace dh

using namespace dh;

extern "C" {

int64_t *dress_delta_fit_hip_strided_flat(p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                                          unsigned int seed, int *hist_size, int keep_multisets,
                                          double **multisets, int64_t *num_subgraphs, int offset, int stride)
{
    return lifted_flat(LIFT_DELTA, g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets,
                       multisets, num_subgraphs, offset, stride);
}

int64_t *dress_delta_fit_hip_flat(p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                                  unsigned int seed, int *hist_size, int keep_multisets, double **multisets,
                                  int64_t *num_subgraphs)
{
    return lifted_flat(LIFT_DELTA, g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets,
                       multisets, num_subgraphs, 0, 1);
}

dress_hist_pair_t *dress_delta_fit_hip_strided(p_dress_graph_t g, int k, int iterations, double epsilon,
                                               int n_samples, unsigned int seed, int *hist_size,
                                               int keep_multisets, double **multisets, int64_t *num_subgraphs,
                                               int offset, int stride)
{
    int flat_len = 0;
    int64_t *flat = lifted_flat(LIFT_DELTA, g, k, iterations, epsilon, n_samples, seed,
                                hist_size ? &flat_len : nullptr, keep_multisets, multisets, num_subgraphs, offset,
                                stride);
    return to_pairs(flat, flat_len, hist_size);
}

dress_hist_pair_t *dress_delta_fit_hip(p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                                       unsigned int seed, int *hist_size, int keep_multisets, double **multisets,
                                       int64_t *num_subgraphs)
{
    return dress_delta_fit_hip_strided(g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets,
                                       multisets, num_subgraphs, 0, 1);
}

int64_t *dress_nabla_fit_hip_strided_flat(p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                                          unsigned int seed, int *hist_size, int keep_multisets,
                                          double **multisets, int64_t *num_tuples, int offset, int stride)
{
    return lifted_flat(LIFT_NABLA, g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets,
                       multisets, num_tuples, offset, stride);
}

int64_t *dress_nabla_fit_hip_flat(p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                                  unsigned int seed, int *hist_size, int keep_multisets, double **multisets,
                                  int64_t *num_tuples)
{
    return lifted_flat(LIFT_NABLA, g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets,
                       multisets, num_tuples, 0, 1);
}

// nabla_dress.c:56-69: the pair form always builds the histogram
dress_hist_pair_t *dress_nabla_fit_hip_strided(p_dress_graph_t g, int k, int iterations, double epsilon,
                                               int n_samples, unsigned int seed, int *hist_size,
                                               int keep_multisets, double **multisets, int64_t *num_tuples,
                                               int offset, int stride)
{
    int flat_len = 0;
    int64_t *flat = lifted_flat(LIFT_NABLA, g, k, iterations, epsilon, n_samples, seed, &flat_len, keep_multisets,
                                multisets, num_tuples, offset, stride);
    return to_pairs(flat, flat_len, hist_size);
}

dress_hist_pair_t *dress_nabla_fit_hip(p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                                       unsigned int seed, int *hist_size, int keep_multisets, double **multisets,
                                       int64_t *num_tuples)
{
    return dress_nabla_fit_hip_strided(g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets,
                                       multisets, num_tuples, 0, 1);
}

// In libdress_hip.so the plain names run on the GPU as well.
dress_hist_pair_t *dress_delta_fit(p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                                   unsigned int seed, int *hist_size, int keep_multisets, double **multisets,
                                   int64_t *num_subgraphs)
{
    return dress_delta_fit_hip(g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets, multisets,
                               num_subgraphs);
}

int64_t *dress_delta_fit_flat(p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                              unsigned int seed, int *hist_size, int keep_multisets, double **multisets,
                              int64_t *num_subgraphs)
{
    return dress_delta_fit_hip_flat(g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets,
                                    multisets, num_subgraphs);
}

dress_hist_pair_t *dress_delta_fit_strided(p_dress_graph_t g, int k, int iterations, double epsilon,
                                           int n_samples, unsigned int seed, int *hist_size, int keep_multisets,
                                           double **multisets, int64_t *num_subgraphs, int offset, int stride)
{
    return dress_delta_fit_hip_strided(g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets,
                                       multisets, num_subgraphs, offset, stride);
}

int64_t *dress_delta_fit_strided_flat(p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                                      unsigned int seed, int *hist_size, int keep_multisets, double **multisets,
                                      int64_t *num_subgraphs, int offset, int stride)
{
    return dress_delta_fit_hip_strided_flat(g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets,
                                            multisets, num_subgraphs, offset, stride);
}

dress_hist_pair_t *dress_nabla_fit(p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                                   unsigned int seed, int *hist_size, int keep_multisets, double **multisets,
                                   int64_t *num_tuples)
{
    return dress_nabla_fit_hip(g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets, multisets,
                               num_tuples);
}

int64_t *dress_nabla_fit_flat(p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                              unsigned int seed, int *hist_size, int keep_multisets, double **multisets,
                              int64_t *num_tuples)
{
    return dress_nabla_fit_hip_flat(g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets,
                                    multisets, num_tuples);
}

int64_t *dress_nabla_fit_strided_flat(p_dress_graph_t g, int k, int iterations, double epsilon, int n_samples,
                                      unsigned int seed, int *hist_size, int keep_multisets, double **multisets,
                                      int64_t *num_tuples, int offset, int stride)
{
    return dress_nabla_fit_hip_strided_flat(g, k, iterations, epsilon, n_samples, seed, hist_size, keep_multisets,
                                            multisets, num_tuples, offset, stride);
}

void dress_hip_set_lifted_mode(int mode)
{
    std::lock_guard<std::mutex> lk(g_lift_mu);
    g_lift_mode = (mode >= 1 && mode <= 3) ? mode : 0;
}

}  // extern "C"

namespace dh {

static thread_local dress_hip_lifted_stats_t t_lstats;

namespace {

// Host parents of a batch, built by worker threads in graph order while the
// caller runs earlier graphs on the device.  wait_built(end) builds graphs
// below `end` that no worker has taken yet and returns once every graph
// below `end` is built.  The destructor lets the workers finish the graph in
// hand and joins them.
// Host threads for work that follows device copies: started on first use,
// shared by every caller, and left waiting on their queue until the
// process ends.
class HostTeam {
public:
    static HostTeam &get()
    {
        static HostTeam *t = new HostTeam();   // lives to process end
        return *t;
    }
    void submit(std::function<void()> f)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.push_back(std::move(f));
        }
        cv_.notify_one();
    }

private:
    HostTeam()
    {
        const int n = std::max(1, std::min(8, (int)std::thread::hardware_concurrency()));
        for (int i = 0; i < n; i++) std::thread([this] { loop(); }).detach();
    }
    void loop()
    {
        for (;;) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return !q_.empty(); });
                f = std::move(q_.front());
                q_.pop_front();
            }
            f();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> q_;
};

// Counts that came down narrowed (1, 2 or 4 bytes each, into pinned
// staging blocks), widened into the int64 result by the host team in
// chunks, each chunk once the event after its copy has passed.  wait()
// returns when every chunk is done; the staging blocks and events go with
// the set.
class CountWidening {
public:
    CountWidening() = default;
    CountWidening(const CountWidening &) = delete;
    CountWidening &operator=(const CountWidening &) = delete;
    ~CountWidening()
    {
        wait();
        for (void *p : staging) pinned_free(p);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
    std::vector<void *> staging;
    std::vector<hipEvent_t> events;
    void add(const void *src, int64_t *dst, int64_t n, int width, hipEvent_t ev)
    {
        constexpr int64_t CHUNK = (int64_t)1 << 18;
        for (int64_t a = 0; a < n; a += CHUNK) {
            const int64_t b = std::min(n, a + CHUNK);
            {
                std::lock_guard<std::mutex> lk(mu_);
                pending_++;
            }
            HostTeam::get().submit([this, src, dst, a, b, width, ev] {
                const bool ok = hipEventSynchronize(ev) == hipSuccess;
                if (ok) {
                    // streaming stores: the result is not read here again,
                    // and ordinary stores would first read every line
                    if (width == 1)
                        for (int64_t i = a; i < b; i++)
                            __builtin_nontemporal_store((int64_t)static_cast<const uint8_t *>(src)[i], dst + i);
                    else if (width == 2)
                        for (int64_t i = a; i < b; i++)
                            __builtin_nontemporal_store((int64_t)static_cast<const uint16_t *>(src)[i], dst + i);
                    else
                        for (int64_t i = a; i < b; i++)
                            __builtin_nontemporal_store((int64_t)static_cast<const uint32_t *>(src)[i], dst + i);
                    __builtin_ia32_sfence();              // before the chunk counts as done
                }
                // (notified under the lock: the set may go as soon as the
                // waiter sees pending_ reach 0)
                std::lock_guard<std::mutex> lk(mu_);
                if (!ok) failed_ = true;
                if (--pending_ == 0) cv_.notify_all();
            });
        }
    }
    bool wait()
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return pending_ == 0; });
        return !failed_;
    }

private:
    std::mutex mu_;
    std::condition_variable cv_;
    int64_t pending_ = 0;
    bool failed_ = false;
};

class ParentPool {
public:
    ParentPool(int n, int threads, std::function<void(int)> fn) : n_(n), fn_(std::move(fn)), done_((size_t)n, 0)
    {
        for (int t = 0; t < threads; t++) pool_.emplace_back([this] { work(); });
    }
    ~ParentPool()
    {
        stop_ = true;
        for (auto &t : pool_) t.join();
    }
    ParentPool(const ParentPool &) = delete;
    ParentPool &operator=(const ParentPool &) = delete;
    void wait_built(int end)
    {
        for (int gi = next_.load(); gi < end; gi = next_.load())
            if (next_.compare_exchange_weak(gi, gi + 1)) build(gi);
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return built_ >= end; });
    }

private:
    void work()
    {
        while (!stop_) {
            const int gi = next_.fetch_add(1);
            if (gi >= n_) return;
            build(gi);
        }
    }
    void build(int gi)
    {
        fn_(gi);
        std::lock_guard<std::mutex> lk(mu_);
        done_[(size_t)gi] = 1;
        while (built_ < n_ && done_[(size_t)built_]) built_++;
        cv_.notify_all();
    }
    const int n_;
    std::function<void(int)> fn_;
    std::vector<char> done_;
    int built_ = 0;
    std::atomic<int> next_{0};
    std::atomic<bool> stop_{false};
    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<std::thread> pool_;
};

// upper bound of the live values of one graph's items (every edge of every
// subgraph), -1 when it does not fit in 62 bits
int64_t batch_value_bound(int kind, int N, int64_t E, int k)
{
    if (k == 0) return E;
    if (k >= N || E == 0) return 0;
    const int64_t c = kind == LIFT_DELTA ? dress_delta_binom(N, k) : dress_nabla_perm_count(N, k);
    if (c <= 0) return -1;
    if (c > ((int64_t)1 << 62) / E) return -1;
    return c * E;
}

// values per group of a pipelined batch (DRESS_HIP_BATCH_GROUP_VALUES
// overrides; 0 keeps a batch in one group).  The 512-graph delta^1 batch
// (9.6M values, 8 groups) took 4.10-4.16 ms a call against 4.24-4.31 with
// 1.8M (6 groups); ramps of growing groups gave 4.12-4.27.
int64_t batch_group_values()
{
    if (const char *e = getenv("DRESS_HIP_BATCH_GROUP_VALUES")) return (int64_t)atoll(e);
    return 1200000;
}

// consecutive graphs in at most max_groups groups of about per_group values
// each: boundaries b[0] = 0 < b[1] < ... < b[G] = n (one group when the
// values fit in one, or per_group <= 0).  With three groups or more the
// first and the last hold half a group each, so that the device starts
// sooner and the last download is shorter.
std::vector<int> batch_groups(const std::vector<int64_t> &bound, int64_t per_group, int max_groups)
{
    const int n = (int)bound.size();
    int64_t total = 0;
    for (int64_t x : bound) total += x;
    int G = 1;
    if (per_group > 0 && total > per_group) G = (int)std::min<int64_t>(max_groups, (total + per_group - 1) / per_group);
    G = std::max(1, std::min(G, n));
    // group g ends where the values reach ends[g] / ends[G] of the total:
    // shares 1, 2, ..., 2, 1 with three groups or more, else equal
    std::vector<int64_t> ends((size_t)G + 1, 0);
    for (int g = 1; g <= G; g++) ends[(size_t)g] = ends[(size_t)g - 1] + ((G >= 3 && (g == 1 || g == G)) ? 1 : 2);
    std::vector<int> b{0};
    int64_t acc = 0;
    for (int i = 0; i + 1 < n && (int)b.size() < G; i++) {
        acc += bound[(size_t)i];
        if ((__int128)acc * ends[(size_t)G] >= (__int128)total * ends[b.size()]) b.push_back(i + 1);
    }
    b.push_back(n);
    return b;
}

}  // namespace

void record_lifted(int reruns, double device_ms, int64_t subgraphs, int64_t values)
{
    t_lstats.subgraphs = subgraphs;
    t_lstats.values = values;
    t_lstats.fp64_reruns = reruns;
    t_lstats.device_ms = device_ms;
}

}  // namespace dh

extern "C" {

void dress_hip_last_lifted_stats(dress_hip_lifted_stats_t *out)
{
    if (out) *out = dh::t_lstats;
}

void dress_hip_lifted_batch_free(dress_hip_lifted_batch_t *r)
{
    if (!r) return;
    pinned_free(r->hist_offset); pinned_free(r->hist_value); pinned_free(r->hist_count);
    pinned_free(r->qhist_offset); pinned_free(r->qhist_key); pinned_free(r->qhist_count);
    pinned_free(r->qbound_offset); pinned_free(r->qbound_key); pinned_free(r->qbound_count);
    memset(r, 0, sizeof(*r));
}

int dress_hip_fit_batch_eligible(int N, int E) { return (N >= 0 && E >= 0 && lds_eligible_size(N, E)) ? 1 : 0; }

void dress_hip_fit_batch_free(dress_hip_fit_batch_t *r)
{
    if (!r) return;
    free(r->iterations); free(r->delta); pinned_free(r->edge_dress); free(r->vertex_offset);
    pinned_free(r->vertex_dress);
    memset(r, 0, sizeof(*r));
}

// Whole-graph fits as k = 0 items of the batched kernel (dress.c:546-608 per
// graph, with NW = 1.0).
int dress_hip_fit_batch(int n_graphs, const int *N, const int64_t *edge_offset, const int *U, const int *V,
                        const double *W, dress_variant_t variant, int max_iterations, double epsilon,
                        dress_hip_precision_t precision, dress_hip_fit_batch_t *out)
{
    using namespace dh;
    clear_error();
    if (!out) { set_error("NULL result"); return -1; }
    memset(out, 0, sizeof(*out));
    if (n_graphs < 0 || (n_graphs > 0 && (!N || !edge_offset)) || (int)variant < 0 || (int)variant > 3) {
        set_error("bad batch arguments");
        return -1;
    }
    auto t0 = std::chrono::steady_clock::now();
    int64_t totE = n_graphs ? edge_offset[n_graphs] - edge_offset[0] : 0, totN = 0;
    for (int gi = 0; gi < n_graphs; gi++) {
        int64_t Eg = edge_offset[gi + 1] - edge_offset[gi];
        if (N[gi] < 0 || Eg < 0 || !lds_eligible_size(N[gi], (int)Eg)) {
            set_error("graph " + std::to_string(gi) + " does not fit the batched kernel");
            return -1;
        }
        if (!check_vertex_ids(N[gi], Eg, U ? U + edge_offset[gi] : nullptr, V ? V + edge_offset[gi] : nullptr,
                              ("dress_hip_fit_batch graph " + std::to_string(gi)).c_str()))
            return -1;
        totN += N[gi];
    }
    out->n_graphs = n_graphs;
    out->iterations = (int *)malloc(((size_t)n_graphs + 1) * sizeof(int));
    out->delta = (double *)malloc(((size_t)n_graphs + 1) * sizeof(double));
    out->edge_dress = (double *)pinned_alloc(((size_t)totE + 1) * sizeof(double));
    out->vertex_offset = (int64_t *)malloc(((size_t)n_graphs + 1) * sizeof(int64_t));
    out->vertex_dress = (double *)pinned_alloc(((size_t)totN + 1) * sizeof(double));
    if (!out->iterations || !out->delta || !out->edge_dress || !out->vertex_offset || !out->vertex_dress) {
        dress_hip_fit_batch_free(out);
        set_error("host allocation of batch results failed");
        return -1;
    }
    out->vertex_offset[0] = 0;
    for (int gi = 0; gi < n_graphs; gi++) out->vertex_offset[gi + 1] = out->vertex_offset[gi] + N[gi];
    if (max_iterations <= 0) {
        // dress.c:546-608 runs no iteration: values stay at their initial state
        for (int gi = 0; gi < n_graphs; gi++) { out->iterations[gi] = max_iterations; out->delta[gi] = 0.0; }
        for (int64_t e = 0; e < totE; e++) out->edge_dress[e] = 1.0;
        for (int64_t v = 0; v < totN; v++) out->vertex_dress[v] = 0.0;
        return 0;
    }
    BatchScope scope;
    hipStream_t s = ctx_stream();
    const bool own_stream = s == nullptr;
    if (own_stream && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        dress_hip_fit_batch_free(out);
        set_error("stream creation failed");
        return -1;
    }
    std::vector<HostParent> parents((size_t)n_graphs);
    std::vector<int> graph_of((size_t)n_graphs), item_parent((size_t)n_graphs), item_verts;
    std::vector<int64_t> item_live((size_t)n_graphs), ms_rows;
    FitOutputs fo;
    std::vector<int> iters;
    std::vector<double> deltas;
    fo.iters = &iters;
    fo.deltas = &deltas;
    fo.vn_ofs.assign((size_t)n_graphs, 0);
    for (int gi = 0; gi < n_graphs; gi++) {
        int64_t eb = edge_offset[gi], ee = edge_offset[gi + 1];
        host_parent_build(N[gi], (int)(ee - eb), U + eb, V + eb, W ? W + eb : nullptr, (int)variant,
                          parents[(size_t)gi]);
        graph_of[(size_t)gi] = gi;
        item_parent[(size_t)gi] = gi;
        item_live[(size_t)gi] = ee - eb;
        fo.vn_ofs[(size_t)gi] = out->vertex_offset[gi];
    }
    LiftJob J;
    J.kind = LIFT_DELTA;
    J.k = 0;
    J.max_it = max_iterations;
    J.eps = epsilon;
    J.mixed = precision == DRESS_HIP_FP32_FP64;
    memset(J.marks, 0, sizeof(J.marks));
    std::vector<const HostParent *> ps;
    for (auto &P : parents) ps.push_back(&P);
    LiftOut lo;
    fo.d_vnorm = nullptr;
    bool ok = dalloc(&fo.d_vnorm, (size_t)totN + 1, "vertex norms") &&
              run_items(ps, graph_of, item_parent, item_verts, item_live, J, nullptr, 0, ms_rows, s, lo, nullptr,
                        nullptr, 0, 0.0, &fo);
    if (ok) {
        ok = (totE == 0 || hipMemcpyAsync(out->edge_dress, lo.d_vals, (size_t)totE * sizeof(double),
                                          hipMemcpyDeviceToHost, s) == hipSuccess) &&
             (totN == 0 || hipMemcpyAsync(out->vertex_dress, fo.d_vnorm, (size_t)totN * sizeof(double),
                                          hipMemcpyDeviceToHost, s) == hipSuccess) &&
             hipStreamSynchronize(s) == hipSuccess;
        if (!ok) set_error("batch result download failed");
    }
    dfree(lo.d_vals);
    dfree(fo.d_vnorm);
    if (own_stream) (void)hipStreamDestroy(s);
    if (!ok) {
        dress_hip_fit_batch_free(out);
        return -1;
    }
    for (int gi = 0; gi < n_graphs; gi++) {
        out->iterations[gi] = iters[(size_t)gi];
        out->delta[gi] = deltas[(size_t)gi];
    }
    out->fp64_reruns = lo.reruns;
    out->device_ms = lo.device_ms;
    out->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

// Many graphs at once: every subgraph of every graph in one persistent launch;
// per graph the exact (value, count) histogram and the quantised one.
int dress_hip_lifted_batch(int kind, int k, int n_graphs, const int *N, const int64_t *edge_offset, const int *U,
                           const int *V, const double *W, dress_variant_t variant, int max_iterations,
                           double epsilon, dress_hip_precision_t precision, double qtol,
                           dress_hip_lifted_batch_t *out)
{
    using namespace dh;
    auto t0 = std::chrono::steady_clock::now();
    PhaseTimer pt("dress_hip_lifted_batch");
    clear_error();
    if (!out) { set_error("NULL result"); return -1; }
    memset(out, 0, sizeof(*out));
    if (n_graphs < 0 || k < 0 || k > LIFT_KMAX || (kind != LIFT_DELTA && kind != LIFT_NABLA)) {
        set_error("bad batch arguments");
        return -1;
    }
    if (n_graphs > 0 && (!N || !edge_offset)) {
        set_error("bad batch arguments");
        return -1;
    }
    for (int gi = 0; gi < n_graphs; gi++) {
        const int64_t Eg = edge_offset[gi + 1] - edge_offset[gi];
        if (N[gi] < 0 || Eg < 0) {
            set_error("graph " + std::to_string(gi) + ": negative size");
            return -1;
        }
        if (!check_vertex_ids(N[gi], Eg, U ? U + edge_offset[gi] : nullptr, V ? V + edge_offset[gi] : nullptr,
                              ("dress_hip_lifted_batch graph " + std::to_string(gi)).c_str()))
            return -1;
    }
    for (int gi = 0; gi < n_graphs; gi++) {
        const int64_t Eg = edge_offset[gi + 1] - edge_offset[gi];
        if (Eg > INT32_MAX || !lds_eligible_size(N[gi], (int)Eg)) {
            set_error("graph " + std::to_string(gi) + " is too large for the batched kernel");
            return -1;
        }
    }
    pt.mark("arguments checked");
    BatchScope scope;
    hipStream_t s = ctx_stream();
    const bool own_stream = s == nullptr;
    if (own_stream && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        set_error("stream creation failed");
        return -1;
    }
    // the floor(d / eps) tables only on request (qtol >= 0)
    const bool quant = qtol >= 0.0;
    // Groups of consecutive graphs, run one after the other: while the
    // device fits a group, host threads build the parents of later groups
    // and a copy stream downloads the histograms of earlier ones.  The
    // result arrays are then sized by the bound on the values before the
    // histograms exist.  Batches with quantised tables, inside a server
    // worker, or of one group's size run as one group.
    std::vector<int64_t> bound((size_t)n_graphs);
    bool bounded = true;
    int64_t cap = 0;
    for (int gi = 0; gi < n_graphs; gi++) {
        bound[(size_t)gi] = batch_value_bound(kind, N[gi], edge_offset[gi + 1] - edge_offset[gi], k);
        if (bound[(size_t)gi] < 0 || cap > ((int64_t)1 << 62) - bound[(size_t)gi]) bounded = false;
        else cap += bound[(size_t)gi];
    }
    // several groups: group g + 1 is launched on s before group g's
    // histogram runs on hs and its download on cs
    hipStream_t cs = nullptr, hs = s;
    std::vector<int> gb{0, n_graphs};
    if (!quant && bounded && scope.on) {
        std::vector<int> b = batch_groups(bound, batch_group_values(), 8);
        hipStream_t c = batch_side_stream(0), h = batch_side_stream(1);
        if (b.size() > 2 && c && h) {
            gb = b;
            cs = c;
            hs = h;
        }
    }
    const int G = (int)gb.size() - 1;
    pt.mark("groups chosen");
    // parents and items of each graph (host worker threads, graph order)
    std::vector<HostParent> parents((size_t)n_graphs);
    std::vector<std::vector<int>> g_verts((size_t)n_graphs);
    std::vector<std::vector<int64_t>> g_live((size_t)n_graphs);
    auto build_one = [&](int gi) {
        int64_t eb = edge_offset[gi], ee = edge_offset[gi + 1];
        int Eg = (int)(ee - eb);
        host_parent_build(N[gi], Eg, U + eb, V + eb, W ? W + eb : nullptr, (int)variant, parents[(size_t)gi]);
        if (k == 0) {                        // the graph itself
            g_live[(size_t)gi].assign(1, Eg);
            return;
        }
        if (k >= N[gi]) return;
        int64_t total = kind == LIFT_DELTA ? dress_delta_binom(N[gi], k) : dress_nabla_perm_count(N[gi], k);
        std::vector<int64_t> sel;
        enumerate_items(kind, N[gi], k, 0, 0, 0, 1, total, sel, g_verts[(size_t)gi]);
        dress_graph_t tmp;
        memset(&tmp, 0, sizeof(tmp));
        tmp.N = N[gi];
        tmp.E = Eg;
        tmp.U = const_cast<int *>(U + eb);
        tmp.V = const_cast<int *>(V + eb);
        g_live[(size_t)gi] = kind == LIFT_DELTA ? delta_live_counts(&tmp, g_verts[(size_t)gi], k, sel.size())
                                                : std::vector<int64_t>(sel.size(), (int64_t)Eg);
    };
    LiftJob J;
    J.kind = kind;
    J.k = (k == 0) ? 0 : k;
    J.max_it = max_iterations;
    J.eps = epsilon;
    J.mixed = precision == DRESS_HIP_FP32_FP64;
    memset(J.marks, 0, sizeof(J.marks));
    if (kind == LIFT_NABLA) marker_weights(J.marks, k);
    out->n_graphs = n_graphs;
    out->hist_offset = (int64_t *)malloc(((size_t)n_graphs + 1) * sizeof(int64_t));
    out->qhist_offset = (int64_t *)malloc(((size_t)n_graphs + 1) * sizeof(int64_t));
    out->qbound_offset = (int64_t *)malloc(((size_t)n_graphs + 1) * sizeof(int64_t));
    bool ok = out->hist_offset && out->qhist_offset && out->qbound_offset;
    // A pipelined group's counts come down at the narrowest width that
    // holds its largest count and are widened on host threads (into a
    // block whose pages have been touched), so the link carries 9 rather
    // than 16 bytes per distinct value when every count is below 256.
    // DRESS_HIP_NARROW_COUNTS=0 downloads them as int64.
    const char *nc = getenv("DRESS_HIP_NARROW_COUNTS");
    const bool narrow_counts = !(nc && !strcmp(nc, "0"));
    if (ok && G > 1) {
        out->hist_value = (double *)pinned_alloc(((size_t)cap + 1) * sizeof(double));
        out->hist_count = (int64_t *)pinned_alloc(((size_t)cap + 1) * sizeof(int64_t), narrow_counts);
        ok = out->hist_value && out->hist_count;
    }
    if (!ok) set_error("host allocation of batch results failed");
    pt.mark("results allocated");
    auto d2h = [](void *dst, const void *src, size_t bytes, hipStream_t st) {
        return bytes == 0 || hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st) == hipSuccess;
    };
    std::vector<void *> held;            // device blocks that downloads still read
    CountWidening widen;
    unsigned long long *qk = nullptr, *qc = nullptr;
    int64_t qcap = 1024, runs_total = 0, subgraphs = 0, values = 0;
    int reruns = 0;
    double device_ms = 0.0;
    unsigned long long work[4] = {0, 0, 0, 0};
    // the run of each group prepared on the host ahead of the device: by a
    // planner thread when there are several groups
    struct GroupPlan {
        RunState R;
        LiftOut lo;
        std::vector<const HostParent *> ps;
        std::vector<int> graph_of, item_parent, item_verts;
        std::vector<int64_t> item_live, ms_rows;
        size_t max_items = 0;                 // most items of one graph
        HistPlan hp;                          // the histogram's merges, when queued without a wait
        bool hist_planned = false;
        bool ok = false;
        std::string err;
    };
    // In a pipelined batch, items leave k_lifted with their values sorted
    // and each graph's histogram merges its items' runs, when that takes at
    // most 8 rounds (256 items per graph); else the values are radix sorted.
    // (The 512-graph delta^1 batch: 4.6 ms a call against 5.5.  In one
    // group the sort's cost inside k_lifted, 0.18 ms, is not hidden and the
    // radix sort is faster: 6.8 against 7.1 ms.)  DRESS_HIP_HIST_MERGE=0
    // keeps the radix sort, =1 merges in one group too.
    const char *hm = getenv("DRESS_HIP_HIST_MERGE");
    const int merge_mode = hm ? atoi(hm) : -1;
    auto merges = [&](const GroupPlan &P) {
        return merge_mode != 0 && (G > 1 || merge_mode == 1) && P.max_items <= 256;
    };
    // each graph's values are one segment of its group's histogram: the
    // items of a graph are contiguous
    auto group_segments = [&](int gr, const GroupPlan &P) {
        const int ga = gb[(size_t)gr], gz = gb[(size_t)gr + 1];
        std::vector<int64_t> seg_ofs((size_t)(gz - ga) + 1, 0);
        size_t it = 0;
        for (int gi = ga; gi < gz; gi++) {
            while (it < P.item_parent.size() && P.item_parent[it] == gi - ga) it++;
            seg_ofs[(size_t)(gi - ga) + 1] = P.lo.val_ofs[it];
        }
        return seg_ofs;
    };
    // With merged runs, each group's histogram is planned with the group
    // and queued on hs behind an event after the group's launches (in
    // FP32->FP64, after the FP64 redo of its ambiguous items), with no wait
    // on the host: it runs beside the next group's launches, and the host
    // waits for its event only to queue the group's downloads.
    // DRESS_HIP_HIST_ASYNC=0 makes the host wait for each group's launches,
    // then run its histogram on hs.
    const char *ha = getenv("DRESS_HIP_HIST_ASYNC");
    const bool async_hist = !(ha && !strcmp(ha, "0")) && G > 1 && hs != s && !quant;
    std::vector<std::unique_ptr<GroupPlan>> plans((size_t)G);
    {
        // host threads building parents: with several groups only the first
        // group's wait for them, and 8 leave the calling thread room (the
        // 512-graph delta^1 batch: 5.5 ms a call against 5.9 with 16, whose
        // times also spread twice as wide); DRESS_HIP_BATCH_THREADS overrides
        int nt = std::max(1, std::min({n_graphs / 32, (int)std::thread::hardware_concurrency(), G > 1 ? 8 : 16}));
        if (const char *e = getenv("DRESS_HIP_BATCH_THREADS")) nt = std::max(1, atoi(e));
        ParentPool pool(n_graphs, nt - 1, build_one);
        auto plan_group = [&](int gr) {
            PhaseTimer ppt("  plan_group");
            std::unique_ptr<GroupPlan> P(new GroupPlan());
            const int ga = gb[(size_t)gr], gz = gb[(size_t)gr + 1];
            pool.wait_built(gz);
            ppt.mark("parents built");
            for (int gi = ga; gi < gz; gi++) {
                P->graph_of.push_back(gi);
                P->ps.push_back(&parents[(size_t)gi]);
                P->max_items = std::max(P->max_items, g_live[(size_t)gi].size());
                for (int64_t x : g_live[(size_t)gi]) {
                    P->item_parent.push_back(gi - ga);
                    P->item_live.push_back(x);
                }
                P->item_verts.insert(P->item_verts.end(), g_verts[(size_t)gi].begin(), g_verts[(size_t)gi].end());
            }
            ppt.mark("items");
            P->ok = run_prepare(P->ps, P->graph_of, P->item_parent, P->item_verts, P->item_live, J, P->ms_rows,
                                nullptr, P->R, P->lo);
            ppt.mark("prepared");
            if (P->ok && async_hist && merges(*P) && gz > ga) {
                const std::vector<int64_t> item_end(P->lo.val_ofs.begin() + 1, P->lo.val_ofs.end());
                const int r = hist_plan(P->hp, group_segments(gr, *P), item_end);
                P->ok = r >= 0;
                P->hist_planned = r > 0;
                ppt.mark("histogram planned");
            }
            if (!P->ok) P->err = dress_hip_last_error();
            // a host-side test hook: DRESS_HIP_BATCH_FAIL_GROUP=g fails group g's
            // preparation, while earlier groups may be on the device
            if (const char *e = getenv("DRESS_HIP_BATCH_FAIL_GROUP"))
                if (P->ok && atoi(e) == gr) {
                    P->ok = false;
                    P->err = "preparation of batch group " + std::to_string(gr) + " failed (DRESS_HIP_BATCH_FAIL_GROUP)";
                }
            return P;
        };
        std::mutex pm;
        std::condition_variable pcv;
        int planned = 0;
        std::atomic<bool> stop_planning{false};
        std::thread planner;
        if (G > 1)
            planner = std::thread([&] {
                for (int gr = 0; gr < G && !stop_planning; gr++) {
                    std::unique_ptr<GroupPlan> P = plan_group(gr);
                    std::lock_guard<std::mutex> lk(pm);
                    plans[(size_t)gr] = std::move(P);
                    planned = gr + 1;
                    pcv.notify_all();
                }
            });
        pt.mark("threads started");
        // group gr's run queued on s once its plan is ready
        auto launch_group = [&](int gr) -> bool {
            if (G > 1) {
                std::unique_lock<std::mutex> lk(pm);
                pcv.wait(lk, [&] { return planned > gr; });
            } else {
                plans[0] = plan_group(0);
            }
            GroupPlan &P = *plans[(size_t)gr];
            if (gr == 0) pt.mark("first group prepared");
            if (!P.ok) {
                set_error(P.err.empty() ? std::string("batch preparation failed") : P.err);
                return false;
            }
            if (quant) {
                while (qcap < 2 * P.lo.val_ofs.back()) qcap <<= 1;
                if (!(dalloc(&qk, (size_t)qcap, "quantised table") && dalloc(&qc, (size_t)qcap, "quantised table")))
                    return false;
                hipLaunchKernelGGL(k_qtable_clear, dim3(grid_for(qcap, 4096)), dim3(TPB), 0, s, qk, qc, qcap);
            }
            return run_launch(P.R, P.ps, P.item_parent, J, nullptr, 0, s, P.lo, qk, qc, quant ? qcap : 0,
                              quant ? qtol : 0.0, nullptr, merges(P));
        };
        std::vector<AsyncHist> ah((size_t)G);
        std::vector<hipEvent_t> hev((size_t)G, nullptr);
        hipEvent_t kev = nullptr;                        // end of a group's launches on s
        if (async_hist && hipEventCreateWithFlags(&kev, hipEventDisableTiming) != hipSuccess) {
            set_error("histogram event failed");
            ok = false;
        }
        // two groups stay queued on s ahead of the one being finished, so
        // the device goes from one group's launches to the next while the
        // host waits for a histogram
        bool chain = ok;                                 // every launch so far queued
        int queued = 0;                                  // groups [0, queued) are on s
        // group g's planned histogram queued on hs, behind the work on s so
        // far when after_s
        auto enqueue_hist = [&](int g, bool after_s) {
            GroupPlan &P = *plans[(size_t)g];
            if (!P.hist_planned) return;
            if (after_s && (hipEventRecord(kev, s) != hipSuccess || hipStreamWaitEvent(hs, kev, 0) != hipSuccess)) {
                set_error("histogram event failed");
                chain = false;
                return;
            }
            if (!hist_enqueue(ah[(size_t)g], P.hp, P.lo.d_vals, P.lo.val_ofs.back(), hs)) chain = false;
            else if (hipEventCreateWithFlags(&hev[(size_t)g], hipEventDisableTiming) != hipSuccess ||
                     hipEventRecord(hev[(size_t)g], hs) != hipSuccess) {
                set_error("histogram event failed");
                chain = false;
            }
        };
        auto queue_group = [&](int g) {
            if (!chain || g >= G) return;
            if (!launch_group(g)) {
                chain = false;
                return;
            }
            queued = g + 1;
            if (!J.mixed) enqueue_hist(g, true);
        };
        // whether group g's plan is ready (then queueing it does not wait)
        auto plan_ready = [&](int g) {
            std::lock_guard<std::mutex> lk(pm);
            return planned > g;
        };
        // group gr's histogram collected (or made on hs and waited for),
        // its downloads queued on cs, then its host state released
        auto finish_group = [&](int gr) -> bool {
            const int ga = gb[(size_t)gr], gz = gb[(size_t)gr + 1];
            GroupPlan &P = *plans[(size_t)gr];
            LiftOut &lo = P.lo;
            double *hv = nullptr;
            int64_t *hc = nullptr, runs = 0;
            void *packed = nullptr;                      // the counts at width bytes each (width < 8)
            int width = 8;
            std::vector<int64_t> run_ofs((size_t)(gz - ga) + 1, 0);
            bool have = false;                           // the histogram is on the device
            bool good = true;
            if (hev[(size_t)gr]) {
                AsyncHist &H = ah[(size_t)gr];
                good = hipEventSynchronize(hev[(size_t)gr]) == hipSuccess;
                if (!good) set_error("batch histogram failed");
                bool failed = true;
                if (good) hist_collect(H, &runs, run_ofs, &width, &failed);
                have = good && !failed;
                if (have) {
                    hv = H.val;
                    hc = H.cnt;
                    packed = H.packed;
                    held.push_back(packed);              // (hv and hc below)
                } else {
                    width = 8;
                }
                hist_release(H, have);
                (void)hipEventDestroy(hev[(size_t)gr]);
                hev[(size_t)gr] = nullptr;
                pt.mark("group histogram");
            }
            // else (and when the packed keys did not fit) on hs, waited for
            if (good && gz > ga && !have) {
                std::vector<int64_t> item_end;
                if (merges(P)) item_end.assign(lo.val_ofs.begin() + 1, lo.val_ofs.end());
                good = exact_histogram_dev(lo.d_vals, lo.val_ofs.back(), group_segments(gr, P), hs, &hv, &hc, &runs,
                                           run_ofs, merges(P) ? &item_end : nullptr);
            }
            if (gr == 0) pt.mark("first group on the device");
            if (good && G == 1) {
                out->hist_value = (double *)pinned_alloc(((size_t)runs + 1) * sizeof(double));
                out->hist_count = (int64_t *)pinned_alloc(((size_t)runs + 1) * sizeof(int64_t));
                good = out->hist_value && out->hist_count;
                if (!good) set_error("host allocation of batch results failed");
            } else if (good && runs_total + runs > cap) {
                set_error("batch histogram larger than its bound");
                good = false;
            }
            if (good) {
                for (int gi = ga; gi < gz; gi++) out->hist_offset[gi] = runs_total + run_ofs[(size_t)(gi - ga)];
                good = d2h(out->hist_value + runs_total, hv, (size_t)runs * sizeof(double), cs ? cs : s);
                if (good && packed && width < 8 && narrow_counts && cs && runs > 0) {
                    // into a pinned staging block (pool-sized, so the copy
                    // is asynchronous), widened after the event (a blocking
                    // one: the host threads that wait for it sleep instead
                    // of spinning beside the calling thread)
                    void *stg =
                        pinned_alloc(std::max<size_t>((size_t)runs * (size_t)width, (size_t)1 << 20), true);
                    hipEvent_t ev = nullptr;
                    if (stg) widen.staging.push_back(stg);
                    good = stg && hipEventCreateWithFlags(&ev, hipEventDisableTiming | hipEventBlockingSync) ==
                                      hipSuccess;
                    if (ev) widen.events.push_back(ev);
                    good = good && d2h(stg, packed, (size_t)runs * (size_t)width, cs) &&
                           hipEventRecord(ev, cs) == hipSuccess;
                    if (good) widen.add(stg, out->hist_count + runs_total, runs, width, ev);
                } else {
                    good = good && d2h(out->hist_count + runs_total, hc, (size_t)runs * sizeof(int64_t), cs ? cs : s);
                }
                if (!good && !has_error()) set_error("batch result download failed");
            }
            held.push_back(hv);
            held.push_back(hc);
            if (!good) return false;
            pt.mark("downloads queued");
            dfree(lo.d_vals);
            lo.d_vals = nullptr;
            runs_total += runs;
            subgraphs += (int64_t)P.item_parent.size();
            values += lo.val_ofs.back();
            reruns += lo.reruns;
            device_ms += lo.device_ms;
            for (int i = 0; i < 4; i++) work[i] += lo.work[i];
            plans[(size_t)gr].reset();
            for (int gi = ga; gi < gz; gi++) {
                parents[(size_t)gi] = HostParent();
                std::vector<int>().swap(g_verts[(size_t)gi]);
            }
            return true;
        };
        queue_group(0);
        queue_group(1);
        // Group gr - 1's histogram is collected and its downloads queued
        // once group gr's launches are done: by then the histogram, queued
        // behind gr - 1's launches, has long finished, and the host waits
        // only for launches.
        for (int gr = 0; ok && gr < G; gr++) {
            GroupPlan &P = *plans[(size_t)gr];
            // (FP32->FP64: the FP64 redo on hs, and the histogram behind it
            // there, since the group's launches are done)
            ok = run_finish(P.R, P.ps, P.graph_of, P.item_parent, J, s, P.lo, qk, qc, quant ? qcap : 0,
                            quant ? qtol : 0.0, nullptr, gr < queued, J.mixed && async_hist ? hs : nullptr);
            pt.mark("group launches done");
            if (ok && J.mixed) {                         // the values are final after the FP64 redo
                enqueue_hist(gr, false);
                if (!chain) ok = false;
            }
            // the group after next: now when its plan is ready, so that s
            // always holds a group; else after the downloads
            bool next_queued = false;
            if (ok && gr + 2 < G && plan_ready(gr + 2)) {
                queue_group(gr + 2);
                next_queued = true;
                pt.mark("next launch queued");
            }
            if (ok && gr > 0) ok = finish_group(gr - 1);
            if (ok && !next_queued) {
                queue_group(gr + 2);
                pt.mark("next launch queued");
            }
            if (!chain) ok = false;                      // a later group failed
        }
        if (ok && G > 0) ok = finish_group(G - 1);
        if (!ok) {
            // queued work may still read the groups' blocks
            (void)hipStreamSynchronize(s);
            if (hs != s) (void)hipStreamSynchronize(hs);
            if (cs) (void)hipStreamSynchronize(cs);
        }
        for (int g = 0; g < G; g++)                      // histograms of groups not reached
            if (hev[(size_t)g]) {
                hist_release(ah[(size_t)g], false);
                (void)hipEventDestroy(hev[(size_t)g]);
            }
        if (kev) (void)hipEventDestroy(kev);
        stop_planning = true;
        if (planner.joinable()) planner.join();
        for (auto &P : plans)
            if (P) {
                dfree(P->lo.d_vals);
                P->lo.d_vals = nullptr;
            }
        plans.clear();
    }
    pt.mark("groups");
    if (ok) out->hist_offset[n_graphs] = runs_total;
    int64_t nq = 0, nb = 0;
    int64_t *qk2 = nullptr, *qc2 = nullptr, *bk2 = nullptr, *bc2 = nullptr, *qofs = nullptr, *bofs = nullptr;
    if (ok && quant) ok = qtable_split(qk, qc, qcap, n_graphs, s, &qk2, &qc2, &bk2, &bc2, &qofs, &bofs, &nq, &nb);
    pt.mark("qtable_split");
    if (pt.on)
        fprintf(stderr, "dress_hip_lifted_batch: %d groups, %lld values, %lld exact pairs, %lld bins, %lld boundary bins\n",
                G, (long long)values, (long long)runs_total, (long long)nq, (long long)nb);
    if (ok) {
        out->qhist_key = (int64_t *)pinned_alloc(((size_t)nq + 1) * sizeof(int64_t));
        out->qhist_count = (int64_t *)pinned_alloc(((size_t)nq + 1) * sizeof(int64_t));
        out->qbound_key = (int64_t *)pinned_alloc(((size_t)nb + 1) * sizeof(int64_t));
        out->qbound_count = (int64_t *)pinned_alloc(((size_t)nb + 1) * sizeof(int64_t));
        if (!out->qhist_key || !out->qhist_count || !out->qbound_key || !out->qbound_count) {
            set_error("host allocation of batch results failed");
            ok = false;
        }
    }
    if (ok) {
        if (!quant)
            for (int gi = 0; gi <= n_graphs; gi++) out->qhist_offset[gi] = out->qbound_offset[gi] = 0;
        ok = d2h(out->qhist_key, qk2, (size_t)nq * sizeof(int64_t), s) &&
             d2h(out->qhist_count, qc2, (size_t)nq * sizeof(int64_t), s) &&
             d2h(out->qbound_key, bk2, (size_t)nb * sizeof(int64_t), s) &&
             d2h(out->qbound_count, bc2, (size_t)nb * sizeof(int64_t), s) &&
             (!quant || d2h(out->qhist_offset, qofs, ((size_t)n_graphs + 1) * sizeof(int64_t), s)) &&
             (!quant || d2h(out->qbound_offset, bofs, ((size_t)n_graphs + 1) * sizeof(int64_t), s));
        if (!ok && !has_error()) set_error("batch result download failed");
    }
    // every download has finished before its device blocks go back
    if (cs && hipStreamSynchronize(cs) != hipSuccess && ok) {
        set_error("batch result download failed");
        ok = false;
    }
    if (hipStreamSynchronize(s) != hipSuccess && ok) {
        set_error("batch result download failed");
        ok = false;
    }
    if (!widen.wait() && ok) {                    // before the results may be freed below
        set_error("batch result download failed");
        ok = false;
    }
    pt.mark("download");
    for (void *p : held) dfree(p);
    dfree(qk2); dfree(qc2); dfree(bk2); dfree(bc2); dfree(qofs); dfree(bofs);
    dfree(qk); dfree(qc);
    if (own_stream) (void)hipStreamDestroy(s);
    if (!ok) {
        dress_hip_lifted_batch_free(out);
        return -1;
    }
    out->subgraphs = subgraphs;
    out->values = values;
    out->fp64_reruns = reruns;
    out->device_ms = device_ms;
    for (int i = 0; i < 4; i++) out->work[i] = (int64_t)work[i];
    out->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    record_lifted(reruns, device_ms, subgraphs, values);
    return 0;
}

}  // extern "C"

extern "C" {

// Host-only: the subgraph indices (and vertex tuples) a lifted call with
// these arguments processes, in processing order.  Returns the count, or -1.
int64_t dress_hip_enumerate_subgraphs(int kind, int N, int k, int n_samples, unsigned int seed, int offset,
                                      int stride, int64_t *sel, int *verts, int64_t cap)
{
    using namespace dh;
    clear_error();
    if (N < 0 || k < 1 || k >= N || stride <= 0 || offset < 0 || (kind != LIFT_DELTA && kind != LIFT_NABLA)) {
        set_error("bad enumeration arguments");
        return -1;
    }
    int64_t total = kind == LIFT_DELTA ? dress_delta_binom(N, k) : dress_nabla_perm_count(N, k);
    std::vector<int64_t> s;
    std::vector<int> v;
    enumerate_items(kind, N, k, n_samples, seed, offset, stride, total, s, v);
    int64_t n = (int64_t)s.size();
    for (int64_t i = 0; i < n && i < cap; i++) {
        if (sel) sel[i] = s[(size_t)i];
        if (verts) memcpy(verts + i * k, &v[(size_t)i * k], (size_t)k * sizeof(int));
    }
    return n;
}

}  // extern "C"
