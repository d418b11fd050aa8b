// dress_serve.hip -- the request server (continuous batching) in front of
// the engine.
//
// Requests (fit, delta^k, nabla^k) are copied into a queue.  Each worker
// thread owns a stream and a device-block cache (ctx_begin), so the work
// of different requests overlaps on the device and no hipFree stalls it.
// A worker takes the oldest request of its lane (batched fits, device-wide
// fits, batched delta/nabla, large delta/nabla) whose estimated footprint
// fits in the memory budget beside the running ones; a request that fits
// the LDS-resident kernel takes
// every other waiting request of the same kind and parameters along, up to
// max_batch, a work cap and the budget, and they run as one batched launch
// (dress_hip_fit_batch, dress_hip_lifted_batch).  A result does not depend
// on what ran beside it; FP64 results equal the single-graph entry points'
// bit for bit.
#include "dress_capi_internal.h"
#include "dress_hip_internal.h"

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dress/hip/dress_hip.h"

struct dress_hip_server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<struct ServeReq *> lanes[4];   // waiting requests, one queue per lane
    int64_t queued() const
    {
        return (int64_t)(lanes[0].size() + lanes[1].size() + lanes[2].size() + lanes[3].size());
    }
    std::deque<dress_hip_response_t> done;
    std::vector<std::thread> workers;
    int device = 0;
    int n_workers = 1;
    int max_batch = 64;
    bool stop = false;
    std::atomic<bool> cancel{false};    // running requests give up at the next subgraph
    int64_t next_id = 0;
    int64_t reserved_large = 0;         // part of st.reserved_bytes held by the large lane
    std::atomic<int> short_load{0};     // queued + running requests outside the large lane
    std::atomic<int> fit_load{0};       // queued + running fit requests
    std::atomic<int> inject{0};         // test hook: HIP error code for the next finished request
    bool device_failed = false;         // a device error left the device unusable
    std::string failure;                // ... and what it was
    int64_t host_budget = 0;            // largest result one request may hold in host memory
    int64_t chunk_values = 0;           // most subgraph values of a request on the device at once
    dress_hip_server_stats_t st{};
};

struct ServeReq {
    int64_t id;
    dress_hip_request_t r;          // pointer fields unused; see U/V/W
    std::vector<int> U, V;
    std::vector<double> W;
    bool weighted;
    bool lds;                       // fits the batched kernel
    int64_t bytes;                  // estimated device footprint
    int64_t work;                   // edges (fit) or subgraph values (delta/nabla)
    double arrival;
};

namespace {

using namespace dh;

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int64_t count_items(int kind, int N, int k)
{
    if (kind == DRESS_HIP_REQ_FIT || k == 0) return 1;
    if (k >= N) return 0;
    return kind == DRESS_HIP_REQ_DELTA ? dress_delta_binom(N, k) : dress_nabla_perm_count(N, k);
}

// Device bytes a request may hold at its peak: graph build (half-edge keys,
// both CSRs, intersection index of at most T_bound entries) and fit state;
// lifted requests add their subgraph values and the histogram assembly
// (values, two key buffers, two segment buffers, run flags and positions,
// unique values, run starts and counts: 72 B per value, 96 with the sort's
// temporary storage) and, beyond the batched kernel, one multi-subgraph
// pass (up to 256 slots of values, half-edge copies and norms; at most
// 4 GiB).
// Delta/nabla requests run in chunks of at most this many subgraph values
// on the device (their histograms merged on the host), so a request's
// device footprint is bounded whatever its size: 32M values take 3.6 GB
// (DRESS_HIP_SERVE_CHUNK_VALUES overrides).
constexpr int64_t SERVE_CHUNK_VALUES = (int64_t)32 << 20;

int64_t footprint(const ServeReq &q, int64_t t_bound, int64_t chunk_values)
{
    const int64_t N = q.r.N, E = q.r.E;
    int64_t fit = 400 * E + 64 * N + 8 * t_bound + (1 << 20);
    if (q.r.kind == DRESS_HIP_REQ_FIT) return fit;
    int64_t items = count_items(q.r.kind, q.r.N, q.r.k);
    double vals = std::min((double)items * (double)E, (double)chunk_values);
    double extra = (double)LIFT_BYTES_PER_VALUE * vals;
    if (!q.lds) extra += std::min(2048.0 * (8.0 * (double)E + (double)N + 8.0), 4294967296.0);
    double b = (double)fit + extra;
    return b > 9.0e18 ? (int64_t)9.0e18 : (int64_t)b;
}

// Host bytes a request's result may take at its peak: a fit's values; a
// delta/nabla histogram of up to one pair per subgraph value (16 B), twice
// while chunk histograms are merged
double host_result_bytes(const ServeReq &q)
{
    if (q.r.kind == DRESS_HIP_REQ_FIT) return 8.0 * ((double)q.r.E + (double)q.r.N);
    return 32.0 * (double)count_items(q.r.kind, q.r.N, q.r.k) * (double)q.r.E;
}

// sum over edges of min(deg u, deg v): an upper bound of the index size
int64_t intersection_bound(int N, int E, const int *U, const int *V)
{
    std::vector<int> deg((size_t)N + 1, 0);
    for (int e = 0; e < E; e++) {
        if (U[e] >= 0 && U[e] < N) deg[(size_t)U[e]]++;
        if (V[e] >= 0 && V[e] < N && V[e] != U[e]) deg[(size_t)V[e]]++;
    }
    int64_t t = 0;
    for (int e = 0; e < E; e++) {
        int u = U[e], v = V[e];
        if (u < 0 || u >= N || v < 0 || v >= N) continue;
        t += std::min(deg[(size_t)u], deg[(size_t)v]);
    }
    return t;
}

bool same_batch(const ServeReq &a, const ServeReq &b)
{
    return a.r.kind == b.r.kind && (a.r.kind == DRESS_HIP_REQ_FIT || a.r.k == b.r.k) && a.r.variant == b.r.variant &&
           a.r.max_iterations == b.r.max_iterations && a.r.epsilon == b.r.epsilon &&
           a.r.precision == b.r.precision;
}

void fail(dress_hip_response_t &out, const std::string &msg)
{
    out.status = -1;
    snprintf(out.error, sizeof(out.error), "%s", msg.c_str());
}

dress_hip_response_t blank_response(const ServeReq &q)
{
    dress_hip_response_t out;
    memset(&out, 0, sizeof(out));
    out.id = q.id;
    out.kind = q.r.kind;
    out.N = q.r.N;
    out.E = q.r.E;
    out.footprint = q.bytes;
    out.arrival = q.arrival;
    out.delta = NAN;
    return out;
}

// dress_init_graph takes ownership of malloc'd arrays
p_dress_graph_t make_graph(const ServeReq &q)
{
    size_t E = (size_t)q.r.E;
    int *U = (int *)malloc((E + 1) * sizeof(int));
    int *V = (int *)malloc((E + 1) * sizeof(int));
    double *W = q.weighted ? (double *)malloc((E + 1) * sizeof(double)) : nullptr;
    if (!U || !V || (q.weighted && !W)) {
        free(U); free(V); free(W);
        return nullptr;
    }
    if (E) {
        memcpy(U, q.U.data(), E * sizeof(int));
        memcpy(V, q.V.data(), E * sizeof(int));
        if (W) memcpy(W, q.W.data(), E * sizeof(double));
    }
    p_dress_graph_t g = dress_init_graph(q.r.N, q.r.E, U, V, W, nullptr, q.r.variant, 0);
    if (!g) { free(U); free(V); free(W); }
    return g;
}

// Device-wide fit of one request: the graph is built on the worker's
// stream and fitted from d = 1 with vertex weights 1.0 -- the arithmetic of
// dress_init_graph + dress_fit_hip_ex -- without filling a host
// dress_graph_t (no CSR download, no uploads of initial values).
void run_fit(const ServeReq &q, dress_hip_response_t &out)
{
    out.edge_dress = (double *)malloc(((size_t)q.r.E + 1) * sizeof(double));
    out.vertex_dress = (double *)malloc(((size_t)q.r.N + 1) * sizeof(double));
    if (!out.edge_dress || !out.vertex_dress) return fail(out, "host allocation of results failed");
    if (q.r.max_iterations <= 0) {
        // dress.c:546-608 runs no iteration
        for (int e = 0; e < q.r.E; e++) out.edge_dress[e] = 1.0;
        for (int v = 0; v < q.r.N; v++) out.vertex_dress[v] = 0.0;
        out.iterations = q.r.max_iterations;
        out.delta = 0.0;
        return;
    }
    clear_error();
    DevGraph *G = new DevGraph();
    FitResult r;
    bool ok = build_graph(G, q.r.N, q.r.E, q.U.data(), q.V.data(), q.weighted ? q.W.data() : nullptr, nullptr,
                          (int)q.r.variant);
    if (ok) release_reference_csr(G);
    ok = ok && build_bins(G) && fit_fill_values(G, 1.0) &&
         fit_run(G, q.r.max_iterations, q.r.epsilon, q.r.precision == DRESS_HIP_FP32_FP64 ? PREC_FP32_FP64 : PREC_FP64,
                 &r) &&
         fit_download(G, out.edge_dress, out.vertex_dress);
    free_graph(G);
    if (!ok) return fail(out, std::string("fit failed: ") + dress_hip_last_error());
    out.iterations = r.iterations;
    out.delta = r.delta;
}

void run_lifted_one(const ServeReq &q, dress_hip_response_t &out)
{
    p_dress_graph_t g = make_graph(q);
    if (!g) return fail(out, std::string("graph build failed: ") + dress_hip_last_error());
    set_thread_precision(q.r.precision == DRESS_HIP_FP32_FP64 ? PREC_FP32_FP64 : PREC_FP64);
    int hs = 0;
    int64_t num = 0;
    dress_hist_pair_t *h =
        q.r.kind == DRESS_HIP_REQ_DELTA
            ? dress_delta_fit_hip(g, q.r.k, q.r.max_iterations, q.r.epsilon, 0, 0, &hs, 0, nullptr, &num)
            : dress_nabla_fit_hip(g, q.r.k, q.r.max_iterations, q.r.epsilon, 0, 0, &hs, 0, nullptr, &num);
    set_thread_precision(-1);
    const char *err = dress_hip_last_error();
    if (!h && err && *err) fail(out, std::string("lifted fit failed: ") + err);
    out.hist = h;
    out.hist_size = hs;
    out.subgraphs = num;
    dress_free_graph(g);
}

void run_lifted_batch(const std::vector<ServeReq *> &qs, std::vector<dress_hip_response_t> &outs)
{
    const int n = (int)qs.size();
    std::vector<int> N((size_t)n);
    std::vector<int64_t> off((size_t)n + 1, 0);
    bool any_w = false;
    for (int i = 0; i < n; i++) {
        N[(size_t)i] = qs[(size_t)i]->r.N;
        off[(size_t)i + 1] = off[(size_t)i] + qs[(size_t)i]->r.E;
        any_w = any_w || qs[(size_t)i]->weighted;
    }
    std::vector<int> U((size_t)off[(size_t)n] + 1), V((size_t)off[(size_t)n] + 1);
    std::vector<double> W(any_w ? (size_t)off[(size_t)n] + 1 : 0);
    for (int i = 0; i < n; i++) {
        const ServeReq &q = *qs[(size_t)i];
        std::copy(q.U.begin(), q.U.end(), U.begin() + off[(size_t)i]);
        std::copy(q.V.begin(), q.V.end(), V.begin() + off[(size_t)i]);
        if (any_w) {
            if (q.weighted) std::copy(q.W.begin(), q.W.end(), W.begin() + off[(size_t)i]);
            else std::fill(W.begin() + off[(size_t)i], W.begin() + off[(size_t)i + 1], 1.0);
        }
    }
    const ServeReq &q0 = *qs[0];
    dress_hip_lifted_batch_t r;
    int kind = q0.r.kind == DRESS_HIP_REQ_DELTA ? 0 : 1;
    int rc = dress_hip_lifted_batch(kind, q0.r.k, n, N.data(), off.data(), U.data(), V.data(),
                                    any_w ? W.data() : nullptr, q0.r.variant, q0.r.max_iterations, q0.r.epsilon,
                                    q0.r.precision, -1.0, &r);
    if (rc != 0) {
        std::string msg = std::string("batched lifted fit failed: ") + dress_hip_last_error();
        for (auto &o : outs) fail(o, msg);
        return;
    }
    for (int i = 0; i < n; i++) {
        dress_hip_response_t &o = outs[(size_t)i];
        int64_t b = r.hist_offset[i], e = r.hist_offset[i + 1];
        o.subgraphs = count_items(qs[(size_t)i]->r.kind, N[(size_t)i], q0.r.k);
        o.batch_size = n;
        o.hist_size = (int)(e - b);
        if (e > b) {
            o.hist = (dress_hist_pair_t *)malloc((size_t)(e - b) * sizeof(dress_hist_pair_t));
            if (!o.hist) { fail(o, "host allocation of results failed"); continue; }
            for (int64_t j = b; j < e; j++) {
                o.hist[j - b].value = r.hist_value[j];
                o.hist[j - b].count = r.hist_count[j];
            }
        }
    }
    dress_hip_lifted_batch_free(&r);
}

// fits small enough for the LDS-resident kernel, many per launch
void run_fit_batch(const std::vector<ServeReq *> &qs, std::vector<dress_hip_response_t> &outs)
{
    const int n = (int)qs.size();
    std::vector<int> N((size_t)n);
    std::vector<int64_t> off((size_t)n + 1, 0);
    bool any_w = false;
    for (int i = 0; i < n; i++) {
        N[(size_t)i] = qs[(size_t)i]->r.N;
        off[(size_t)i + 1] = off[(size_t)i] + qs[(size_t)i]->r.E;
        any_w = any_w || qs[(size_t)i]->weighted;
    }
    std::vector<int> U((size_t)off[(size_t)n] + 1), V((size_t)off[(size_t)n] + 1);
    std::vector<double> W(any_w ? (size_t)off[(size_t)n] + 1 : 0);
    for (int i = 0; i < n; i++) {
        const ServeReq &q = *qs[(size_t)i];
        std::copy(q.U.begin(), q.U.end(), U.begin() + off[(size_t)i]);
        std::copy(q.V.begin(), q.V.end(), V.begin() + off[(size_t)i]);
        if (any_w) {
            if (q.weighted) std::copy(q.W.begin(), q.W.end(), W.begin() + off[(size_t)i]);
            else std::fill(W.begin() + off[(size_t)i], W.begin() + off[(size_t)i + 1], 1.0);
        }
    }
    const ServeReq &q0 = *qs[0];
    dress_hip_fit_batch_t r;
    if (dress_hip_fit_batch(n, N.data(), off.data(), U.data(), V.data(), any_w ? W.data() : nullptr, q0.r.variant,
                            q0.r.max_iterations, q0.r.epsilon, q0.r.precision, &r) != 0) {
        std::string msg = std::string("batched fit failed: ") + dress_hip_last_error();
        for (auto &o : outs) fail(o, msg);
        return;
    }
    for (int i = 0; i < n; i++) {
        dress_hip_response_t &o = outs[(size_t)i];
        const int Eg = qs[(size_t)i]->r.E, Ng = N[(size_t)i];
        o.iterations = r.iterations[i];
        o.delta = r.delta[i];
        o.edge_dress = (double *)malloc(((size_t)Eg + 1) * sizeof(double));
        o.vertex_dress = (double *)malloc(((size_t)Ng + 1) * sizeof(double));
        if (!o.edge_dress || !o.vertex_dress) { fail(o, "host allocation of results failed"); continue; }
        if (Eg) memcpy(o.edge_dress, r.edge_dress + off[(size_t)i], (size_t)Eg * sizeof(double));
        if (Ng) memcpy(o.vertex_dress, r.vertex_dress + r.vertex_offset[i], (size_t)Ng * sizeof(double));
    }
    dress_hip_fit_batch_free(&r);
}

// Worker lanes.  Batched fits, device-wide fits, batched delta/nabla and
// the delta/nabla requests too large for the batched kernel (a subgraph at
// a time: long) each have workers of their own, so short requests never
// queue behind long launches of another kind.  A worker serves its own lane
// first and, when that is empty, helps with any lane but the large one;
// only large-lane workers take large requests.
enum Lane { LANE_FIT_BATCHED = 0, LANE_FIT = 1, LANE_LIFT_BATCHED = 2, LANE_LIFT_LARGE = 3 };

int lane_of(const ServeReq *q)
{
    if (q->r.kind == DRESS_HIP_REQ_FIT) return q->lds ? LANE_FIT_BATCHED : LANE_FIT;
    return q->lds ? LANE_LIFT_BATCHED : LANE_LIFT_LARGE;
}

// lane of worker i: one per lane first, then device-wide fits twice as
// often as the batched lanes; the large delta/nabla lane keeps one worker
int worker_lane(int i)
{
    static const int first[4] = {LANE_FIT_BATCHED, LANE_FIT, LANE_LIFT_BATCHED, LANE_LIFT_LARGE};
    static const int cycle[4] = {LANE_FIT, LANE_FIT, LANE_LIFT_BATCHED, LANE_FIT_BATCHED};
    return i < 4 ? first[i] : cycle[(i - 4) % 4];
}

// the large delta/nabla lane holds at most this share of the budget, so the
// short requests of the other lanes always find room
constexpr double LARGE_BUDGET_SHARE = 0.5;

// most work one batched launch takes on: edges of its graphs for fits,
// subgraph values for delta/nabla (keeps each launch to tens of ms)
constexpr int64_t FIT_BATCH_EDGES = 4 << 20;
constexpr int64_t LIFT_BATCH_VALUES = 16 << 20;

// Under the lock: takes the next request(s) for a worker, or returns false.
// The worker's own lane first: its oldest request that fits the budget,
// looking at most SCAN_LIMIT deep; otherwise the oldest such request at the
// head of another lane (not the large lane, unless the worker may take
// it).  A batchable request takes waiting requests of its lane with the
// same parameters along, looking at most 4 x max_batch deep.
constexpr size_t SCAN_LIMIT = 256;

bool take(dress_hip_server *S, int lane, bool may_take_large, std::vector<ServeReq *> &out, int64_t &bytes)
{
    if (S->queued() == 0 || S->device_failed) return false;
    const int64_t room = S->st.budget_bytes - S->st.reserved_bytes;
    const int64_t large_room = (int64_t)(LARGE_BUDGET_SHARE * (double)S->st.budget_bytes) - S->reserved_large;
    const bool idle = S->st.running == 0;
    auto admissible = [&](const ServeReq *q) {
        if (idle) return true;
        return q->bytes <= room && (lane_of(q) != LANE_LIFT_LARGE || q->bytes <= large_room);
    };
    auto first_admissible = [&](int l) {
        auto &dq = S->lanes[l];
        size_t seen = 0;
        for (auto it = dq.begin(); it != dq.end() && seen < SCAN_LIMIT; ++it, ++seen)
            if (admissible(*it)) return it;
        return dq.end();
    };
    int from = -1;
    auto pick = S->lanes[lane].end();
    {
        auto it = first_admissible(lane);
        if (it != S->lanes[lane].end()) { from = lane; pick = it; }
    }
    if (from < 0) {
        for (int l = 0; l < 4; l++) {
            if (l == lane || (l == LANE_LIFT_LARGE && !may_take_large)) continue;
            auto it = first_admissible(l);
            if (it == S->lanes[l].end()) continue;
            if (from < 0 || (*it)->arrival < (*pick)->arrival) { from = l; pick = it; }
        }
    }
    if (from < 0) return false;
    auto &dq = S->lanes[from];
    ServeReq *first = *pick;
    dq.erase(pick);
    out.push_back(first);
    bytes = first->bytes;
    if (first->lds) {
        const int64_t cap = first->r.kind == DRESS_HIP_REQ_FIT ? FIT_BATCH_EDGES : LIFT_BATCH_VALUES;
        int64_t work = first->work;
        size_t seen = 0;
        const size_t depth = 4 * (size_t)std::max(S->max_batch, 1);
        for (auto it = dq.begin(); it != dq.end() && (int)out.size() < S->max_batch && seen < depth; ++seen) {
            ServeReq *q = *it;
            if (q->lds && same_batch(*q, *first) && bytes + q->bytes <= std::max(room, first->bytes) &&
                work + q->work <= cap) {
                out.push_back(q);
                bytes += q->bytes;
                work += q->work;
                it = dq.erase(it);
            } else {
                ++it;
            }
        }
    }
    S->st.reserved_bytes += bytes;
    if (lane_of(first) == LANE_LIFT_LARGE) S->reserved_large += bytes;
    S->st.peak_reserved_bytes = std::max(S->st.peak_reserved_bytes, S->st.reserved_bytes);
    S->st.running += (int64_t)out.size();
    S->st.queued = S->queued();
    return true;
}

// HIP errors after which the device (its context) stays unusable: every
// later call on it fails the same way
bool sticky_error(hipError_t e)
{
    switch (e) {
    case hipErrorIllegalAddress:
    case hipErrorLaunchFailure:
    case hipErrorAssert:
    case hipErrorLaunchTimeOut:
    case hipErrorECCNotCorrectable:
    case hipErrorContextIsDestroyed:
    case hipErrorRuntimeMemory:     // HSA queue errors (a queue abort, an
    case hipErrorRuntimeOther:      // invalid allocation) end in these
        return true;
    default:
        return false;
    }
}

// After a failed request: is the device still usable?  Returns "" when it
// is, else what is wrong.  `injected` (test hook) stands for the probe's
// result; otherwise the worker's stream is synchronised and a small block
// is allocated, written and freed.
std::string device_error_check(hipStream_t stream, hipError_t injected)
{
    hipError_t e = injected;
    if (e == hipSuccess) {
        e = hipStreamSynchronize(stream);
        if (e == hipSuccess) e = hipGetLastError();
        void *p = nullptr;
        if (e == hipSuccess) e = hipMalloc(&p, 256);
        if (e == hipSuccess) e = hipMemsetAsync(p, 0, 256, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (p) (void)hipFree(p);
        if (!sticky_error(e)) {
            (void)hipGetLastError();
            return std::string();
        }
    } else if (!sticky_error(e)) {
        return std::string();
    }
    return std::string("device error ") + std::to_string((int)e) + " (" + hipGetErrorString(e) + ")";
}

// Under the lock: the device is unusable; every queued request is answered
// with the cause, later submissions are refused, workers take nothing more.
void stop_after_device_error(dress_hip_server *S, const std::string &why)
{
    S->device_failed = true;
    S->failure = why;
    S->st.device_failed = 1;
    const double t = now_s();
    for (auto &dq : S->lanes) {
        for (ServeReq *q : dq) {
            dress_hip_response_t o = blank_response(*q);
            o.admitted = o.finished = t;
            fail(o, "server stopped after " + why);
            S->st.failed++;
            if (lane_of(q) != LANE_LIFT_LARGE) S->short_load--;
            if (q->r.kind == DRESS_HIP_REQ_FIT) S->fit_load--;
            S->done.push_back(o);
            delete q;
        }
        dq.clear();
    }
    S->st.queued = 0;
    fprintf(stderr, "libdress_hip: request server stopped after %s\n", why.c_str());
}

void worker_main(dress_hip_server *S, int index)
{
    hipSetDevice(S->device);
    hipStream_t stream = nullptr;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) stream = nullptr;
    if (stream) ctx_begin(stream);
    set_thread_cancel(&S->cancel);
    const int lane = worker_lane(index);
    // the large delta/nabla lane fills a quarter of the device at most, so
    // the short requests of the other lanes keep finding free compute units
    if (lane == LANE_LIFT_LARGE) {
        set_thread_grid_share(25);
        set_thread_pressure(&S->short_load);      // and yields to them between passes
        set_thread_max_slots(64);                 // short launches to yield between
    } else if (lane == LANE_LIFT_BATCHED) {
        // batched delta/nabla launches fill half the device while fits wait
        set_thread_grid_share(50);
        set_thread_pressure(&S->fit_load);
    }
    // without a large lane (fewer than four workers) the last worker takes
    // large requests as well
    const bool may_take_large = lane == LANE_LIFT_LARGE || (S->n_workers < 4 && index == S->n_workers - 1);
    set_thread_lifted_values(S->chunk_values);
    for (;;) {
        std::vector<ServeReq *> batch;
        int64_t bytes = 0;
        double admitted = 0.0;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            while (!S->stop && !take(S, lane, may_take_large, batch, bytes)) S->cv_work.wait(lk);
            if (batch.empty()) break;                   // stopping
            admitted = now_s();
        }
        std::vector<dress_hip_response_t> outs;
        for (ServeReq *q : batch) {
            outs.push_back(blank_response(*q));
            outs.back().admitted = admitted;
            outs.back().batch_size = (int)batch.size();
        }
        bool chunked = false;
        if (!stream) {
            for (auto &o : outs) fail(o, "stream creation failed");
        } else if (batch[0]->r.kind == DRESS_HIP_REQ_FIT) {
            if (batch[0]->lds) run_fit_batch(batch, outs);
            else run_fit(*batch[0], outs[0]);
        } else if (batch.size() == 1) {
            chunked = (double)count_items(batch[0]->r.kind, batch[0]->r.N, batch[0]->r.k) * (double)batch[0]->r.E >
                      (double)S->chunk_values;
            run_lifted_one(*batch[0], outs[0]);
        } else {
            run_lifted_batch(batch, outs);
        }
        // the test hook: this request fails as if the device reported `code`
        const int injected = S->inject.exchange(0);
        if (injected) {
            const std::string msg = std::string("device error (injected): ") +
                                    hipGetErrorString((hipError_t)injected);
            for (auto &o : outs) fail(o, msg);
        }
        bool failed = false;
        for (auto &o : outs) failed = failed || o.status != 0;
        // a failure may come from a device error that leaves the device
        // unusable (every later call fails): check, and stop the server
        std::string dead;
        if (failed) dead = device_error_check(stream, (hipError_t)injected);
        // large transient blocks go back to the device before the budget
        // they were reserved under is released, once this worker holds
        // more than 16 GiB of them
        if (dead.empty()) ctx_trim((size_t)256 << 20, (size_t)16 << 30);
        const double fin = now_s();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->st.reserved_bytes -= bytes;
            if (lane_of(batch[0]) == LANE_LIFT_LARGE) S->reserved_large -= bytes;
            S->st.running -= (int64_t)batch.size();
            if (batch[0]->lds) {
                S->st.lifted_launches++;
                S->st.lifted_requests += (int64_t)batch.size();
            }
            if (lane_of(batch[0]) != LANE_LIFT_LARGE) S->short_load -= (int)batch.size();
            if (batch[0]->r.kind == DRESS_HIP_REQ_FIT) S->fit_load -= (int)batch.size();
            for (auto &o : outs) {
                o.finished = fin;
                if (o.status == 0) S->st.completed++;
                else S->st.failed++;
                S->done.push_back(o);
            }
            if (chunked) S->st.chunked_requests++;
            if (!dead.empty() && !S->device_failed) stop_after_device_error(S, dead);
        }
        for (ServeReq *q : batch) delete q;
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
    set_thread_cancel(nullptr);
    set_thread_pressure(nullptr);
    set_thread_lifted_values(0);
    if (stream) {
        ctx_end();
        hipStreamDestroy(stream);
    }
}

}  // namespace

extern "C" {

double dress_hip_clock(void) { return now_s(); }

dress_hip_server_t *dress_hip_server_create(int workers, int64_t budget_bytes, int max_batch)
{
    clear_error();
    if (workers < 1 || workers > 64) {
        set_error("dress_hip_server_create: workers must be in 1..64");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        hipGetLastError();
        set_error("dress_hip_server_create: no GPU");
        return nullptr;
    }
    auto *S = new dress_hip_server();
    hipGetDevice(&S->device);
    if (budget_bytes <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            hipGetLastError();
            fr = (size_t)8 << 30;
        }
        budget_bytes = (int64_t)(fr / 2);
    }
    S->st.budget_bytes = budget_bytes;
    S->host_budget = (int64_t)32 << 30;
    S->chunk_values = SERVE_CHUNK_VALUES;
    if (const char *e = getenv("DRESS_HIP_SERVE_CHUNK_VALUES")) {
        long long v = atoll(e);
        if (v > 0) S->chunk_values = (int64_t)v;
    }
    if (const char *e = getenv("DRESS_HIP_SERVE_HOST_BYTES")) {
        long long v = atoll(e);
        if (v > 0) S->host_budget = (int64_t)v;
    }
    S->max_batch = max_batch < 1 ? 1 : max_batch;
    S->n_workers = workers;
    for (int i = 0; i < workers; i++) S->workers.emplace_back(worker_main, S, i);
    return S;
}

int64_t dress_hip_server_submit(dress_hip_server_t *S, const dress_hip_request_t *req)
{
    clear_error();
    if (!S || !req) { set_error("dress_hip_server_submit: NULL argument"); return -1; }
    if (req->N < 0 || req->E < 0 || (req->E > 0 && (!req->U || !req->V)) ||
        (req->kind != DRESS_HIP_REQ_FIT && req->kind != DRESS_HIP_REQ_DELTA && req->kind != DRESS_HIP_REQ_NABLA) ||
        (req->kind != DRESS_HIP_REQ_FIT && req->k < 0) || (int)req->variant < 0 || (int)req->variant > 3) {
        set_error("dress_hip_server_submit: bad request");
        return -1;
    }
    for (int e = 0; e < req->E; e++)
        if (req->U[e] < 0 || req->U[e] >= req->N || req->V[e] < 0 || req->V[e] >= req->N) {
            set_error("dress_hip_server_submit: vertex id out of range");
            return -1;
        }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->device_failed) {
            set_error("dress_hip_server_submit: the server stopped after " + S->failure);
            return -1;
        }
    }
    auto *q = new ServeReq();
    q->r = *req;
    q->r.U = q->r.V = nullptr;
    q->r.W = nullptr;
    q->U.assign(req->U, req->U + req->E);
    q->V.assign(req->V, req->V + req->E);
    q->weighted = req->W != nullptr;
    if (q->weighted) q->W.assign(req->W, req->W + req->E);
    q->arrival = req->arrival >= 0 ? req->arrival : now_s();
    q->lds = lds_eligible_size(req->N, req->E) && (req->kind != DRESS_HIP_REQ_FIT || req->max_iterations > 0);
    q->bytes = footprint(*q, intersection_bound(req->N, req->E, req->U, req->V), S->chunk_values);
    q->work = req->kind == DRESS_HIP_REQ_FIT ? (int64_t)req->E
                                             : count_items(req->kind, req->N, req->k) * (int64_t)req->E;
    std::lock_guard<std::mutex> lk(S->mu);
    q->id = S->next_id++;
    S->st.submitted++;
    const int64_t limit = lane_of(q) == LANE_LIFT_LARGE
                              ? (int64_t)(LARGE_BUDGET_SHARE * (double)S->st.budget_bytes)
                              : S->st.budget_bytes;
    // a delta/nabla result counts its pairs in an int (hist_size = pairs);
    // every subgraph value may be distinct
    const double values = q->r.kind == DRESS_HIP_REQ_FIT
                              ? 0.0
                              : (double)count_items(q->r.kind, q->r.N, q->r.k) * (double)q->r.E;
    const double hb = host_result_bytes(*q);
    if (hb > (double)S->host_budget || values > (double)INT32_MAX) {
        dress_hip_response_t o = blank_response(*q);
        o.admitted = o.finished = now_s();
        fail(o, "request result of up to " + std::to_string((int64_t)hb) +
                    " B exceeds the server's host budget of " + std::to_string(S->host_budget) +
                    " B per result (or the int hist_size of the API)");
        S->st.rejected++;
        S->st.rejected_host++;
        S->done.push_back(o);
        int64_t id = q->id;
        delete q;
        S->cv_done.notify_all();
        return id;
    }
    if (q->bytes > limit) {
        dress_hip_response_t o = blank_response(*q);
        o.admitted = o.finished = now_s();
        fail(o, "request footprint " + std::to_string(q->bytes) + " B exceeds the server budget of " +
                    std::to_string(limit) + " B for its kind");
        S->st.rejected++;
        S->done.push_back(o);
        int64_t id = q->id;
        delete q;
        S->cv_done.notify_all();
        return id;
    }
    if (lane_of(q) != LANE_LIFT_LARGE) S->short_load++;
    if (q->r.kind == DRESS_HIP_REQ_FIT) S->fit_load++;
    S->lanes[lane_of(q)].push_back(q);
    S->st.queued = S->queued();
    S->st.max_queued = std::max(S->st.max_queued, S->st.queued);
    S->cv_work.notify_one();
    return q->id;
}

int dress_hip_server_next(dress_hip_server_t *S, double timeout_s, dress_hip_response_t *out)
{
    if (!S || !out) return 0;
    std::unique_lock<std::mutex> lk(S->mu);
    auto until = std::chrono::steady_clock::now() +
                 std::chrono::microseconds((int64_t)(std::max(timeout_s, 0.0) * 1e6));
    while (S->done.empty())
        if (S->cv_done.wait_until(lk, until) == std::cv_status::timeout && S->done.empty()) return 0;
    *out = S->done.front();
    S->done.pop_front();
    return 1;
}

void dress_hip_response_free(dress_hip_response_t *r)
{
    if (!r) return;
    free(r->edge_dress);
    free(r->vertex_dress);
    free(r->hist);
    r->edge_dress = r->vertex_dress = nullptr;
    r->hist = nullptr;
}

void dress_hip_server_stats(dress_hip_server_t *S, dress_hip_server_stats_t *out)
{
    if (!S || !out) return;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
}

int dress_hip_server_inject_error(dress_hip_server_t *S, int hip_error)
{
    if (!S) return -1;
    S->inject = hip_error;
    return 0;
}

void dress_hip_server_destroy(dress_hip_server_t *S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cancel = true;
        for (auto &dq : S->lanes) {
            for (ServeReq *q : dq) delete q;
            dq.clear();
        }
        S->st.queued = 0;
    }
    S->cv_work.notify_all();
    for (auto &t : S->workers) t.join();
    for (auto &o : S->done) dress_hip_response_free(&o);
    delete S;
}

}  // extern "C"
