/*
 * dress_oracle.c -- CPU oracle for the DRESS hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py as the checker / the timed CPU baseline.
 * The product library (libdress_hip.so) never links or calls this file.
 *
 * A restatement of libdress v0.8.3 (velicast/dress-graph) in FP64:
 *   graph construction   libdress/src/dress.c:103-296, 378-427
 *   intersection index   libdress/src/dress.c:307-363
 *   sort + KBN sum       libdress/src/dress.c:51-66
 *   fit iteration        libdress/src/dress.c:452-608
 *   OpenMP fit           libdress/src/omp/dress_omp.c:117-184
 *   virtual-edge query   libdress/src/dress.c:651-739
 *   delta^k driver       libdress/src/delta_dress_impl.c:95-169, 233-380
 *   nabla^k driver       libdress/src/nabla_dress_impl.c:33-91, 142-311
 *   histogram            libdress/src/dress_histogram.c (exact double keys,
 *                        sorted ascending; restated here as sort + run-length)
 *
 * Build with -ffp-contract=off and no -march (the reference is compiled for
 * generic x86-64, so it contains no fused multiply-adds).
 */
#include "dress_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* construction                                                        */
/* ------------------------------------------------------------------ */

/* one adjacency entry while building: neighbour, input edge, weight */
typedef struct { int x; int e; double w; } ora_half_t;

/* Rows are ordered by (neighbour, input edge).  The reference sorts by the
 * neighbour only with qsort (dress.c:25-30, 155, 287); glibc's qsort is a
 * stable merge sort here, and entries are inserted in input-edge order, so
 * ties end up ordered by edge index, which is what this comparator fixes. */
static int ora_half_cmp(const void *a, const void *b)
{
    const ora_half_t *p = (const ora_half_t *)a, *q = (const ora_half_t *)b;
    if (p->x != q->x) return p->x < q->x ? -1 : 1;
    if (p->e != q->e) return p->e < q->e ? -1 : 1;
    return 0;
}

/* Binary search for neighbour v in row u of the raw adjacency; the probe
 * sequence matches dress.c:70-87 so duplicate edges resolve identically. */
static int ora_find_raw(int u, int v, const int *off, const ora_half_t *rows)
{
    int lo = off[u], hi = off[u + 1] - 1;
    while (lo <= hi) {
        int mid = lo + (hi - lo) / 2;
        int x = rows[mid].x;
        if (x == v) return mid;
        if (x < v) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

static void ora_sort_rows(const int *off, ora_half_t *rows, int N)
{
    int i;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1024)
#endif
    for (i = 0; i < N; i++) {
        int n = off[i + 1] - off[i];
        if (n > 1) qsort(rows + off[i], (size_t)n, sizeof(ora_half_t), ora_half_cmp);
    }
}

/* dress.c:103-160 then dress.c:172-296 */
static void ora_build_adjacency(p_dress_graph_t g)
{
    int N = g->N, E = g->E, e, u;
    dress_variant_t var = g->variant;
    const int *U = g->U, *V = g->V;

    /* raw adjacency: undirected stores both directions, else out-edges */
    int *roff = (int *)calloc((size_t)N + 1, sizeof(int));
    for (e = 0; e < E; e++) {
        roff[U[e] + 1]++;
        if (var == DRESS_VARIANT_UNDIRECTED) roff[V[e] + 1]++;
    }
    for (u = 0; u < N; u++) roff[u + 1] += roff[u];
    ora_half_t *raw = (ora_half_t *)malloc(((size_t)roff[N] + 1) * sizeof(ora_half_t));
    int *fill = (int *)calloc((size_t)N, sizeof(int));
    for (e = 0; e < E; e++) {
        double w = g->W ? g->W[e] : 1.0;
        int a = U[e], b = V[e];
        ora_half_t h1 = { b, e, w };
        raw[roff[a] + fill[a]++] = h1;
        if (var == DRESS_VARIANT_UNDIRECTED) {
            ora_half_t h2 = { a, e, w };
            raw[roff[b] + fill[b]++] = h2;
        }
    }
    ora_sort_rows(roff, raw, N);

    /* variant adjacency sizes */
    int *cnt = (int *)calloc((size_t)N + 1, sizeof(int));
    for (e = 0; e < E; e++) {
        int a = U[e], b = V[e];
        switch (var) {
        case DRESS_VARIANT_UNDIRECTED: cnt[a]++; cnt[b]++; break;
        case DRESS_VARIANT_DIRECTED:
            cnt[a]++;
            if (ora_find_raw(b, a, roff, raw) < 0) cnt[b]++;
            break;
        case DRESS_VARIANT_FORWARD: cnt[a]++; break;
        case DRESS_VARIANT_BACKWARD: cnt[b]++; break;
        }
    }
    int *off = (int *)malloc(((size_t)N + 1) * sizeof(int));
    off[0] = 0;
    int maxdeg = 0;
    for (u = 0; u < N; u++) {
        if (cnt[u] > maxdeg) maxdeg = cnt[u];
        off[u + 1] = off[u] + cnt[u];
    }
    g->max_degree = maxdeg;
    int S = off[N];
    ora_half_t *rows = (ora_half_t *)malloc(((size_t)S + 1) * sizeof(ora_half_t));
    memset(fill, 0, (size_t)N * sizeof(int));

    /* walk the raw rows in vertex order, as dress.c:228-265 does */
    for (u = 0; u < N; u++) {
        int i;
        for (i = roff[u]; i < roff[u + 1]; i++) {
            int v = raw[i].x, eid = raw[i].e;
            double w = raw[i].w;
            if (var == DRESS_VARIANT_UNDIRECTED) {
                ora_half_t h = { v, eid, 2.0 * w };
                rows[off[u] + fill[u]++] = h;
            } else if (var == DRESS_VARIANT_DIRECTED) {
                int r = ora_find_raw(v, u, roff, raw);
                ora_half_t h = { v, eid, w };
                if (r >= 0) h.w += raw[r].w;
                rows[off[u] + fill[u]++] = h;
                if (r < 0) {
                    ora_half_t hin = { u, eid, w };
                    rows[off[v] + fill[v]++] = hin;
                }
            } else if (var == DRESS_VARIANT_FORWARD) {
                ora_half_t h = { v, eid, w };
                rows[off[u] + fill[u]++] = h;
            } else {
                ora_half_t hin = { u, eid, w };
                rows[off[v] + fill[v]++] = hin;
            }
        }
    }
    ora_sort_rows(off, rows, N);

    g->adj_offset = off;
    g->adj_target = (int *)malloc(((size_t)S + 1) * sizeof(int));
    g->adj_edge_idx = (int *)malloc(((size_t)S + 1) * sizeof(int));
    g->edge_weight = (double *)malloc(((size_t)E + 1) * sizeof(double));
    for (int i = 0; i < S; i++) {
        g->adj_target[i] = rows[i].x;
        g->adj_edge_idx[i] = rows[i].e;
        g->edge_weight[rows[i].e] = rows[i].w;
    }
    free(rows); free(cnt); free(fill); free(raw); free(roff);
}

/* dress.c:307-363: multiset intersection of rows U[e] and V[e] by a sorted
 * merge; entry k of edge e records the edges (u,x) and (v,x). */
static void ora_build_intercepts(p_dress_graph_t g)
{
    int E = g->E, e;
    int *io = (int *)malloc(((size_t)E + 1) * sizeof(int));
    io[0] = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (e = 0; e < E; e++) {
        int a = g->adj_offset[g->U[e]], ae = g->adj_offset[g->U[e] + 1];
        int b = g->adj_offset[g->V[e]], be = g->adj_offset[g->V[e] + 1];
        int c = 0;
        while (a < ae && b < be) {
            int x = g->adj_target[a], y = g->adj_target[b];
            if (x == y) { c++; a++; b++; }
            else if (x < y) a++;
            else b++;
        }
        io[e + 1] = c;
    }
    for (e = 0; e < E; e++) io[e + 1] += io[e];
    int T = io[E];
    int *ux = (int *)malloc(((size_t)T + 1) * sizeof(int));
    int *vx = (int *)malloc(((size_t)T + 1) * sizeof(int));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (e = 0; e < E; e++) {
        int a = g->adj_offset[g->U[e]], ae = g->adj_offset[g->U[e] + 1];
        int b = g->adj_offset[g->V[e]], be = g->adj_offset[g->V[e] + 1];
        int k = io[e];
        while (a < ae && b < be) {
            int x = g->adj_target[a], y = g->adj_target[b];
            if (x == y) {
                ux[k] = g->adj_edge_idx[a];
                vx[k] = g->adj_edge_idx[b];
                k++; a++; b++;
            } else if (x < y) a++;
            else b++;
        }
    }
    g->intercept_offset = io;
    g->intercept_edge_ux = ux;
    g->intercept_edge_vx = vx;
}

/* dress.c:378-427.  Edge values start at 1.0 (dress.c:418-421). */
p_dress_graph_t ora_init_graph(int N, int E, int *U, int *V, double *W,
                               double *NW, dress_variant_t variant,
                               int precompute_intercepts)
{
    p_dress_graph_t g = (p_dress_graph_t)calloc(1, sizeof(dress_graph_t));
    g->variant = variant;
    g->N = N; g->E = E; g->U = U; g->V = V; g->W = W;
    g->precompute_intercepts = precompute_intercepts;
    g->vertex_dress = (double *)calloc((size_t)N + 1, sizeof(double));
    if (NW) g->NW = NW;
    else {
        g->NW = (double *)malloc(((size_t)N + 1) * sizeof(double));
        for (int i = 0; i < N; i++) g->NW[i] = 1.0;
    }
    ora_build_adjacency(g);
    g->edge_dress = (double *)malloc(((size_t)E + 1) * sizeof(double));
    g->edge_dress_next = (double *)malloc(((size_t)E + 1) * sizeof(double));
    for (int e = 0; e < E; e++) g->edge_dress[e] = g->edge_dress_next[e] = 1.0;
    if (precompute_intercepts) ora_build_intercepts(g);
    return g;
}

p_dress_graph_t ora_graph_from_parts(int N, int E, dress_variant_t variant,
                                     int *U, int *V, double *W, double *NW,
                                     int *adj_offset, int *adj_target,
                                     int *adj_edge_idx, double *edge_weight,
                                     int *intercept_offset, int *ux, int *vx)
{
    p_dress_graph_t g = (p_dress_graph_t)calloc(1, sizeof(dress_graph_t));
    g->variant = variant;
    g->N = N; g->E = E; g->U = U; g->V = V; g->W = W;
    g->adj_offset = adj_offset; g->adj_target = adj_target;
    g->adj_edge_idx = adj_edge_idx; g->edge_weight = edge_weight;
    int maxdeg = 0;
    for (int u = 0; u < N; u++) {
        int d = adj_offset[u + 1] - adj_offset[u];
        if (d > maxdeg) maxdeg = d;
    }
    g->max_degree = maxdeg;
    g->vertex_dress = (double *)calloc((size_t)N + 1, sizeof(double));
    if (NW) g->NW = NW;
    else {
        g->NW = (double *)malloc(((size_t)N + 1) * sizeof(double));
        for (int i = 0; i < N; i++) g->NW[i] = 1.0;
    }
    g->edge_dress = (double *)malloc(((size_t)E + 1) * sizeof(double));
    g->edge_dress_next = (double *)malloc(((size_t)E + 1) * sizeof(double));
    for (int e = 0; e < E; e++) g->edge_dress[e] = g->edge_dress_next[e] = 1.0;
    g->precompute_intercepts = intercept_offset != NULL;
    g->intercept_offset = intercept_offset;
    g->intercept_edge_ux = ux;
    g->intercept_edge_vx = vx;
    return g;
}

/* dress.c:615-643 */
void ora_free_graph(p_dress_graph_t g)
{
    if (!g) return;
    free(g->adj_offset); free(g->adj_target); free(g->adj_edge_idx);
    free(g->W); free(g->edge_weight); free(g->edge_dress);
    free(g->edge_dress_next); free(g->vertex_dress); free(g->NW);
    if (g->precompute_intercepts) {
        free(g->intercept_offset); free(g->intercept_edge_ux);
        free(g->intercept_edge_vx);
    }
    free(g->U); free(g->V); free(g);
}

void ora_free(void *p) { free(p); }

/* ------------------------------------------------------------------ */
/* sums                                                                */
/* ------------------------------------------------------------------ */

static int ora_dbl_cmp(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* dress.c:51-66: ascending sort, then Neumaier-compensated running sum. */
double ora_kbn_sorted_sum(double *buf, int n)
{
    if (n <= 0) return 0.0;
    qsort(buf, (size_t)n, sizeof(double), ora_dbl_cmp);
    double s = buf[0], c = 0.0;
    for (int i = 1; i < n; i++) {
        double x = buf[i];
        double t = s + x;
        c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
        s = t;
    }
    return s + c;
}

/* ------------------------------------------------------------------ */
/* fit                                                                 */
/* ------------------------------------------------------------------ */

#define ORA_STACK 4096

/* dress.c:555-574: norm of u from its own self term and its row. */
static void ora_vertex_norm(p_dress_graph_t g, int u)
{
    int b = g->adj_offset[u], n = g->adj_offset[u + 1] - b;
    double st[ORA_STACK];
    double *buf = (n + 1 <= ORA_STACK) ? st : (double *)malloc(((size_t)n + 1) * sizeof(double));
    buf[0] = 4.0 * g->NW[u];
    for (int i = 0; i < n; i++) {
        int e = g->adj_edge_idx[b + i];
        buf[i + 1] = g->edge_weight[e] * g->edge_dress[e];
    }
    g->vertex_dress[u] = sqrt(ora_kbn_sorted_sum(buf, n + 1));
    if (buf != st) free(buf);
}

/* dress.c:452-529: terms of edge e (index or merge walk), self/cross term,
 * sorted KBN numerator over the product of the two norms. */
static double ora_edge_update(p_dress_graph_t g, int e)
{
    int u = g->U[e], v = g->V[e];
    double den = g->vertex_dress[u] * g->vertex_dress[v];
    int cap;
    if (g->precompute_intercepts)
        cap = g->intercept_offset[e + 1] - g->intercept_offset[e] + 1;
    else {
        int du = g->adj_offset[u + 1] - g->adj_offset[u];
        int dv = g->adj_offset[v + 1] - g->adj_offset[v];
        cap = (du < dv ? du : dv) + 1;
    }
    double st[ORA_STACK];
    double *buf = (cap <= ORA_STACK) ? st : (double *)malloc((size_t)cap * sizeof(double));
    int n = 0;
    if (g->precompute_intercepts) {
        for (int k = g->intercept_offset[e]; k < g->intercept_offset[e + 1]; k++) {
            int a = g->intercept_edge_ux[k], b = g->intercept_edge_vx[k];
            buf[n++] = g->edge_weight[a] * g->edge_dress[a]
                     + g->edge_weight[b] * g->edge_dress[b];
        }
    } else {
        int a = g->adj_offset[u], ae = g->adj_offset[u + 1];
        int b = g->adj_offset[v], be = g->adj_offset[v + 1];
        while (a < ae && b < be) {
            int x = g->adj_target[a], y = g->adj_target[b];
            if (x == y) {
                int ea = g->adj_edge_idx[a], eb = g->adj_edge_idx[b];
                buf[n++] = g->edge_weight[ea] * g->edge_dress[ea]
                         + g->edge_weight[eb] * g->edge_dress[eb];
                a++; b++;
            } else if (x < y) a++;
            else b++;
        }
    }
    double own = g->edge_weight[e] * g->edge_dress[e];
    if (g->variant == DRESS_VARIANT_FORWARD || g->variant == DRESS_VARIANT_BACKWARD)
        buf[n++] = 4.0 * g->NW[u] + own;
    else
        buf[n++] = 4.0 * g->NW[u] + 4.0 * g->NW[v] + 2.0 * own;
    double num = ora_kbn_sorted_sum(buf, n);
    if (buf != st) free(buf);
    double d = den > 0.0 ? num / den : 0.0;
    g->edge_dress_next[e] = d;
    return d;
}

/* dress.c:546-608 (sequential) and omp/dress_omp.c:117-184 (parallel):
 * vertex phase, edge phase, swap, strict max|delta| < epsilon test.
 * *iterations is the 0-based index of the converging iteration, or
 * max_iterations; *delta is the max change of the last iteration run. */
static void ora_fit_impl(p_dress_graph_t g, int max_iterations, double epsilon,
                         int *iterations, double *delta, int par)
{
    for (int it = 0; it < max_iterations; it++) {
        double md = 0.0;
        int u, e;
        if (par) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 512)
#endif
            for (u = 0; u < g->N; u++) ora_vertex_norm(g, u);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 512) reduction(max:md)
#endif
            for (e = 0; e < g->E; e++) {
                double prev = g->edge_dress[e];
                double d = fabs(prev - ora_edge_update(g, e));
                if (d > md) md = d;
            }
        } else {
            for (u = 0; u < g->N; u++) ora_vertex_norm(g, u);
            for (e = 0; e < g->E; e++) {
                double prev = g->edge_dress[e];
                double d = fabs(prev - ora_edge_update(g, e));
                if (d > md) md = d;
            }
        }
        double *t = g->edge_dress;
        g->edge_dress = g->edge_dress_next;
        g->edge_dress_next = t;
        if (delta) *delta = md;
        if (md < epsilon) {
            if (iterations) *iterations = it;
            return;
        }
    }
    if (iterations) *iterations = max_iterations;
}

void ora_fit(p_dress_graph_t g, int max_iterations, double epsilon,
             int *iterations, double *delta)
{
    ora_fit_impl(g, max_iterations, epsilon, iterations, delta, 0);
}

void ora_fit_omp(p_dress_graph_t g, int max_iterations, double epsilon,
                 int *iterations, double *delta)
{
    ora_fit_impl(g, max_iterations, epsilon, iterations, delta, 1);
}

/* dress.c:651-739: existing edge -> its value; otherwise a 1-D fixed point
 * of the virtual edge against the frozen fitted state. */
double ora_get(const p_dress_graph_t g, int u, int v, int max_iterations,
               double epsilon, double edge_weight)
{
    int lo = g->adj_offset[u], hi = g->adj_offset[u + 1] - 1;
    while (lo <= hi) {
        int mid = lo + (hi - lo) / 2;
        int nb = g->adj_target[mid];
        if (nb == v) return g->edge_dress[g->adj_edge_idx[mid]];
        if (nb < v) lo = mid + 1; else hi = mid - 1;
    }
    int du = g->adj_offset[u + 1] - g->adj_offset[u];
    int dv = g->adj_offset[v + 1] - g->adj_offset[v];
    int cap = (du < dv ? du : dv) + 1;
    double st[ORA_STACK];
    double *buf = (cap <= ORA_STACK) ? st : (double *)malloc((size_t)cap * sizeof(double));
    int n = 0;
    int a = g->adj_offset[u], ae = g->adj_offset[u + 1];
    int b = g->adj_offset[v], be = g->adj_offset[v + 1];
    while (a < ae && b < be) {
        int x = g->adj_target[a], y = g->adj_target[b];
        if (x == y) {
            int ea = g->adj_edge_idx[a], eb = g->adj_edge_idx[b];
            buf[n++] = g->edge_weight[ea] * g->edge_dress[ea]
                     + g->edge_weight[eb] * g->edge_dress[eb];
            a++; b++;
        } else if (x < y) a++;
        else b++;
    }
    int dir = (g->variant == DRESS_VARIANT_FORWARD || g->variant == DRESS_VARIANT_BACKWARD);
    buf[n++] = dir ? 4.0 * g->NW[u] : 4.0 * g->NW[u] + 4.0 * g->NW[v];
    double A = ora_kbn_sorted_sum(buf, n);
    if (buf != st) free(buf);
    double cross = dir ? 1.0 : 2.0;
    double cw = (g->variant == DRESS_VARIANT_UNDIRECTED) ? 2.0 * edge_weight : edge_weight;
    double du2 = g->vertex_dress[u] * g->vertex_dress[u];
    double dv2 = g->vertex_dress[v] * g->vertex_dress[v];
    if (du2 <= 0.0 || dv2 <= 0.0) return 0.0;
    double d = 1.0;
    for (int it = 0; it < max_iterations; it++) {
        double wd = cw * d;
        double den = sqrt(du2 + wd) * sqrt(dv2 + wd);
        if (den <= 0.0) return 0.0;
        double nx = (A + cross * wd) / den;
        if (fabs(nx - d) <= epsilon) { d = nx; break; }
        d = nx;
    }
    return d;
}

/* ------------------------------------------------------------------ */
/* value multiset -> flat exact histogram                              */
/* ------------------------------------------------------------------ */

typedef struct { double *v; size_t n, cap; } ora_vals_t;

static void ora_vals_push(ora_vals_t *a, const double *x, size_t n)
{
    if (a->n + n > a->cap) {
        size_t c = a->cap ? a->cap : 1024;
        while (c < a->n + n) c *= 2;
        a->v = (double *)realloc(a->v, c * sizeof(double));
        a->cap = c;
    }
    memcpy(a->v + a->n, x, n * sizeof(double));
    a->n += n;
}

/* dress_histogram.c:182-213: distinct values ascending with their counts,
 * packed as [value bits, count, ...]; NULL and *len = 0 when empty. */
static int64_t *ora_flatten(ora_vals_t *a, int *len)
{
    if (a->n == 0) { if (len) *len = 0; return NULL; }
    qsort(a->v, a->n, sizeof(double), ora_dbl_cmp);
    size_t distinct = 0;
    for (size_t i = 0; i < a->n; i++)
        if (i == 0 || !(a->v[i] == a->v[i - 1])) distinct++;
    int64_t *out = (int64_t *)malloc(2 * distinct * sizeof(int64_t));
    size_t j = 0;
    for (size_t i = 0; i < a->n; i++) {
        if (i == 0 || !(a->v[i] == a->v[i - 1])) {
            memcpy(&out[2 * j], &a->v[i], sizeof(double));
            out[2 * j + 1] = 1;
            j++;
        } else {
            out[2 * (j - 1) + 1]++;
        }
    }
    if (len) *len = (int)(2 * distinct);
    return out;
}

/* ------------------------------------------------------------------ */
/* delta^k / nabla^k drivers                                           */
/* ------------------------------------------------------------------ */

/* delta_dress_impl.c:194-204 */
int64_t ora_binom(int n, int k)
{
    if (k < 0 || k > n) return 0;
    if (k == 0 || k == n) return 1;
    if (k > n - k) k = n - k;
    int64_t r = 1;
    for (int i = 0; i < k; i++) r = r * (n - i) / (i + 1);
    return r;
}

/* nabla_dress_impl.c:111-118 */
int64_t ora_perm_count(int n, int k)
{
    if (k < 0 || k > n) return 0;
    int64_t r = 1;
    for (int i = 0; i < k; i++) r *= (n - i);
    return r;
}

/* delta_dress_impl.h:21-25: the LCG of the sampled drivers */
static int ora_rand(unsigned int *s)
{
    *s = *s * 1103515245u + 12345u;
    return (int)((*s >> 16) & 0x7fff);
}

static int ora_i64_cmp(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* delta_dress_impl.c:215-219 and 296-316 (same in nabla_dress_impl.c:
 * 124-128, 208-228): n_samples draws of ((r1 << 31) | r2) mod count,
 * seeded with seed + offset, sorted, deduplicated. */
int ora_sample_picks(int64_t count, int64_t n_samples, unsigned int seed,
                     int offset, int64_t *out)
{
    unsigned int s = seed + (unsigned int)offset;
    for (int64_t i = 0; i < n_samples; i++) {
        int64_t hi = (int64_t)ora_rand(&s);
        int64_t lo = (int64_t)ora_rand(&s);
        int64_t r = (hi << 31) | lo;
        out[i] = (r & 0x7FFFFFFFFFFFFFFFLL) % count;
    }
    qsort(out, (size_t)n_samples, sizeof(int64_t), ora_i64_cmp);
    int64_t u = 0;
    for (int64_t i = 0; i < n_samples; i++)
        if (i == 0 || out[i] != out[i - 1]) out[u++] = out[i];
    return (int)u;
}

/* delta_dress_impl.c:95-169: subgraph without the vertices in del[0..k)
 * (sorted), surviving vertices renumbered in order, vertex weights dropped,
 * raw weights kept.  NULL when no edge survives. */
static p_dress_graph_t ora_delta_subgraph(p_dress_graph_t g, const int *del, int k,
                                          int *edge_map)
{
    int N = g->N, E = g->E;
    int *map = (int *)malloc(((size_t)N + 1) * sizeof(int));
    int di = 0, id = 0;
    for (int v = 0; v < N; v++) {
        if (di < k && v == del[di]) { map[v] = -1; di++; }
        else map[v] = id++;
    }
    int sE = 0;
    for (int e = 0; e < E; e++)
        if (map[g->U[e]] >= 0 && map[g->V[e]] >= 0) sE++;
    if (sE == 0) {
        if (edge_map) for (int e = 0; e < E; e++) edge_map[e] = -1;
        free(map);
        return NULL;
    }
    int *sU = (int *)malloc((size_t)sE * sizeof(int));
    int *sV = (int *)malloc((size_t)sE * sizeof(int));
    double *sW = g->W ? (double *)malloc((size_t)sE * sizeof(double)) : NULL;
    int j = 0;
    for (int e = 0; e < E; e++) {
        int a = map[g->U[e]], b = map[g->V[e]];
        if (a >= 0 && b >= 0) {
            sU[j] = a; sV[j] = b;
            if (sW) sW[j] = g->W[e];
            if (edge_map) edge_map[e] = j;
            j++;
        } else if (edge_map) edge_map[e] = -1;
    }
    free(map);
    return ora_init_graph(N - k, sE, sU, sV, sW, NULL, g->variant,
                          g->precompute_intercepts);
}

/* nabla_dress_impl.c:61-91: same topology, vertex weight 1.0 except the
 * tuple, which carries the marker weights. */
static p_dress_graph_t ora_nabla_marked(p_dress_graph_t g, const int *tuple, int k,
                                        const double *mark)
{
    int N = g->N, E = g->E;
    int *sU = (int *)malloc(((size_t)E + 1) * sizeof(int));
    int *sV = (int *)malloc(((size_t)E + 1) * sizeof(int));
    memcpy(sU, g->U, (size_t)E * sizeof(int));
    memcpy(sV, g->V, (size_t)E * sizeof(int));
    double *sW = NULL;
    if (g->W) {
        sW = (double *)malloc(((size_t)E + 1) * sizeof(double));
        memcpy(sW, g->W, (size_t)E * sizeof(double));
    }
    double *nw = (double *)malloc(((size_t)N + 1) * sizeof(double));
    for (int v = 0; v < N; v++) nw[v] = 1.0;
    for (int i = 0; i < k; i++) nw[tuple[i]] = mark[i];
    return ora_init_graph(N, E, sU, sV, sW, nw, g->variant, g->precompute_intercepts);
}

/* nabla_dress_impl.c:33-46: sqrt of the first 16 primes, then sqrt(2i+3). */
static void ora_marker_weights(double *out, int k)
{
    static const int primes[16] = { 2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53 };
    for (int i = 0; i < k; i++)
        out[i] = (i < 16) ? sqrt((double)primes[i]) : sqrt((double)(2 * i + 3));
}

/* Advance a k-combination of {0..N-1} to its lexicographic successor
 * (the order of the DFS at delta_dress_impl.c:326-371).  0 at the end. */
static int ora_next_combo(int *c, int k, int N)
{
    int i = k - 1;
    while (i >= 0 && c[i] == N - k + i) i--;
    if (i < 0) return 0;
    c[i]++;
    for (int j = i + 1; j < k; j++) c[j] = c[j - 1] + 1;
    return 1;
}

/* Advance an ordered k-tuple of distinct vertices to its lexicographic
 * successor (the order of the DFS at nabla_dress_impl.c:238-301). */
static int ora_next_tuple(int *t, int k, int N)
{
    char *used = (char *)calloc((size_t)N + 1, 1);
    for (int i = 0; i < k; i++) used[t[i]] = 1;
    for (int i = k - 1; i >= 0; i--) {
        used[t[i]] = 0;
        int x = t[i] + 1;
        while (x < N && used[x]) x++;
        if (x < N) {
            t[i] = x; used[x] = 1;
            int y = 0;
            for (int j = i + 1; j < k; j++) {
                while (used[y]) y++;
                t[j] = y; used[y] = 1;
            }
            free(used);
            return 1;
        }
    }
    free(used);
    return 0;
}

typedef struct {
    int is_nabla;
    p_dress_graph_t g;
    int k, iterations;
    double epsilon;
    const double *mark;
} ora_job_t;

/* fit one subgraph / marked graph; returns its graph (caller frees) */
static p_dress_graph_t ora_run_item(const ora_job_t *J, const int *items, int *edge_map)
{
    p_dress_graph_t s = J->is_nabla
        ? ora_nabla_marked(J->g, items, J->k, J->mark)
        : ora_delta_subgraph(J->g, items, J->k, edge_map);
    if (s) ora_fit(s, J->iterations, J->epsilon, NULL, NULL);
    return s;
}

static int64_t *ora_lifted_fit(int is_nabla, p_dress_graph_t g, int k, int iterations,
                               double epsilon, int n_samples, unsigned int seed,
                               int *hist_size, int keep_multisets,
                               double **multisets, int64_t *num_items,
                               int offset, int stride, int use_omp)
{
    int N = g->N, E = g->E;
    int do_hist = hist_size != NULL;
    ora_vals_t vals = { NULL, 0, 0 };
    int64_t total = (k == 0) ? 1 : (is_nabla ? ora_perm_count(N, k) : ora_binom(N, k));
    if (num_items) *num_items = total;
    double *ms = NULL;
    if (keep_multisets && multisets) {
        size_t len = (size_t)total * (size_t)E;
        ms = (double *)malloc((len ? len : 1) * sizeof(double));
        if (ms) {
            for (size_t i = 0; i < len; i++) ms[i] = NAN;
            *multisets = ms;
        }
    }
    int64_t *flat = NULL;
    if (k == 0) {
        /* delta_dress_impl.c:268-286: fit g itself, in place */
        if (offset == 0) {
            ora_fit(g, iterations, epsilon, NULL, NULL);
            if (do_hist) ora_vals_push(&vals, g->edge_dress, (size_t)E);
            if (ms) memcpy(ms, g->edge_dress, (size_t)E * sizeof(double));
        }
        flat = do_hist ? ora_flatten(&vals, hist_size) : NULL;
        free(vals.v);
        return flat;
    }
    if (k >= N) {
        flat = do_hist ? ora_flatten(&vals, hist_size) : NULL;
        return flat;
    }

    double *mark = NULL;
    if (is_nabla) {
        mark = (double *)malloc((size_t)k * sizeof(double));
        ora_marker_weights(mark, k);
    }
    ora_job_t J = { is_nabla, g, k, iterations, epsilon, mark };

    int64_t per = (total + stride - 1 - offset) / stride;
    int64_t limit = (n_samples == 0) ? per : (n_samples < per ? n_samples : per);
    int sampling = (n_samples > 0 && limit < per);
    int64_t *picks = NULL;
    int npicks = 0;
    if (sampling) {
        picks = (int64_t *)malloc((size_t)limit * sizeof(int64_t));
        npicks = ora_sample_picks(per, limit, seed, offset, picks);
    }

    /* list the indices s to process, with their item tuples */
    int64_t nsel = 0;
    int64_t *sel = (int64_t *)malloc((size_t)(sampling ? npicks : per) * sizeof(int64_t) + 8);
    int *items = (int *)malloc(((size_t)(sampling ? npicks : per) + 1) * (size_t)k * sizeof(int));
    int *cur = (int *)malloc((size_t)k * sizeof(int));
    for (int i = 0; i < k; i++) cur[i] = is_nabla ? i : i;
    int64_t s = 0, my = 0;
    int pi = 0;
    do {
        if (s % stride == offset) {
            int take = !sampling || (pi < npicks && my == picks[pi]);
            if (take && sampling) pi++;
            if (take) {
                sel[nsel] = s;
                memcpy(items + (size_t)nsel * k, cur, (size_t)k * sizeof(int));
                nsel++;
            }
            my++;
            if (sampling && pi >= npicks) break;
        }
        s++;
    } while (is_nabla ? ora_next_tuple(cur, k, N) : ora_next_combo(cur, k, N));

    if (use_omp && !sampling) {
#ifdef _OPENMP
        int nt = omp_get_max_threads();
#else
        int nt = 1;
#endif
        ora_vals_t *tv = (ora_vals_t *)calloc((size_t)nt, sizeof(ora_vals_t));
#ifdef _OPENMP
#pragma omp parallel
#endif
        {
#ifdef _OPENMP
            int tid = omp_get_thread_num();
#else
            int tid = 0;
#endif
            int *emap = ms ? (int *)malloc(((size_t)E + 1) * sizeof(int)) : NULL;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int64_t i = 0; i < nsel; i++) {
                p_dress_graph_t sg = ora_run_item(&J, items + (size_t)i * k, emap);
                if (sg) {
                    if (do_hist) ora_vals_push(&tv[tid], sg->edge_dress, (size_t)sg->E);
                    if (ms) {
                        double *row = ms + (size_t)sel[i] * E;
                        for (int e = 0; e < E; e++)
                            row[e] = is_nabla ? sg->edge_dress[e]
                                              : (emap[e] >= 0 ? sg->edge_dress[emap[e]] : NAN);
                    }
                    ora_free_graph(sg);
                }
            }
            free(emap);
        }
        for (int t = 0; t < nt; t++) {
            if (tv[t].n) ora_vals_push(&vals, tv[t].v, tv[t].n);
            free(tv[t].v);
        }
        free(tv);
    } else {
        int *emap = ms ? (int *)malloc(((size_t)E + 1) * sizeof(int)) : NULL;
        for (int64_t i = 0; i < nsel; i++) {
            p_dress_graph_t sg = ora_run_item(&J, items + (size_t)i * k, emap);
            if (sg) {
                if (do_hist) ora_vals_push(&vals, sg->edge_dress, (size_t)sg->E);
                if (ms) {
                    double *row = ms + (size_t)sel[i] * E;
                    for (int e = 0; e < E; e++)
                        row[e] = is_nabla ? sg->edge_dress[e]
                                          : (emap[e] >= 0 ? sg->edge_dress[emap[e]] : NAN);
                }
                ora_free_graph(sg);
            }
        }
        free(emap);
    }
    free(cur); free(items); free(sel); free(picks); free(mark);
    flat = do_hist ? ora_flatten(&vals, hist_size) : NULL;
    free(vals.v);
    return flat;
}

int64_t *ora_delta_fit_flat(p_dress_graph_t g, int k, int iterations,
                            double epsilon, int n_samples, unsigned int seed,
                            int *hist_size, int keep_multisets,
                            double **multisets, int64_t *num_subgraphs,
                            int offset, int stride, int use_omp)
{
    return ora_lifted_fit(0, g, k, iterations, epsilon, n_samples, seed, hist_size,
                          keep_multisets, multisets, num_subgraphs, offset, stride,
                          use_omp);
}

int64_t *ora_nabla_fit_flat(p_dress_graph_t g, int k, int iterations,
                            double epsilon, int n_samples, unsigned int seed,
                            int *hist_size, int keep_multisets,
                            double **multisets, int64_t *num_tuples,
                            int offset, int stride, int use_omp)
{
    return ora_lifted_fit(1, g, k, iterations, epsilon, n_samples, seed, hist_size,
                          keep_multisets, multisets, num_tuples, offset, stride,
                          use_omp);
}
