/*
 * dress_gen.c -- deterministic synthetic graph generators (host C).
 *
 * Workload synthesis for the benchmarks and tests; not part of the fit path.
 * Random numbers come from splitmix64 seeded with the caller's seed, so a
 * (generator, parameters, seed) triple always yields the same edge list.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

static uint64_t gen_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* uniform integer in [0, bound) without modulo bias */
static uint64_t gen_below(uint64_t *s, uint64_t bound)
{
    uint64_t lim = UINT64_MAX - UINT64_MAX % bound;
    uint64_t r;
    do { r = gen_next(s); } while (r >= lim);
    return r % bound;
}

/*
 * Random d-regular simple graph on n vertices: configuration model (a
 * uniform random perfect matching of the n*d stubs), rejecting the whole
 * matching as soon as it produces a self-loop or a repeated edge.  The
 * accepted graph is uniform over simple d-regular graphs.
 * Writes n*d/2 edges (u < v, sorted by (u, v)) to U, V.
 * Returns the number of attempts, or -1 after max_attempts failures.
 */
int dress_gen_random_regular(int n, int d, uint64_t seed, int max_attempts,
                             int *U, int *V)
{
    if (n <= 0 || d < 0 || d >= n || ((int64_t)n * d) % 2) return -1;
    int m = n * d;
    int *stub = (int *)malloc((size_t)m * sizeof(int));
    unsigned char *adj = (unsigned char *)calloc((size_t)n * n, 1);
    uint64_t s = seed;
    for (int att = 1; att <= max_attempts; att++) {
        for (int v = 0, k = 0; v < n; v++)
            for (int j = 0; j < d; j++) stub[k++] = v;
        memset(adj, 0, (size_t)n * n);
        int ok = 1;
        for (int i = 0; i < m; i += 2) {
            int j = i + 1 + (int)gen_below(&s, (uint64_t)(m - i - 1));
            int t = stub[i + 1]; stub[i + 1] = stub[j]; stub[j] = t;
            int a = stub[i], b = stub[i + 1];
            if (a == b || adj[(size_t)a * n + b]) { ok = 0; break; }
            adj[(size_t)a * n + b] = adj[(size_t)b * n + a] = 1;
        }
        if (!ok) continue;
        int k = 0;
        for (int a = 0; a < n; a++)
            for (int b = a + 1; b < n; b++)
                if (adj[(size_t)a * n + b]) { U[k] = a; V[k] = b; k++; }
        free(stub); free(adj);
        return att;
    }
    free(stub); free(adj);
    return -1;
}

/*
 * Random d-regular simple graph by a switch chain: start from the circulant
 * graph C_n(1..d/2) (d even), then perform swaps_per_edge * (n*d/2) random
 * double-edge swaps {a-b, c-e} -> {a-e, c-b}, skipping any that would create
 * a self-loop or a repeated edge.  Used where whole-matching rejection is
 * infeasible (d = 12 accepts one matching in ~1e15).
 * Writes n*d/2 edges (u < v, sorted) to U, V.  Returns 0, or -1 on bad input.
 */
int dress_gen_regular_switch(int n, int d, uint64_t seed, int swaps_per_edge,
                             int *U, int *V)
{
    if (n <= 0 || d <= 0 || d % 2 || d >= n) return -1;
    int m = n * d / 2;
    int *eu = (int *)malloc((size_t)m * sizeof(int));
    int *ev = (int *)malloc((size_t)m * sizeof(int));
    unsigned char *adj = (unsigned char *)calloc((size_t)n * n, 1);
    int k = 0;
    for (int a = 0; a < n; a++)
        for (int j = 1; j <= d / 2; j++) {
            int b = (a + j) % n;
            eu[k] = a; ev[k] = b; k++;
            adj[(size_t)a * n + b] = adj[(size_t)b * n + a] = 1;
        }
    uint64_t s = seed;
    int64_t swaps = (int64_t)swaps_per_edge * m;
    for (int64_t it = 0; it < swaps; it++) {
        int i = (int)gen_below(&s, (uint64_t)m), j = (int)gen_below(&s, (uint64_t)m);
        if (i == j) continue;
        int a = eu[i], b = ev[i], c = eu[j], e = ev[j];
        if (gen_next(&s) & 1) { int t = c; c = e; e = t; }
        /* new edges a-e and c-b */
        if (a == e || c == b) continue;
        if (adj[(size_t)a * n + e] || adj[(size_t)c * n + b]) continue;
        adj[(size_t)a * n + b] = adj[(size_t)b * n + a] = 0;
        adj[(size_t)c * n + e] = adj[(size_t)e * n + c] = 0;
        adj[(size_t)a * n + e] = adj[(size_t)e * n + a] = 1;
        adj[(size_t)c * n + b] = adj[(size_t)b * n + c] = 1;
        eu[i] = a; ev[i] = e; eu[j] = c; ev[j] = b;
    }
    k = 0;
    for (int a = 0; a < n; a++)
        for (int b = a + 1; b < n; b++)
            if (adj[(size_t)a * n + b]) { U[k] = a; V[k] = b; k++; }
    free(eu); free(ev); free(adj);
    return 0;
}

/* uniform double in [0, 1) */
static double gen_unit(uint64_t *s) { return (double)(gen_next(s) >> 11) * (1.0 / 9007199254740992.0); }

/*
 * G(n, p): every pair {w, v}, w < v, is an edge independently with
 * probability p, drawn by geometric skipping over the pairs in (v, w)
 * order (Batagelj and Brandes, Phys. Rev. E 71, 036113, 2005), so the cost
 * is O(n + m).  Edges are written as U = w < V = v, ordered by (v, w).
 * W, if not NULL, receives float32 weights uniform in (0.1, 1] widened to
 * double.  Returns the edge count, or -1 when more than cap edges were
 * drawn (nothing beyond cap is written).
 */
int64_t dress_gen_gnp(int n, double p, uint64_t seed, int64_t cap, int *U, int *V, double *W)
{
    if (n < 0 || !(p >= 0.0)) return -1;
    uint64_t s = seed;
    int64_t m = 0;
    if (n < 2 || p <= 0.0) return 0;
    if (p >= 1.0) {
        for (int v = 1; v < n; v++)
            for (int w = 0; w < v; w++) {
                if (m >= cap) return -1;
                U[m] = w;
                V[m] = v;
                if (W) W[m] = (double)(float)(1.0 - 0.9 * gen_unit(&s));
                m++;
            }
        return m;
    }
    const double lq = log(1.0 - p);
    int64_t v = 1, w = -1;
    while (v < n) {
        double r = gen_unit(&s);
        w += 1 + (int64_t)floor(log(1.0 - r) / lq);
        while (w >= v && v < n) {
            w -= v;
            v++;
        }
        if (v < n) {
            if (m >= cap) return -1;
            U[m] = (int)w;
            V[m] = (int)v;
            if (W) W[m] = (double)(float)(1.0 - 0.9 * gen_unit(&s));
            m++;
        }
    }
    return m;
}
