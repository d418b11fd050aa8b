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
