/*
 * spike_rcm.c -- reverse Cuthill-McKee reordering to turn a general sparse matrix into a banded
 * one before the SPIKE solve (the reference declares reorder_rcm as a stub, spike_algebra.c:23-30,
 * and calls cuSOLVER's symrcm only from a standalone helper, spike_cuda.c:898-970).
 *
 * Algorithm (Cuthill & McKee 1969; George & Liu): on the symmetrized pattern |A| + |A|^T, start
 * from a pseudo-peripheral vertex (repeated BFS to the farthest minimum-degree vertex), visit
 * vertices breadth first taking unvisited neighbours in increasing degree, restart on every
 * connected component, and reverse the final order.
 */
#include <stdlib.h>
#include <string.h>

#include "spike_ext.h"

typedef struct { int n; int *ptr; int *adj; int *deg; } sym_graph;

static int cmp_deg_ctx_deg;
static const int *cmp_deg_arr;
static int cmp_by_degree(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    const int dx = cmp_deg_arr[x], dy = cmp_deg_arr[y];
    if (dx != dy) return dx < dy ? -1 : 1;
    return x < y ? -1 : (x > y);
}

static int build_graph(int n, const int *rowptr, const int *colind, sym_graph *g)
{
    g->n = n;
    g->deg = (int *)calloc((size_t)n + 1, sizeof(int));
    int *cnt = (int *)calloc((size_t)n + 1, sizeof(int));
    if (!g->deg || !cnt) { free(cnt); return -1; }
    for (int r = 0; r < n; ++r)
        for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            int c = colind[k];
            if (c == r) continue;
            cnt[r]++; cnt[c]++;
        }
    g->ptr = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int64_t tot = 0;
    for (int r = 0; r < n; ++r) tot += cnt[r];
    g->adj = (int *)malloc(sizeof(int) * (size_t)(tot > 0 ? tot : 1));
    if (!g->ptr || !g->adj) { free(cnt); return -1; }
    g->ptr[0] = 0;
    for (int r = 0; r < n; ++r) g->ptr[r + 1] = g->ptr[r] + cnt[r];
    memset(cnt, 0, sizeof(int) * (size_t)n);
    for (int r = 0; r < n; ++r)
        for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            int c = colind[k];
            if (c == r) continue;
            g->adj[g->ptr[r] + cnt[r]++] = c;
            g->adj[g->ptr[c] + cnt[c]++] = r;
        }
    free(cnt);
    /* de-duplicate each adjacency list */
    int w = 0;
    int *seen = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) seen[i] = -1;
    int *nptr = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    nptr[0] = 0;
    for (int r = 0; r < n; ++r) {
        for (int k = g->ptr[r]; k < g->ptr[r + 1]; ++k) {
            int c = g->adj[k];
            if (seen[c] != r) { seen[c] = r; g->adj[w++] = c; }
        }
        nptr[r + 1] = w;
        g->deg[r] = nptr[r + 1] - nptr[r];
    }
    free(g->ptr);
    g->ptr = nptr;
    free(seen);
    return 0;
}

/* BFS from s restricted to unvisited-in-order vertices; returns the last level's min-degree
 * vertex and the eccentricity (levels - 1). */
static int bfs_far(const sym_graph *g, int s, const char *done, int *level, int *queue, int *ecc)
{
    for (int i = 0; i < g->n; ++i) level[i] = -1;
    int h = 0, t = 0;
    queue[t++] = s;
    level[s] = 0;
    int last = s;
    while (h < t) {
        int v = queue[h++];
        last = v;
        for (int k = g->ptr[v]; k < g->ptr[v + 1]; ++k) {
            int u = g->adj[k];
            if (level[u] < 0 && !done[u]) { level[u] = level[v] + 1; queue[t++] = u; }
        }
    }
    int L = level[last], best = last;
    for (int i = 0; i < t; ++i) {
        int v = queue[i];
        if (level[v] == L && g->deg[v] < g->deg[best]) best = v;
    }
    *ecc = L;
    return best;
}

int spike_rcm_order(int n, const int *rowptr, const int *colind, int *perm, int *ku, int *kl)
{
    if (n <= 0 || !rowptr || !colind || !perm) return SPIKE_ERR_INVALID_ARG;
    sym_graph g;
    memset(&g, 0, sizeof g);
    if (build_graph(n, rowptr, colind, &g)) return SPIKE_ERR_OOM;
    char *done = (char *)calloc((size_t)n, 1);
    int *level = (int *)malloc(sizeof(int) * (size_t)n);
    int *queue = (int *)malloc(sizeof(int) * (size_t)n);
    int *order = (int *)malloc(sizeof(int) * (size_t)n);
    int *nb = (int *)malloc(sizeof(int) * (size_t)n);
    if (!done || !level || !queue || !order || !nb) {
        free(done); free(level); free(queue); free(order); free(nb);
        free(g.ptr); free(g.adj); free(g.deg);
        return SPIKE_ERR_OOM;
    }
    cmp_deg_arr = g.deg;
    (void)cmp_deg_ctx_deg;
    int cnt = 0;
    for (int seed = 0; seed < n; ++seed) {
        if (done[seed]) continue;
        /* pseudo-peripheral start inside this component */
        int s = seed, ecc = -1, e2;
        for (int it = 0; it < 8; ++it) {
            int far = bfs_far(&g, s, done, level, queue, &e2);
            if (e2 <= ecc) break;
            ecc = e2;
            s = far;
        }
        int h = cnt;
        order[cnt++] = s;
        done[s] = 1;
        while (h < cnt) {
            int v = order[h++];
            int m = 0;
            for (int k = g.ptr[v]; k < g.ptr[v + 1]; ++k) {
                int u = g.adj[k];
                if (!done[u]) { done[u] = 1; nb[m++] = u; }
            }
            qsort(nb, (size_t)m, sizeof(int), cmp_by_degree);
            for (int i = 0; i < m; ++i) order[cnt++] = nb[i];
        }
    }
    for (int i = 0; i < n; ++i) perm[i] = order[n - 1 - i];
    /* bandwidth of P A P^T */
    int *inv = level;
    for (int i = 0; i < n; ++i) inv[perm[i]] = i;
    int u = 0, l = 0;
    for (int r = 0; r < n; ++r)
        for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            int d = inv[colind[k]] - inv[r];
            if (d > u) u = d;
            if (-d > l) l = -d;
        }
    if (ku) *ku = u;
    if (kl) *kl = l;
    free(done); free(level); free(queue); free(order); free(nb);
    free(g.ptr); free(g.adj); free(g.deg);
    return SPIKE_SUCCESS;
}

int spike_permute_csr(int dtype, int n, const int *rowptr, const int *colind, const void *aij, const int *perm,
                      int *rowptr_out, int *colind_out, void *aij_out)
{
    if (n <= 0 || !rowptr || !perm || !rowptr_out) return SPIKE_ERR_INVALID_ARG;
    const size_t hs = dtype == SPIKE_DTYPE_Z ? 16 : 8;
    int *inv = (int *)malloc(sizeof(int) * (size_t)n);
    if (!inv) return SPIKE_ERR_OOM;
    for (int i = 0; i < n; ++i) inv[perm[i]] = i;
    rowptr_out[0] = 0;
    for (int i = 0; i < n; ++i) rowptr_out[i + 1] = rowptr_out[i] + (rowptr[perm[i] + 1] - rowptr[perm[i]]);
    for (int i = 0; i < n; ++i) {
        const int r = perm[i];
        int q = rowptr_out[i];
        /* insertion into sorted column order */
        for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const int c = inv[colind[k]];
            int pos = q;
            while (pos > rowptr_out[i] && colind_out[pos - 1] > c) {
                colind_out[pos] = colind_out[pos - 1];
                memcpy((char *)aij_out + (size_t)pos * hs, (char *)aij_out + (size_t)(pos - 1) * hs, hs);
                --pos;
            }
            colind_out[pos] = c;
            memcpy((char *)aij_out + (size_t)pos * hs, (const char *)aij + (size_t)k * hs, hs);
            ++q;
        }
    }
    free(inv);
    return SPIKE_SUCCESS;
}
