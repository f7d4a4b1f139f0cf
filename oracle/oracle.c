/*
 * oracle.c -- CPU restatement of single-clustering v0.7.0's Leiden core and kNN Gaussian
 * builder, for testing the HIP implementation. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Plain C11 (POSIX threads only in the kd-tree kNN, which restates the reference's threaded
 * search); everything else is single-threaded. Each function cites the reference file:line it
 * restates.
 */
#define _POSIX_C_SOURCE 200809L
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <time.h>

#define UNASSIGNED 0xFFFFFFFFu

/* ------------------------------------------------------------------------------------------ */
/* RNG: ChaCha8 (rand_chacha 0.9.0) seeded by rand_core 0.9.3's seed_from_u64 (PCG32 expansion),
 * with rand 0.9.1's shuffle (IncreasingUniform) and random::<f64>(). The block function is
 * pinned by the RFC 7539 vectors in tests; the wrapper streams are unpinned.                  */

void or_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, int rounds,
                     uint32_t out[16]) {
    uint32_t s[16], x[16];
    s[0] = 0x61707865u; s[1] = 0x3320646eu; s[2] = 0x79622d32u; s[3] = 0x6b206574u;
    for (int i = 0; i < 8; i++) s[4 + i] = key[i];
    s[12] = (uint32_t)counter; s[13] = (uint32_t)(counter >> 32);
    s[14] = (uint32_t)stream;  s[15] = (uint32_t)(stream >> 32);
    memcpy(x, s, sizeof(s));
#define ROTL(v, c) (((v) << (c)) | ((v) >> (32 - (c))))
#define QR(a, b, c, d) \
    x[a] += x[b]; x[d] ^= x[a]; x[d] = ROTL(x[d], 16); \
    x[c] += x[d]; x[b] ^= x[c]; x[b] = ROTL(x[b], 12); \
    x[a] += x[b]; x[d] ^= x[a]; x[d] = ROTL(x[d], 8);  \
    x[c] += x[d]; x[b] ^= x[c]; x[b] = ROTL(x[b], 7);
    for (int r = 0; r < rounds; r += 2) {
        QR(0, 4, 8, 12) QR(1, 5, 9, 13) QR(2, 6, 10, 14) QR(3, 7, 11, 15)
        QR(0, 5, 10, 15) QR(1, 6, 11, 12) QR(2, 7, 8, 13) QR(3, 4, 9, 14)
    }
#undef QR
#undef ROTL
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

typedef struct {
    uint32_t key[8];
    uint64_t counter;
    uint32_t buf[64];
    int idx;
} rng_t;

static void rng_refill(rng_t *r) {
    for (int b = 0; b < 4; b++) or_chacha_block(r->key, r->counter + (uint64_t)b, 0, 8, r->buf + 16 * b);
    r->counter += 4;
}

/* rand_core 0.9 SeedableRng::seed_from_u64: PCG32 output words fill the 32-byte seed. */
static void rng_seed_u64(rng_t *r, uint64_t state) {
    const uint64_t MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
    for (int i = 0; i < 8; i++) {
        state = state * MUL + INC;
        uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
        uint32_t rot = (uint32_t)(state >> 59);
        r->key[i] = (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    r->counter = 0;
    r->idx = 64;
}

static uint32_t rng_u32(rng_t *r) {
    if (r->idx >= 64) { rng_refill(r); r->idx = 0; }
    return r->buf[r->idx++];
}

/* rand_core BlockRng::next_u64: two consecutive words, low first, straddling a refill. */
static uint64_t rng_u64(rng_t *r) {
    if (r->idx < 63) {
        uint64_t lo = r->buf[r->idx], hi = r->buf[r->idx + 1];
        r->idx += 2;
        return (hi << 32) | lo;
    } else if (r->idx >= 64) {
        rng_refill(r);
        r->idx = 2;
        return ((uint64_t)r->buf[1] << 32) | r->buf[0];
    } else {
        uint64_t lo = r->buf[63];
        rng_refill(r);
        r->idx = 1;
        return ((uint64_t)r->buf[0] << 32) | lo;
    }
}

/* rand 0.9 StandardUniform for f64: 53 high bits times 2^-53. */
static double rng_f64(rng_t *r) {
    return (double)(rng_u64(r) >> 11) * (1.0 / 9007199254740992.0);
}

/* rand 0.9 UniformInt<u32>::sample_single(0, bound): widening multiply, one extra draw when
 * the low half falls in the biased zone. */
static uint32_t rng_below_u32(rng_t *r, uint32_t bound) {
    uint32_t range = bound; /* [0, bound) == inclusive [0, bound-1], range = bound */
    if (range == 0) return rng_u32(r);
    uint64_t m = (uint64_t)rng_u32(r) * range;
    uint32_t result = (uint32_t)(m >> 32), lo = (uint32_t)m;
    if (lo > (uint32_t)(0u - range)) {
        uint64_t m2 = (uint64_t)rng_u32(r) * range;
        uint32_t new_hi = (uint32_t)(m2 >> 32);
        uint32_t sum = lo + new_hi;
        if (sum < lo) result += 1;
    }
    return result;
}

/* rand 0.9 SliceRandom::shuffle -> partial_shuffle(len) with IncreasingUniform. */
static void rng_shuffle(rng_t *r, uint64_t *a, uint64_t n) {
    if (n <= 1) return;
    uint32_t cn = 0;               /* IncreasingUniform::new(rng, m = 0) */
    uint32_t chunk = 0;
    uint32_t chunk_remaining = 1;  /* n == 0 */
    for (uint64_t i = 0; i < n; i++) {
        uint32_t next_n = cn + 1;
        uint32_t next_rem;
        if (chunk_remaining == 0) {
            /* calculate_bound_u32(next_n) */
            uint64_t product = next_n;
            uint32_t current = next_n + 1;
            for (;;) {
                uint64_t p = product * current;
                if (p > 0xFFFFFFFFull) break;
                product = p;
                current++;
            }
            uint32_t count = current - next_n;
            chunk = rng_below_u32(r, (uint32_t)product);
            next_rem = count - 1;
        } else {
            next_rem = chunk_remaining - 1;
        }
        uint32_t res;
        if (next_rem == 0) {
            res = chunk;
        } else {
            res = chunk % next_n;
            chunk /= next_n;
        }
        chunk_remaining = next_rem;
        cn = next_n;
        uint64_t t = a[i]; a[i] = a[res]; a[res] = t;
    }
}

void or_chacha8_seed_u64_stream(uint64_t seed, uint32_t *out, uint64_t count) {
    rng_t r;
    rng_seed_u64(&r, seed);
    for (uint64_t i = 0; i < count; i++) out[i] = rng_u32(&r);
}

/* ------------------------------------------------------------------------------------------ */
/* CSRNetwork                                                                                  */

void or_config_default(or_config *c) {
    /* LeidenConfig::default (leiden/mod.rs:92-106) */
    c->objective = 0;
    c->resolution = 1.0;
    c->has_seed = 1;
    c->seed = 42;
    c->n_iterations = 2;
    c->max_levels = 64;
    c->tolerance = 1e-9;
    c->refine = 1;
    c->refine_randomness = 0.01;
    c->patience = 2;
    c->has_max_community_weight = 0;
    c->max_community_weight = 0.0;
}

void or_graph_free(or_graph *g) {
    if (!g) return;
    free(g->ptrs); free(g->nbrs); free(g->wts); free(g->node_w); free(g->strength);
    free(g);
}

typedef struct { uint32_t u; float w; } pair_uw;

static int cmp_pair_u(const void *a, const void *b) {
    uint32_t x = ((const pair_uw *)a)->u, y = ((const pair_uw *)b)->u;
    return (x > y) - (x < y);
}

/* CSRNetwork::finish (csr_network.rs:262-351) + check_degree_sum (:712-723).
 * Takes ownership of ptrs/nbrs/wts/node_w. */
static int graph_finish(uint64_t n, uint64_t *ptrs, uint32_t *nbrs, float *wts, double *node_w,
                        or_graph **out) {
    uint64_t max_deg = 0;
    for (uint64_t v = 0; v < n; v++) {
        uint64_t d = ptrs[v + 1] - ptrs[v];
        if (d > max_deg) max_deg = d;
    }
    pair_uw *scratch = (pair_uw *)malloc((max_deg ? max_deg : 1) * sizeof(pair_uw));
    double *strengths = (double *)calloc(n ? n : 1, sizeof(double));
    double total = 0.0;
    uint64_t edge_count = 0, write = 0;
    int any_loops = 0;
    for (uint64_t v = 0; v < n; v++) {
        uint64_t lo = ptrs[v], hi = ptrs[v + 1];
        ptrs[v] = write;
        uint64_t len = hi - lo;
        for (uint64_t i = 0; i < len; i++) { scratch[i].u = nbrs[lo + i]; scratch[i].w = wts[lo + i]; }
        /* sort_unstable_by_key on u: equal keys are merged below, so stability is moot */
        qsort(scratch, len, sizeof(pair_uw), cmp_pair_u);
        uint64_t i = 0;
        while (i < len) {
            uint32_t u = scratch[i].u;
            double w = 0.0;
            while (i < len && scratch[i].u == u) { w += (double)scratch[i].w; i++; }
            if (w == 0.0) continue;
            nbrs[write] = u;
            wts[write] = (float)w;
            write++;
            double stored = (double)(float)w;
            if ((uint64_t)u == v) { any_loops = 1; strengths[v] += 2.0 * stored; }
            else strengths[v] += stored;
            if (v <= (uint64_t)u) { total += stored; edge_count++; }
        }
    }
    ptrs[n] = write;
    free(scratch);

    or_graph *g = (or_graph *)calloc(1, sizeof(or_graph));
    g->n = n; g->ptrs = ptrs; g->nbrs = nbrs; g->wts = wts; g->node_w = node_w;
    g->strength = strengths; g->total_weight = total; g->edge_count = edge_count;
    g->any_self_loops = any_loops;

    double sum = 0.0;
    for (uint64_t v = 0; v < n; v++) sum += strengths[v];
    double expected = 2.0 * total;
    double tol = 1e-6 * fmax(fabs(expected), 1.0);
    if (fabs(sum - expected) > tol) { or_graph_free(g); return OR_ASYMMETRIC_GRAPH; }
    *out = g;
    return OR_OK;
}

static int validate_node_weights(const double *w, uint64_t n) {
    for (uint64_t v = 0; v < n; v++)
        if (!isfinite(w[v]) || w[v] < 0.0) return OR_INVALID_NODE_WEIGHT;
    return OR_OK;
}

/* CSRNetwork::from_csr_parts (csr_network.rs:185-256). Inputs are copied. */
int or_graph_from_csr_parts(uint64_t n, const uint64_t *ptrs_in, const uint32_t *nbrs_in,
                            const float *wts_in, uint64_t nnz, const double *node_w_in,
                            or_graph **out) {
    if (n > 0xFFFFFFFFull) return OR_TOO_MANY_NODES;
    if (ptrs_in[n] != nnz) return OR_INVALID_CSR;
    for (uint64_t v = 0; v < n; v++)
        if (ptrs_in[v] > ptrs_in[v + 1]) return OR_INVALID_CSR;
    for (uint64_t e = 0; e < nnz; e++) {
        if ((uint64_t)nbrs_in[e] >= n) return OR_NODE_INDEX_OUT_OF_RANGE;
        if (!isfinite(wts_in[e])) return OR_NON_FINITE_WEIGHT;
        if (wts_in[e] < 0.0f) return OR_NEGATIVE_WEIGHT;
    }
    if (node_w_in) {
        int st = validate_node_weights(node_w_in, n);
        if (st) return st;
    }
    uint64_t *ptrs = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    uint32_t *nbrs = (uint32_t *)malloc((nnz ? nnz : 1) * sizeof(uint32_t));
    float *wts = (float *)malloc((nnz ? nnz : 1) * sizeof(float));
    double *nw = (double *)malloc((n ? n : 1) * sizeof(double));
    memcpy(ptrs, ptrs_in, (n + 1) * sizeof(uint64_t));
    if (nnz) { memcpy(nbrs, nbrs_in, nnz * sizeof(uint32_t)); memcpy(wts, wts_in, nnz * sizeof(float)); }
    for (uint64_t v = 0; v < n; v++) nw[v] = node_w_in ? node_w_in[v] : 1.0;
    return graph_finish(n, ptrs, nbrs, wts, nw, out);
}

/* CSRNetwork::from_edges_with_node_weights (csr_network.rs:120-167). */
int or_graph_from_edges(uint64_t n, const or_edge *edges, uint64_t m, const double *node_w_in,
                        or_graph **out) {
    if (node_w_in) {
        int st = validate_node_weights(node_w_in, n);
        if (st) return st;
    }
    if (n > 0xFFFFFFFFull) return OR_TOO_MANY_NODES;
    uint64_t *counts = (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < m; i++) {
        if (edges[i].from >= n || edges[i].to >= n) { free(counts); return OR_NODE_INDEX_OUT_OF_RANGE; }
        if (!isfinite(edges[i].w)) { free(counts); return OR_NON_FINITE_WEIGHT; }
        if (edges[i].w < 0.0) { free(counts); return OR_NEGATIVE_WEIGHT; }
        counts[edges[i].from]++;
        if (edges[i].from != edges[i].to) counts[edges[i].to]++;
    }
    uint64_t *ptrs = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    ptrs[0] = 0;
    for (uint64_t v = 0; v < n; v++) ptrs[v + 1] = ptrs[v] + counts[v];
    uint64_t slots = ptrs[n];
    uint32_t *nbrs = (uint32_t *)malloc((slots ? slots : 1) * sizeof(uint32_t));
    float *wts = (float *)malloc((slots ? slots : 1) * sizeof(float));
    for (uint64_t v = 0; v < n; v++) counts[v] = ptrs[v];
    for (uint64_t i = 0; i < m; i++) {
        float w = (float)edges[i].w;
        uint64_t a = edges[i].from, b = edges[i].to;
        nbrs[counts[a]] = (uint32_t)b; wts[counts[a]] = w; counts[a]++;
        if (a != b) { nbrs[counts[b]] = (uint32_t)a; wts[counts[b]] = w; counts[b]++; }
    }
    free(counts);
    double *nw = (double *)malloc((n ? n : 1) * sizeof(double));
    for (uint64_t v = 0; v < n; v++) nw[v] = node_w_in ? node_w_in[v] : 1.0;
    return graph_finish(n, ptrs, nbrs, wts, nw, out);
}

/* CSRNetwork::self_loop_weight (csr_network.rs:452-462). */
double or_self_loop_weight(const or_graph *g, uint64_t v) {
    if (!g->any_self_loops) return 0.0;
    uint64_t lo = g->ptrs[v], hi = g->ptrs[v + 1];
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (g->nbrs[mid] == (uint32_t)v) return (double)g->wts[mid];
        if (g->nbrs[mid] < (uint32_t)v) lo = mid + 1; else hi = mid;
    }
    return 0.0;
}

static int edge_weight(const or_graph *g, uint64_t from, uint64_t to, double *w) {
    uint64_t a = from, b = to;
    if (g->ptrs[to + 1] - g->ptrs[to] < g->ptrs[from + 1] - g->ptrs[from]) { a = to; b = from; }
    uint64_t lo = g->ptrs[a], hi = g->ptrs[a + 1];
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (g->nbrs[mid] == (uint32_t)b) { *w = (double)g->wts[mid]; return 1; }
        if (g->nbrs[mid] < (uint32_t)b) lo = mid + 1; else hi = mid;
    }
    return 0;
}

/* CSRNetwork::validate_symmetry (csr_network.rs:627-646). */
int or_graph_validate_symmetry(const or_graph *g) {
    for (uint64_t v = 0; v < g->n; v++) {
        for (uint64_t e = g->ptrs[v]; e < g->ptrs[v + 1]; e++) {
            uint64_t u = g->nbrs[e];
            double w = (double)g->wts[e], back;
            if (!edge_weight(g, u, v, &back)) return OR_INVALID_CSR;
            if (fabs(back - w) > 1e-6 * fmax(fabs(w), 1.0)) return OR_INVALID_CSR;
        }
    }
    return OR_OK;
}

typedef struct { uint32_t a, b; float w; } coarse_pair;

/* CSRNetwork::aggregate (csr_network.rs:516-608). */
or_graph *or_graph_aggregate(const or_graph *g, const uint64_t *groups, uint64_t ng) {
    uint64_t n = g->n;
    double *new_nw = (double *)calloc(ng ? ng : 1, sizeof(double));
    for (uint64_t v = 0; v < n; v++) new_nw[groups[v]] += g->node_w[v];

    uint64_t *starts = (uint64_t *)calloc(ng + 1, sizeof(uint64_t));
    for (uint64_t v = 0; v < n; v++) starts[groups[v] + 1]++;
    for (uint64_t c = 0; c < ng; c++) starts[c + 1] += starts[c];
    uint32_t *members = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint64_t *cursor = (uint64_t *)malloc((ng + 1) * sizeof(uint64_t));
    memcpy(cursor, starts, (ng + 1) * sizeof(uint64_t));
    for (uint64_t v = 0; v < n; v++) members[cursor[groups[v]]++] = (uint32_t)v;

    double *acc = (double *)calloc(ng ? ng : 1, sizeof(double));
    uint64_t *touched = (uint64_t *)malloc((ng ? ng : 1) * sizeof(uint64_t));
    uint64_t n_touched = 0;
    uint64_t *counts = (uint64_t *)calloc(ng ? ng : 1, sizeof(uint64_t));
    uint64_t cap = 1024, np = 0;
    coarse_pair *pairs = (coarse_pair *)malloc(cap * sizeof(coarse_pair));

    for (uint64_t c = 0; c < ng; c++) {
        double loops = 0.0;
        for (uint64_t i = starts[c]; i < starts[c + 1]; i++) {
            uint64_t v = members[i];
            for (uint64_t e = g->ptrs[v]; e < g->ptrs[v + 1]; e++) {
                uint64_t u = g->nbrs[e];
                double w = (double)g->wts[e];
                if (u == v) loops += w;
                uint64_t h = groups[u];
                if (h < c) continue;
                if (acc[h] == 0.0) touched[n_touched++] = h;
                acc[h] += w;
            }
        }
        for (uint64_t t = 0; t < n_touched; t++) {
            uint64_t h = touched[t];
            double w = acc[h];
            acc[h] = 0.0;
            if (w <= 0.0) continue;
            if (np == cap) { cap *= 2; pairs = (coarse_pair *)realloc(pairs, cap * sizeof(coarse_pair)); }
            if (h == c) {
                pairs[np].a = (uint32_t)c; pairs[np].b = (uint32_t)c; pairs[np].w = (float)(0.5 * (w + loops));
                counts[c]++;
            } else {
                pairs[np].a = (uint32_t)c; pairs[np].b = (uint32_t)h; pairs[np].w = (float)w;
                counts[c]++; counts[h]++;
            }
            np++;
        }
        n_touched = 0;
    }

    uint64_t *ptrs = (uint64_t *)malloc((ng + 1) * sizeof(uint64_t));
    ptrs[0] = 0;
    for (uint64_t c = 0; c < ng; c++) ptrs[c + 1] = ptrs[c] + counts[c];
    uint64_t slots = ptrs[ng];
    uint32_t *nbrs = (uint32_t *)malloc((slots ? slots : 1) * sizeof(uint32_t));
    float *wts = (float *)malloc((slots ? slots : 1) * sizeof(float));
    memcpy(cursor, ptrs, (ng + 1) * sizeof(uint64_t));
    for (uint64_t p = 0; p < np; p++) {
        uint32_t a = pairs[p].a, b = pairs[p].b;
        nbrs[cursor[a]] = b; wts[cursor[a]] = pairs[p].w; cursor[a]++;
        if (a != b) { nbrs[cursor[b]] = a; wts[cursor[b]] = pairs[p].w; cursor[b]++; }
    }
    free(starts); free(members); free(cursor); free(acc); free(touched); free(counts); free(pairs);
    or_graph *out = NULL;
    int st = graph_finish(ng, ptrs, nbrs, wts, new_nw, &out);
    if (st != OR_OK) {
        fprintf(stderr, "oracle aggregate: finish failed (%d)\n", st);
        abort();
    }
    return out;
}

/* ------------------------------------------------------------------------------------------ */
/* Partition (leiden/partition.rs:26-333)                                                      */

typedef struct {
    uint64_t n;
    uint32_t *memb;
    uint64_t slots, cap;
    uint32_t *size;
    double *weight, *strength, *internal;
    uint64_t *free_list;
    uint64_t n_free;
    double total_weight, total_node_weight;
} partition;

static void part_grow(partition *p, uint64_t need) {
    if (need <= p->cap) return;
    uint64_t nc = p->cap ? p->cap : 16;
    while (nc < need) nc *= 2;
    p->size = (uint32_t *)realloc(p->size, nc * sizeof(uint32_t));
    p->weight = (double *)realloc(p->weight, nc * sizeof(double));
    p->strength = (double *)realloc(p->strength, nc * sizeof(double));
    p->internal = (double *)realloc(p->internal, nc * sizeof(double));
    p->free_list = (uint64_t *)realloc(p->free_list, nc * sizeof(uint64_t));
    p->cap = nc;
}

static double total_node_weight(const or_graph *g) {
    double s = 0.0;
    for (uint64_t v = 0; v < g->n; v++) s += g->node_w[v];
    return s;
}

/* Partition::from_membership (partition.rs:75-112). */
static void part_from_membership(partition *p, const or_graph *g, const uint64_t *memb) {
    memset(p, 0, sizeof(*p));
    uint64_t n = g->n;
    uint64_t slots = 0;
    for (uint64_t v = 0; v < n; v++) if (memb[v] + 1 > slots) slots = memb[v] + 1;
    p->n = n;
    p->memb = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    part_grow(p, slots + 1);
    p->slots = slots;
    for (uint64_t c = 0; c < slots; c++) { p->size[c] = 0; p->weight[c] = 0; p->strength[c] = 0; p->internal[c] = 0; }
    p->total_weight = g->total_weight;
    p->total_node_weight = total_node_weight(g);
    for (uint64_t v = 0; v < n; v++) {
        uint64_t c = memb[v];
        p->memb[v] = (uint32_t)c;
        p->size[c] += 1;
        p->weight[c] += g->node_w[v];
        p->strength[c] += g->strength[v];
    }
    for (uint64_t v = 0; v < n; v++) {
        uint64_t c = memb[v];
        for (uint64_t e = g->ptrs[v]; e < g->ptrs[v + 1]; e++) {
            uint64_t u = g->nbrs[e];
            if (memb[u] != c) continue;
            if (u == v || v < u) p->internal[c] += (double)g->wts[e];
        }
    }
    p->n_free = 0;
    for (uint64_t c = slots; c-- > 0;)
        if (p->size[c] == 0) p->free_list[p->n_free++] = c;
}

/* Partition::singleton (partition.rs:53-69). */
static void part_singleton(partition *p, const or_graph *g) {
    memset(p, 0, sizeof(*p));
    uint64_t n = g->n;
    p->n = n;
    p->memb = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    part_grow(p, n + 1);
    p->slots = n;
    for (uint64_t v = 0; v < n; v++) {
        p->memb[v] = (uint32_t)v;
        p->size[v] = 1;
        p->weight[v] = g->node_w[v];
        p->strength[v] = g->strength[v];
        p->internal[v] = or_self_loop_weight(g, v);
    }
    p->n_free = 0;
    p->total_weight = g->total_weight;
    p->total_node_weight = total_node_weight(g);
}

static void part_free(partition *p) {
    free(p->memb); free(p->size); free(p->weight); free(p->strength); free(p->internal);
    free(p->free_list);
    memset(p, 0, sizeof(*p));
}

static void part_copy(partition *dst, const partition *src) {
    memset(dst, 0, sizeof(*dst));
    dst->n = src->n;
    dst->memb = (uint32_t *)malloc((src->n ? src->n : 1) * sizeof(uint32_t));
    memcpy(dst->memb, src->memb, src->n * sizeof(uint32_t));
    part_grow(dst, src->cap ? src->cap : 1);
    dst->slots = src->slots;
    memcpy(dst->size, src->size, src->slots * sizeof(uint32_t));
    memcpy(dst->weight, src->weight, src->slots * sizeof(double));
    memcpy(dst->strength, src->strength, src->slots * sizeof(double));
    memcpy(dst->internal, src->internal, src->slots * sizeof(double));
    memcpy(dst->free_list, src->free_list, src->n_free * sizeof(uint64_t));
    dst->n_free = src->n_free;
    dst->total_weight = src->total_weight;
    dst->total_node_weight = src->total_node_weight;
}

static inline double p_weight(const partition *p, uint64_t c) { return c < p->slots ? p->weight[c] : 0.0; }
static inline double p_strength(const partition *p, uint64_t c) { return c < p->slots ? p->strength[c] : 0.0; }
static inline uint64_t p_size(const partition *p, uint64_t c) { return c < p->slots ? p->size[c] : 0; }

static inline uint64_t part_empty_community(const partition *p) {
    return p->n_free ? p->free_list[p->n_free - 1] : p->slots;
}

/* Partition::remove_node (partition.rs:201-224). */
static uint64_t part_remove(partition *p, const or_graph *g, uint64_t v, double w_own) {
    uint64_t c = p->memb[v];
    p->size[c] -= 1;
    p->weight[c] -= g->node_w[v];
    p->strength[c] -= g->strength[v];
    p->internal[c] -= w_own + or_self_loop_weight(g, v);
    if (p->size[c] == 0) {
        p->weight[c] = 0.0; p->strength[c] = 0.0; p->internal[c] = 0.0;
        p->free_list[p->n_free++] = c;
    }
    p->memb[v] = UNASSIGNED;
    return c;
}

/* Partition::insert_node (partition.rs:230-260). */
static void part_insert(partition *p, const or_graph *g, uint64_t v, uint64_t c, double w_c) {
    if (c == p->slots) {
        part_grow(p, p->slots + 1);
        p->size[c] = 0; p->weight[c] = 0.0; p->strength[c] = 0.0; p->internal[c] = 0.0;
        p->slots++;
    } else if (p->size[c] == 0) {
        for (uint64_t i = p->n_free; i-- > 0;) {
            if (p->free_list[i] == c) {
                p->free_list[i] = p->free_list[p->n_free - 1];
                p->n_free--;
                break;
            }
        }
    }
    p->size[c] += 1;
    p->weight[c] += g->node_w[v];
    p->strength[c] += g->strength[v];
    p->internal[c] += w_c + or_self_loop_weight(g, v);
    p->memb[v] = (uint32_t)c;
}

/* Partition::renumber (partition.rs:265-296); returns the new id table (caller frees). */
static void part_renumber(partition *p) {
    uint64_t slots = p->slots;
    uint64_t *new_id = (uint64_t *)malloc((slots ? slots : 1) * sizeof(uint64_t));
    uint64_t next = 0;
    for (uint64_t c = 0; c < slots; c++) new_id[c] = p->size[c] > 0 ? next++ : UINT64_MAX;
    for (uint64_t c = 0; c < slots; c++) {
        uint64_t t = new_id[c];
        if (t == UINT64_MAX || t == c) continue;
        p->size[t] = p->size[c]; p->weight[t] = p->weight[c];
        p->strength[t] = p->strength[c]; p->internal[t] = p->internal[c];
    }
    p->slots = next;
    p->n_free = 0;
    for (uint64_t v = 0; v < p->n; v++)
        if (p->memb[v] != UNASSIGNED) p->memb[v] = (uint32_t)new_id[p->memb[v]];
    free(new_id);
}

static uint64_t part_community_count(const partition *p) { return p->slots - p->n_free; }

/* ------------------------------------------------------------------------------------------ */
/* Objectives (leiden/objective.rs:72-184)                                                     */

typedef struct { double strength, node_weight, self_loop; } insert_ctx;

static inline double delta_insert(int objective, double res, const partition *p,
                                  const insert_ctx *ctx, uint64_t c, double w_c) {
    if (objective == 0) {
        double four_m = 4.0 * p->total_weight;
        if (four_m == 0.0) return 0.0;
        double k_c = p_strength(p, c), k = ctx->strength;
        return (w_c + ctx->self_loop) - res * k * (2.0 * k_c + k) / four_m;
    } else {
        double n_c = p_weight(p, c), a = ctx->node_weight;
        return (w_c + ctx->self_loop) - res * a * (2.0 * n_c + a - 1.0) / 2.0;
    }
}

static double part_quality(int objective, double res, const partition *p) {
    double q = 0.0;
    if (objective == 0) {
        double four_m = 4.0 * p->total_weight;
        if (four_m == 0.0) return 0.0;
        for (uint64_t c = 0; c < p->slots; c++) {
            double k_c = p->strength[c];
            q += p->internal[c] - res * k_c * k_c / four_m;
        }
    } else {
        for (uint64_t c = 0; c < p->slots; c++) {
            double n_c = p->weight[c];
            q += p->internal[c] - res * n_c * (n_c - 1.0) / 2.0;
        }
    }
    return q;
}

/* modularity (objective.rs:221-244). */
double or_modularity(const or_graph *g, const uint64_t *labels, double res) {
    double m = g->total_weight;
    if (m == 0.0) return 0.0;
    uint64_t nc = 0;
    for (uint64_t v = 0; v < g->n; v++) if (labels[v] + 1 > nc) nc = labels[v] + 1;
    double *internal = (double *)calloc(nc ? nc : 1, sizeof(double));
    double *strength = (double *)calloc(nc ? nc : 1, sizeof(double));
    for (uint64_t v = 0; v < g->n; v++) {
        uint64_t c = labels[v];
        strength[c] += g->strength[v];
        for (uint64_t e = g->ptrs[v]; e < g->ptrs[v + 1]; e++) {
            uint64_t u = g->nbrs[e];
            if (labels[u] == c && (u == v || v < u)) internal[c] += (double)g->wts[e];
        }
    }
    double two_m = 2.0 * m, q = 0.0;
    for (uint64_t c = 0; c < nc; c++) {
        double f = strength[c] / two_m;
        q += 2.0 * internal[c] / two_m - res * (f * f);
    }
    free(internal); free(strength);
    return q;
}

double or_quality(const or_graph *g, const uint64_t *labels, int objective, double res) {
    partition p;
    part_from_membership(&p, g, labels);
    double q = part_quality(objective, res, &p);
    part_free(&p);
    return q;
}

/* ------------------------------------------------------------------------------------------ */
/* NeighborWeights + MoveScratch (partition.rs:343-551)                                        */

typedef struct {
    uint64_t *stamp;
    double *weight;
    uint64_t cap;
    uint64_t generation;
    uint64_t *touched;
    uint64_t n_touched, touched_cap;
    /* MoveScratch */
    uint64_t *order;
    uint64_t *queue;
    uint64_t q_head, q_len, q_cap;
    unsigned char *queued;
} scratch_t;

static void nw_ensure(scratch_t *s, uint64_t slots) {
    if (s->cap >= slots) return;
    s->stamp = (uint64_t *)realloc(s->stamp, slots * sizeof(uint64_t));
    s->weight = (double *)realloc(s->weight, slots * sizeof(double));
    for (uint64_t i = s->cap; i < slots; i++) { s->stamp[i] = 0; s->weight[i] = 0.0; }
    s->cap = slots;
    if (s->touched_cap < slots) {
        s->touched = (uint64_t *)realloc(s->touched, slots * sizeof(uint64_t));
        s->touched_cap = slots;
    }
}

static void nw_begin(scratch_t *s) {
    s->n_touched = 0;
    s->generation += 1;  /* u64 wrap is unreachable in practice; restated for completeness */
    if (s->generation == 0) {
        for (uint64_t i = 0; i < s->cap; i++) s->stamp[i] = 0;
        s->generation = 1;
    }
}

static inline void nw_add(scratch_t *s, uint64_t c, double w) {
    if (c >= s->cap) nw_ensure(s, c + 1);
    if (s->stamp[c] != s->generation) {
        s->stamp[c] = s->generation;
        s->weight[c] = w;
        s->touched[s->n_touched++] = c;
    } else {
        s->weight[c] += w;
    }
}

static inline double nw_weight_to(const scratch_t *s, uint64_t c) {
    return (c < s->cap && s->stamp[c] == s->generation) ? s->weight[c] : 0.0;
}

/* NeighborWeights::collect (partition.rs:426-440). */
static void nw_collect(scratch_t *s, const or_graph *g, const uint32_t *memb, uint64_t v) {
    nw_begin(s);
    for (uint64_t e = g->ptrs[v]; e < g->ptrs[v + 1]; e++) {
        uint64_t u = g->nbrs[e];
        if (u == v) continue;
        uint32_t c = memb[u];
        if (c == UNASSIGNED) continue;
        nw_add(s, c, (double)g->wts[e]);
    }
}

/* NeighborWeights::collect_constrained (partition.rs:447-468). */
static void nw_collect_constrained(scratch_t *s, const or_graph *g, const uint32_t *memb,
                                   uint64_t v, const uint32_t *constraint) {
    nw_begin(s);
    uint32_t own = constraint[v];
    for (uint64_t e = g->ptrs[v]; e < g->ptrs[v + 1]; e++) {
        uint64_t u = g->nbrs[e];
        if (u == v || constraint[u] != own) continue;
        uint32_t c = memb[u];
        if (c == UNASSIGNED) continue;
        nw_add(s, c, (double)g->wts[e]);
    }
}

static void scratch_init(scratch_t *s, uint64_t n, uint64_t slots) {
    memset(s, 0, sizeof(*s));
    nw_ensure(s, slots ? slots : 1);
    s->order = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    s->q_cap = n ? n : 1;
    s->queue = (uint64_t *)malloc(s->q_cap * sizeof(uint64_t));
    s->queued = (unsigned char *)malloc(n ? n : 1);
}

static void scratch_free(scratch_t *s) {
    free(s->stamp); free(s->weight); free(s->touched); free(s->order); free(s->queue); free(s->queued);
}

static void scratch_ensure_n(scratch_t *s, uint64_t n) {
    if (s->q_cap < n) {
        s->q_cap = n;
        s->queue = (uint64_t *)realloc(s->queue, n * sizeof(uint64_t));
        s->order = (uint64_t *)realloc(s->order, n * sizeof(uint64_t));
        s->queued = (unsigned char *)realloc(s->queued, n);
    }
}

/* MoveScratch::seed_queue (partition.rs:512-522). The queue is a ring of capacity n: a node is
 * queued at most once, so it never holds more than n entries. */
static void scratch_seed_queue(scratch_t *s, uint64_t n, rng_t *r) {
    scratch_ensure_n(s, n);
    for (uint64_t i = 0; i < n; i++) s->order[i] = i;
    rng_shuffle(r, s->order, n);
    for (uint64_t i = 0; i < n; i++) s->queue[i] = s->order[i];
    s->q_head = 0; s->q_len = n;
    memset(s->queued, 1, n);
}

static int scratch_pop(scratch_t *s, uint64_t n, uint64_t *v) {
    if (s->q_len == 0) return 0;
    *v = s->queue[s->q_head];
    s->q_head = (s->q_head + 1) % n;
    s->q_len--;
    s->queued[*v] = 0;
    return 1;
}

static void scratch_push(scratch_t *s, uint64_t n, uint64_t v) {
    if (!s->queued[v]) {
        s->queued[v] = 1;
        s->queue[(s->q_head + s->q_len) % n] = v;
        s->q_len++;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* local_move (local_move.rs:39-122)                                                           */

static double local_move(const or_graph *g, partition *p, int objective, double res,
                         scratch_t *s, rng_t *r, int has_max, double max_w) {
    uint64_t n = g->n;
    if (n == 0) return 0.0;
    scratch_seed_queue(s, n, r);
    double threshold = 1e-12 * fmax(fabs(p->total_weight), 1.0);
    double total_gain = 0.0;
    uint64_t moves = 0, budget = n * 100;
    uint64_t v;
    while (scratch_pop(s, n, &v)) {
        insert_ctx ctx = { g->strength[v], g->node_w[v], or_self_loop_weight(g, v) };
        nw_ensure(s, p->slots + 1);
        nw_collect(s, g, p->memb, v);
        uint64_t old = p->memb[v];
        double w_old = nw_weight_to(s, old);
        part_remove(p, g, v, w_old);
        double base = delta_insert(objective, res, p, &ctx, old, w_old);
        uint64_t best = old;
        double best_gain = base;
        for (uint64_t i = 0; i < s->n_touched; i++) {
            uint64_t c = s->touched[i];
            if (c == old) continue;
            if (has_max && !(p_weight(p, c) + ctx.node_weight <= max_w)) continue;
            double gain = delta_insert(objective, res, p, &ctx, c, nw_weight_to(s, c));
            if (gain > best_gain || (gain == best_gain && c < best)) { best = c; best_gain = gain; }
        }
        uint64_t empty = part_empty_community(p);
        if (empty != old && (!has_max || p_weight(p, empty) + ctx.node_weight <= max_w)) {
            double gain = delta_insert(objective, res, p, &ctx, empty, 0.0);
            if (gain > best_gain || (gain == best_gain && empty < best)) { best = empty; best_gain = gain; }
        }
        int accept = best != old && best_gain - base > threshold && moves < budget;
        uint64_t target = accept ? best : old;
        part_insert(p, g, v, target, nw_weight_to(s, target));
        if (!accept) continue;
        total_gain += best_gain - base;
        moves++;
        for (uint64_t e = g->ptrs[v]; e < g->ptrs[v + 1]; e++) {
            uint64_t u = g->nbrs[e];
            if (u != v && !s->queued[u] && p->memb[u] != target) scratch_push(s, n, u);
        }
    }
    return total_gain;
}

/* ------------------------------------------------------------------------------------------ */
/* refine + select (refine.rs:29-153)                                                          */

typedef struct { uint64_t c; double rel; } cand_t;

static int select_cand(const cand_t *cands, uint64_t nc, double best_rel, double randomness,
                       rng_t *r, double *wbuf, uint64_t *out) {
    if (nc == 0) return 0;
    if (randomness <= 0.0) {
        uint64_t best = UINT64_MAX;
        for (uint64_t i = 0; i < nc; i++)
            if (cands[i].rel >= best_rel && cands[i].c < best) best = cands[i].c;
        *out = best;
        return 1;
    }
    double total = 0.0;
    for (uint64_t i = 0; i < nc; i++) { wbuf[i] = exp((cands[i].rel - best_rel) / randomness); total += wbuf[i]; }
    if (!isfinite(total) || total <= 0.0) {
        uint64_t best = UINT64_MAX;
        for (uint64_t i = 0; i < nc; i++)
            if (cands[i].rel >= best_rel && cands[i].c < best) best = cands[i].c;
        *out = best;
        return 1;
    }
    double threshold = rng_f64(r) * total;
    for (uint64_t i = 0; i < nc; i++) {
        threshold -= wbuf[i];
        if (threshold <= 0.0) { *out = cands[i].c; return 1; }
    }
    *out = cands[nc - 1].c;
    return 1;
}

static void refine(const or_graph *g, const uint32_t *constraint, int objective, double res,
                   scratch_t *s, rng_t *r, int has_max, double max_w, double randomness,
                   partition *refined) {
    uint64_t n = g->n;
    part_singleton(refined, g);
    if (n == 0) return;
    double mean_w = g->edge_count > 0 ? g->total_weight / (double)g->edge_count : 1.0;
    randomness = randomness * mean_w;
    scratch_ensure_n(s, n);
    for (uint64_t i = 0; i < n; i++) s->order[i] = i;
    rng_shuffle(r, s->order, n);
    uint64_t *order = (uint64_t *)malloc(n * sizeof(uint64_t));
    memcpy(order, s->order, n * sizeof(uint64_t));
    cand_t *cands = (cand_t *)malloc((n + 1) * sizeof(cand_t));
    double *wbuf = (double *)malloc((n + 1) * sizeof(double));
    for (uint64_t oi = 0; oi < n; oi++) {
        uint64_t v = order[oi];
        uint64_t own = refined->memb[v];
        if (p_size(refined, own) != 1) continue;
        insert_ctx ctx = { g->strength[v], g->node_w[v], or_self_loop_weight(g, v) };
        nw_ensure(s, refined->slots + 1);
        nw_collect_constrained(s, g, refined->memb, v, constraint);
        double w_own = nw_weight_to(s, own);
        part_remove(refined, g, v, w_own);
        double base = delta_insert(objective, res, refined, &ctx, own, w_own);
        uint64_t nc = 0;
        double best_rel = 0.0;
        for (uint64_t i = 0; i < s->n_touched; i++) {
            uint64_t c = s->touched[i];
            if (c == own) continue;
            if (has_max && p_weight(refined, c) + ctx.node_weight > max_w) continue;
            double rel = delta_insert(objective, res, refined, &ctx, c, nw_weight_to(s, c)) - base;
            if (rel >= 0.0) {
                cands[nc].c = c; cands[nc].rel = rel; nc++;
                if (rel > best_rel) best_rel = rel;
            }
        }
        uint64_t target = own;
        if (!select_cand(cands, nc, best_rel, randomness, r, wbuf, &target)) target = own;
        part_insert(refined, g, v, target, nw_weight_to(s, target));
    }
    free(order); free(cands); free(wbuf);
}

/* ------------------------------------------------------------------------------------------ */
/* run_once + leiden (leiden/mod.rs:168-289)                                                   */

static double run_once(const or_graph *graph, int objective, double res, rng_t *rng,
                       const or_config *cfg, const uint64_t *initial, uint64_t *labels_out,
                       uint64_t *n_clusters_out) {
    uint64_t n = graph->n;
    rng_t level_rng;
    rng_seed_u64(&level_rng, rng_u64(rng));

    const or_graph *level = graph;
    or_graph *owned = NULL;
    partition part;
    part_from_membership(&part, level, initial);
    scratch_t s;
    scratch_init(&s, n, (part.slots > n ? part.slots : n) + 1);
    double total_improvement = 0.0;
    uint64_t *node_at_level = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    for (uint64_t v = 0; v < n; v++) node_at_level[v] = v;

    for (uint64_t lvl = 0; lvl < cfg->max_levels; lvl++) {
        double imp = local_move(level, &part, objective, res, &s, &level_rng,
                                cfg->has_max_community_weight, cfg->max_community_weight);
        total_improvement += imp;

        partition collapse;
        if (cfg->refine) {
            refine(level, part.memb, objective, res, &s, &level_rng, cfg->has_max_community_weight,
                   cfg->max_community_weight, cfg->refine_randomness, &collapse);
        } else {
            part_copy(&collapse, &part);
        }
        part_renumber(&collapse);
        uint64_t ln = level->n;
        uint64_t *groups = (uint64_t *)malloc((ln ? ln : 1) * sizeof(uint64_t));
        for (uint64_t v = 0; v < ln; v++) groups[v] = collapse.memb[v];
        /* VectorGrouping::from_assignments + normalize_groups: already consecutive here */
        uint64_t ng = 0;
        for (uint64_t v = 0; v < ln; v++) if (groups[v] + 1 > ng) ng = groups[v] + 1;
        or_graph *coarse = or_graph_aggregate(level, groups, ng);

        if (coarse->n >= level->n) {
            or_graph_free(coarse);
            free(groups);
            part_free(&collapse);
            break;
        }
        uint64_t *coarse_memb = (uint64_t *)calloc(coarse->n ? coarse->n : 1, sizeof(uint64_t));
        for (uint64_t v = 0; v < ln; v++) coarse_memb[collapse.memb[v]] = part.memb[v];
        for (uint64_t v = 0; v < n; v++) node_at_level[v] = collapse.memb[node_at_level[v]];

        if (owned) or_graph_free(owned);
        owned = coarse;
        level = coarse;
        part_free(&part);
        part_from_membership(&part, level, coarse_memb);
        free(coarse_memb);
        free(groups);
        part_free(&collapse);
    }
    part_renumber(&part);
    for (uint64_t v = 0; v < n; v++) labels_out[v] = part.memb[node_at_level[v]];
    *n_clusters_out = part_community_count(&part);
    part_free(&part);
    scratch_free(&s);
    free(node_at_level);
    if (owned) or_graph_free(owned);
    return total_improvement;
}

static int config_validate(const or_config *c) {
    double r = c->resolution;
    if (!isfinite(r) || r < 0.0) return OR_INVALID_CONFIG;
    if (c->n_iterations == 0 || c->max_levels == 0 || c->patience == 0) return OR_INVALID_CONFIG;
    if (!isfinite(c->refine_randomness) || c->refine_randomness < 0.0) return OR_INVALID_CONFIG;
    if (c->has_max_community_weight &&
        (!isfinite(c->max_community_weight) || c->max_community_weight <= 0.0))
        return OR_INVALID_CONFIG;
    return OR_OK;
}

int or_leiden(const or_graph *g, const or_config *cfg, uint64_t *labels_out,
              uint64_t *n_clusters_out) {
    int st = config_validate(cfg);
    if (st) return st;
    rng_t rng;
    uint64_t seed = cfg->seed;
    if (!cfg->has_seed) {
        FILE *f = fopen("/dev/urandom", "rb");
        if (!f || fread(&seed, sizeof(seed), 1, f) != 1) seed = 0x9E3779B97F4A7C15ull;
        if (f) fclose(f);
    }
    rng_seed_u64(&rng, seed);
    uint64_t n = g->n;
    if (n == 0) { *n_clusters_out = 0; return OR_OK; }
    uint64_t *labels = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *next = (uint64_t *)malloc(n * sizeof(uint64_t));
    for (uint64_t v = 0; v < n; v++) labels[v] = v;
    uint64_t n_clusters = n, stagnant = 0;
    for (uint64_t it = 0; it < cfg->n_iterations; it++) {
        uint64_t clusters = 0;
        double imp = run_once(g, cfg->objective, cfg->resolution, &rng, cfg, labels, next, &clusters);
        memcpy(labels, next, n * sizeof(uint64_t));
        n_clusters = clusters;
        double scale = fmax(fabs(g->total_weight), 1.0);
        if (imp <= cfg->tolerance * scale) {
            stagnant++;
            if (stagnant >= cfg->patience) break;
        } else {
            stagnant = 0;
        }
    }
    memcpy(labels_out, labels, n * sizeof(uint64_t));
    *n_clusters_out = n_clusters;
    free(labels); free(next);
    return OR_OK;
}

/* One local-moving pass from a given membership (local_move.rs:39-122), for unit tests. */
double or_local_move_once(const or_graph *g, uint64_t *membership, int objective, double res,
                          uint64_t seed, int has_max, double max_w) {
    partition p;
    part_from_membership(&p, g, membership);
    scratch_t s;
    scratch_init(&s, g->n, p.slots + 1);
    rng_t r;
    rng_seed_u64(&r, seed);
    double gain = local_move(g, &p, objective, res, &s, &r, has_max, max_w);
    for (uint64_t v = 0; v < g->n; v++) membership[v] = p.memb[v];
    scratch_free(&s);
    part_free(&p);
    return gain;
}

/* ------------------------------------------------------------------------------------------ */
/* Metrics: disconnected_community_count (tests/common/mod.rs:171-194), nmi (:133-168),
 * ARI (examples/real_data.rs:50-75).                                                          */

uint64_t or_normalize_groups(const uint64_t *labels, uint64_t n, uint64_t *out) {
    /* group_count = max + 1 (from_assignments, grouping.rs:154-172) */
    uint64_t group_count = 0;
    for (uint64_t v = 0; v < n; v++)
        if (labels[v] + 1 > group_count) group_count = labels[v] + 1;
    /* sizes, then ascending new ids for the groups in use (normalize_groups, :226-254) */
    uint64_t *sizes = (uint64_t *)calloc(group_count ? group_count : 1, sizeof(uint64_t));
    for (uint64_t v = 0; v < n; v++) sizes[labels[v]]++;
    uint64_t next = 0;
    for (uint64_t c = 0; c < group_count; c++) sizes[c] = sizes[c] ? next++ : UINT64_MAX;
    for (uint64_t v = 0; v < n; v++) out[v] = sizes[labels[v]];
    free(sizes);
    return next;
}

uint64_t or_disconnected_community_count(const or_graph *g, const uint64_t *labels) {
    uint64_t n = g->n, nc = 0;
    for (uint64_t v = 0; v < n; v++) if (labels[v] + 1 > nc) nc = labels[v] + 1;
    unsigned char *seen = (unsigned char *)calloc(n ? n : 1, 1);
    uint64_t *comps = (uint64_t *)calloc(nc ? nc : 1, sizeof(uint64_t));
    uint64_t *stack = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    for (uint64_t v = 0; v < n; v++) {
        if (seen[v]) continue;
        comps[labels[v]]++;
        uint64_t top = 0;
        stack[top++] = v;
        seen[v] = 1;
        while (top) {
            uint64_t u = stack[--top];
            for (uint64_t e = g->ptrs[u]; e < g->ptrs[u + 1]; e++) {
                uint64_t w = g->nbrs[e];
                if (!seen[w] && labels[w] == labels[u]) { seen[w] = 1; stack[top++] = w; }
            }
        }
    }
    uint64_t bad = 0;
    for (uint64_t c = 0; c < nc; c++) if (comps[c] > 1) bad++;
    free(seen); free(comps); free(stack);
    return bad;
}

double or_nmi(const uint64_t *a, const uint64_t *b, uint64_t n) {
    if (n == 0) return 1.0;
    uint64_t ka = 0, kb = 0;
    for (uint64_t i = 0; i < n; i++) { if (a[i] + 1 > ka) ka = a[i] + 1; if (b[i] + 1 > kb) kb = b[i] + 1; }
    double *joint = (double *)calloc(ka * kb, sizeof(double));
    double *pa = (double *)calloc(ka, sizeof(double)), *pb = (double *)calloc(kb, sizeof(double));
    for (uint64_t i = 0; i < n; i++) { joint[a[i] * kb + b[i]] += 1.0; pa[a[i]] += 1.0; pb[b[i]] += 1.0; }
    double N = (double)n, mi = 0.0;
    for (uint64_t i = 0; i < ka; i++)
        for (uint64_t j = 0; j < kb; j++) {
            double p = joint[i * kb + j] / N;
            if (p > 0.0) mi += p * log(p / ((pa[i] / N) * (pb[j] / N)));
        }
    double ha = 0.0, hb = 0.0;
    for (uint64_t i = 0; i < ka; i++) if (pa[i] > 0.0) ha -= (pa[i] / N) * log(pa[i] / N);
    for (uint64_t j = 0; j < kb; j++) if (pb[j] > 0.0) hb -= (pb[j] / N) * log(pb[j] / N);
    free(joint); free(pa); free(pb);
    return ha + hb == 0.0 ? 1.0 : 2.0 * mi / (ha + hb);
}

typedef struct { uint64_t a, b; } pair_ab;
static int cmp_ab(const void *x, const void *y) {
    const pair_ab *p = (const pair_ab *)x, *q = (const pair_ab *)y;
    if (p->a != q->a) return (p->a > q->a) - (p->a < q->a);
    return (p->b > q->b) - (p->b < q->b);
}

double or_ari(const uint64_t *a, const uint64_t *b, uint64_t n) {
    if (n < 2) return 1.0;
    pair_ab *ps = (pair_ab *)malloc(n * sizeof(pair_ab));
    for (uint64_t i = 0; i < n; i++) { ps[i].a = a[i]; ps[i].b = b[i]; }
    qsort(ps, n, sizeof(pair_ab), cmp_ab);
    double sum_ij = 0.0;
    for (uint64_t i = 0; i < n;) {
        uint64_t j = i;
        while (j < n && ps[j].a == ps[i].a && ps[j].b == ps[i].b) j++;
        double c = (double)(j - i);
        sum_ij += c * (c - 1.0) / 2.0;
        i = j;
    }
    uint64_t ka = 0, kb = 0;
    for (uint64_t i = 0; i < n; i++) { if (a[i] + 1 > ka) ka = a[i] + 1; if (b[i] + 1 > kb) kb = b[i] + 1; }
    double *ca = (double *)calloc(ka, sizeof(double)), *cb = (double *)calloc(kb, sizeof(double));
    for (uint64_t i = 0; i < n; i++) { ca[a[i]] += 1.0; cb[b[i]] += 1.0; }
    double sa = 0.0, sb = 0.0;
    for (uint64_t i = 0; i < ka; i++) sa += ca[i] * (ca[i] - 1.0) / 2.0;
    for (uint64_t j = 0; j < kb; j++) sb += cb[j] * (cb[j] - 1.0) / 2.0;
    double total = (double)n * ((double)n - 1.0) / 2.0;
    double expected = sa * sb / total;
    double maxv = 0.5 * (sa + sb);
    free(ps); free(ca); free(cb);
    if (maxv - expected == 0.0) return 1.0;
    return (sum_ij - expected) / (maxv - expected);
}

/* ------------------------------------------------------------------------------------------ */
/* kNN (neighborhood/mod.rs). The reference's exact search is kiddo's kd-tree
 * (knn_search_large_batches :233-314: k+1 nearest including self, squared distance); this
 * restatement is a brute force with an explicit tie rule (FP64 d2, then index ascending, self
 * excluded). Floating-point contraction is disabled so the GPU FP64 rescore can match bits. */

static double d2_f64(const float *x, const float *y, uint32_t d) {
    double s = 0.0;
    for (uint32_t t = 0; t < d; t++) {
        double diff = (double)x[t] - (double)y[t];
        double sq = diff * diff;
        s = s + sq;
    }
    return s;
}

void or_knn_bruteforce(const float *data, uint64_t n, uint32_t d, uint32_t k,
                       const uint64_t *rows, uint64_t n_rows, uint32_t *idx_out,
                       double *d2_out) {
    uint64_t nq = rows ? n_rows : n;
    double *bd = (double *)malloc((k + 1) * sizeof(double));
    uint32_t *bi = (uint32_t *)malloc((k + 1) * sizeof(uint32_t));
    for (uint64_t qi = 0; qi < nq; qi++) {
        uint64_t i = rows ? rows[qi] : qi;
        uint32_t cnt = 0;
        const float *x = data + i * d;
        for (uint64_t j = 0; j < n; j++) {
            if (j == i) continue;
            double dd = d2_f64(x, data + j * d, d);
            if (cnt == k && !(dd < bd[k - 1])) continue; /* equal d2: lower index already kept */
            uint32_t pos = cnt < k ? cnt : k - 1;
            while (pos > 0 && dd < bd[pos - 1]) { bd[pos] = bd[pos - 1]; bi[pos] = bi[pos - 1]; pos--; }
            bd[pos] = dd; bi[pos] = (uint32_t)j;
            if (cnt < k) cnt++;
        }
        for (uint32_t t = 0; t < k; t++) {
            idx_out[qi * k + t] = t < cnt ? bi[t] : UNASSIGNED;
            d2_out[qi * k + t] = t < cnt ? bd[t] : INFINITY;
        }
    }
    free(bd); free(bi);
}

/* ---- exact kNN by kd-tree (the reference's n <= 250 000 path) ------------------------------
 * knn_arrayd_kiddo_gaussian (neighborhood/mod.rs:51-98) builds a kiddo 5.2.2 KdTree over the
 * points and knn_search_large_batches (neighborhood/mod.rs:233-314) queries nearest_n(k + 1)
 * for every row, self included, on min(64, threads) workers over contiguous row chunks. This
 * restatement builds a balanced tree (split dimension cycling with depth, median split, leaves of
 * at most 32 points, kiddo's default bucket size) and answers with the same exact result as
 * or_knn_bruteforce: k nearest by FP64 squared distance, ties to the lower index, self excluded.
 * A subtree is skipped only when its splitting plane is strictly farther than the current k-th
 * distance: every term of the FP64 sum is non-negative and rounding is monotone, so a computed
 * distance is never below its computed plane term and pruning never drops a point that ties. */
#include <pthread.h>

#define KD_LEAF 32u

typedef struct {
    const float *data;
    uint32_t d;
    uint32_t *perm;  /* point order: every node owns a contiguous range */
    uint32_t *lo, *hi, *left, *right, *dim;
    float *split;
    uint32_t n_nodes;
} kd_tree;

static uint32_t kd_build_rec(kd_tree *t, uint32_t lo, uint32_t hi, uint32_t depth) {
    uint32_t id = t->n_nodes++;
    t->lo[id] = lo;
    t->hi[id] = hi;
    t->left[id] = t->right[id] = UNASSIGNED;
    if (hi - lo <= KD_LEAF) return id;
    uint32_t dm = depth % t->d, mid = lo + (hi - lo) / 2;
    /* quickselect of the median on dimension dm (ties ordered by index, so the split is deterministic) */
    uint32_t a = lo, b = hi - 1;
    while (a < b) {
        uint32_t pv = t->perm[a + (b - a) / 2];
        float pvx = t->data[(uint64_t)pv * t->d + dm];
        uint32_t i = a, j = b;
        while (i <= j) {
            for (;;) {
                uint32_t q = t->perm[i];
                float x = t->data[(uint64_t)q * t->d + dm];
                if (x < pvx || (x == pvx && q < pv)) i++; else break;
            }
            for (;;) {
                uint32_t q = t->perm[j];
                float x = t->data[(uint64_t)q * t->d + dm];
                if (x > pvx || (x == pvx && q > pv)) j--; else break;
            }
            if (i <= j) {
                uint32_t tmp = t->perm[i]; t->perm[i] = t->perm[j]; t->perm[j] = tmp;
                i++;
                if (j == 0) break;
                j--;
            }
        }
        if (mid <= j) b = j;
        else if (mid >= i) a = i;
        else break;
    }
    t->dim[id] = dm;
    t->split[id] = t->data[(uint64_t)t->perm[mid] * t->d + dm];
    uint32_t l = kd_build_rec(t, lo, mid, depth + 1);
    uint32_t r = kd_build_rec(t, mid, hi, depth + 1);
    t->left[id] = l;
    t->right[id] = r;
    return id;
}

typedef struct {
    double *bd;
    uint32_t *bi;
    uint32_t cnt, k;
} kd_best;

static void kd_offer(kd_best *b, double dd, uint32_t j) {
    if (b->cnt == b->k && !(dd < b->bd[b->k - 1] || (dd == b->bd[b->k - 1] && j < b->bi[b->k - 1]))) return;
    uint32_t pos = b->cnt < b->k ? b->cnt : b->k - 1;
    while (pos > 0 && (dd < b->bd[pos - 1] || (dd == b->bd[pos - 1] && j < b->bi[pos - 1]))) {
        b->bd[pos] = b->bd[pos - 1];
        b->bi[pos] = b->bi[pos - 1];
        pos--;
    }
    b->bd[pos] = dd;
    b->bi[pos] = j;
    if (b->cnt < b->k) b->cnt++;
}

static void kd_search(const kd_tree *t, uint32_t node, const float *x, uint32_t self, kd_best *b) {
    if (t->left[node] == UNASSIGNED) {
        for (uint32_t p = t->lo[node]; p < t->hi[node]; p++) {
            uint32_t j = t->perm[p];
            if (j == self) continue;
            kd_offer(b, d2_f64(x, t->data + (uint64_t)j * t->d, t->d), j);
        }
        return;
    }
    uint32_t dm = t->dim[node];
    double diff = (double)x[dm] - (double)t->split[node];
    /* points equal to the split value can sit on either side: the near side is searched first
     * and the far side is pruned only by a strictly larger plane distance */
    uint32_t near = diff < 0.0 ? t->left[node] : t->right[node];
    uint32_t far = diff < 0.0 ? t->right[node] : t->left[node];
    kd_search(t, near, x, self, b);
    double plane = diff * diff;
    if (b->cnt < b->k || !(plane > b->bd[b->k - 1])) kd_search(t, far, x, self, b);
}

typedef struct {
    const kd_tree *t;
    const uint64_t *rows;
    uint64_t q0, q1;
    uint32_t k;
    uint32_t *idx_out;
    double *d2_out;
} kd_job;

static void *kd_worker(void *arg) {
    kd_job *jb = (kd_job *)arg;
    kd_best b;
    b.k = jb->k;
    b.bd = (double *)malloc(jb->k * sizeof(double));
    b.bi = (uint32_t *)malloc(jb->k * sizeof(uint32_t));
    for (uint64_t qi = jb->q0; qi < jb->q1; qi++) {
        uint64_t i = jb->rows ? jb->rows[qi] : qi;
        b.cnt = 0;
        kd_search(jb->t, 0, jb->t->data + i * jb->t->d, (uint32_t)i, &b);
        for (uint32_t s = 0; s < jb->k; s++) {
            jb->idx_out[qi * jb->k + s] = s < b.cnt ? b.bi[s] : UNASSIGNED;
            jb->d2_out[qi * jb->k + s] = s < b.cnt ? b.bd[s] : INFINITY;
        }
    }
    free(b.bd);
    free(b.bi);
    return NULL;
}

int or_knn_kdtree(const float *data, uint64_t n, uint32_t d, uint32_t k, const uint64_t *rows,
                  uint64_t n_rows, uint32_t threads, uint32_t *idx_out, double *d2_out,
                  double *build_seconds) {
    if (n == 0 || d == 0 || k == 0 || n > 0xFFFFFFF0ull) return 1;
    kd_tree t;
    memset(&t, 0, sizeof(t));
    t.data = data;
    t.d = d;
    uint64_t cap = 2 * (n / KD_LEAF + 1) * 2 + 4;
    t.perm = (uint32_t *)malloc(n * sizeof(uint32_t));
    t.lo = (uint32_t *)malloc(cap * sizeof(uint32_t));
    t.hi = (uint32_t *)malloc(cap * sizeof(uint32_t));
    t.left = (uint32_t *)malloc(cap * sizeof(uint32_t));
    t.right = (uint32_t *)malloc(cap * sizeof(uint32_t));
    t.dim = (uint32_t *)calloc(cap, sizeof(uint32_t));
    t.split = (float *)calloc(cap, sizeof(float));
    for (uint64_t i = 0; i < n; i++) t.perm[i] = (uint32_t)i;
    struct timespec ts0, ts1;
    clock_gettime(CLOCK_MONOTONIC, &ts0);
    kd_build_rec(&t, 0, (uint32_t)n, 0);
    clock_gettime(CLOCK_MONOTONIC, &ts1);
    if (build_seconds) *build_seconds = (double)(ts1.tv_sec - ts0.tv_sec) + 1e-9 * (double)(ts1.tv_nsec - ts0.tv_nsec);
    uint64_t nq = rows ? n_rows : n;
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64; /* min(64, threads), neighborhood/mod.rs:245 */
    uint64_t chunk = (nq + threads - 1) / threads;
    pthread_t *tid = (pthread_t *)malloc(threads * sizeof(pthread_t));
    kd_job *jobs = (kd_job *)malloc(threads * sizeof(kd_job));
    uint32_t started = 0;
    for (uint32_t w = 0; w < threads; w++) {
        uint64_t q0 = (uint64_t)w * chunk, q1 = q0 + chunk < nq ? q0 + chunk : nq;
        if (q0 >= q1) break;
        jobs[w] = (kd_job){&t, rows, q0, q1, k, idx_out, d2_out};
        if (pthread_create(&tid[w], NULL, kd_worker, &jobs[w]) != 0) {
            kd_worker(&jobs[w]);
            tid[w] = 0;
        }
        started = w + 1;
    }
    for (uint32_t w = 0; w < started; w++)
        if (tid[w]) pthread_join(tid[w], NULL);
    free(tid); free(jobs);
    free(t.perm); free(t.lo); free(t.hi); free(t.left); free(t.right); free(t.dim); free(t.split);
    return 0;
}

typedef struct { uint32_t r, c; float w; } trip_t;
static int cmp_trip(const void *x, const void *y) {
    const trip_t *p = (const trip_t *)x, *q = (const trip_t *)y;
    if (p->r != q->r) return (p->r > q->r) - (p->r < q->r);
    return (p->c > q->c) - (p->c < q->c);
}

static int cmp_double(const void *x, const void *y) {
    double a = *(const double *)x, b = *(const double *)y;
    return (a > b) - (a < b);
}

/* Gaussian weight (neighborhood/mod.rs:142-153), evaluated in f64 and stored as f32. */
static float gauss_w(double si, double sj, double d2) {
    double si2 = si * si, sj2 = sj * sj;
    double den = si2 + sj2;
    if (!(den > 0.0)) return 0.0f;
    double num = 2.0 * si * sj;
    double w = sqrt(num / den) * exp(-d2 / den);
    return (float)w;
}

uint64_t or_knn_gaussian(const uint32_t *idx, const double *d2, uint64_t n, uint32_t k, int mode,
                         uint64_t *ptrs_out, uint32_t *nbrs_out, float *wts_out) {
    double *sigma = (double *)malloc((n ? n : 1) * sizeof(double));
    double *tmp = (double *)malloc((k + 1) * sizeof(double));
    for (uint64_t i = 0; i < n; i++) {
        if (mode == 0) {
            double dk = d2[i * k + (k - 1)];
            sigma[i] = isfinite(dk) ? sqrt(dk) : 1.0;
        } else {
            /* sigma_i = sqrt(upper median of the nonzero d2 in the k+1 list incl. self) (:105-124) */
            uint32_t cnt = 0;
            for (uint32_t t = 0; t < k; t++) if (d2[i * k + t] > 0.0 && isfinite(d2[i * k + t])) tmp[cnt++] = d2[i * k + t];
            if (cnt == 0) sigma[i] = 1.0;
            else { qsort(tmp, cnt, sizeof(double), cmp_double); sigma[i] = sqrt(tmp[cnt / 2]); }
        }
    }
    trip_t *tr = (trip_t *)malloc((2 * n * k + 1) * sizeof(trip_t));
    uint64_t nt = 0;
    for (uint64_t i = 0; i < n; i++) {
        for (uint32_t t = 0; t < k; t++) {
            uint32_t j = idx[i * k + t];
            if (j == UNASSIGNED) continue;
            float w = gauss_w(sigma[i], sigma[j], d2[i * k + t]);
            if (mode == 1) {
                if (!((uint64_t)i <= (uint64_t)j)) continue;
                if (!(w > 1e-14f)) continue;
                tr[nt++] = (trip_t){ (uint32_t)i, j, w };
                if ((uint64_t)j != i) tr[nt++] = (trip_t){ j, (uint32_t)i, w };
            } else {
                /* union: emit both directions, duplicates collapse below (weights identical) */
                uint32_t a = (uint32_t)i < j ? (uint32_t)i : j, b = (uint32_t)i < j ? j : (uint32_t)i;
                float wc = gauss_w(sigma[a], sigma[b], d2[i * k + t]);
                if (!(wc > 0.0f)) continue;
                tr[nt++] = (trip_t){ a, b, wc };
                tr[nt++] = (trip_t){ b, a, wc };
            }
        }
    }
    qsort(tr, nt, sizeof(trip_t), cmp_trip);
    uint64_t out = 0;
    for (uint64_t v = 0; v <= n; v++) ptrs_out[v] = 0;
    for (uint64_t t = 0; t < nt; t++) {
        if (t > 0 && tr[t].r == tr[t - 1].r && tr[t].c == tr[t - 1].c) {
            /* the same pair from both endpoints' lists: max of equal weights */
            if (tr[t].w > wts_out[out - 1]) wts_out[out - 1] = tr[t].w;
            continue;
        }
        nbrs_out[out] = tr[t].c;
        wts_out[out] = tr[t].w;
        ptrs_out[tr[t].r + 1]++;
        out++;
    }
    for (uint64_t v = 0; v < n; v++) ptrs_out[v + 1] += ptrs_out[v];
    free(sigma); free(tmp); free(tr);
    return out;
}
