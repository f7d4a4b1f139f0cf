/*
 * oracle.h -- CPU restatement of single-clustering's Leiden core and kNN Gaussian builder.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (libscgpu) links, loads or calls this
 * code. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only
 * as the checker / the timed CPU baseline.
 *
 * Every function restates a file:line of the reference (SingleRust/single-clustering v0.7.0),
 * cited on the definition in oracle.c. Pinning status of each part:
 *   - CSR ingest / finish / aggregate: pinned by the reference's unit KATs
 *     (csr_network.rs:748-995) and by brute-force definitions in tests/test_oracle_graph.py.
 *   - Rb / Cpm quality and modularity: pinned by the committed leidenalg fixture
 *     (160 memberships, quality x2 to 1e-6, igraph modularity to 1e-7) and the karate optimum.
 *   - local_move / refine / leiden: pinned by the reference's invariant, known-answer,
 *     exhaustive-optimality and fixture-statistics tests, ported to tests/test_oracle_leiden.py.
 *   - ChaCha8 block function: pinned by RFC 7539 / zero-key keystream vectors. The
 *     rand_core seed_from_u64 expansion, rand 0.9 shuffle and random::<f64>() streams are
 *     restated from their published algorithms but are "parity unpinned": no Rust toolchain
 *     exists here to produce their vectors, so exact equality of labels with the Rust crate
 *     for a given seed is NOT claimed.
 *   - Brute-force kNN (FP64) and the Gaussian connectivity: pinned by a scipy/numpy
 *     brute-force check and the closed form on tiny inputs.
 */
#ifndef SC_ORACLE_H
#define SC_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: same order as ClusteringError's variants (error.rs:7-68), 0 = ok. */
enum {
    OR_OK = 0,
    OR_NODE_INDEX_OUT_OF_RANGE = 1,
    OR_NON_FINITE_WEIGHT = 2,
    OR_NEGATIVE_WEIGHT = 3,
    OR_INVALID_NODE_WEIGHT = 4,
    OR_NODE_WEIGHT_LENGTH_MISMATCH = 5,
    OR_MEMBERSHIP_LENGTH_MISMATCH = 6,
    OR_INVALID_CONFIG = 7,
    OR_TOO_MANY_NODES = 8,
    OR_INVALID_CSR = 9,
    OR_ASYMMETRIC_GRAPH = 10,
    OR_ALLOC = 100
};

typedef struct or_graph {
    uint64_t n;
    uint64_t *ptrs;     /* n + 1 */
    uint32_t *nbrs;     /* ptrs[n] */
    float *wts;         /* ptrs[n] */
    double *node_w;     /* n */
    double *strength;   /* n */
    double total_weight;
    uint64_t edge_count;
    int any_self_loops;
} or_graph;

typedef struct or_edge {
    uint64_t from, to;
    double w;
} or_edge;

typedef struct or_config {
    int objective;            /* 0 = Rb, 1 = Cpm */
    double resolution;
    int has_seed;
    uint64_t seed;
    uint64_t n_iterations;
    uint64_t max_levels;
    double tolerance;
    int refine;
    double refine_randomness;
    uint64_t patience;
    int has_max_community_weight;
    double max_community_weight;
} or_config;

void or_config_default(or_config *cfg);

int or_graph_from_csr_parts(uint64_t n, const uint64_t *ptrs, const uint32_t *nbrs,
                            const float *wts, uint64_t nnz, const double *node_w,
                            or_graph **out);
int or_graph_from_edges(uint64_t n, const or_edge *edges, uint64_t n_edges,
                        const double *node_w, or_graph **out);
void or_graph_free(or_graph *g);
int or_graph_validate_symmetry(const or_graph *g);
double or_self_loop_weight(const or_graph *g, uint64_t v);

/* groups: consecutive group ids in [0, n_groups) for every node. */
or_graph *or_graph_aggregate(const or_graph *g, const uint64_t *groups, uint64_t n_groups);

double or_modularity(const or_graph *g, const uint64_t *labels, double resolution);
/* Rb::quality / Cpm::quality of Partition::from_membership(labels). */
double or_quality(const or_graph *g, const uint64_t *labels, int objective, double resolution);

int or_leiden(const or_graph *g, const or_config *cfg, uint64_t *labels_out,
              uint64_t *n_clusters_out);

/* Single-level building blocks, for tests: local moving from a given membership. */
double or_local_move_once(const or_graph *g, uint64_t *membership, int objective,
                          double resolution, uint64_t seed, int has_max, double max_w);

uint64_t or_disconnected_community_count(const or_graph *g, const uint64_t *labels);
/* VectorGrouping::from_assignments + normalize_groups (grouping.rs:154-172, 226-254), which
 * Partition::renumber_into_grouping (partition.rs:265-306) feeds: ids in use compacted to
 * 0..count in ascending order. Returns count. */
uint64_t or_normalize_groups(const uint64_t *labels, uint64_t n, uint64_t *out);
double or_nmi(const uint64_t *a, const uint64_t *b, uint64_t n);
double or_ari(const uint64_t *a, const uint64_t *b, uint64_t n);

/* RNG pieces (exposed for the known-answer tests of the block function). */
void or_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, int rounds,
                     uint32_t out[16]);
void or_chacha8_seed_u64_stream(uint64_t seed, uint32_t *out, uint64_t count);

/* Exact kNN: k nearest by FP64 squared distance, ties to the lower index, self excluded.
 * rows: NULL for all rows, or a list of n_rows query rows. */
void or_knn_bruteforce(const float *data, uint64_t n, uint32_t d, uint32_t k,
                       const uint64_t *rows, uint64_t n_rows, uint32_t *idx_out,
                       double *d2_out);

/* Exact kNN through a kd-tree on `threads` workers (the reference's kd-tree path,
 * neighborhood/mod.rs:51-98, 233-314): same result as or_knn_bruteforce. build_seconds gets the
 * tree construction time (nullable). Returns nonzero on invalid sizes. */
int or_knn_kdtree(const float *data, uint64_t n, uint32_t d, uint32_t k, const uint64_t *rows,
                  uint64_t n_rows, uint32_t threads, uint32_t *idx_out, double *d2_out,
                  double *build_seconds);

/* Gaussian connectivity over a kNN result (k neighbours per row, self excluded).
 * mode 0: sigma_i = distance to the k-th neighbour, union (max) symmetrization.
 * mode 1: reference-compatible: sigma_i = upper median of nonzero distances over the k+1
 *         list that includes self at distance 0, pair {i<j} kept iff j in knn(i).
 * Output: full symmetric CSR (rows sorted) written to caller buffers sized 2*n*k;
 * returns nnz. */
uint64_t or_knn_gaussian(const uint32_t *idx, const double *d2, uint64_t n, uint32_t k, int mode,
                         uint64_t *ptrs_out, uint32_t *nbrs_out, float *wts_out);

#ifdef __cplusplus
}
#endif
#endif
