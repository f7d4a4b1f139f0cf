/*
 * scgpu.h -- C ABI of the MI355X (gfx950) Leiden + kNN-graph library.
 *
 * Drop-in boundary for SingleRust/single-clustering v0.7.0. Each entry point names the Rust
 * item it replaces (file:line in that crate). Plain pointers and sizes only; no exceptions cross
 * this boundary; every fallible call returns an sc_status and fills an optional sc_error.
 *
 * Graph handles are immutable after construction and device-resident; concurrent sc_leiden
 * calls on one handle are safe (each call owns its workspace and stream).
 *
 * Size limits: fewer than 2^31 nodes and fewer than 2^32 - 1 stored adjacency entries per
 * graph (for a batch: per union of all jobs; for kNN graphs: 2 * n * k). Larger inputs return
 * SC_TOO_MANY_NODES or SC_INVALID_ARGUMENT before any device work.
 */
#ifndef SCGPU_H
#define SCGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One code per ClusteringError variant, in declaration order (src/error.rs:7-68), plus
 * device / argument / memory failures that the CPU crate has no counterpart for. */
typedef enum sc_status {
    SC_OK = 0,
    SC_NODE_INDEX_OUT_OF_RANGE = 1,      /* ClusteringError::NodeIndexOutOfRange      */
    SC_NON_FINITE_WEIGHT = 2,            /* ClusteringError::NonFiniteWeight          */
    SC_NEGATIVE_WEIGHT = 3,              /* ClusteringError::NegativeWeight           */
    SC_INVALID_NODE_WEIGHT = 4,          /* ClusteringError::InvalidNodeWeight        */
    SC_NODE_WEIGHT_LENGTH_MISMATCH = 5,  /* ClusteringError::NodeWeightLengthMismatch */
    SC_MEMBERSHIP_LENGTH_MISMATCH = 6,   /* ClusteringError::MembershipLengthMismatch */
    SC_INVALID_CONFIG = 7,               /* ClusteringError::InvalidConfig            */
    SC_TOO_MANY_NODES = 8,               /* ClusteringError::TooManyNodes             */
    SC_INVALID_CSR = 9,                  /* ClusteringError::InvalidCsr               */
    SC_ASYMMETRIC_GRAPH = 10,            /* ClusteringError::AsymmetricGraph          */
    SC_DEVICE_ERROR = 100,
    SC_INVALID_ARGUMENT = 101,
    SC_OUT_OF_MEMORY = 102
} sc_status;

/* Payload of the failing variant (error.rs:7-68); unused fields are zero. msg holds the
 * variant's Display text (error.rs:70-113) or the device error string. */
typedef struct sc_error {
    int32_t code;
    uint64_t node;
    uint64_t n_nodes;
    uint64_t edge_u;
    uint64_t edge_v;
    double weight;
    uint64_t got;
    uint64_t expected;
    double degree_sum;
    double expected_sum;
    char msg[256];
} sc_error;

/* ObjectiveKind (leiden/objective.rs:188-199). */
typedef enum sc_objective { SC_RB = 0, SC_CPM = 1 } sc_objective;

/* LeidenConfig, field by field (leiden/mod.rs:49-90). Option<T> fields become has_x + x. */
typedef struct sc_leiden_config {
    int32_t objective;            /* sc_objective */
    double resolution;
    int32_t has_seed;             /* 0: seed from the OS (ChaCha8Rng::from_os_rng) */
    uint64_t seed;
    uint64_t n_iterations;
    uint64_t max_levels;
    double tolerance;
    int32_t refine;
    double refine_randomness;
    uint64_t patience;
    int32_t has_max_community_weight;
    double max_community_weight;
} sc_leiden_config;

typedef struct sc_graph sc_graph;

/* Per-call counters of the last sc_leiden on this thread (no reference counterpart; used by
 * the benchmark to compute algorithmic bytes and by tests to inspect convergence). */
typedef struct sc_leiden_stats {
    uint64_t passes;
    uint64_t levels;
    uint64_t move_iterations;
    uint64_t refine_rounds;
    uint64_t vertex_visits;        /* local-move + refine vertex evaluations */
    uint64_t edge_visits;          /* adjacency entries read by those evaluations */
    uint64_t moves;
    double algorithmic_bytes;      /* minimum unique HBM bytes of every executed pass */
    double move_kernel_ms;         /* summed device time of the local-move decide kernels */
    double move_kernel_bytes;      /* algorithmic bytes of those kernels */
    uint64_t move_kernel_launches;
    double total_ms;
    /* host wall time per phase: [0] degree buckets, [1] local moving, [2] refinement,
     * [3] renumbering + coarse membership, [4] aggregation, [5] quality evaluation */
    double phase_ms[6];
} sc_leiden_stats;

/* kNN options: sigma rule and symmetrization (neighborhood/mod.rs:104-168). */
typedef struct sc_knn_opts {
    int32_t precision;   /* 0 = BF16 MFMA candidates + exact rescore, 1 = FP32 candidates + exact rescore */
    int32_t mode;        /* 0 = sigma: k-th neighbour distance, union (max) symmetrization;
                            1 = reference-compatible: sigma = upper median, lower-row emission */
} sc_knn_opts;

typedef struct sc_knn_stats {
    uint64_t n;
    uint64_t candidates_rescored;
    uint64_t rows_fallback;        /* rows whose certified candidate set overflowed */
    double tile_kernel_ms;
    double total_ms;
    double flops;                  /* 2 * pairs_computed * d (d unpadded) */
    double pairs_computed;         /* point pairs the distance tiles computed, padding included:
                                      the brute force's npad^2, or fewer with block pruning */
} sc_knn_stats;

/* ---- library ------------------------------------------------------------------------- */
const char *sc_version(void);
/* Releases cached device memory; optional. */
void sc_shutdown(void);
/* No counterpart in the reference (diagnostic). Device bytes held by the calling thread's
 * allocator: blocks in use (live) and blocks cached for reuse. A block allocated on one thread
 * and released on another (a graph built on one thread and freed on another) leaves the
 * allocating thread's pool and is returned to the device at once. Either pointer may be NULL. */
void sc_pool_bytes(uint64_t *live_out, uint64_t *cached_out);
/* Records device time of the hot kernels into sc_leiden_stats (adds HIP events around each
 * local-move decide launch). Also switched on by the environment variable SCGPU_PROFILE=1. */
void sc_set_profiling(int32_t enable);
/* No counterpart in the reference. Selects how local moving and refinement are launched:
 * 1 (default) = one cooperative persistent launch per phase of a level, with grid barriers
 * between sub-sweeps; 0 = one launch per kernel of each sub-sweep; 2 = mode 0's launches of one
 * iteration captured into a hipGraph once per level and phase, and the graph replayed per
 * iteration (the iteration index lives in device memory). All give identical labels. The
 * environment variable SCGPU_SWEEP_MODE=0/1/2 (or SCGPU_PERSISTENT=0) selects it at start-up.
 * Returns SC_INVALID_ARGUMENT for other values. */
sc_status sc_set_sweep_mode(int32_t mode);
/* No counterpart in the reference. Online autotuner of launch parameters that never change
 * results (workgroups per persistent sweep phase, kNN column splits): the first calls for a
 * shape try each candidate once, later calls use the fastest. The table is process-wide and
 * mutex-guarded; SCGPU_AUTOTUNE=0 disables it. The report has one line per tuned shape (best
 * time per candidate in ms, then the chosen index); returns its length. */
size_t sc_autotune_report(char *buf, size_t len);
void sc_autotune_reset(void);

/* LeidenConfig::default (leiden/mod.rs:92-106). */
sc_leiden_config sc_leiden_config_default(void);

/* ---- graph ----------------------------------------------------------------------------- */
/* CSRNetwork::from_csr_parts (src/network/csr_network.rs:185-256). Host arrays are borrowed
 * for the call and copied to HBM; rows may be unsorted, duplicates are summed, zero sums are
 * dropped; the degree-sum check (:712-723) runs. node_weights may be NULL (unit weights). */
sc_status sc_graph_from_csr_parts(const uint64_t *node_ptrs, uint64_t n_nodes,
                                  const uint32_t *neighbors, const float *weights, uint64_t nnz,
                                  const double *node_weights, sc_graph **out, sc_error *err);

/* Zero-copy analogue of from_csr_parts for device-resident buffers (ownership is not taken;
 * the arrays are read during the call, and the graph keeps its own compacted copy). */
sc_status sc_graph_from_csr_parts_device(const uint64_t *d_node_ptrs, uint64_t n_nodes,
                                         const uint32_t *d_neighbors, const float *d_weights,
                                         uint64_t nnz, const double *d_node_weights,
                                         sc_graph **out, sc_error *err);

/* The real-data example's reader (examples/real_data.rs:22-44) + from_csr_parts (:129-136):
 * <dir>/indptr.bin (u64), indices.bin (u32) and data.bin (f32), raw little-endian, the
 * connectivities matrix as tools/export_h5ad.py:106-108 writes it; unit node weights. A file
 * that cannot be read is SC_INVALID_ARGUMENT, file lengths that do not fit together are
 * SC_INVALID_CSR (both before any device work); the arrays are then checked as
 * sc_graph_from_csr_parts checks them. */
sc_status sc_graph_from_csr_export(const char *dir, sc_graph **out, sc_error *err);

/* CSRNetwork::from_edges_with_node_weights (csr_network.rs:120-167). Each undirected edge once;
 * (v, v, w) is a self-loop. node_weights may be NULL. */
sc_status sc_graph_from_edges(uint64_t n_nodes, const uint64_t *edge_from, const uint64_t *edge_to,
                              const double *edge_weight, uint64_t n_edges,
                              const double *node_weights, sc_graph **out, sc_error *err);

/* CSRNetwork::from_csr_matrix / from_csr_matrix_with_node_weights (csr_network.rs:360-391): a
 * sparse adjacency matrix (n_rows x n_rows, CSR arrays: row_ptrs[n_rows + 1], usize column
 * indices and f64 values, nnz entries) read as symmetric: only entries with row <= col, zeros
 * skipped, a diagonal entry is a self-loop; then from_edges_with_node_weights. node_weights may
 * be NULL (unit weights). Row offsets that do not start at 0, end at nnz or decrease are
 * SC_INVALID_CSR (a CsrMatrix cannot hold them). */
sc_status sc_graph_from_csr_matrix(uint64_t n_rows, const uint64_t *row_ptrs, const uint64_t *col_indices,
                                   const double *values, uint64_t nnz, const double *node_weights,
                                   sc_graph **out, sc_error *err);

/* CSRNetwork::to_csr_matrix (csr_network.rs:610-618): row_ptrs[n + 1], col_indices[nnz] (usize)
 * and values[nnz] (f64) of the stored adjacency. Any output pointer may be NULL. */
sc_status sc_graph_to_csr_matrix(const sc_graph *g, uint64_t *row_ptrs, uint64_t *col_indices, double *values,
                                 sc_error *err);

void sc_graph_free(sc_graph *g);

/* Accessors (csr_network.rs:407-508). */
uint64_t sc_graph_node_count(const sc_graph *g);
uint64_t sc_graph_edge_count(const sc_graph *g);
uint64_t sc_graph_nnz(const sc_graph *g);
double sc_graph_total_weight(const sc_graph *g);
int32_t sc_graph_has_self_loops(const sc_graph *g);

/* CSRNetwork::to_csr_parts / node_weights / strengths (csr_network.rs:484-495, :428-430):
 * copies the device arrays back. Any output pointer may be NULL. */
sc_status sc_graph_to_csr_parts(const sc_graph *g, uint64_t *node_ptrs, uint32_t *neighbors,
                                float *weights, double *node_weights, double *strengths,
                                sc_error *err);

/* CSRNetwork::validate_symmetry (csr_network.rs:627-646). */
sc_status sc_graph_validate_symmetry(const sc_graph *g, sc_error *err);

/* CSRNetwork::aggregate (csr_network.rs:516-608). groups: consecutive ids in [0, n_groups)
 * for every node (VectorGrouping::from_assignments after normalization, grouping.rs:154-172). */
sc_status sc_graph_aggregate(const sc_graph *g, const uint64_t *groups, sc_graph **out,
                             sc_error *err);

/* Partition::renumber_into_grouping / VectorGrouping::from_assignments (partition.rs:265-306,
 * grouping.rs:154-172, 226-254): the ids in use compacted to 0..n_groups in ascending order, on
 * the device with the compaction every level of sc_leiden uses. labels below 2^32 - 1. */
sc_status sc_renumber(const uint64_t *labels, uint64_t n, uint64_t *out, uint64_t *n_groups_out,
                      sc_error *err);

/* ---- clustering ------------------------------------------------------------------------ */
/* leiden (leiden/mod.rs:168-204). labels_out: n entries, consecutive from 0 (Clustering). */
sc_status sc_leiden(const sc_graph *g, const sc_leiden_config *cfg, uint64_t *labels_out,
                    uint64_t *n_clusters_out, sc_error *err);

/* Batched leiden: every graph at every resolution, as one run. Equivalent to the loop
 *   for g in graphs, for r in resolutions: leiden(g, cfg.with_resolution(r))
 * (leiden/mod.rs:168-204 once per job; LeidenConfig::with_resolution, mod.rs:108-112) with each
 * job's labels equal to that standalone call's. Job j = g * n_resolutions + r; labels_out holds
 * the jobs' labels one after another (job j's n_g entries, consecutive from 0), n_clusters_out
 * one count per job. The jobs run as one disjoint-union graph, so every local-moving and
 * refinement phase of a level is one persistent launch over all jobs. */
sc_status sc_leiden_batch(const sc_graph *const *graphs, uint64_t n_graphs, const double *resolutions,
                          uint64_t n_resolutions, const sc_leiden_config *cfg, uint64_t *labels_out,
                          uint64_t *n_clusters_out, sc_error *err);

/* ---- learned active-vertex filter (no counterpart in the reference) --------------------- */
/* An 8 -> 64 -> 64 -> 1 perceptron (ReLU) predicts whether an active vertex will move when
 * local moving decides it; parameters: sc_mlp_param_count() floats in the order
 * W1[8][64], b1[64], W2[64][64], b2[64], w3[64], b3. Features (each about [0, 1]): degree, the
 * own-community weight fraction and best-gain margin of the vertex's last decision on this
 * level, neighbour moves since that decision, whether it has decided on this level yet,
 * iteration, level, own community size. */
uint32_t sc_mlp_param_count(void);
/* He-uniform weights, zero biases, from a seed. */
void sc_mlp_init(uint64_t seed, float *w);
/* Mean loss (weighted BCE on the logit, positives weighted by pos_weight) and its gradient
 * over n samples X[n][8], y[n] in {0, 1}; the HIP backward pass. */
sc_status sc_mlp_grad(const float *w, const float *X, const uint8_t *y, uint64_t n, float pos_weight,
                      float *grad_out, double *loss_out, sc_error *err);
/* Minibatch Adam with HIP forward/backward kernels; loss_hist (nullable) gets each epoch's loss. */
sc_status sc_mlp_train(const float *X, const uint8_t *y, uint64_t n, uint32_t epochs, uint32_t batch, float lr,
                       float pos_weight, uint64_t seed, float *w_inout, double *loss_hist, sc_error *err);
/* Logits of n samples. bf16 == 0: FP32 weights (the training arithmetic); bf16 != 0: the
 * filter's inference, BF16-rounded weights on the wave-level MFMA path. */
sc_status sc_mlp_predict(const float *w, const float *X, uint64_t n, int32_t bf16, float *out, sc_error *err);

/* Runs leiden() and records up to cap local-moving decisions as training samples
 * (features x_out[cap][8], moved y_out[cap]); n_out = samples written. */
sc_status sc_leiden_trace(const sc_graph *g, const sc_leiden_config *cfg, uint64_t cap, float *x_out,
                          uint8_t *y_out, uint64_t *n_out, sc_error *err);

typedef struct {
    uint64_t skipped;           /* filter decisions "stays" (vertex-steps not decided) */
    uint64_t judged;            /* vertices the first plain (judge) iteration decided */
    uint64_t judge_moves;       /* of them, vertices that moved */
    uint64_t skipped_at_judge;  /* vertices whose last evaluation before the judge was a skip */
    uint64_t skipped_moved;     /* of them, vertices the judge iteration moved: the filter's errors */
    double skip_error_rate;     /* skipped_moved / skipped_at_judge */
} sc_filter_stats;

/* leiden() whose local moving skips active vertices the model gives a move probability below
 * threshold: a skipped vertex leaves the active set as if it had decided to stay (a neighbour's
 * move re-activates it), until nothing moves. Then the judge: judge == 1 activates every vertex
 * (a full plain sweep), judge == 2 only the vertices whose last evaluation was a skip, and plain
 * iterations run to convergence, so every level ends in a converged plain local-moving state;
 * stats counts the skipped vertices the judge moved. judge == 0: no judge runs and only
 * stats->skipped is counted. */
sc_status sc_leiden_filtered(const sc_graph *g, const sc_leiden_config *cfg, const float *weights,
                             double threshold, int32_t judge, uint64_t *labels_out, uint64_t *n_clusters_out,
                             sc_filter_stats *stats, sc_error *err);

/* Counters of the last sc_leiden on the calling thread. */
sc_status sc_leiden_last_stats(sc_leiden_stats *out);

/* modularity (leiden/objective.rs:221-244). Infallible in the crate; returns NaN on a device
 * error or a label vector shorter than the graph cannot be detected, so callers pass n labels. */
double sc_modularity(const sc_graph *g, const uint64_t *labels, double resolution);

/* Rb::quality / Cpm::quality of Partition::from_membership(labels)
 * (objective.rs:108-119, :171-178; partition.rs:75-112). */
sc_status sc_quality(const sc_graph *g, const uint64_t *labels, int32_t objective,
                     double resolution, double *out, sc_error *err);

/* No counterpart in the reference: for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc). Same result as sc_quality, bit for bit (the
 * same kernels and summation order), but d_labels is a device pointer (n entries) and every
 * working buffer lives in one file-scope device array of the library (~32 MiB per GPU, held
 * until the process ends): the call makes no hipMalloc / hipFree and creates no stream. It runs
 * on the null stream and synchronizes before returning. Limit: labels below 2^20 (1,048,576
 * communities); a larger label returns SC_INVALID_ARGUMENT (payload got = the label) before the
 * per-community work. Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (from several host threads) race and must be serialized by the caller. */
sc_status sc_quality_noalloc(const sc_graph *g, const uint64_t *d_labels, int32_t objective,
                             double resolution, double *out, sc_error *err);

/* disconnected_community_count (tests/common/mod.rs:171-194): communities with more than one
 * connected component using same-label edges only. */
sc_status sc_disconnected_community_count(const sc_graph *g, const uint64_t *labels,
                                          uint64_t *out, sc_error *err);

/* ---- kNN ------------------------------------------------------------------------------- */
/* Exact k nearest neighbours of every row (self excluded, FP64 distance, ties to the lower
 * index). Replaces the search of knn_arrayd_adaptive (neighborhood/mod.rs:592-625), whose
 * kd-tree path is exact and whose HNSW path is approximate. data: n x d row-major f32.
 * idx_out: n*k, d2_out: n*k squared distances (either may be NULL). */
sc_status sc_knn(const float *data, uint64_t n, uint32_t d, uint32_t k, const sc_knn_opts *opts,
                 uint32_t *idx_out, double *d2_out, sc_knn_stats *stats, sc_error *err);

/* kNN + Gaussian connectivity + symmetrization + ingest, device-resident end to end: the
 * connectivities of knn_arrayd_kiddo_gaussian (neighborhood/mod.rs:104-181) fed straight into
 * CSRNetwork::from_csr_parts. */
sc_status sc_knn_graph(const float *data, uint64_t n, uint32_t d, uint32_t k,
                       const sc_knn_opts *opts, sc_graph **out, sc_knn_stats *stats,
                       sc_error *err);

/* Same as sc_knn_graph for points already resident in device memory (d_data: n x d f32). */
sc_status sc_knn_graph_device(const float *d_data, uint64_t n, uint32_t d, uint32_t k,
                              const sc_knn_opts *opts, sc_graph **out, sc_knn_stats *stats,
                              sc_error *err);

/* Gaussian connectivity of a given kNN result (neighborhood/mod.rs:104-168), as a full
 * symmetric CSR with sorted rows. Capacity of neighbors/weights: 2*n*k. */
sc_status sc_knn_gaussian_csr(const uint32_t *idx, const double *d2, uint64_t n, uint32_t k,
                              int32_t mode, uint64_t *node_ptrs, uint32_t *neighbors,
                              float *weights, uint64_t *nnz_out, sc_error *err);

#ifdef __cplusplus
}
#endif
#endif
