// graph.hip -- device-resident CSRNetwork: ingest, finish, accessors, symmetry validation.
//
// Restates src/network/csr_network.rs (from_csr_parts :185-256, from_edges :120-167,
// finish :262-351, check_degree_sum :712-723, validate_symmetry :627-646) as HIP kernels.
#include "common.hpp"

#include <cstring>

#include <algorithm>
#include <cmath>

namespace scgpu {

constexpr uint64_t NONE64 = ~0ull;

// ---- validation (from_csr_parts :191-253) --------------------------------------------------
__global__ void k_check_ptrs(const uint64_t *ptrs, uint64_t n, unsigned long long *first_bad) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x)
        if (ptrs[v] > ptrs[v + 1]) atomicMin(first_bad, (unsigned long long)v);
}

__global__ void k_check_entries(const uint32_t *nbrs, const float *wts, uint64_t nnz, uint64_t n,
                                unsigned long long *first_bad) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nnz;
         e += (uint64_t)gridDim.x * blockDim.x) {
        float w = wts[e];
        if ((uint64_t)nbrs[e] >= n || !isfinite(w) || w < 0.0f)
            atomicMin(first_bad, (unsigned long long)e);
    }
}

__global__ void k_check_node_w(const double *nw, uint64_t n, unsigned long long *first_bad) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        double x = nw[v];
        if (!isfinite(x) || x < 0.0) atomicMin(first_bad, (unsigned long long)v);
    }
}

__global__ void k_fill_f64(double *p, uint64_t n, double x) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = x;
}

static uint64_t first_bad(void (*launch)(unsigned long long *, hipStream_t), hipStream_t s) {
    DevBuf<unsigned long long> d(1);
    unsigned long long init = NONE64, out = 0;
    HIP_OK(hipMemcpyAsync(d.get(), &init, sizeof(init), hipMemcpyHostToDevice, s));
    launch(d.get(), s);
    to_host(&out, d.get(), 1, s);
    sync_stream(s);
    return out;
}

// Validates device-resident CSR input in the reference's order of checks and throws the
// matching ClusteringError.
static void validate_csr_device(const uint64_t *d_ptrs, uint64_t n, const uint32_t *d_nbrs,
                                const float *d_wts, uint64_t nnz, const double *d_nw,
                                hipStream_t s) {
    check_node_count(n);
    uint64_t last = 0;
    to_host(&last, d_ptrs + n, 1, s);
    sync_stream(s);
    if (last != nnz)
        throw ScError(SC_INVALID_CSR, "malformed CSR input: node_ptrs ends at " + std::to_string(last) +
                                          " but there are " + std::to_string(nnz) + " entries");
    unsigned block = 256;
    uint64_t bad = 0;
    {
        DevBuf<unsigned long long> d(1);
        unsigned long long init = NONE64;
        HIP_OK(hipMemcpyAsync(d.get(), &init, sizeof(init), hipMemcpyHostToDevice, s));
        if (n) hipLaunchKernelGGL(k_check_ptrs, dim3(grid_for(n, block, 4096)), dim3(block), 0, s, d_ptrs, n, d.get());
        KERNEL_OK();
        to_host(reinterpret_cast<unsigned long long *>(&bad), d.get(), 1, s);
        sync_stream(s);
    }
    if (bad != NONE64)
        throw ScError(SC_INVALID_CSR, "malformed CSR input: node_ptrs is not non-decreasing at " +
                                          std::to_string(bad));
    {
        DevBuf<unsigned long long> d(1);
        unsigned long long init = NONE64;
        HIP_OK(hipMemcpyAsync(d.get(), &init, sizeof(init), hipMemcpyHostToDevice, s));
        if (nnz) hipLaunchKernelGGL(k_check_entries, dim3(grid_for(nnz, block, 8192)), dim3(block), 0, s, d_nbrs, d_wts, nnz, n, d.get());
        KERNEL_OK();
        to_host(reinterpret_cast<unsigned long long *>(&bad), d.get(), 1, s);
        sync_stream(s);
    }
    if (bad != NONE64) {
        uint32_t u = 0;
        float w = 0.0f;
        to_host(&u, d_nbrs + bad, 1, s);
        to_host(&w, d_wts + bad, 1, s);
        sync_stream(s);
        // the reference reports edge (0, u) for these, not the real row (csr_network.rs:228-238)
        if ((uint64_t)u >= n) {
            ScError e(SC_NODE_INDEX_OUT_OF_RANGE, "node index " + std::to_string(u) +
                                                      " is out of range for a graph with " +
                                                      std::to_string(n) + " nodes");
            e.payload.node = u;
            e.payload.n_nodes = n;
            throw e;
        }
        if (!std::isfinite(w)) {
            ScError e(SC_NON_FINITE_WEIGHT, "edge (0, " + std::to_string(u) + ") has a non-finite weight");
            e.payload.edge_u = 0;
            e.payload.edge_v = u;
            throw e;
        }
        char buf[160];
        std::snprintf(buf, sizeof(buf), "edge (0, %u) has negative weight %g; weights must be non-negative", u, (double)w);
        ScError e(SC_NEGATIVE_WEIGHT, buf);
        e.payload.edge_u = 0;
        e.payload.edge_v = u;
        e.payload.weight = w;
        throw e;
    }
    if (d_nw) {
        DevBuf<unsigned long long> d(1);
        unsigned long long init = NONE64;
        HIP_OK(hipMemcpyAsync(d.get(), &init, sizeof(init), hipMemcpyHostToDevice, s));
        if (n) hipLaunchKernelGGL(k_check_node_w, dim3(grid_for(n, block, 4096)), dim3(block), 0, s, d_nw, n, d.get());
        KERNEL_OK();
        to_host(reinterpret_cast<unsigned long long *>(&bad), d.get(), 1, s);
        sync_stream(s);
        if (bad != NONE64) {
            double w = 0.0;
            to_host(&w, d_nw + bad, 1, s);
            sync_stream(s);
            char buf[160];
            std::snprintf(buf, sizeof(buf), "node %llu has invalid weight %g; must be finite and non-negative",
                          (unsigned long long)bad, w);
            ScError e(SC_INVALID_NODE_WEIGHT, buf);
            e.payload.node = bad;
            e.payload.weight = w;
            throw e;
        }
    }
}

// ---- finish (:262-351) -----------------------------------------------------------------------
// Per row: is it strictly increasing with no zero weight? Rows that are need no sort/merge.
__global__ void k_row_clean(const uint64_t *ptrs, const uint32_t *nbrs, const float *wts,
                            uint64_t n, unsigned int *dirty, unsigned int *max_deg) {
    unsigned int md = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo = ptrs[v], hi = ptrs[v + 1];
        bool bad = false;
        for (uint64_t e = lo; e < hi; e++) {
            if (wts[e] == 0.0f || (e > lo && nbrs[e] <= nbrs[e - 1])) { bad = true; break; }
        }
        if (bad) atomicOr(dirty, 1u);
        uint64_t d = hi - lo;
        md = max(md, (unsigned int)(d > 0xFFFFFFFFull ? 0xFFFFFFFFu : d));
    }
    md = block_max(md);
    if (threadIdx.x == 0 && md) atomicMax(max_deg, md);
}

// Row statistics of a sorted CSR, G lanes per row: strength (self-loop counted twice), the
// self-loop weight, the row's share (u >= v) of total weight and edge count, the largest degree
// and whether any loop exists. The reference sums each row sequentially in ascending-neighbour
// order in f64; the lanes' tree sum equals that whenever the exactness bound of f64_quantum
// holds (f32 weights of similar magnitude always meet it), otherwise lane 0 redoes the row
// sequentially.
template <int G>
__global__ void __launch_bounds__(256) k_row_stats(const uint64_t *ptrs, const uint32_t *nbrs, const float *wts,
                                                   uint64_t n, double *strength, double *self_loop,
                                                   double *row_total, unsigned long long *edge_count,
                                                   unsigned int *flags, unsigned int *bucket_counts) {
    constexpr int NG = 256 / G;
    const int lane = threadIdx.x % G;
    __shared__ unsigned int s_bc[NUM_BUCKETS];
    if (threadIdx.x < NUM_BUCKETS) s_bc[threadIdx.x] = 0;
    __syncthreads();
    unsigned long long local_edges = 0;
    unsigned int max_deg = 0, loop_seen = 0;
    for (uint64_t v0 = (uint64_t)blockIdx.x * NG; v0 < n; v0 += (uint64_t)gridDim.x * NG) {
        const uint64_t v = v0 + threadIdx.x / G;
        uint64_t lo = 0, hi = 0;
        if (v < n) {
            lo = ptrs[v];
            hi = ptrs[v + 1];
        }
        double st = 0.0, tot = 0.0, loop = 0.0, sabs = 0.0;
        int q = 4096;
        for (uint64_t e = lo + lane; e < hi; e += G) {
            const uint64_t u = nbrs[e];
            const double w = (double)wts[e];
            q = min(q, f64_quantum(w));
            sabs += 2.0 * fabs(w);
            if (u == v) {
                st += 2.0 * w;
                loop = w;
                loop_seen = 1u;
            } else {
                st += w;
            }
            if (v <= u) {
                tot += w;
                local_edges++;
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            st += __shfl_xor(st, off, G);
            tot += __shfl_xor(tot, off, G);
            loop += __shfl_xor(loop, off, G);
            sabs += __shfl_xor(sabs, off, G);
            q = min(q, __shfl_xor(q, off, G));
        }
        if (v < n && lane == 0) {
            if (!(sabs < ldexp(1.0, min(53 + q, 1000)))) {
                st = 0.0;
                tot = 0.0;
                for (uint64_t e = lo; e < hi; e++) {
                    const uint64_t u = nbrs[e];
                    const double w = (double)wts[e];
                    st += u == v ? 2.0 * w : w;
                    if (v <= u) tot += w;
                }
            }
            strength[v] = st;
            self_loop[v] = loop;
            row_total[v] = tot;
            max_deg = max(max_deg, (unsigned int)min<uint64_t>(hi - lo, 0xFFFFFFFFull));
            atomicAdd(&s_bc[bucket_of(hi - lo)], 1u);
        }
    }
    // one atomic per workgroup for each counter (same-address atomics serialize)
    local_edges = block_sum(local_edges);
    max_deg = block_max(max_deg);
    loop_seen = block_max(loop_seen);
    if (threadIdx.x == 0) {
        if (local_edges) atomicAdd(edge_count, local_edges);
        if (loop_seen) atomicOr(&flags[0], 1u);
        if (max_deg) atomicMax(&flags[2], max_deg);
    }
    __syncthreads();
    if (threadIdx.x < NUM_BUCKETS && s_bc[threadIdx.x]) atomicAdd(&bucket_counts[threadIdx.x], s_bc[threadIdx.x]);
}

void graph_finish_sorted(sc_graph &g, hipStream_t s, bool count_entries) {
    uint64_t n = g.n;
    g.strength.resize(n);
    g.self_loop.resize(n);
    g.row_total.resize(n);
    DevBuf<double> &row_total = g.row_total;
    // [0..1] edge count, then flags: [2] any self-loop, [3] unused, [4] max degree, [5] unused,
    // [6..] vertices per degree class
    constexpr int NW = 6 + NUM_BUCKETS;
    DevBuf<unsigned int> cnt(NW);
    HIP_OK(hipMemsetAsync(cnt.get(), 0, NW * sizeof(unsigned int), s));
    unsigned long long *ec_d = reinterpret_cast<unsigned long long *>(cnt.get());
    unsigned int *flags = cnt.get() + 2;
    if (n) {
        hipLaunchKernelGGL((k_row_stats<16>), dim3(grid_for(n, 16, 1024)), dim3(256), 0, s, g.ptrs.get(),
                           g.nbrs.get(), g.wts.get(), n, g.strength.get(), g.self_loop.get(), row_total.get(), ec_d,
                           flags, cnt.get() + 6);
        KERNEL_OK();
    }
    unsigned int h[NW] = {0};
    to_host(h, cnt.get(), NW, s);
    uint64_t entries = g.nnz;
    if (count_entries) to_host(&entries, g.ptrs.get() + n, 1, s);
    const double *in[4] = {row_total.get(), g.strength.get(), g.strength.get(), g.node_w.get()};
    const bool is_max[4] = {false, false, true, false};
    double r[4];
    device_reduce4_f64(in, is_max, n, r, s);  // synchronises the stream
    g.nnz = entries;
    unsigned long long ec = 0;
    std::memcpy(&ec, h, sizeof(ec));
    g.edge_count = ec;
    g.any_self_loops = h[2] != 0;
    g.max_degree = h[4];
    for (int i = 0; i < NUM_BUCKETS; i++) g.bucket_counts[i] = h[6 + i];
    g.has_bucket_counts = n <= 0xFFFFFFFFull;
    g.total_weight = r[0];
    const double deg_sum = r[1];
    g.max_strength = n ? r[2] : 0.0;
    g.total_node_weight = r[3];
    double expected = 2.0 * g.total_weight;
    double tol = 1e-6 * std::max(std::fabs(expected), 1.0);
    if (std::fabs(deg_sum - expected) > tol) {
        char buf[256];
        std::snprintf(buf, sizeof(buf),
                      "adjacency is not symmetric: \xce\xa3 strength = %.17g, expected 2 \xc2\xb7 total_weight = %.17g. "
                      "A k-NN graph must be symmetrised before use.", deg_sum, expected);
        ScError e(SC_ASYMMETRIC_GRAPH, buf);
        e.payload.degree_sum = deg_sum;
        e.payload.expected_sum = expected;
        throw e;
    }
}

__global__ void k_expand_keys(const uint64_t *ptrs, const uint32_t *nbrs, uint64_t n, uint64_t *keys) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        for (uint64_t e = ptrs[v]; e < ptrs[v + 1]; e++) keys[e] = (v << 32) | (uint64_t)nbrs[e];
    }
}

__global__ void k_run_heads(const uint64_t *keys, uint64_t nnz, uint32_t *head) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nnz;
         e += (uint64_t)gridDim.x * blockDim.x)
        head[e] = (e == 0 || keys[e] != keys[e - 1]) ? 1u : 0u;
}

// One thread per run of equal (row, col): f64 sum in sorted order, zero sums dropped.
__global__ void k_run_merge(const uint64_t *keys, const float *vals, const uint32_t *head,
                            const uint32_t *run_idx, uint64_t nnz, uint32_t *keep, double *run_w,
                            uint64_t *run_key) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nnz;
         e += (uint64_t)gridDim.x * blockDim.x) {
        if (!head[e]) continue;
        uint64_t key = keys[e];
        double w = 0.0;
        for (uint64_t f = e; f < nnz && keys[f] == key; f++) w += (double)vals[f];
        uint32_t r = run_idx[e];
        run_w[r] = w;
        run_key[r] = key;
        keep[r] = (w != 0.0) ? 1u : 0u;
    }
}

__global__ void k_run_emit(const uint64_t *run_key, const double *run_w, const uint32_t *keep,
                           const uint32_t *out_idx, uint64_t runs, uint32_t *nbrs, float *wts,
                           unsigned long long *row_count) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < runs;
         r += (uint64_t)gridDim.x * blockDim.x) {
        if (!keep[r]) continue;
        uint32_t o = out_idx[r];
        nbrs[o] = (uint32_t)(run_key[r] & 0xFFFFFFFFull);
        wts[o] = (float)run_w[r];
        atomicAdd(row_count + (run_key[r] >> 32), 1ull);
    }
}

void graph_finish_unsorted(sc_graph &g, hipStream_t s) {
    uint64_t n = g.n, nnz = g.nnz;
    // fast path: every row already strictly increasing with no zero weights
    DevBuf<unsigned int> flags(2);
    HIP_OK(hipMemsetAsync(flags.get(), 0, 2 * sizeof(unsigned int), s));
    if (n) {
        hipLaunchKernelGGL(k_row_clean, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, g.ptrs.get(),
                           g.nbrs.get(), g.wts.get(), n, flags.get(), flags.get() + 1);
        KERNEL_OK();
    }
    unsigned int fl[2];
    to_host(fl, flags.get(), 2, s);
    sync_stream(s);
    if (fl[0] == 0) {
        graph_finish_sorted(g, s);
        return;
    }
    // sort entries by (row, col); stable, so duplicates keep their input order
    DevBuf<uint64_t> keys(nnz);
    DevBuf<float> vals(nnz);
    HIP_OK(hipMemcpyAsync(vals.get(), g.wts.get(), nnz * sizeof(float), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_expand_keys, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, g.ptrs.get(),
                       g.nbrs.get(), n, keys.get());
    KERNEL_OK();
    radix_sort_pairs(keys, vals, nnz, 32 + bits_for(n), s);
    DevBuf<uint32_t> head(nnz), run_idx(nnz);
    hipLaunchKernelGGL(k_run_heads, dim3(grid_for(nnz, 256, 8192)), dim3(256), 0, s, keys.get(), nnz, head.get());
    KERNEL_OK();
    exclusive_scan_u32(head.get(), run_idx.get(), nnz, s);
    uint32_t last_idx = 0, last_head = 0;
    to_host(&last_idx, run_idx.get() + nnz - 1, 1, s);
    to_host(&last_head, head.get() + nnz - 1, 1, s);
    sync_stream(s);
    uint64_t runs = (uint64_t)last_idx + last_head;
    DevBuf<uint32_t> keep(runs), out_idx(runs);
    DevBuf<double> run_w(runs);
    DevBuf<uint64_t> run_key(runs);
    hipLaunchKernelGGL(k_run_merge, dim3(grid_for(nnz, 256, 8192)), dim3(256), 0, s, keys.get(), vals.get(),
                       head.get(), run_idx.get(), nnz, keep.get(), run_w.get(), run_key.get());
    KERNEL_OK();
    exclusive_scan_u32(keep.get(), out_idx.get(), runs, s);
    uint32_t lk = 0, lo = 0;
    to_host(&lk, keep.get() + runs - 1, 1, s);
    to_host(&lo, out_idx.get() + runs - 1, 1, s);
    sync_stream(s);
    uint64_t out_nnz = (uint64_t)lo + lk;
    DevBuf<uint32_t> nbrs(out_nnz);
    DevBuf<float> wts(out_nnz);
    DevBuf<unsigned long long> row_count(n + 1);
    HIP_OK(hipMemsetAsync(row_count.get(), 0, (n + 1) * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_run_emit, dim3(grid_for(runs, 256, 8192)), dim3(256), 0, s, run_key.get(), run_w.get(),
                       keep.get(), out_idx.get(), runs, nbrs.get(), wts.get(), row_count.get());
    KERNEL_OK();
    DevBuf<uint64_t> ptrs(n + 1);
    exclusive_scan_u64(reinterpret_cast<uint64_t *>(row_count.get()), ptrs.get(), n + 1, s);
    g.ptrs = std::move(ptrs);
    g.nbrs = std::move(nbrs);
    g.wts = std::move(wts);
    g.nnz = out_nnz;
    graph_finish_sorted(g, s);
}

// ---- validate_symmetry (:627-646) ------------------------------------------------------------
__device__ __forceinline__ bool row_find(const uint64_t *ptrs, const uint32_t *nbrs, const float *wts,
                                         uint64_t row, uint32_t target, float *w) {
    uint64_t lo = ptrs[row], hi = ptrs[row + 1];
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        uint32_t x = nbrs[mid];
        if (x == target) { *w = wts[mid]; return true; }
        if (x < target) lo = mid + 1; else hi = mid;
    }
    return false;
}

// edge_weight(u, v) searches the shorter of the two rows (csr_network.rs:465-478); restated
// exactly, including that a search of row v finds the entry (v, u) itself.
__global__ void k_symmetry(const uint64_t *ptrs, const uint32_t *nbrs, const float *wts, uint64_t n,
                           unsigned long long *first_bad) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t dv = ptrs[v + 1] - ptrs[v];
        for (uint64_t e = ptrs[v]; e < ptrs[v + 1]; e++) {
            uint32_t u = nbrs[e];
            double w = (double)wts[e];
            uint64_t du = ptrs[u + 1] - ptrs[u];
            float back = 0.0f;
            bool found = (du <= dv) ? row_find(ptrs, nbrs, wts, u, (uint32_t)v, &back)
                                    : row_find(ptrs, nbrs, wts, v, u, &back);
            if (!found || fabs((double)back - w) > 1e-6 * fmax(fabs(w), 1.0)) {
                atomicMin(first_bad, (unsigned long long)e);
                break;
            }
        }
    }
}

}  // namespace scgpu

using namespace scgpu;

static void upload_graph(sc_graph &g, const uint64_t *ptrs, uint64_t n, const uint32_t *nbrs,
                         const float *wts, uint64_t nnz, const double *node_w, bool device_src,
                         hipStream_t s) {
    g.n = n;
    g.nnz = nnz;
    g.ptrs.resize(n + 1);
    g.nbrs.resize(nnz);
    g.wts.resize(nnz);
    g.node_w.resize(n);
    hipMemcpyKind kind = device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIP_OK(hipMemcpyAsync(g.ptrs.get(), ptrs, (n + 1) * sizeof(uint64_t), kind, s));
    if (nnz) {
        HIP_OK(hipMemcpyAsync(g.nbrs.get(), nbrs, nnz * sizeof(uint32_t), kind, s));
        HIP_OK(hipMemcpyAsync(g.wts.get(), wts, nnz * sizeof(float), kind, s));
    }
    if (node_w) {
        if (n) HIP_OK(hipMemcpyAsync(g.node_w.get(), node_w, n * sizeof(double), kind, s));
    } else if (n) {
        hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, g.node_w.get(), n, 1.0);
        KERNEL_OK();
    }
}

static sc_status build_from_csr(const uint64_t *ptrs, uint64_t n, const uint32_t *nbrs,
                                const float *wts, uint64_t nnz, const double *node_w,
                                bool device_src, sc_graph **out, sc_error *err) {
    if (!out || (!ptrs) || (nnz && (!nbrs || !wts))) {
        if (err) { std::memset(err, 0, sizeof(*err)); err->code = SC_INVALID_ARGUMENT;
                   std::snprintf(err->msg, sizeof(err->msg), "null argument"); }
        return SC_INVALID_ARGUMENT;
    }
    *out = nullptr;
    sc_graph *g = nullptr;
    try {
        check_node_count(n);
        check_entry_count(nnz);
        g = new sc_graph();
        hipStream_t s = thread_stream();
        upload_graph(*g, ptrs, n, nbrs, wts, nnz, node_w, device_src, s);
        validate_csr_device(g->ptrs.get(), n, g->nbrs.get(), g->wts.get(), nnz,
                            node_w ? g->node_w.get() : nullptr, s);
        graph_finish_unsorted(*g, s);
        sync_stream(s);
        *out = g;
        return SC_OK;
    } catch (...) {
        delete g;
        return status_from_exception(err);
    }
}

// ---- scanpy CSR export --------------------------------------------------------------------
// <dir>/indptr.bin (u64), indices.bin (u32) and data.bin (f32), raw little-endian, as the
// export tool writes a connectivities matrix (tools/export_h5ad.py:106-108) and the real-data
// example reads it (examples/real_data.rs:22-44, 129-136).
static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "the export files are little-endian");

static std::vector<unsigned char> read_whole_file(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw ScError(SC_INVALID_ARGUMENT, path + ": cannot open");
    std::vector<unsigned char> buf;
    unsigned char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) throw ScError(SC_INVALID_ARGUMENT, path + ": read error");
    return buf;
}

template <typename T>
static std::vector<T> read_array_file(const std::string &path) {
    const std::vector<unsigned char> raw = read_whole_file(path);
    if (raw.size() % sizeof(T))
        throw ScError(SC_INVALID_CSR, "malformed CSR input: " + path + " holds " + std::to_string(raw.size()) +
                                          " bytes, not a whole number of " + std::to_string(sizeof(T)) +
                                          "-byte entries");
    std::vector<T> out(raw.size() / sizeof(T));
    if (!out.empty()) std::memcpy(out.data(), raw.data(), raw.size());
    return out;
}

extern "C" {

sc_status sc_graph_from_csr_export(const char *dir, sc_graph **out, sc_error *err) {
    std::vector<uint64_t> ptrs;
    std::vector<uint32_t> nbrs;
    std::vector<float> wts;
    try {
        if (!dir || !out) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        const std::string d(dir);
        ptrs = read_array_file<uint64_t>(d + "/indptr.bin");
        nbrs = read_array_file<uint32_t>(d + "/indices.bin");
        wts = read_array_file<float>(d + "/data.bin");
        if (ptrs.empty())
            throw ScError(SC_INVALID_CSR, "malformed CSR input: " + d + "/indptr.bin is empty (it holds n + 1 row offsets)");
        if (nbrs.size() != wts.size())
            throw ScError(SC_INVALID_CSR, "malformed CSR input: indices.bin has " + std::to_string(nbrs.size()) +
                                              " entries but data.bin has " + std::to_string(wts.size()));
    } catch (...) {
        return status_from_exception(err);
    }
    // the row offsets, ids and weights are checked as from_csr_parts checks them
    return build_from_csr(ptrs.data(), ptrs.size() - 1, nbrs.data(), wts.data(), nbrs.size(), nullptr, false, out,
                          err);
}

sc_status sc_graph_from_csr_parts(const uint64_t *node_ptrs, uint64_t n_nodes,
                                  const uint32_t *neighbors, const float *weights, uint64_t nnz,
                                  const double *node_weights, sc_graph **out, sc_error *err) {
    return build_from_csr(node_ptrs, n_nodes, neighbors, weights, nnz, node_weights, false, out, err);
}

sc_status sc_graph_from_csr_parts_device(const uint64_t *d_node_ptrs, uint64_t n_nodes,
                                         const uint32_t *d_neighbors, const float *d_weights,
                                         uint64_t nnz, const double *d_node_weights,
                                         sc_graph **out, sc_error *err) {
    return build_from_csr(d_node_ptrs, n_nodes, d_neighbors, d_weights, nnz, d_node_weights, true, out, err);
}

sc_status sc_graph_from_edges(uint64_t n, const uint64_t *from, const uint64_t *to,
                              const double *ew, uint64_t m, const double *node_weights,
                              sc_graph **out, sc_error *err) {
    // Host-side scatter exactly as the reference (two passes, :133-164), then device finish.
    try {
        if (!out || (m && (!from || !to || !ew))) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        if (node_weights) {
            for (uint64_t v = 0; v < n; v++) {
                double x = node_weights[v];
                if (!std::isfinite(x) || x < 0.0) {
                    char buf[160];
                    std::snprintf(buf, sizeof(buf), "node %llu has invalid weight %g; must be finite and non-negative",
                                  (unsigned long long)v, x);
                    ScError e(SC_INVALID_NODE_WEIGHT, buf);
                    e.payload.node = v;
                    e.payload.weight = x;
                    throw e;
                }
            }
        }
        check_node_count(n);
        std::vector<uint64_t> counts(n, 0);
        for (uint64_t i = 0; i < m; i++) {
            uint64_t a = from[i], b = to[i];
            double w = ew[i];
            if (a >= n || b >= n) {
                uint64_t bad = a >= n ? a : b;
                ScError e(SC_NODE_INDEX_OUT_OF_RANGE, "node index " + std::to_string(bad) +
                                                          " is out of range for a graph with " +
                                                          std::to_string(n) + " nodes");
                e.payload.node = bad;
                e.payload.n_nodes = n;
                throw e;
            }
            if (!std::isfinite(w)) {
                ScError e(SC_NON_FINITE_WEIGHT, "edge (" + std::to_string(a) + ", " + std::to_string(b) +
                                                    ") has a non-finite weight");
                e.payload.edge_u = a;
                e.payload.edge_v = b;
                throw e;
            }
            if (w < 0.0) {
                char buf[160];
                std::snprintf(buf, sizeof(buf), "edge (%llu, %llu) has negative weight %g; weights must be non-negative",
                              (unsigned long long)a, (unsigned long long)b, w);
                ScError e(SC_NEGATIVE_WEIGHT, buf);
                e.payload.edge_u = a;
                e.payload.edge_v = b;
                e.payload.weight = w;
                throw e;
            }
            counts[a]++;
            if (a != b) counts[b]++;
        }
        std::vector<uint64_t> ptrs(n + 1, 0);
        for (uint64_t v = 0; v < n; v++) ptrs[v + 1] = ptrs[v] + counts[v];
        uint64_t slots = ptrs[n];
        check_entry_count(slots);
        std::vector<uint32_t> nbrs(slots);
        std::vector<float> wts(slots);
        std::vector<uint64_t> cur(ptrs.begin(), ptrs.end() - 1);
        for (uint64_t i = 0; i < m; i++) {
            uint64_t a = from[i], b = to[i];
            float w = (float)ew[i];
            nbrs[cur[a]] = (uint32_t)b; wts[cur[a]] = w; cur[a]++;
            if (a != b) { nbrs[cur[b]] = (uint32_t)a; wts[cur[b]] = w; cur[b]++; }
        }
        sc_graph *g = new sc_graph();
        try {
            hipStream_t s = thread_stream();
            upload_graph(*g, ptrs.data(), n, nbrs.data(), wts.data(), slots, node_weights, false, s);
            graph_finish_unsorted(*g, s);
            sync_stream(s);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

// CSRNetwork::from_csr_matrix_with_node_weights (csr_network.rs:360-391): the triplets of a
// sparse matrix in row order, upper triangle and diagonal only (row <= col), zero entries
// skipped, become an edge list for from_edges_with_node_weights. The matrix arrives as CSR
// arrays (usize column indices, f64 values); its row offsets are checked as a CsrMatrix's are.
sc_status sc_graph_from_csr_matrix(uint64_t n_rows, const uint64_t *row_ptrs, const uint64_t *col_indices,
                                   const double *values, uint64_t nnz, const double *node_weights, sc_graph **out,
                                   sc_error *err) {
    std::vector<uint64_t> a, b;
    std::vector<double> w;
    try {
        if (!out || !row_ptrs || (nnz && (!col_indices || !values))) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        if (row_ptrs[0] != 0 || row_ptrs[n_rows] != nnz)
            throw ScError(SC_INVALID_CSR, "malformed CSR input: row offsets must start at 0 and end at the entry count");
        for (uint64_t r = 0; r < n_rows; r++)
            if (row_ptrs[r + 1] < row_ptrs[r])
                throw ScError(SC_INVALID_CSR, "malformed CSR input: row offsets must be non-decreasing");
        for (uint64_t r = 0; r < n_rows; r++)
            for (uint64_t e = row_ptrs[r]; e < row_ptrs[r + 1]; e++)
                if (r <= col_indices[e] && values[e] != 0.0) {
                    a.push_back(r);
                    b.push_back(col_indices[e]);
                    w.push_back(values[e]);
                }
    } catch (...) {
        return status_from_exception(err);
    }
    return sc_graph_from_edges(n_rows, a.data(), b.data(), w.data(), a.size(), node_weights, out, err);
}

// CSRNetwork::to_csr_matrix (csr_network.rs:610-618): the stored adjacency as a symmetric CSR
// matrix with usize column indices and f64 values. Any output pointer may be NULL.
sc_status sc_graph_to_csr_matrix(const sc_graph *g, uint64_t *row_ptrs, uint64_t *col_indices, double *values,
                                 sc_error *err) {
    try {
        if (!g) throw ScError(SC_INVALID_ARGUMENT, "null graph");
        std::vector<uint32_t> nb(col_indices ? g->nnz : 0);
        std::vector<float> wt(values ? g->nnz : 0);
        hipStream_t s = thread_stream();
        if (row_ptrs) to_host(row_ptrs, g->ptrs.get(), g->n + 1, s);
        if (col_indices) to_host(nb.data(), g->nbrs.get(), g->nnz, s);
        if (values) to_host(wt.data(), g->wts.get(), g->nnz, s);
        sync_stream(s);
        for (uint64_t e = 0; col_indices && e < g->nnz; e++) col_indices[e] = nb[e];
        for (uint64_t e = 0; values && e < g->nnz; e++) values[e] = (double)wt[e];
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

void sc_graph_free(sc_graph *g) {
    if (!g) return;
    (void)hipStreamSynchronize(thread_stream());
    delete g;
}

uint64_t sc_graph_node_count(const sc_graph *g) { return g ? g->n : 0; }
uint64_t sc_graph_edge_count(const sc_graph *g) { return g ? g->edge_count : 0; }
uint64_t sc_graph_nnz(const sc_graph *g) { return g ? g->nnz : 0; }
double sc_graph_total_weight(const sc_graph *g) { return g ? g->total_weight : 0.0; }
int32_t sc_graph_has_self_loops(const sc_graph *g) { return g && g->any_self_loops ? 1 : 0; }

sc_status sc_graph_to_csr_parts(const sc_graph *g, uint64_t *node_ptrs, uint32_t *neighbors,
                                float *weights, double *node_weights, double *strengths,
                                sc_error *err) {
    try {
        if (!g) throw ScError(SC_INVALID_ARGUMENT, "null graph");
        hipStream_t s = thread_stream();
        if (node_ptrs) to_host(node_ptrs, g->ptrs.get(), g->n + 1, s);
        if (neighbors) to_host(neighbors, g->nbrs.get(), g->nnz, s);
        if (weights) to_host(weights, g->wts.get(), g->nnz, s);
        if (node_weights) to_host(node_weights, g->node_w.get(), g->n, s);
        if (strengths) to_host(strengths, g->strength.get(), g->n, s);
        sync_stream(s);
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

sc_status sc_graph_validate_symmetry(const sc_graph *g, sc_error *err) {
    try {
        if (!g) throw ScError(SC_INVALID_ARGUMENT, "null graph");
        hipStream_t s = thread_stream();
        DevBuf<unsigned long long> d(1);
        unsigned long long init = NONE64, bad = 0;
        HIP_OK(hipMemcpyAsync(d.get(), &init, sizeof(init), hipMemcpyHostToDevice, s));
        if (g->n)
            hipLaunchKernelGGL(k_symmetry, dim3(grid_for(g->n, 256, 8192)), dim3(256), 0, s, g->ptrs.get(),
                               g->nbrs.get(), g->wts.get(), g->n, d.get());
        KERNEL_OK();
        to_host(reinterpret_cast<unsigned long long *>(&bad), d.get(), 1, s);
        sync_stream(s);
        if (bad == NONE64) return SC_OK;
        // locate the row of entry `bad` on the host (rare path)
        std::vector<uint64_t> ptrs(g->n + 1);
        to_host(ptrs.data(), g->ptrs.get(), g->n + 1, s);
        uint32_t u = 0;
        float w = 0.0f;
        to_host(&u, g->nbrs.get() + bad, 1, s);
        to_host(&w, g->wts.get() + bad, 1, s);
        sync_stream(s);
        uint64_t v = (uint64_t)(std::upper_bound(ptrs.begin(), ptrs.end(), bad) - ptrs.begin()) - 1;
        ScError e(SC_INVALID_CSR, "malformed CSR input: edge (" + std::to_string(v) + ", " + std::to_string(u) +
                                      ") has no matching reverse edge of weight " + std::to_string(w));
        e.payload.edge_u = v;
        e.payload.edge_v = u;
        e.payload.weight = w;
        throw e;
    } catch (...) {
        return status_from_exception(err);
    }
}

}  // extern "C"
