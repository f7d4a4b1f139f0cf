// quality.hip -- modularity, Rb/Cpm quality and the community-connectivity check.
//
// modularity: leiden/objective.rs:221-244. Rb::quality :108-119, Cpm::quality :171-178 over
// Partition::from_membership (partition.rs:75-112). disconnected_community_count:
// tests/common/mod.rs:171-194.
//
// Each vertex sums its own internal contribution in row order (neighbours u with the same label
// and u == v or v < u: every edge once, loops once), exactly the reference's inner loop. The
// per-community totals (internal weight, strength, node weight) are accumulated as int64 fixed
// point with power-of-two scales chosen so no sum exceeds 2^61: integer atomics are exact and
// associative, so the result does not depend on scheduling (deterministic) and carries a
// relative error below n * 2^-61. The sum over communities is a fixed-shape tree.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace scgpu {

__global__ void k_max_label(const uint64_t *labels, uint64_t n, unsigned long long *max_label) {
    unsigned long long local = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t l = labels[v];
        if (l > local) local = l;
    }
    local = block_max(local);
    if (threadIdx.x == 0) atomicMax(max_label, local);
}

// G lanes per vertex stride its row; the lane partial sums combine in a fixed shuffle tree.
template <int G>
__global__ void __launch_bounds__(256) k_vertex_terms(const uint64_t *labels, uint64_t n, const uint64_t *ptrs,
                                                      const uint32_t *nbrs, const float *wts, const double *strength,
                                                      const double *node_w, int s_in, int sK, int sW,
                                                      long long *in_fx, long long *K_fx, long long *W_fx,
                                                      const uint32_t *vjob = nullptr, const int *jscale = nullptr) {
    constexpr int NG = 256 / G;
    constexpr uint32_t TAB = 1024;  // per-workgroup label table (few labels recur within a block)
    __shared__ unsigned long long t_key[TAB];
    __shared__ long long t_in[TAB], t_K[TAB], t_W[TAB];
    for (uint32_t i = threadIdx.x; i < TAB; i += blockDim.x) {
        t_key[i] = ~0ull;
        t_in[i] = t_K[i] = t_W[i] = 0;
    }
    __syncthreads();
    const int lane = threadIdx.x % G;
    for (uint64_t v = blockIdx.x * (uint64_t)NG + threadIdx.x / G; v < n; v += (uint64_t)gridDim.x * NG) {
        const uint64_t c = labels[v];
        double acc = 0.0;
        // EB entries per lane in flight (neighbour ids, then their labels, as independent
        // loads); each lane still adds its entries e, e + G, e + 2G, ... in that order
        constexpr int EB = 4;
        const uint64_t end = ptrs[v + 1];
        for (uint64_t e0 = ptrs[v] + lane; e0 < end; e0 += (uint64_t)EB * G) {
            uint64_t u[EB];
            float w[EB];
#pragma unroll
            for (int j = 0; j < EB; j++) {
                const uint64_t e = e0 + (uint64_t)j * G;
                u[j] = e < end ? nbrs[e] : v;
                w[j] = e < end ? wts[e] : 0.0f;
            }
            bool same[EB];
#pragma unroll
            for (int j = 0; j < EB; j++) same[j] = labels[u[j]] == c;
#pragma unroll
            for (int j = 0; j < EB; j++)
                if (e0 + (uint64_t)j * G < end && same[j] && (u[j] == v || v < u[j])) acc += (double)w[j];
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
        if (lane == 0) {
            int si = s_in, sk = sK, sw = sW;
            if (vjob) {  // batched run: the scales of this vertex's job
                const uint32_t j = vjob[v];
                si = jscale[3 * j];
                sk = jscale[3 * j + 1];
                sw = jscale[3 * j + 2];
            }
            const long long fi = acc != 0.0 ? to_fx(acc, si) : 0, fk = to_fx(strength[v], sk),
                            fw = to_fx(node_w[v], sw);
            // integer sums are order independent: pre-add in LDS, one global atomic per label
            uint32_t h = hash32((uint32_t)c ^ (uint32_t)(c >> 32)) & (TAB - 1);
            bool done = false;
            for (uint32_t probe = 0; probe < 32 && !done; probe++, h = (h + 1) & (TAB - 1)) {
                const unsigned long long prev = atomicCAS(&t_key[h], ~0ull, (unsigned long long)c);
                if (prev == ~0ull || prev == c) {
                    atomicAdd((unsigned long long *)&t_in[h], (unsigned long long)fi);
                    atomicAdd((unsigned long long *)&t_K[h], (unsigned long long)fk);
                    atomicAdd((unsigned long long *)&t_W[h], (unsigned long long)fw);
                    done = true;
                }
            }
            if (!done) {
                atomicAdd((unsigned long long *)&in_fx[c], (unsigned long long)fi);
                atomicAdd((unsigned long long *)&K_fx[c], (unsigned long long)fk);
                atomicAdd((unsigned long long *)&W_fx[c], (unsigned long long)fw);
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < TAB; i += blockDim.x) {
        const unsigned long long c = t_key[i];
        if (c == ~0ull) continue;
        if (t_in[i]) atomicAdd((unsigned long long *)&in_fx[c], (unsigned long long)t_in[i]);
        atomicAdd((unsigned long long *)&K_fx[c], (unsigned long long)t_K[i]);
        atomicAdd((unsigned long long *)&W_fx[c], (unsigned long long)t_W[i]);
    }
}

// mode 0: modularity term 2*in/2m - gamma*(K/2m)^2; mode 1: Rb quality term; mode 2: Cpm.
__global__ void k_community_terms(const long long *in_fx, const long long *K_fx, const long long *W_fx,
                                  uint64_t nl, int s_in, int sK, int sW, int mode, double gamma, double m,
                                  double *term, const uint32_t *ljob = nullptr, const int *jscale = nullptr,
                                  const double *jgamma = nullptr, const double *jm = nullptr) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < nl;
         c += (uint64_t)gridDim.x * blockDim.x) {
        if (ljob) {  // batched run: the label's job supplies scales, resolution and m
            const uint32_t j = ljob[c];
            s_in = jscale[3 * j];
            sK = jscale[3 * j + 1];
            sW = jscale[3 * j + 2];
            gamma = jgamma[j];
            m = jm[j];
        }
        double internal = from_fx(in_fx[c], s_in), st = from_fx(K_fx[c], sK), nw = from_fx(W_fx[c], sW);
        double t;
        if (mode == 0) {
            double two_m = 2.0 * m;
            double f = st / two_m;
            t = 2.0 * internal / two_m - gamma * (f * f);
        } else if (mode == 1) {
            double four_m = 4.0 * m;
            t = (four_m == 0.0) ? 0.0 : internal - gamma * st * st / four_m;
        } else {
            t = internal - gamma * nw * (nw - 1.0) / 2.0;
        }
        term[c] = t;
    }
}

double quality_device(const sc_graph &g, const uint64_t *d_labels, int mode, double gamma, hipStream_t s) {
    uint64_t n = g.n;
    if (n == 0) return 0.0;
    if (mode == 0 && g.total_weight == 0.0) return 0.0;
    DevBuf<unsigned long long> mx(1);
    HIP_OK(hipMemsetAsync(mx.get(), 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_max_label, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, d_labels, n, mx.get());
    KERNEL_OK();
    unsigned long long max_label = 0;
    to_host(&max_label, mx.get(), 1, s);
    sync_stream(s);
    if (max_label >= 0xFFFFFFFFull) throw ScError(SC_INVALID_ARGUMENT, "labels must be below 2^32 - 1");
    const uint64_t nl = max_label + 1;
    const int s_in = fixed_scale_for(g.total_weight + 1e-300);
    const int sK = fixed_scale_for(2.0 * g.total_weight + 1e-300);
    const int sW = fixed_scale_for(g.total_node_weight + 1e-300);
    DevBuf<long long> acc(3 * nl);
    HIP_OK(hipMemsetAsync(acc.get(), 0, 3 * nl * sizeof(long long), s));
    hipLaunchKernelGGL((k_vertex_terms<8>), dim3(grid_for(n, 32, 1024)), dim3(256), 0, s, d_labels, n, g.ptrs.get(),
                       g.nbrs.get(), g.wts.get(), g.strength.get(), g.node_w.get(), s_in, sK, sW, acc.get(),
                       acc.get() + nl, acc.get() + 2 * nl);
    KERNEL_OK();
    DevBuf<double> term(nl);
    hipLaunchKernelGGL(k_community_terms, dim3(grid_for(nl, 256, 4096)), dim3(256), 0, s, acc.get(), acc.get() + nl,
                       acc.get() + 2 * nl, nl, s_in, sK, sW, mode, gamma, g.total_weight, term.get());
    KERNEL_OK();
    return device_sum_f64(term.get(), nl, s);
}

// Quality of the singleton partition (label v for vertex v): a vertex's only same-label entry is
// its self-loop (rows hold each neighbour once), whose weight is self_loop[v], so the per-label
// totals are the vertex's own values, stored without atomics. Same totals and sum as
// quality_device on labels 0 .. n - 1.
__global__ void k_singleton_terms(uint64_t n, const double *self_loop, const double *strength, const double *node_w,
                                  int s_in, int sK, int sW, long long *in_fx, long long *K_fx, long long *W_fx) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const double loop = self_loop[v];
        in_fx[v] = loop != 0.0 ? to_fx(loop, s_in) : 0;
        K_fx[v] = to_fx(strength[v], sK);
        W_fx[v] = to_fx(node_w[v], sW);
    }
}

double quality_singletons_device(const sc_graph &g, int mode, double gamma, hipStream_t s) {
    const uint64_t n = g.n;
    if (n == 0) return 0.0;
    if (mode == 0 && g.total_weight == 0.0) return 0.0;
    const int s_in = fixed_scale_for(g.total_weight + 1e-300);
    const int sK = fixed_scale_for(2.0 * g.total_weight + 1e-300);
    const int sW = fixed_scale_for(g.total_node_weight + 1e-300);
    DevBuf<long long> acc(3 * n);
    hipLaunchKernelGGL(k_singleton_terms, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, n, g.self_loop.get(),
                       g.strength.get(), g.node_w.get(), s_in, sK, sW, acc.get(), acc.get() + n, acc.get() + 2 * n);
    KERNEL_OK();
    DevBuf<double> term(n);
    hipLaunchKernelGGL(k_community_terms, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, acc.get(), acc.get() + n,
                       acc.get() + 2 * n, n, s_in, sK, sW, mode, gamma, g.total_weight, term.get());
    KERNEL_OK();
    return device_sum_f64(term.get(), n, s);
}

// ---- batched runs: one quality per job -----------------------------------------------------
__global__ void k_label_jobs(const uint64_t *labels, const uint32_t *vjob, uint64_t n, uint32_t *ljob) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        ljob[labels[v]] = vjob[v];
}

// Smallest label of each job (one workgroup per job over its vertex range; ~0 for a job
// without vertices).
__global__ void k_job_min_label(const uint64_t *labels, const uint64_t *vranges, uint64_t *lranges) {
    const uint32_t j = blockIdx.x;
    unsigned long long mn = ~0ull;
    for (uint64_t v = vranges[j] + threadIdx.x; v < vranges[j + 1]; v += blockDim.x)
        mn = min(mn, (unsigned long long)labels[v]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mn = min(mn, __shfl_xor(mn, off));
    __shared__ unsigned long long sh[16];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mn;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned w = 1; w < blockDim.x / 64; w++) mn = min(mn, sh[w]);
        lranges[j] = mn;
    }
}

// Quality of each job of a batched run: labels are consecutive over the union and contiguous
// per job (ascending), so job j's communities are labels lranges[j] .. lranges[j+1). Per job
// the fixed-point scales, resolution and m are the job's own, and the sum over its communities
// uses the association order quality_device uses for a graph of that job alone.
void quality_jobs_device(const sc_graph &g, const uint64_t *d_labels, int mode, const uint32_t *vjob,
                         const uint64_t *d_vranges, uint32_t J, const double *h_gamma, const double *h_tw,
                         const double *h_tnw, double *h_q, hipStream_t s) {
    const uint64_t n = g.n;
    for (uint32_t j = 0; j < J; j++) h_q[j] = 0.0;
    if (n == 0 || J == 0) return;
    DevBuf<unsigned long long> mx(1);
    HIP_OK(hipMemsetAsync(mx.get(), 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_max_label, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, d_labels, n, mx.get());
    KERNEL_OK();
    DevBuf<uint64_t> lr(J + 1);
    hipLaunchKernelGGL(k_job_min_label, dim3(J), dim3(256), 0, s, d_labels, d_vranges, lr.get());
    KERNEL_OK();
    unsigned long long max_label = 0;
    std::vector<uint64_t> hlr(J + 1);
    to_host(&max_label, mx.get(), 1, s);
    to_host(hlr.data(), lr.get(), J, s);
    sync_stream(s);
    const uint64_t nl = max_label + 1;
    // a job without vertices has no smallest label (the kernel leaves ~0): its label range is
    // the empty range at the next job's start, so the ranges stay ascending
    hlr[J] = nl;
    for (uint32_t j = J; j-- > 0;)
        if (hlr[j] > hlr[j + 1]) hlr[j] = hlr[j + 1];
    to_device(lr.get(), hlr.data(), J + 1, s);
    std::vector<int> js(3 * (size_t)J);
    std::vector<double> jm(J);
    for (uint32_t j = 0; j < J; j++) {
        js[3 * j] = fixed_scale_for(h_tw[j] + 1e-300);
        js[3 * j + 1] = fixed_scale_for(2.0 * h_tw[j] + 1e-300);
        js[3 * j + 2] = fixed_scale_for(h_tnw[j] + 1e-300);
        jm[j] = h_tw[j];
    }
    DevBuf<int> d_js(3 * (size_t)J);
    DevBuf<double> d_g(J), d_m(J), d_q(J);
    to_device(d_js.get(), js.data(), 3 * (size_t)J, s);
    to_device(d_g.get(), h_gamma, J, s);
    to_device(d_m.get(), jm.data(), J, s);
    DevBuf<uint32_t> ljob(nl);
    hipLaunchKernelGGL(k_label_jobs, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, d_labels, vjob, n, ljob.get());
    KERNEL_OK();
    DevBuf<long long> acc(3 * nl);
    HIP_OK(hipMemsetAsync(acc.get(), 0, 3 * nl * sizeof(long long), s));
    hipLaunchKernelGGL((k_vertex_terms<8>), dim3(grid_for(n, 32, 1024)), dim3(256), 0, s, d_labels, n, g.ptrs.get(),
                       g.nbrs.get(), g.wts.get(), g.strength.get(), g.node_w.get(), 0, 0, 0, acc.get(),
                       acc.get() + nl, acc.get() + 2 * nl, vjob, (const int *)d_js.get());
    KERNEL_OK();
    DevBuf<double> term(nl);
    hipLaunchKernelGGL(k_community_terms, dim3(grid_for(nl, 256, 4096)), dim3(256), 0, s, acc.get(), acc.get() + nl,
                       acc.get() + 2 * nl, nl, 0, 0, 0, mode, 0.0, 0.0, term.get(), (const uint32_t *)ljob.get(),
                       (const int *)d_js.get(), (const double *)d_g.get(), (const double *)d_m.get());
    KERNEL_OK();
    device_segment_reduce_f64(term.get(), lr.get(), J, false, d_q.get(), s);
    to_host(h_q, d_q.get(), J, s);
    sync_stream(s);
}

// ---- connected components restricted to same-label edges --------------------------------
// Parents always point to a smaller id, so following them terminates at a root.
__device__ __forceinline__ uint32_t cc_find(const uint32_t *comp, uint32_t v) {
    const volatile uint32_t *c = comp;
    uint32_t p = c[v];
    while (p != v) {
        v = p;
        p = c[v];
    }
    return p;
}

__global__ void k_cc_init(uint32_t *comp, uint64_t n) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x)
        comp[v] = (uint32_t)v;
}

// Hook: for every same-label edge join the two roots, the larger root pointing to the smaller.
__global__ void k_cc_hook(const uint64_t *ptrs, const uint32_t *nbrs, const uint32_t *labels,
                          uint64_t n, uint32_t *comp, unsigned int *changed) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lv = labels[v];
        for (uint64_t e = ptrs[v]; e < ptrs[v + 1]; e++) {
            uint32_t u = nbrs[e];
            if (u == v || labels[u] != lv) continue;
            uint32_t a = cc_find(comp, (uint32_t)v), b = cc_find(comp, u);
            while (a != b) {
                uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
                // hook only while hi is still a root; otherwise re-find and retry
                uint32_t old = atomicCAS(&comp[hi], hi, lo);
                if (old == hi) {
                    *changed = 1u;
                    break;
                }
                a = cc_find(comp, hi);
                b = cc_find(comp, lo);
            }
        }
    }
}

__global__ void k_cc_compress(uint32_t *comp, uint64_t n) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x)
        comp[v] = cc_find(comp, (uint32_t)v);
}

__global__ void k_cc_count(const uint32_t *comp, const uint32_t *labels, uint64_t n, uint32_t *ncomp) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x)
        if (comp[v] == v) atomicAdd(&ncomp[labels[v]], 1u);
}

__global__ void k_cc_bad(const uint32_t *ncomp, uint64_t nl, unsigned long long *bad) {
    unsigned long long local = 0;
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < nl;
         c += (uint64_t)gridDim.x * blockDim.x)
        if (ncomp[c] > 1) local++;
    if (local) atomicAdd(bad, local);
}

uint64_t disconnected_count_device(const sc_graph &g, const uint32_t *d_labels32, uint64_t n_labels,
                                   hipStream_t s) {
    uint64_t n = g.n;
    if (n == 0) return 0;
    DevBuf<uint32_t> comp(n);
    DevBuf<unsigned int> changed(1);
    hipLaunchKernelGGL(k_cc_init, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, comp.get(), n);
    KERNEL_OK();
    for (int it = 0; it < 1000; it++) {
        HIP_OK(hipMemsetAsync(changed.get(), 0, sizeof(unsigned int), s));
        hipLaunchKernelGGL(k_cc_hook, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, g.ptrs.get(), g.nbrs.get(),
                           d_labels32, n, comp.get(), changed.get());
        KERNEL_OK();
        hipLaunchKernelGGL(k_cc_compress, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, comp.get(), n);
        KERNEL_OK();
        unsigned int ch = 0;
        to_host(&ch, changed.get(), 1, s);
        sync_stream(s);
        if (!ch) break;
    }
    DevBuf<uint32_t> ncomp(n_labels);
    HIP_OK(hipMemsetAsync(ncomp.get(), 0, n_labels * sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_cc_count, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, comp.get(), d_labels32, n,
                       ncomp.get());
    KERNEL_OK();
    DevBuf<unsigned long long> bad(1);
    HIP_OK(hipMemsetAsync(bad.get(), 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_cc_bad, dim3(grid_for(n_labels, 256, 4096)), dim3(256), 0, s, ncomp.get(), n_labels,
                       bad.get());
    KERNEL_OK();
    unsigned long long out = 0;
    to_host(&out, bad.get(), 1, s);
    sync_stream(s);
    return out;
}

// ---- allocation-free quality ----------------------------------------------------------------
// A variant of sc_quality for callers that cannot allocate device memory. Its working memory is
// one file-scope device array per GPU, laid out as: the largest label (8 bytes), the partials
// of the fixed-shape sum, the per-community fixed-point totals (in, K, W: 3 x 8 bytes per
// community) and the per-community terms (8 bytes per community).
constexpr uint64_t NOALLOC_MAX_COMMUNITIES = 1ull << 20;
constexpr size_t NA_MAX_OFF = 0;
constexpr size_t NA_PART_OFF = 256;
constexpr size_t NA_ACC_OFF = (NA_PART_OFF + RED_PARTIALS * sizeof(double) + 255) / 256 * 256;
constexpr size_t NA_TERM_OFF = NA_ACC_OFF + 3 * NOALLOC_MAX_COMMUNITIES * sizeof(long long);
constexpr size_t NA_BYTES = NA_TERM_OFF + NOALLOC_MAX_COMMUNITIES * sizeof(double);  // ~32 MiB

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// Device address of this GPU's copy of the array (the array's name is a host-side stand-in and
// must not reach a kernel): looked up once per device, with that device current.
static unsigned char *scratch_address() {
    static std::mutex mu;
    static unsigned char *addr[64] = {nullptr};
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw ScError(SC_DEVICE_ERROR, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        addr[dev] = static_cast<unsigned char *>(p);
    }
    return addr[dev];
}

__global__ void k_zero_u64(unsigned long long *p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// quality_device's computation (same kernels, grids and summation order, so the result is
// bit-identical) with every buffer in the scratch array. Zeroing is done by kernels, the only
// copies are two 8-byte readbacks into host memory, and the stream is synchronized twice.
static double quality_noalloc(const sc_graph &g, const uint64_t *d_labels, int mode, double gamma, hipStream_t s) {
    const uint64_t n = g.n;
    if (n == 0) return 0.0;
    unsigned char *base = scratch_address();
    unsigned long long *mx = reinterpret_cast<unsigned long long *>(base + NA_MAX_OFF);
    double *part = reinterpret_cast<double *>(base + NA_PART_OFF);
    long long *acc = reinterpret_cast<long long *>(base + NA_ACC_OFF);
    double *term = reinterpret_cast<double *>(base + NA_TERM_OFF);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, mx, (uint64_t)1);
    KERNEL_OK();
    hipLaunchKernelGGL(k_max_label, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, d_labels, n, mx);
    KERNEL_OK();
    unsigned long long max_label = 0;
    HIP_OK(hipMemcpyAsync(&max_label, mx, sizeof(max_label), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (max_label >= NOALLOC_MAX_COMMUNITIES) {
        ScError e(SC_INVALID_ARGUMENT, "label " + std::to_string(max_label) +
                                           " exceeds the allocation-free limit: labels must be below 2^20 (1048576)");
        e.payload.got = max_label;
        e.payload.expected = NOALLOC_MAX_COMMUNITIES - 1;
        throw e;
    }
    const uint64_t nl = max_label + 1;
    const int s_in = fixed_scale_for(g.total_weight + 1e-300);
    const int sK = fixed_scale_for(2.0 * g.total_weight + 1e-300);
    const int sW = fixed_scale_for(g.total_node_weight + 1e-300);
    hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(3 * nl, 256, 4096)), dim3(256), 0, s,
                       reinterpret_cast<unsigned long long *>(acc), 3 * nl);
    KERNEL_OK();
    hipLaunchKernelGGL((k_vertex_terms<8>), dim3(grid_for(n, 32, 1024)), dim3(256), 0, s, d_labels, n, g.ptrs.get(),
                       g.nbrs.get(), g.wts.get(), g.strength.get(), g.node_w.get(), s_in, sK, sW, acc, acc + nl,
                       acc + 2 * nl);
    KERNEL_OK();
    hipLaunchKernelGGL(k_community_terms, dim3(grid_for(nl, 256, 4096)), dim3(256), 0, s, acc, acc + nl,
                       acc + 2 * nl, nl, s_in, sK, sW, mode, gamma, g.total_weight, term);
    KERNEL_OK();
    device_sum_f64_async(term, nl, part, s);
    double q = 0.0;
    HIP_OK(hipMemcpyAsync(&q, part + RED_BLOCKS, sizeof(q), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return q;
}

}  // namespace scgpu

using namespace scgpu;

extern "C" {

sc_status sc_quality_noalloc(const sc_graph *g, const uint64_t *d_labels, int32_t objective, double resolution,
                             double *out, sc_error *err) {
    try {
        if (!g || !out || (!d_labels && g->n)) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        if (objective != SC_RB && objective != SC_CPM) throw ScError(SC_INVALID_ARGUMENT, "unknown objective");
        // the default (null) stream: nothing is created on first use
        *out = quality_noalloc(*g, d_labels, objective == SC_RB ? 1 : 2, resolution, hipStream_t(0));
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

double sc_modularity(const sc_graph *g, const uint64_t *labels, double resolution) {
    try {
        if (!g || (!labels && g->n)) return NAN;
        hipStream_t s = thread_stream();
        DevBuf<uint64_t> d(g->n);
        to_device(d.get(), labels, g->n, s);
        double q = quality_device(*g, d.get(), 0, resolution, s);
        sync_stream(s);
        return q;
    } catch (...) {
        stage_drop();
        return NAN;
    }
}

sc_status sc_quality(const sc_graph *g, const uint64_t *labels, int32_t objective, double resolution,
                     double *out, sc_error *err) {
    try {
        if (!g || !out || (!labels && g->n)) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        if (objective != SC_RB && objective != SC_CPM) throw ScError(SC_INVALID_ARGUMENT, "unknown objective");
        hipStream_t s = thread_stream();
        DevBuf<uint64_t> d(g->n);
        to_device(d.get(), labels, g->n, s);
        *out = quality_device(*g, d.get(), objective == SC_RB ? 1 : 2, resolution, s);
        sync_stream(s);
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

sc_status sc_disconnected_community_count(const sc_graph *g, const uint64_t *labels, uint64_t *out,
                                          sc_error *err) {
    try {
        if (!g || !out || (!labels && g->n)) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        uint64_t n = g->n;
        uint64_t maxl = 0;
        for (uint64_t v = 0; v < n; v++) maxl = std::max(maxl, labels[v]);
        if (n && maxl >= 0xFFFFFFFFull) throw ScError(SC_INVALID_ARGUMENT, "labels must be below 2^32 - 1");
        std::vector<uint32_t> l32(n);
        for (uint64_t v = 0; v < n; v++) l32[v] = (uint32_t)labels[v];
        hipStream_t s = thread_stream();
        DevBuf<uint32_t> d(n);
        to_device(d.get(), l32.data(), n, s);
        *out = disconnected_count_device(*g, d.get(), n ? maxl + 1 : 0, s);
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

}  // extern "C"
