// common.hpp -- shared host/device infrastructure of libscgpu (gfx950 only).
//
// Error model: internal code throws ScError; every extern "C" entry point catches it and turns
// it into an sc_status + sc_error, so no exception crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/scgpu.h"

namespace scgpu {

struct ScError {
    sc_status code;
    std::string msg;
    sc_error payload;
    ScError(sc_status c, std::string m) : code(c), msg(std::move(m)) {
        std::memset(&payload, 0, sizeof(payload));
    }
};

[[noreturn]] void throw_hip(hipError_t e, const char *what, const char *file, int line);

#define HIP_OK(expr)                                                   \
    do {                                                               \
        hipError_t _e = (expr);                                        \
        if (_e != hipSuccess) ::scgpu::throw_hip(_e, #expr, __FILE__, __LINE__); \
    } while (0)

#define KERNEL_OK() HIP_OK(hipGetLastError())

void fill_error(sc_error *err, const ScError &e);

// Call only inside a catch (...) block of an extern "C" entry point: maps the exception in
// flight to an sc_status and fills err. ScError keeps its code; std::bad_alloc (host memory)
// becomes SC_OUT_OF_MEMORY; anything else SC_DEVICE_ERROR with its text. No exception crosses
// the C ABI.
sc_status status_from_exception(sc_error *err) noexcept;

// ---- caching device allocator -----------------------------------------------------------
// hipMalloc is slow and synchronizing; blocks are cached by rounded size and reused. A block
// is returned to the cache only after the owning call has synchronized its stream.
class DevicePool {
  public:
    // The calling thread's pool.
    static DevicePool &get();
    void *alloc(size_t bytes);
    // Returns a block to its owner: cached for reuse when the calling thread owns the pool,
    // hipFree'd at once otherwise.
    static void return_block(DevicePool *owner, void *p);
    void trim();
    size_t cached_bytes();
    size_t live_bytes();
    // Frees every block, cached or live (process-exit teardown only).
    void release_all();

  private:
    void release(void *p);
    void release_foreign(void *p);
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> live_;
    size_t cached_ = 0;
    size_t live_bytes_ = 0;
};

template <typename T>
class DevBuf {
  public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { resize(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_), pool_(o.pool_) { o.p_ = nullptr; o.n_ = 0; o.pool_ = nullptr; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            reset();
            p_ = o.p_; n_ = o.n_; pool_ = o.pool_;
            o.p_ = nullptr; o.n_ = 0; o.pool_ = nullptr;
        }
        return *this;
    }
    ~DevBuf() { reset(); }
    void resize(size_t n) {
        reset();
        DevicePool &pool = DevicePool::get();
        p_ = static_cast<T *>(pool.alloc((n ? n : 1) * sizeof(T)));
        n_ = n;
        pool_ = &pool;
    }
    void reset() {
        if (p_) DevicePool::return_block(pool_, p_);
        p_ = nullptr;
        n_ = 0;
        pool_ = nullptr;
    }
    T *get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

  private:
    T *p_ = nullptr;
    size_t n_ = 0;
    DevicePool *pool_ = nullptr;  // the pool (of the allocating thread) the block belongs to
};

// Stream owned by the calling thread (default priority), created lazily.
hipStream_t thread_stream();

// Waits until the stream's work is done. It polls first (a blocking wait can put the thread
// to sleep, and waking it costs tens of microseconds per small readback) and blocks only when
// the work takes longer than a couple of milliseconds.
void sync_stream(hipStream_t s);

// Profiling switch: when on, the Leiden driver records events around its decide kernels.
bool profiling_enabled();
void set_profiling(int on);

// ---- fixed point -----------------------------------------------------------------------
// Community totals and neighbour-community weights are accumulated as int64 fixed point with
// a per-level power-of-two scale, so atomic accumulation is order independent (deterministic)
// while keeping ~2^-61 relative resolution of the largest possible sum.
inline int fixed_scale_for(double max_sum) {
    if (!(max_sum > 0.0)) return 40;
    int e;
    std::frexp(max_sum, &e);  // max_sum < 2^e
    return 61 - e;
}

__device__ __forceinline__ long long to_fx(double x, int scale) {
    return __double2ll_rn(ldexp(x, scale));
}

__device__ __forceinline__ double from_fx(long long x, int scale) {
    return ldexp((double)x, -scale);
}

// ---- hashing / counter-based randomness ---------------------------------------------------
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

__host__ __device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Uniform in (0, 1) from 53 random bits, never exactly 0 or 1.
__device__ __forceinline__ double u01_open(uint64_t bits) {
    return ((double)(bits >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// ---- launch helpers ---------------------------------------------------------------------
inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 65535u * 16u) {
    uint64_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// Stored entries of any graph on the device stay below this bound: entry positions, run
// indices and scans over entries are 32-bit. check_entry_count throws SC_INVALID_ARGUMENT
// for larger graphs before any device work.
constexpr uint64_t MAX_ENTRIES = 0xFFFFFFFFull;
void check_entry_count(uint64_t nnz);
// Nodes (points, jobs' vertices together) of any device run: below 2^31, so 32-bit vertex
// loops (index + grid stride) never wrap and community ids, which span [0, 2n), stay below
// 2^32 - 1. Above u32::MAX the reference's TooManyNodes message, between the two the device
// limit; both SC_TOO_MANY_NODES with n_nodes set.
constexpr uint64_t MAX_NODES = (1ull << 31) - 1;
void check_node_count(uint64_t n);

// Device-wide helpers implemented with rocprim in util.hip.
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, hipStream_t s);
void exclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, hipStream_t s);
void radix_sort_pairs_u64_f32(uint64_t *keys, float *vals, uint64_t n, int end_bit, hipStream_t s);
void radix_sort_pairs_u64_u32(uint64_t *keys, uint32_t *vals, uint64_t n, int end_bit, hipStream_t s);
void radix_sort_pairs_u32_u32(uint32_t *keys, uint32_t *vals, uint64_t n, int end_bit, hipStream_t s);
// Sorts the first n pairs; the buffers may come back exchanged for the sort's second buffers
// (of size n), so callers must re-read .get() afterwards.
void radix_sort_pairs(DevBuf<uint64_t> &keys, DevBuf<float> &vals, uint64_t n, int end_bit, hipStream_t s);
void radix_sort_pairs(DevBuf<uint32_t> &keys, DevBuf<uint32_t> &vals, uint64_t n, int end_bit, hipStream_t s);
void radix_sort_pairs(DevBuf<uint32_t> &keys, DevBuf<uint64_t> &vals, uint64_t n, int end_bit, hipStream_t s);
// Deterministic sum of n doubles (fixed two-level tree: RED_BLOCKS strided partial sums of
// RED_THREADS threads, then one block).
constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;
constexpr uint64_t RED_PARTIALS = RED_BLOCKS + 1;
// The same sum (identical association order) into caller-provided partials of RED_PARTIALS
// doubles, n >= 1: the result is left in part[RED_BLOCKS]; no allocation, no synchronization.
void device_sum_f64_async(const double *in, uint64_t n, double *part, hipStream_t s);
double device_sum_f64(const double *in, uint64_t n, hipStream_t s);
double device_max_f64(const double *in, uint64_t n, hipStream_t s);
// The same four fixed-shape reductions as device_sum_f64 / device_max_f64 (identical
// association order) in one launch pair and one synchronisation: out[i] = sum (or max, when
// is_max[i]) of in[i][0..n). Null inputs give 0.
void device_reduce4_f64(const double *const in[4], const bool is_max[4], uint64_t n, double out[4], hipStream_t s);
// Per-segment versions of the same fixed-shape reduction: out[j] = sum (or max) of
// in[ranges[j] .. ranges[j+1]) with exactly the association order device_sum_f64 /
// device_max_f64 use for an array of that length (one launch pair for all segments).
void device_segment_reduce_f64(const double *in, const uint64_t *d_ranges, uint32_t nseg, bool is_max, double *d_out,
                               hipStream_t s);

// Workgroup reductions (blockDim.x a multiple of 64, at most 1024); every thread must call them.
// The result is valid in thread 0.
template <typename T, bool MAX>
__device__ __forceinline__ T block_reduce(T v) {
    __shared__ T sh[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const T o = __shfl_xor(v, off);
        v = MAX ? (o > v ? o : v) : v + o;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned w = 1; w < blockDim.x / 64; w++) v = MAX ? (sh[w] > v ? sh[w] : v) : v + sh[w];
    }
    return v;
}
template <typename T>
__device__ __forceinline__ T block_sum(T v) { return block_reduce<T, false>(v); }
template <typename T>
__device__ __forceinline__ T block_max(T v) { return block_reduce<T, true>(v); }

// Exponent of the lowest set bit of x (x = m * 2^q with m odd), or a large value for 0. A sum
// of terms that are all multiples of 2^q, whose absolute values add up to less than 2^(53+q),
// has every partial sum exactly representable in f64: any association order then yields the
// sequential result bit for bit.
__device__ __forceinline__ int f64_quantum(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x) & 0x7FFFFFFFFFFFFFFFull;
    if (b == 0ull) return 4096;
    const int ex = (int)(b >> 52);
    const unsigned long long frac = b & 0xFFFFFFFFFFFFFull;
    const unsigned long long mant = ex ? (frac | (1ull << 52)) : frac;
    return (ex ? ex : 1) - 1075 + __builtin_ctzll(mant);
}

inline int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b) != 0) b++;
    return b;
}

template <typename T>
void to_device(T *dst, const T *src, size_t n, hipStream_t s) {
    if (n) HIP_OK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}
// Small device-to-host readbacks go through a per-thread pinned staging buffer (a copy into
// pageable memory waits inside the copy call); the values reach dst when sync_stream returns.
bool stage_d2h(void *dst, const void *src, size_t bytes, hipStream_t s);
void stage_drop();

// dst is valid after the next sync_stream(s).
template <typename T>
void to_host(T *dst, const T *src, size_t n, hipStream_t s) {
    if (!n) return;
    if (stage_d2h(dst, src, n * sizeof(T), s)) return;
    HIP_OK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

}  // namespace scgpu

// The device-resident graph behind sc_graph (CSRNetworkData, csr_network.rs:39-60), plus the
// per-node self-loop weight, cached so kernels never binary-search a row for it.
namespace scgpu {
// Degree classes of the sweeps' decide step (the LDS table layout each class uses); the last
// class keeps its tables in global memory.
constexpr int NUM_BUCKETS = 8;
constexpr int GTAB_BUCKET = 7;

__host__ __device__ __forceinline__ int bucket_of(uint64_t d) {
    if (d <= 8) return 0;
    if (d <= 16) return 1;
    if (d <= 32) return 2;
    if (d <= 128) return 3;
    if (d <= 512) return 4;
    if (d <= 1024) return 5;
    if (d <= 2048) return 6;
    return GTAB_BUCKET;
}

// Upper degree bounds of the LDS degree classes 0..6 (class 7 takes the rest). Each bound is at
// most the class's capacity (the defaults, which bucket_of uses): a vertex may go to a class
// with more lanes and a larger table than its degree needs, never to a smaller one, so any
// bounds give identical decisions.
struct BucketBounds {
    uint32_t hi[NUM_BUCKETS - 1];
};
constexpr BucketBounds DEFAULT_BUCKET_BOUNDS{{8, 16, 32, 128, 512, 1024, 2048}};

__host__ __device__ __forceinline__ int bucket_of(uint64_t d, const BucketBounds &b) {
#pragma unroll
    for (int k = 0; k < NUM_BUCKETS - 1; k++)
        if (d <= b.hi[k]) return k;
    return GTAB_BUCKET;
}
}  // namespace scgpu

struct sc_graph {
    uint64_t n = 0;
    uint64_t nnz = 0;
    scgpu::DevBuf<uint64_t> ptrs;      // n + 1
    scgpu::DevBuf<uint32_t> nbrs;      // nnz, strictly increasing per row
    scgpu::DevBuf<float> wts;          // nnz
    scgpu::DevBuf<double> node_w;      // n
    scgpu::DevBuf<double> strength;    // n (self-loops twice)
    scgpu::DevBuf<double> self_loop;   // n
    scgpu::DevBuf<double> row_total;   // n: the row's share (u >= v) of total_weight
    double total_weight = 0.0;
    uint64_t edge_count = 0;
    bool any_self_loops = false;
    double total_node_weight = 0.0;
    double max_strength = 0.0;
    uint32_t max_degree = 0;
    // vertices per degree class (bucket_of), counted when the graph was finished
    bool has_bucket_counts = false;
    uint32_t bucket_counts[scgpu::NUM_BUCKETS] = {0};
};

namespace scgpu {
// Builds the cached aggregates of a graph whose rows are already sorted, merged and free of
// zero weights (CSRNetwork::finish's second half, csr_network.rs:313-351): strengths,
// self-loops, total weight, edge count, and the degree-sum check. count_entries: g.nnz is not
// known on the host yet and is read from ptrs[n] with the other row statistics (the nbrs/wts
// buffers may then be larger than the entries they hold).
void graph_finish_sorted(sc_graph &g, hipStream_t s, bool count_entries = false);
// Full finish from arbitrary rows: sort, merge duplicates in f64, drop zero sums, compact.
void graph_finish_unsorted(sc_graph &g, hipStream_t s);
}  // namespace scgpu
