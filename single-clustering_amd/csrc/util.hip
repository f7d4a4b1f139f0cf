// util.hip -- allocator, streams, error plumbing and device-wide primitives.
#include "common.hpp"

#include <algorithm>
#include <vector>
#include <tuple>
#include <chrono>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <atomic>
#include <cstdlib>

namespace scgpu {

void throw_hip(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "HIP error %d (%s) at %s:%d in %s", (int)e,
                  hipGetErrorString(e), file, line, what);
    sc_status code = (e == hipErrorOutOfMemory) ? SC_OUT_OF_MEMORY : SC_DEVICE_ERROR;
    throw ScError(code, buf);
}

void fill_error(sc_error *err, const ScError &e) {
    stage_drop();  // the failed call's pending readbacks point into frames that are gone
    if (!err) return;
    *err = e.payload;
    err->code = (int32_t)e.code;
    std::snprintf(err->msg, sizeof(err->msg), "%s", e.msg.c_str());
}

sc_status status_from_exception(sc_error *err) noexcept {
    try {
        throw;
    } catch (const ScError &e) {
        fill_error(err, e);
        return e.code;
    } catch (const std::bad_alloc &) {
        fill_error(err, ScError(SC_OUT_OF_MEMORY, "host memory allocation failed"));
        return SC_OUT_OF_MEMORY;
    } catch (const std::exception &x) {
        fill_error(err, ScError(SC_DEVICE_ERROR, std::string("internal error: ") + x.what()));
        return SC_DEVICE_ERROR;
    } catch (...) {
        fill_error(err, ScError(SC_DEVICE_ERROR, "internal error: unknown exception"));
        return SC_DEVICE_ERROR;
    }
}

// ---- process-wide registry of the library's HIP objects ------------------------------------
// Streams, pinned staging buffers and device pools are created per host thread. They are
// registered here so that a thread's own objects are released when the thread ends, and so that
// everything still held is released once, in order, by an atexit handler that runs before the
// HIP runtime's own teardown (it is registered after the runtime's first use). After that
// teardown no HIP call is made: blocks released later are dropped.
namespace {
struct Registry {
    std::mutex mu;
    std::vector<DevicePool *> pools;
    std::vector<hipStream_t> streams;
    std::vector<void *> pinned;
    bool torn_down = false;
};
Registry &registry() {
    static Registry *r = new Registry();  // never destroyed: used by the atexit handler
    return *r;
}
std::once_flag g_atexit_once;

void teardown_at_exit() {
    Registry &r = registry();
    std::vector<DevicePool *> pools;
    std::vector<hipStream_t> streams;
    std::vector<void *> pinned;
    {
        std::lock_guard<std::mutex> lk(r.mu);
        if (r.torn_down) return;
        r.torn_down = true;
        pools.swap(r.pools);
        streams.swap(r.streams);
        pinned.swap(r.pinned);
    }
    (void)hipDeviceSynchronize();
    for (DevicePool *p : pools) p->release_all();
    for (hipStream_t st : streams) (void)hipStreamDestroy(st);
    for (void *b : pinned) (void)hipHostFree(b);
    (void)hipGetLastError();
}

// Called after each successful creation of a HIP object: the handler then runs before the
// runtime's own exit-time teardown, which was registered when the runtime started.
void ensure_teardown_registered() {
    std::call_once(g_atexit_once, [] { std::atexit(teardown_at_exit); });
}

bool torn_down() {
    Registry &r = registry();
    std::lock_guard<std::mutex> lk(r.mu);
    return r.torn_down;
}

template <typename T>
void unregister(std::vector<T> &v, T x) {
    for (size_t i = 0; i < v.size(); i++)
        if (v[i] == x) { v[i] = v.back(); v.pop_back(); return; }
}

// Ends with its host thread (before the process-exit handlers for the main thread): releases
// the thread's cached blocks, its stream and its staging buffer.
struct ThreadObjects {
    DevicePool *pool = nullptr;
    hipStream_t stream = nullptr;
    char *stage = nullptr;
    ~ThreadObjects() {
        Registry &r = registry();
        bool live;
        {
            std::lock_guard<std::mutex> lk(r.mu);
            live = !r.torn_down;
            if (live) {
                if (stream) unregister(r.streams, stream);
                if (stage) unregister(r.pinned, (void *)stage);
            }
        }
        if (!live) return;
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (stage) (void)hipHostFree(stage);
        // the pool object stays (blocks of live graphs still name it as their owner); its cached
        // blocks are freed now
        if (pool) pool->trim();
        (void)hipGetLastError();
    }
};
thread_local ThreadObjects t_objs;
}  // namespace

// One pool per host thread: every block is recycled on the stream of the thread that freed it,
// so stream order alone protects reuse (work of another thread never sees a recycled block). A
// block freed on another thread than the one that allocated it (a graph handle built on one
// thread and freed on another) is hipFree'd at once instead: hipFree waits for the device, so
// no queued work of the owning thread can still use it.
DevicePool &DevicePool::get() {
    DevicePool *&p = t_objs.pool;
    if (!p) {
        p = new DevicePool();  // never destroyed: blocks of live graphs keep a pointer to it
        Registry &r = registry();
        std::lock_guard<std::mutex> lk(r.mu);
        r.pools.push_back(p);
    }
    return *p;
}

static size_t round_size(size_t b) {
    if (b <= 4096) return 4096;
    if (b <= (size_t(1) << 20)) {
        size_t r = 4096;
        while (r < b) r <<= 1;
        return r;
    }
    const size_t mb = size_t(1) << 20;
    return (b + mb - 1) / mb * mb;
}

void *DevicePool::alloc(size_t bytes) {
    size_t sz = round_size(bytes);
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = free_.lower_bound(sz);
        // reuse a cached block when it is not more than 25% larger than needed
        if (it != free_.end() && it->first <= sz + sz / 4) {
            void *p = it->second;
            size_t have = it->first;
            free_.erase(it);
            cached_ -= have;
            live_[p] = have;
            live_bytes_ += have;
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();
        HIP_OK(hipMalloc(&p, sz));
    }
    ensure_teardown_registered();
    std::lock_guard<std::mutex> lk(mu_);
    live_[p] = sz;
    live_bytes_ += sz;
    return p;
}

void DevicePool::release(void *p) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = live_.find(p);
    if (it == live_.end()) return;
    free_.emplace(it->second, p);
    cached_ += it->second;
    live_bytes_ -= it->second;
    live_.erase(it);
}

void DevicePool::release_foreign(void *p) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = live_.find(p);
        if (it == live_.end()) return;
        live_bytes_ -= it->second;
        live_.erase(it);
    }
    if (!torn_down()) (void)hipFree(p);
}

void DevicePool::trim() {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto &kv : free_) (void)hipFree(kv.second);
    free_.clear();
    cached_ = 0;
}

void DevicePool::release_all() {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto &kv : free_) (void)hipFree(kv.second);
    for (auto &kv : live_) (void)hipFree(kv.first);
    free_.clear();
    live_.clear();
    cached_ = 0;
    live_bytes_ = 0;
}

void DevicePool::return_block(DevicePool *owner, void *p) {
    if (!p || !owner) return;
    if (torn_down()) return;  // the exit handler freed every block already
    if (owner == &get()) owner->release(p);
    else owner->release_foreign(p);
}

size_t DevicePool::cached_bytes() {
    std::lock_guard<std::mutex> lk(mu_);
    return cached_;
}

size_t DevicePool::live_bytes() {
    std::lock_guard<std::mutex> lk(mu_);
    return live_bytes_;
}

namespace {
struct HostStage {
    bool failed = false;
    size_t used = 0;
    hipStream_t stream = nullptr;
    std::vector<std::tuple<void *, size_t, size_t>> pending;  // (dst, offset, bytes)
};
thread_local HostStage t_stage;
constexpr size_t STAGE_BYTES = 1 << 16;
constexpr size_t STAGE_MAX_COPY = 4096;

void stage_flush(hipStream_t s) {
    HostStage &st = t_stage;
    if (st.pending.empty() || st.stream != s) return;
    for (auto &p : st.pending) std::memcpy(std::get<0>(p), t_objs.stage + std::get<1>(p), std::get<2>(p));
    st.pending.clear();
    st.used = 0;
    st.stream = nullptr;
}
}  // namespace

bool stage_d2h(void *dst, const void *src, size_t bytes, hipStream_t s) {
    HostStage &st = t_stage;
    if (bytes > STAGE_MAX_COPY || st.failed) return false;
    if (!t_objs.stage) {
        char *b = nullptr;
        if (hipHostMalloc((void **)&b, STAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            st.failed = true;
            return false;
        }
        t_objs.stage = b;
        ensure_teardown_registered();
        Registry &r = registry();
        std::lock_guard<std::mutex> lk(r.mu);
        r.pinned.push_back(b);
    }
    if (!st.pending.empty() && st.stream != s) return false;
    const size_t off = (st.used + 15) & ~size_t(15);
    if (off + bytes > STAGE_BYTES) return false;
    HIP_OK(hipMemcpyAsync(t_objs.stage + off, src, bytes, hipMemcpyDeviceToHost, s));
    st.pending.emplace_back(dst, off, bytes);
    st.used = off + bytes;
    st.stream = s;
    return true;
}

// Readbacks of a call that failed before its sync are dropped (their destinations are gone).
void stage_drop() {
    t_stage.pending.clear();
    t_stage.used = 0;
    t_stage.stream = nullptr;
}

void sync_stream(hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) HIP_OK(e);
        if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(2000)) {
            HIP_OK(hipStreamSynchronize(s));
            break;
        }
    }
    stage_flush(s);
}

hipStream_t thread_stream() {
    hipStream_t &s = t_objs.stream;
    if (!s) {
        HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        ensure_teardown_registered();
        Registry &r = registry();
        std::lock_guard<std::mutex> lk(r.mu);
        r.streams.push_back(s);
    }
    return s;
}

static std::atomic<int> g_profiling{-1};

bool profiling_enabled() {
    int v = g_profiling.load();
    if (v < 0) {
        const char *e = std::getenv("SCGPU_PROFILE");
        v = (e && e[0] == '1') ? 1 : 0;
        g_profiling.store(v);
    }
    return v == 1;
}

void set_profiling(int on) { g_profiling.store(on ? 1 : 0); }

// ---- rocprim wrappers ----------------------------------------------------------------------
// Item counts are passed as size_t (rocprim's own type), so counts of 2^31 and more are not
// truncated. The callers' 32-bit outputs (run indices, positions) stay in range because every
// graph is limited to fewer than 2^32 - 1 stored entries (check_entry_count).
template <typename T>
static void exclusive_scan(const T *in, T *out, uint64_t n, hipStream_t s) {
    if (n == 0) return;
    size_t tmp = 0;
    HIP_OK(rocprim::exclusive_scan(nullptr, tmp, in, out, T(0), (size_t)n, rocprim::plus<T>(), s));
    DevBuf<unsigned char> t(tmp);
    HIP_OK(rocprim::exclusive_scan(t.get(), tmp, in, out, T(0), (size_t)n, rocprim::plus<T>(), s));
}

void exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, hipStream_t s) { exclusive_scan(in, out, n, s); }
void exclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, hipStream_t s) { exclusive_scan(in, out, n, s); }

void check_node_count(uint64_t n) {
    if (n <= MAX_NODES) return;
    ScError e(SC_TOO_MANY_NODES,
              n > 0xFFFFFFFFull
                  ? "graph has " + std::to_string(n) + " nodes, exceeding the 4294967295 addressable by u32 adjacency"
                  : "graph has " + std::to_string(n) + " nodes, exceeding the device limit of 2147483647 (2^31 - 1) nodes");
    e.payload.n_nodes = n;
    throw e;
}

void check_entry_count(uint64_t nnz) {
    if (nnz >= (uint64_t)MAX_ENTRIES) {
        ScError e(SC_INVALID_ARGUMENT, "a graph of " + std::to_string(nnz) +
                                           " stored entries exceeds the device limit of 2^32 - 2 entries");
        e.payload.got = nnz;
        e.payload.expected = (uint64_t)MAX_ENTRIES - 1;
        throw e;
    }
}

// rocprim sorts up to 2^20 items of 4-byte or wider keys by block sort + merge passes (about
// 2 log2(n / 1024) launches), whatever the key width. The keys here are group ids or packed
// ids of few bits: up to 24 bits the onesweep radix sort needs at most three digit passes
// (8 bits each) plus a histogram and a scan launch, and is used for every size above one
// block. Both orders are stable, so the result is the same.
using ShortKeySort = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                                rocprim::default_config, 0>;

// Stable sort of n pairs by the low end_bit bits of the keys, ping-ponging between (keys,
// vals) and (k2, v2); returns true when the result is in (k2, v2).
template <typename K, typename V>
static bool sort_pairs_db(K *keys, V *vals, K *k2, V *v2, uint64_t n, int end_bit, hipStream_t s) {
    size_t tmp = 0;
    if (end_bit <= 24) {
        rocprim::double_buffer<K> kb(keys, k2);
        rocprim::double_buffer<V> vb(vals, v2);
        HIP_OK(rocprim::radix_sort_pairs<ShortKeySort>(nullptr, tmp, kb, vb, (size_t)n, 0u, (unsigned)end_bit, s));
        DevBuf<unsigned char> t(tmp);
        HIP_OK(rocprim::radix_sort_pairs<ShortKeySort>(t.get(), tmp, kb, vb, (size_t)n, 0u, (unsigned)end_bit, s));
        return kb.current() != keys;  // keys and values always move together
    }
    rocprim::double_buffer<K> kb(keys, k2);
    rocprim::double_buffer<V> vb(vals, v2);
    HIP_OK(rocprim::radix_sort_pairs(nullptr, tmp, kb, vb, (size_t)n, 0u, (unsigned)end_bit, s));
    DevBuf<unsigned char> t(tmp);
    HIP_OK(rocprim::radix_sort_pairs(t.get(), tmp, kb, vb, (size_t)n, 0u, (unsigned)end_bit, s));
    return kb.current() != keys;
}

template <typename K, typename V>
static void sort_pairs(K *keys, V *vals, uint64_t n, int end_bit, hipStream_t s) {
    if (n <= 1) return;
    DevBuf<K> k2(n);
    DevBuf<V> v2(n);
    if (sort_pairs_db(keys, vals, k2.get(), v2.get(), n, end_bit, s)) {
        HIP_OK(hipMemcpyAsync(keys, k2.get(), n * sizeof(K), hipMemcpyDeviceToDevice, s));
        HIP_OK(hipMemcpyAsync(vals, v2.get(), n * sizeof(V), hipMemcpyDeviceToDevice, s));
    }
    // the temporaries return to the pool; keep them alive until the copies above are queued
    // (stream order guarantees later users of the recycled blocks run after the copies)
}

void radix_sort_pairs_u64_f32(uint64_t *keys, float *vals, uint64_t n, int end_bit, hipStream_t s) {
    sort_pairs(keys, vals, n, end_bit, s);
}

// Sorts the first n pairs; the buffers may come back swapped with the sort's second buffers
// (their sizes become n), which saves the copy back when the last pass wrote the other buffer.
template <typename K, typename V>
static void sort_pairs_bufs(DevBuf<K> &keys, DevBuf<V> &vals, uint64_t n, int end_bit, hipStream_t s) {
    if (n <= 1) return;
    DevBuf<K> k2(n);
    DevBuf<V> v2(n);
    if (sort_pairs_db(keys.get(), vals.get(), k2.get(), v2.get(), n, end_bit, s)) {
        std::swap(keys, k2);
        std::swap(vals, v2);
    }
}

void radix_sort_pairs(DevBuf<uint64_t> &keys, DevBuf<float> &vals, uint64_t n, int end_bit, hipStream_t s) {
    sort_pairs_bufs(keys, vals, n, end_bit, s);
}
void radix_sort_pairs(DevBuf<uint32_t> &keys, DevBuf<uint32_t> &vals, uint64_t n, int end_bit, hipStream_t s) {
    sort_pairs_bufs(keys, vals, n, end_bit, s);
}
void radix_sort_pairs(DevBuf<uint32_t> &keys, DevBuf<uint64_t> &vals, uint64_t n, int end_bit, hipStream_t s) {
    sort_pairs_bufs(keys, vals, n, end_bit, s);
}
void radix_sort_pairs_u64_u32(uint64_t *keys, uint32_t *vals, uint64_t n, int end_bit, hipStream_t s) {
    sort_pairs(keys, vals, n, end_bit, s);
}
void radix_sort_pairs_u32_u32(uint32_t *keys, uint32_t *vals, uint64_t n, int end_bit, hipStream_t s) {
    sort_pairs(keys, vals, n, end_bit, s);
}

// ---- deterministic reductions -----------------------------------------------------------
// Fixed shape: 1024 blocks x 256 threads of strided partial sums, then one block. The
// association order depends only on n, so results are bit-reproducible.

template <bool MAX>
__global__ void __launch_bounds__(RED_THREADS) k_reduce_stage(const double *in, uint64_t n, double *out) {
    __shared__ double sh[RED_THREADS];
    double acc = MAX ? -INFINITY : 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * RED_THREADS) {
        double x = in[i];
        acc = MAX ? fmax(acc, x) : acc + x;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double o = sh[threadIdx.x + w];
            sh[threadIdx.x] = MAX ? fmax(sh[threadIdx.x], o) : sh[threadIdx.x] + o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

template <bool MAX>
static void reduce_f64_launch(const double *in, uint64_t n, double *part, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_stage<MAX>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, in, n, part);
    KERNEL_OK();
    hipLaunchKernelGGL(k_reduce_stage<MAX>, dim3(1), dim3(RED_THREADS), 0, s, part, (uint64_t)RED_BLOCKS,
                       part + RED_BLOCKS);
    KERNEL_OK();
}

template <bool MAX>
static double reduce_f64(const double *in, uint64_t n, hipStream_t s) {
    if (n == 0) {
        sync_stream(s);  // like the general path: the stream is synchronized on return
        return MAX ? -INFINITY : 0.0;
    }
    DevBuf<double> part(RED_PARTIALS);
    reduce_f64_launch<MAX>(in, n, part.get(), s);
    double r = 0.0;
    to_host(&r, part.get() + RED_BLOCKS, 1, s);
    sync_stream(s);
    return r;
}

void device_sum_f64_async(const double *in, uint64_t n, double *part, hipStream_t s) {
    reduce_f64_launch<false>(in, n, part, s);
}

struct Reduce4Args {
    const double *in[4];
    int is_max[4];
};

// blockIdx.y selects the array; per array the same strided partials and tree as k_reduce_stage.
__global__ void __launch_bounds__(RED_THREADS) k_reduce4_stage(Reduce4Args a, uint64_t n, double *out, int final) {
    __shared__ double sh[RED_THREADS];
    const int y = blockIdx.y;
    const bool mx = a.is_max[y] != 0;
    const double *in = final ? out + (size_t)y * RED_BLOCKS : a.in[y];
    const uint64_t len = final ? (uint64_t)RED_BLOCKS : n;
    double acc = mx ? -INFINITY : 0.0;
    if (in) {
        for (uint64_t i = (uint64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < len;
             i += (uint64_t)gridDim.x * RED_THREADS) {
            double x = in[i];
            acc = mx ? fmax(acc, x) : acc + x;
        }
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double o = sh[threadIdx.x + w];
            sh[threadIdx.x] = mx ? fmax(sh[threadIdx.x], o) : sh[threadIdx.x] + o;
        }
        __syncthreads();
    }
    // partials at out[y * RED_BLOCKS + block], results at out[4 * RED_BLOCKS + y]
    if (threadIdx.x == 0) out[final ? (size_t)4 * RED_BLOCKS + y : (size_t)y * RED_BLOCKS + blockIdx.x] = sh[0];
}

void device_reduce4_f64(const double *const in[4], const bool is_max[4], uint64_t n, double out[4], hipStream_t s) {
    if (n == 0) {
        for (int i = 0; i < 4; i++) out[i] = is_max[i] ? -INFINITY : 0.0;
        sync_stream(s);  // callers rely on this call having synchronized (and flushed readbacks)
        return;
    }
    Reduce4Args a;
    for (int i = 0; i < 4; i++) {
        a.in[i] = in[i];
        a.is_max[i] = is_max[i] ? 1 : 0;
    }
    DevBuf<double> part(4 * RED_BLOCKS + 4);
    hipLaunchKernelGGL(k_reduce4_stage, dim3(RED_BLOCKS, 4), dim3(RED_THREADS), 0, s, a, n, part.get(), 0);
    KERNEL_OK();
    hipLaunchKernelGGL(k_reduce4_stage, dim3(1, 4), dim3(RED_THREADS), 0, s, a, n, part.get(), 1);
    KERNEL_OK();
    to_host(out, part.get() + (size_t)4 * RED_BLOCKS, 4, s);
    sync_stream(s);
}

// Segment j is in[ranges[j] .. ranges[j+1]); blockIdx.y = segment. Stage 1 writes
// part[j * RED_BLOCKS + block], stage 2 reduces those into out[j], each with the strided
// partials and tree of k_reduce_stage for an array of the segment's length.
template <bool MAX>
__global__ void __launch_bounds__(RED_THREADS) k_segment_reduce(const double *in, const uint64_t *ranges,
                                                                 double *part, double *out, int final) {
    __shared__ double sh[RED_THREADS];
    const uint32_t j = blockIdx.y;
    const uint64_t lo = ranges[j], len = ranges[j + 1] - ranges[j];
    const double *src = final ? part + (size_t)j * RED_BLOCKS : in + lo;
    const uint64_t n = final ? (len ? (uint64_t)RED_BLOCKS : 0) : len;
    double acc = MAX ? -INFINITY : 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RED_THREADS) {
        double x = src[i];
        acc = MAX ? fmax(acc, x) : acc + x;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double o = sh[threadIdx.x + w];
            sh[threadIdx.x] = MAX ? fmax(sh[threadIdx.x], o) : sh[threadIdx.x] + o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (final) out[j] = len ? sh[0] : (MAX ? -INFINITY : 0.0);
        else part[(size_t)j * RED_BLOCKS + blockIdx.x] = sh[0];
    }
}

void device_segment_reduce_f64(const double *in, const uint64_t *d_ranges, uint32_t nseg, bool is_max, double *d_out,
                               hipStream_t s) {
    if (nseg == 0) return;
    DevBuf<double> part((size_t)nseg * RED_BLOCKS);
    if (is_max) {
        hipLaunchKernelGGL(k_segment_reduce<true>, dim3(RED_BLOCKS, nseg), dim3(RED_THREADS), 0, s, in, d_ranges,
                           part.get(), d_out, 0);
        KERNEL_OK();
        hipLaunchKernelGGL(k_segment_reduce<true>, dim3(1, nseg), dim3(RED_THREADS), 0, s, in, d_ranges, part.get(),
                           d_out, 1);
    } else {
        hipLaunchKernelGGL(k_segment_reduce<false>, dim3(RED_BLOCKS, nseg), dim3(RED_THREADS), 0, s, in, d_ranges,
                           part.get(), d_out, 0);
        KERNEL_OK();
        hipLaunchKernelGGL(k_segment_reduce<false>, dim3(1, nseg), dim3(RED_THREADS), 0, s, in, d_ranges, part.get(),
                           d_out, 1);
    }
    KERNEL_OK();
}

double device_sum_f64(const double *in, uint64_t n, hipStream_t s) { return reduce_f64<false>(in, n, s); }
double device_max_f64(const double *in, uint64_t n, hipStream_t s) { return reduce_f64<true>(in, n, s); }

}  // namespace scgpu

extern "C" {

const char *sc_version(void) { return "scgpu 0.1.0 (gfx950)"; }

void sc_set_profiling(int32_t enable) { scgpu::set_profiling(enable); }

void sc_shutdown(void) {
    (void)hipDeviceSynchronize();
    scgpu::DevicePool::get().trim();
}

void sc_pool_bytes(uint64_t *live_out, uint64_t *cached_out) {
    scgpu::DevicePool &p = scgpu::DevicePool::get();
    if (live_out) *live_out = p.live_bytes();
    if (cached_out) *cached_out = p.cached_bytes();
}

}

// ---- online autotuner ---------------------------------------------------------------------
#include "tune.hpp"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>

namespace scgpu {
namespace {
struct TuneEntry {
    std::vector<double> best;  // best time seen per candidate (ms), < 0 = not measured
    std::vector<int> count;    // measurements per candidate
};
std::mutex g_tune_mu;
std::map<std::string, TuneEntry> g_tune;
}  // namespace

bool tune_enabled() {
    static const bool on = [] {
        const char *e = std::getenv("SCGPU_AUTOTUNE");
        return !(e && e[0] == '0');
    }();
    return on;
}

TuneChoice tune_pick(const std::string &key, int n_candidates, int reps) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    TuneEntry &e = g_tune[key];
    if ((int)e.best.size() != n_candidates) {
        e.best.assign(n_candidates, -1.0);
        e.count.assign(n_candidates, 0);
    }
    int least = 0;
    for (int i = 1; i < n_candidates; i++)
        if (e.count[i] < e.count[least]) least = i;
    if (e.count[least] < reps) return TuneChoice{least, true};
    int b = 0;
    for (int i = 1; i < n_candidates; i++)
        if (e.best[i] < e.best[b]) b = i;
    return TuneChoice{b, false};
}

void tune_record(const std::string &key, int index, double ms) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    TuneEntry &e = g_tune[key];
    if (index < 0 || index >= (int)e.best.size()) return;
    e.count[index]++;
    if (e.best[index] < 0.0 || ms < e.best[index]) e.best[index] = ms;
}

std::string tune_report() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::string out;
    for (const auto &kv : g_tune) {
        out += kv.first + ":";
        int b = -1;
        for (int i = 0; i < (int)kv.second.best.size(); i++) {
            char buf[48];
            std::snprintf(buf, sizeof(buf), " %.3f", kv.second.best[i]);
            out += buf;
            if (kv.second.best[i] >= 0.0 && (b < 0 || kv.second.best[i] < kv.second.best[b])) b = i;
        }
        out += " -> " + std::to_string(b) + "\n";
    }
    return out;
}

void tune_reset() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
}

}  // namespace scgpu

extern "C" {

size_t sc_autotune_report(char *buf, size_t len) {
    const std::string r = scgpu::tune_report();
    if (buf && len) {
        const size_t c = std::min(len - 1, r.size());
        std::memcpy(buf, r.data(), c);
        buf[c] = 0;
    }
    return r.size();
}

void sc_autotune_reset(void) { scgpu::tune_reset(); }

}
