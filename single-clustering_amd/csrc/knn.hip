// knn.hip -- exact kNN on MFMA and the Gaussian connectivity graph builder.
//
// Replaces the neighbour search of neighborhood/mod.rs (kiddo kd-tree :51-202 / HNSW :359-578)
// with a brute-force contraction on the matrix cores, and restates the sigma / Gaussian weight /
// CSR steps (:104-181, :208-227) on the GPU.
//
// Pipeline:
//  1. prep: f32 rows -> zero-padded split-BF16 (hi, lo) copies and a padded f32 copy, f32 norms.
//  2. tile: 128 query rows per workgroup (two 16-row tiles per wave), points streamed through
//     LDS in stages of 64; the next stage is prefetched into registers while the current one is
//     on the matrix cores. s = |y|^2 - 2 x.y per pair (|x|^2 is constant per row); with BF16 the
//     dot product is xh.yh + xh.yl + xl.yh on v_mfma_f32_16x16x32_bf16 (FP32 accumulate), with
//     FP32 it is v_mfma_f32_16x16x4_f32. Each row keeps its 2k smallest s in an LDS list; a pair
//     touches the list only when it beats the row's current 2k-th value. Insertions run one
//     leader lane per row per round, so different rows insert in parallel.
//  3. rescore: the candidates get exact FP64 squared distances (sequential sum, no
//     contraction), sorted by (d2, index). A row is certified exact when the smallest distance
//     any non-candidate could have (2k-th approximate value minus a rigorous error bound) is
//     strictly above the k-th exact distance; other rows are recomputed by an FP64 scan.
//  4. gaussian: sigma_i from the k-th (or median) distance, w_ij in f64 stored as f32,
//     symmetrized, sorted into CSR and ingested without leaving HBM.
#include "common.hpp"
#include "tune.hpp"

#include <cstdlib>
#include <string>
#include <vector>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

namespace scgpu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int KNN_BM = 256;                  // query rows per workgroup (64 per wave, one per lane)
constexpr int KNN_TN = 64;                   // column splits are whole multiples of this many points
constexpr int KNN_MAXK2 = 64;                // candidates per row and split at most
constexpr int KNN_MAXCAND = 128;             // candidates per row over all splits
constexpr uint32_t KNONE = 0xFFFFFFFFu;

__device__ __forceinline__ unsigned short bf16_bits_rn(float x) {
    uint32_t u = __float_as_uint(x);
    // round to nearest even (inputs are finite)
    uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(r >> 16);
}

__device__ __forceinline__ float bf16_bits_to_f32(unsigned short b) {
    return __uint_as_float(((uint32_t)b) << 16);
}

// One thread per element of the padded npad x Kp layout, coalesced: split-BF16 copies (bf16
// path) or the f32 copy (fp32 path). Row j holds point j's coordinates, then its squared norm
// ny[j] in the two slots t = d, d + 1 (bf16: hi and lo parts of the f32 norm, lo copy zero; f32:
// the norm and zero), zeros beyond. As the B operand of the tile (column points) a row's dot
// product with A = (-2 x, 1, 1, 0 ...) is the s-value |y|^2 - 2 x.y itself; padding rows hold
// zeros and an infinite norm, so they never pass a threshold. With perm, row j of the layout is
// point perm[j], or padding where perm[j] == KNONE (the grouped layout of the pruned search,
// whose rows number npad with perm); without it rows j < n are the points.
__global__ void k_knn_prep(const float *X, const uint32_t *perm, const float *ny, uint32_t n, uint32_t d, uint32_t Kp,
                           uint64_t total, unsigned short *Xh, unsigned short *Xl, float *Xf) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t j = (uint32_t)(e / Kp), t = (uint32_t)(e % Kp);
        const uint32_t src = perm ? perm[j] : (j < n ? j : KNONE);
        if (t < d) {
            const float v = src != KNONE ? X[(uint64_t)src * d + t] : 0.0f;
            if (Xh) {
                const unsigned short h = bf16_bits_rn(v);
                Xh[e] = h;
                Xl[e] = bf16_bits_rn(v - bf16_bits_to_f32(h));
            } else {
                Xf[e] = v;
            }
        } else if (t < d + 2) {
            const float y = ny[j];  // +inf for padding rows
            if (Xh) {
                const unsigned short h = bf16_bits_rn(y);
                const bool finite = y < INFINITY;
                Xh[e] = t == d ? h : (finite ? bf16_bits_rn(y - bf16_bits_to_f32(h)) : (unsigned short)0);
                Xl[e] = 0;
            } else {
                Xf[e] = t == d ? y : 0.0f;
            }
        } else {
            if (Xh) {
                Xh[e] = 0;
                Xl[e] = 0;
            } else {
                Xf[e] = 0.0f;
            }
        }
    }
}

// Per padded point: f32 norm (+inf for padding points, so that they never pass a candidate
// threshold), exact FP64 norm, and the largest norm (one atomic per wave). With perm, ny follows
// the layout's order (ny[j] of point perm[j]) and nx64 the points' own order. One wave per 64
// layout rows: their coordinates go to LDS with reads along each row (lane t reads coordinate
// t), then each lane sums its own row in coordinate order (the same f32 and FP64 chains as one
// thread per row reading its row).
__global__ void __launch_bounds__(256) k_knn_norms(const float *X, const uint32_t *perm, uint32_t n, uint32_t d,
                                                   uint32_t npad, float *ny, double *nx64, unsigned int *ymax_bits) {
    extern __shared__ float sx[];  // [waves][64][d + 1]
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t W = blockDim.x >> 6;
    float *const my = sx + (size_t)wave * 64 * (d + 1);
    for (uint32_t j0 = (blockIdx.x * W + wave) * 64; j0 < npad; j0 += gridDim.x * W * 64) {  // wave-uniform
        const uint32_t j = j0 + lane;
        const uint32_t src = j < npad ? (perm ? perm[j] : (j < n ? j : KNONE)) : KNONE;
#pragma unroll 4
        for (int r = 0; r < 64; r++) {
            const uint32_t sr = __shfl(src, r);
            for (uint32_t t = lane; t < d; t += 64) my[r * (d + 1) + t] = sr != KNONE ? X[(uint64_t)sr * d + t] : 0.0f;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        float m = 0.0f;
        if (src != KNONE) {
            const float *x = my + lane * (d + 1);
            float s = 0.0f;
            double s64 = 0.0;
            for (uint32_t t = 0; t < d; t++) {
                const float v = x[t];
                s += v * v;
                const double dv = (double)v;
                s64 = __dadd_rn(s64, __dmul_rn(dv, dv));
            }
            ny[j] = s;
            nx64[src] = s64;
            m = sqrtf(s) * 1.0001f;
        } else if (j < npad) {
            ny[j] = INFINITY;
            if (!perm) nx64[j] = 0.0;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if (lane == 0 && m > 0.0f) atomicMax(ymax_bits, __float_as_uint(m));
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");  // my is rewritten by the next rows
    }
}

// KM candidates kept per row and split (16, 32 or 64). NKC: K chunks of 32 (K padded to 32 * NKC).
// grid.x = row blocks of KNN_BM rows, grid.y = column splits.
//
// Each wave owns 64 query rows (four 16-row MFMA tiles). Hits are found in the MFMA output
// layout itself: every lane compares its 16 values (column lane & 15, rows 4 * (lane >> 4) + i of
// each row tile) with a register copy of those rows' thresholds, and appends a hit to its row's
// LDS buffer at a slot taken from the compare ballot. Lane r of the wave keeps row r's KM smallest values sorted in
// registers; when any buffer could overflow on the next sub-tile the wave merges every buffer
// into its list with a branch-free compare-exchange network and republishes the thresholds.
template <int KM>
__device__ __forceinline__ void list_insert_net(float (&lv)[KM], uint32_t (&li)[KM], float x, uint32_t xi) {
#pragma unroll
    for (int p = 0; p < KM; p++) {
        const bool sw = x < lv[p];
        const float t = lv[p];
        const uint32_t ti = li[p];
        lv[p] = sw ? x : t;
        li[p] = sw ? xi : ti;
        x = sw ? t : x;
        xi = sw ? ti : xi;
    }
}

// The least f32 value at or above the double v (v may be infinite).
__device__ __forceinline__ float f32_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, INFINITY);
    return f;
}

template <int NKC> struct KnnTileShape {
    static constexpr int TN = NKC <= 2 ? 64 : 32;      // points per LDS stage
    static constexpr int BCAP = 20;                     // buffered hits per row between merges
};

template <bool BF16, int NKC>
constexpr size_t knn_tile_stage_bytes() {
    constexpr int KP = 32 * NKC;
    constexpr int TN = KnnTileShape<NKC>::TN;
    constexpr size_t STRIDE = BF16 ? (KP + 8) : (KP + 4);
    return BF16 ? 2 * TN * STRIDE * 2 : TN * STRIDE * 4;
}

// Two stage buffers (one barrier per stage: a stage is written while the previous one is read)
// where they fit in 64 KB next to the thresholds and hit buffers; otherwise one.
template <bool BF16, int NKC>
constexpr int knn_tile_stage_bufs() {
    constexpr size_t rest = KNN_BM * 4 * 2 + (size_t)KNN_BM * KnnTileShape<NKC>::BCAP * 8;
    return 2 * knn_tile_stage_bytes<BF16, NKC>() + rest <= 64 * 1024 ? 2 : 1;
}

template <bool BF16, int NKC>
constexpr size_t knn_tile_lds_bytes() {
    return knn_tile_stage_bufs<BF16, NKC>() * knn_tile_stage_bytes<BF16, NKC>() + KNN_BM * 4 * 2 +
           (size_t)KNN_BM * KnnTileShape<NKC>::BCAP * 8;
}

// f32 values as unsigned integers in the same order (negative values below positive ones), for
// atomicMin on thresholds that may be negative.
__device__ __forceinline__ uint32_t f32_ord(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float f32_unord(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// Row-owner merge: lane r folds row r's buffered hits into its sorted list, clears the buffer
// and republishes the row's threshold; then every lane reloads the 16 thresholds it compares
// against in the MFMA layout. With a shared threshold (gthr_row, the row's slot shared by its
// splits) the row publishes its own with an atomic minimum and takes the smaller of the two.
template <int KM, int BCAP>
__device__ __forceinline__ void knn_merge_rows(float (&lv)[KM], uint32_t (&li)[KM], float (&thr_r)[4][4],
                                               uint32_t (&cnt4)[4], uint32_t *sCnt, float *sThr, uint2 *sBuf,
                                               int wave, int lane, float tseed, float tmargin, uint32_t kk,
                                               uint32_t *gthr_row) {
    const int rl = wave * 64 + lane;
    const int q = lane >> 4;
    if ((lane & 15) == 0) {
#pragma unroll
        for (int rt = 0; rt < 4; rt++)
#pragma unroll
            for (int i = 0; i < 4; i++) sCnt[wave * 64 + rt * 16 + 4 * q + i] = (cnt4[rt] >> (8 * i)) & 0xFFu;
    }
    __threadfence_block();
    const uint32_t c = sCnt[rl];
    for (uint32_t p = 0; __any(p < c); p++) {
        const uint2 e = p < c ? sBuf[rl * BCAP + p] : make_uint2(__float_as_uint(INFINITY), KNONE);
        list_insert_net<KM>(lv, li, __uint_as_float(e.x), e.y);
    }
    // the threshold: the list's last value, the seed, and the row's k-th value plus the margin
    // (2.02 times the error bound; the sum rounded up): a point at or above it is strictly
    // farther than k points of this list (each within the bound of its value), so it cannot be
    // among the k nearest
    float lk = INFINITY;
#pragma unroll
    for (int p = 0; p < KM; p++) lk = (uint32_t)p + 1u == kk ? lv[p] : lk;
    const float kthr = lk < INFINITY ? f32_up((double)lk + (double)tmargin) : INFINITY;
    float t = fminf(fminf(lv[KM - 1], tseed), kthr);
    // every split's threshold is one of these bounds (a point at or above it is strictly farther
    // than k points of that split's list, which only improves), so the smallest over the
    // row's splits holds for all of them
    if (gthr_row) t = fminf(t, f32_unord(atomicMin(gthr_row, f32_ord(t))));
    sThr[rl] = t;
    __threadfence_block();
#pragma unroll
    for (int rt = 0; rt < 4; rt++) {
        const f32x4 t = *reinterpret_cast<const f32x4 *>(sThr + wave * 64 + rt * 16 + 4 * q);
#pragma unroll
        for (int i = 0; i < 4; i++) thr_r[rt][i] = t[i];
        cnt4[rt] = 0;
    }
}

// MFMA products of one 16-column sub-tile against the wave's four row tiles.
template <bool BF16, int NKC, int STRIDE, int TN>
__device__ __forceinline__ void knn_subtile_mfma(const unsigned short *sBh, const float *sBf, uint32_t lane_off,
                                                 const bf16x8 (&ah)[4][NKC], const bf16x8 (&al)[4][NKC],
                                                 const float (&af)[4][8 * NKC], f32x4 (&acc)[4]) {
#pragma unroll
    for (int rt = 0; rt < 4; rt++) acc[rt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kc = 0; kc < NKC; kc++) {
        if (BF16) {
            const uint32_t off = lane_off + kc * 32;
            const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(sBh + off);
            const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(sBh + TN * STRIDE + off);
#pragma unroll
            for (int rt = 0; rt < 4; rt++) {
                acc[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt][kc], bh, acc[rt], 0, 0, 0);
                acc[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt][kc], bl, acc[rt], 0, 0, 0);
                acc[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rt][kc], bh, acc[rt], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const float b = sBf[lane_off + kc * 32 + q * 4];
#pragma unroll
                for (int rt = 0; rt < 4; rt++)
                    acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[rt][kc * 8 + q], b, acc[rt], 0, 0, 0);
            }
        }
    }
}

// Threshold test of one sub-tile's values in the MFMA layout. A hit's buffer slot is its row's
// count plus the number of hits of lower lanes of the same row (16-lane group) in the ballot.
// Byte i of cnt4[rt] counts the buffered hits of row 16 rt + i of this lane's quarter.
// Returns true when some row could overflow its buffer on the next sub-tile. c0s: the sub-tile's
// first column.
template <int BCAP>
__device__ __forceinline__ bool knn_subtile_hits(const f32x4 (&acc)[4], uint32_t c0s, uint32_t row0,
                                                 uint32_t rl_base, uint64_t qmask, const float (&thr_r)[4][4],
                                                 uint32_t (&cnt4)[4], uint2 *sBuf) {
    // hits are rare once the thresholds have settled: the 16 compare masks (one per value of
    // the lane's rows, wave-wide) are taken once; a value whose mask is empty costs no more
    // vector work, and a sub-tile without any hit returns after the masks
    uint64_t m[4][4];
    uint64_t anym = 0;
#pragma unroll
    for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            m[rt][i] = __ballot(acc[rt][i] < thr_r[rt][i]);
            anym |= m[rt][i];
        }
    if (!anym) return false;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t col = c0s + (lane & 15u);
    // a row never lists itself: only a sub-tile that meets the block's own rows can hold one
    const bool selfp = c0s < row0 + KNN_BM && c0s + 16u > row0;
    bool need = false;
#pragma unroll
    for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint64_t mm = m[rt][i];
            if (mm) {
                const uint32_t rl = rl_base + rt * 16 + i;
                bool hit = acc[rt][i] < thr_r[rt][i];  // this lane's bit of mm
                if (selfp) {
                    hit = hit && col != row0 + rl;
                    mm = __ballot(hit);
                }
                const uint64_t gm = mm & qmask;
                const uint32_t below =
                    __builtin_amdgcn_mbcnt_hi((uint32_t)(gm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)gm, 0u));
                const uint32_t c = (cnt4[rt] >> (8 * i)) & 0xFFu;
                const uint32_t add = (uint32_t)__popcll(gm);
                if (hit) sBuf[rl * BCAP + c + below] = make_uint2(__float_as_uint(acc[rt][i]), col);
                cnt4[rt] += add << (8 * i);
                need |= c + add > (uint32_t)(BCAP - 16);
            }
        }
    return need;
}

// Column stages: the split's contiguous range of cols_per_split points, or, for the pruned
// search (blist non-null), every splits-th entry, from the split's index on, of this row
// block's list of column blocks of TN points: blist[boff[row block] .. + bcnt[row block]).
template <bool BF16, int NKC, int KM>
__global__ void __launch_bounds__(256, (KM == 16 && NKC <= 2) ? 2 : 1) k_knn_tile(const unsigned short *Xh, const unsigned short *Xl,
                                                  const float *Xf, uint32_t d, uint32_t n, uint32_t npad,
                                                  uint32_t cols_per_split, uint32_t row_blocks, uint32_t splits,
                                                  const uint32_t *blist, const uint32_t *boff, const uint32_t *bcnt,
                                                  const uint32_t *rb_order,
                                                  const float *seed_s, const float *margin_s, uint32_t kk,
                                                  uint32_t *gthr, uint32_t *cand_idx, float *cand_s) {
    constexpr int KP = 32 * NKC;
    constexpr int TN = KnnTileShape<NKC>::TN;
    constexpr int BCAP = KnnTileShape<NKC>::BCAP;
    constexpr int STRIDE = BF16 ? (KP + 8) : (KP + 4);            // staged point stride (elements)
    constexpr int STAGE_BYTES = (int)knn_tile_stage_bytes<BF16, NKC>();
    constexpr int NBUF = knn_tile_stage_bufs<BF16, NKC>();
    constexpr int CHUNKS = BF16 ? 2 * TN * KP / 8 : TN * KP / 4;  // 16-byte chunks per stage
    constexpr int CPT = CHUNKS / 256;                             // chunks per thread
    static_assert(CHUNKS % 256 == 0, "stage must split evenly over the workgroup");
    static_assert(STAGE_BYTES % 16 == 0, "threshold array must stay 16-byte aligned");

    extern __shared__ unsigned char smem[];
    float *sThr = reinterpret_cast<float *>(smem + NBUF * STAGE_BYTES);  // per block row
    uint32_t *sCnt = reinterpret_cast<uint32_t *>(sThr + KNN_BM); // per block row
    uint2 *sBuf = reinterpret_cast<uint2 *>(sCnt + KNN_BM);       // [row][BCAP] (s bits, column)

    const uint32_t nwg = row_blocks * splits;
    uint32_t split, rb;
    if (rb_order) {
        // pruned lists differ in length by the size of the clusters they cover: the row blocks
        // run longest list first (rb_order), all splits of one row block together, so the
        // shortest lists fill the end of the grid; consecutive workgroups go to different XCDs
        const uint32_t lg = blockIdx.x;
        if (lg >= nwg) return;  // grid padding (the whole workgroup leaves before any barrier)
        rb = rb_order[lg / splits];
        split = lg % splits;
    } else {
        // XCD-aware order: workgroup w runs on XCD w mod 8, and each XCD takes a contiguous
        // range of the (split, row block) space, so the workgroups resident on one XCD at a time
        // are neighbouring row blocks of one split, whose columns coincide
        const uint32_t chunk = (nwg + 7u) / 8u;
        const uint32_t lg = (blockIdx.x & 7u) * chunk + (blockIdx.x >> 3);
        if (lg >= nwg) return;
        split = lg / row_blocks;
        rb = lg % row_blocks;
    }
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const int quarter = lane >> 4;
    const uint32_t row0 = rb * KNN_BM;
    const uint32_t wrow0 = row0 + wave * 64;
    uint32_t cbeg = 0, ns = 0, jb = 0;
    const uint32_t jstep = blist ? splits : 1u;
    if (blist) {
        // splits take every splits-th block of the list: each split sees a sample of the near
        // blocks, so no split's list is filled by the nearest points alone (which would leave
        // the certificate no margin)
        const uint32_t L0 = boff[rb], L = bcnt[rb];
        jb = L0 + split;
        ns = L > split ? (L - split + splits - 1) / splits : 0u;
    } else {
        cbeg = split * cols_per_split;
        const uint32_t cend = min(npad, cbeg + cols_per_split);
        ns = cend > cbeg ? (cend - cbeg) / TN : 0;
    }

    // a row's threshold starts at its seed (s-space bound above which no k-nearest point lies,
    // pruned search) or +inf, and never rises above it
    const float tseed = seed_s ? seed_s[wrow0 + lane] : INFINITY;
    const float tmargin = margin_s[wrow0 + lane];  // twice the row's error bound, rounded up
    uint32_t *const grow = gthr ? gthr + wrow0 + lane : nullptr;  // the row's shared threshold
    sThr[tid] = tseed;

    // A fragments of the wave's four row tiles, in registers for the whole sweep: the query row
    // as (-2 x, 1, 1, 0, ...) against the column rows (y, |y|^2 hi, |y|^2 lo, 0, ...), so each
    // product is s = |y|^2 - 2 x.y (scaling by -2 is exact in bf16 and f32)
    bf16x8 ah[4][NKC], al[4][NKC];
    float af[4][8 * NKC];
#pragma unroll
    for (int rt = 0; rt < 4; rt++) {
        const uint32_t arow = wrow0 + rt * 16 + (lane & 15);
#pragma unroll
        for (int kc = 0; kc < NKC; kc++) {
            if (BF16) {
                const uint64_t off = (uint64_t)arow * KP + kc * 32 + 8 * quarter;
                union {
                    bf16x8 v;
                    unsigned short u[8];
                } h, l;
                h.v = *reinterpret_cast<const bf16x8 *>(Xh + off);
                l.v = *reinterpret_cast<const bf16x8 *>(Xl + off);
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const uint32_t t = kc * 32 + 8 * quarter + e;
                    if (t < d) {
                        h.u[e] = bf16_bits_rn(-2.0f * bf16_bits_to_f32(h.u[e]));
                        l.u[e] = bf16_bits_rn(-2.0f * bf16_bits_to_f32(l.u[e]));
                    } else {
                        h.u[e] = t < d + 2 ? (unsigned short)0x3F80u : (unsigned short)0;  // 1.0 at the norm slots
                        l.u[e] = 0;
                    }
                }
                ah[rt][kc] = h.v;
                al[rt][kc] = l.v;
            } else {
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const uint32_t t = kc * 32 + q * 4 + quarter;
                    const float v = Xf[(uint64_t)arow * KP + t];
                    af[rt][kc * 8 + q] = t < d ? -2.0f * v : (t == d ? 1.0f : 0.0f);
                }
            }
        }
    }

    // per-thread staging slots: chunk q = tid + 256 * c of the stage
    uint32_t src_off[CPT], dst_off[CPT];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const uint32_t q = tid + 256 * c;
        if (BF16) {
            constexpr uint32_t per_pt = KP / 8, per_arr = TN * KP / 8;
            const uint32_t which = q / per_arr, r = q % per_arr, p = r / per_pt, ch = r % per_pt;
            src_off[c] = which * 0x80000000u | (p * KP + ch * 8);  // element offset within the stage, array bit
            dst_off[c] = which * (TN * STRIDE) + p * STRIDE + ch * 8;
        } else {
            constexpr uint32_t per_pt = KP / 4;
            const uint32_t p = q / per_pt, ch = q % per_pt;
            src_off[c] = p * KP + ch * 4;
            dst_off[c] = p * STRIDE + ch * 4;
        }
    }
    u32x4 pre[CPT];
    // loads stage C0 of the points into pre[] (a macro keeps the offsets in VGPRs)
#define KNN_PREFETCH(C0)                                                                                   \
    do {                                                                                                   \
        _Pragma("unroll") for (int c = 0; c < CPT; c++) {                                                  \
            if (BF16) {                                                                                    \
                const unsigned short *base_ = (src_off[c] & 0x80000000u) ? Xl : Xh;                        \
                pre[c] = *reinterpret_cast<const u32x4 *>(base_ + (uint64_t)(C0) * KP + (src_off[c] & 0x7FFFFFFFu)); \
            } else {                                                                                       \
                pre[c] = *reinterpret_cast<const u32x4 *>(Xf + (uint64_t)(C0) * KP + src_off[c]);          \
            }                                                                                              \
        }                                                                                                  \
    } while (0)

    float lv[KM];
    uint32_t li[KM];
#pragma unroll
    for (int p = 0; p < KM; p++) {
        lv[p] = INFINITY;
        li[p] = KNONE;
    }
    float thr_r[4][4];
    uint32_t cnt4[4];
    __syncthreads();  // the seeded thresholds are in sThr
#pragma unroll
    for (int rt = 0; rt < 4; rt++) {
#pragma unroll
        for (int i = 0; i < 4; i++) thr_r[rt][i] = sThr[wave * 64 + rt * 16 + 4 * quarter + i];
        cnt4[rt] = 0;
    }

    const uint32_t lane_base = (uint32_t)(lane & 15) * STRIDE + (BF16 ? 8u * quarter : (uint32_t)quarter);
    // block-local row of this lane's value i of row tile rt, minus 16 rt + i
    const uint32_t rl_base = wave * 64 + 4 * quarter;
    const uint64_t qmask = 0xFFFFull << (16 * quarter);
    // List entries arrive in batches of 64 stages, one entry per lane (lst_cur: stages
    // [64 b, 64 b + 64) of the current batch b, lst_nxt: the batch after it, loaded a whole batch
    // ahead); a stage's first column is read out of the batch register with a lane read. A list
    // entry loaded per stage would be waited for at the next stage's prefetch, which exposes one
    // memory latency per stage.
    auto list_batch = [&](uint32_t b) -> uint32_t {
        const uint32_t i = 64u * b + (uint32_t)lane;
        return i < ns ? blist[jb + i * jstep] : 0u;
    };
    uint32_t lst_cur = 0, lst_nxt = 0;
    if (blist && ns) {
        lst_cur = list_batch(0);
        lst_nxt = list_batch(1);
    }
    // first column of stage i, for i in the current batch (nxt false) or the next (nxt true)
    auto stage_col = [&](uint32_t i, bool nxt) -> uint32_t {
        if (!blist) return cbeg + i * TN;
        return (uint32_t)__builtin_amdgcn_readlane((int)(nxt ? lst_nxt : lst_cur), (int)(i & 63u)) * TN;
    };
    if (ns) KNN_PREFETCH(stage_col(0, false));
    for (uint32_t st = 0; st < ns; st++) {
        if (blist && st && (st & 63u) == 0) {
            lst_cur = lst_nxt;
            lst_nxt = list_batch((st >> 6) + 1);
        }
        const uint32_t c0 = stage_col(st, false);
        const uint32_t c0n = st + 1 < ns ? stage_col(st + 1, ((st + 1) & 63u) == 0) : 0u;
        // with one buffer, wait until the previous stage is consumed; with two, the barrier of
        // the previous stage already ordered the reads of the buffer written here
        if (NBUF == 1) __syncthreads();
        const uint32_t boff_e = NBUF == 2 ? (st & 1u) * (STAGE_BYTES / (BF16 ? 2 : 4)) : 0u;
        unsigned short *const sBh = reinterpret_cast<unsigned short *>(smem) + (BF16 ? boff_e : 0u);
        float *const sBf = reinterpret_cast<float *>(smem) + (BF16 ? 0u : boff_e);
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            if (BF16) *reinterpret_cast<u32x4 *>(sBh + dst_off[c]) = pre[c];
            else *reinterpret_cast<u32x4 *>(sBf + dst_off[c]) = pre[c];
        }
        if (st + 1 < ns) KNN_PREFETCH(c0n);  // lands while this stage computes
        __syncthreads();

        // two accumulator sets: the products of sub-tile t + 1 are issued before the threshold
        // test of sub-tile t, so the matrix cores and the vector unit overlap
        f32x4 acc0[4], acc1[4];
        knn_subtile_mfma<BF16, NKC, STRIDE, TN>(sBh, sBf, lane_base, ah, al, af, acc0);
#pragma unroll
        for (int sub = 0; sub < TN / 16; sub += 2) {
            knn_subtile_mfma<BF16, NKC, STRIDE, TN>(sBh, sBf, (uint32_t)(sub + 1) * (16u * STRIDE) + lane_base, ah,
                                                    al, af, acc1);
            bool need = knn_subtile_hits<BCAP>(acc0, c0 + sub * 16, row0, rl_base, qmask, thr_r, cnt4, sBuf);
            if (__any(need)) knn_merge_rows<KM, BCAP>(lv, li, thr_r, cnt4, sCnt, sThr, sBuf, wave, lane, tseed, tmargin, kk, grow);
            if (sub + 2 < TN / 16)
                knn_subtile_mfma<BF16, NKC, STRIDE, TN>(sBh, sBf, (uint32_t)(sub + 2) * (16u * STRIDE) + lane_base,
                                                        ah, al, af, acc0);
            need = knn_subtile_hits<BCAP>(acc1, c0 + (sub + 1) * 16, row0, rl_base, qmask, thr_r, cnt4, sBuf);
            if (__any(need)) knn_merge_rows<KM, BCAP>(lv, li, thr_r, cnt4, sCnt, sThr, sBuf, wave, lane, tseed, tmargin, kk, grow);
        }
    }
#undef KNN_PREFETCH
    knn_merge_rows<KM, BCAP>(lv, li, thr_r, cnt4, sCnt, sThr, sBuf, wave, lane, tseed, tmargin, kk, grow);
    // write this split's candidate list: cand[(row * splits + split) * KM + t]
    const uint32_t my_row = wrow0 + lane;
    if (my_row < n) {
        const uint64_t o = ((uint64_t)my_row * splits + split) * KM;
#pragma unroll
        for (int p = 0; p < KM; p++) {
            cand_idx[o + p] = li[p];
            cand_s[o + p] = lv[p];
        }
    }
}

// Sequential FP64 sum of squared differences (this file is built without contraction, so the
// multiply and the add round separately, as in the CPU restatement).
__device__ __forceinline__ double exact_d2(const float *X, uint32_t d, uint32_t i, uint32_t j) {
    const float *a = X + (uint64_t)i * d;
    const float *b = X + (uint64_t)j * d;
    double s = 0.0;
    uint32_t t = 0;
    // eight coordinates of both points are loaded together, then added in order t, t + 1, ...
    for (; t + 8 <= d; t += 8) {
        float av[8], bv[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            av[q] = a[t + q];
            bv[q] = b[t + q];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const double diff = (double)av[q] - (double)bv[q];
            s = __dadd_rn(s, __dmul_rn(diff, diff));
        }
    }
    for (; t < d; t++) {
        double diff = (double)a[t] - (double)b[t];
        s = __dadd_rn(s, __dmul_rn(diff, diff));
    }
    return s;
}

// LW lanes per row (a wave holds 64 / LW rows; LW >= the row's candidates KM * splits, 32 or
// 64): FP64 rescore of the candidates of all splits, top-k by (d2, index), and the certificate.
// With |s_approx - s| <= eps for every candidate (s = |y|^2 - 2 x.y), the k-th smallest exact
// value among the candidates is at most S_k + eps, where S_k is the k-th smallest approximate
// value; a candidate above S_k + 2 eps can therefore not be among the k nearest candidates and is
// not rescored. A row's ballots are masked to its lanes.
constexpr int RS_WAVES = 4;

template <int LW>
__global__ void __launch_bounds__(64 * RS_WAVES)
k_knn_rescore(const float *X, uint32_t n, uint32_t d, uint32_t k, uint32_t KM, uint32_t splits,
              const uint32_t *cand_idx, const float *cand_s, const double *nx64, double eps_coef, double eps_const,
              uint32_t *idx_out, double *d2_out, int fallback_all, uint32_t *fallback, unsigned int *n_fallback,
              unsigned long long *n_rescored) {
    constexpr int RPW = 64 / LW;          // rows per wave
    constexpr int HH = LW == 64 ? 2 : 1;  // candidates per lane: up to KNN_MAXCAND (64 lanes), 32 (32)
    __shared__ float sS[RS_WAVES * RPW][KNN_MAXCAND];
    __shared__ uint32_t sI[RS_WAVES * RPW][KNN_MAXCAND];
    __shared__ double sD[RS_WAVES * RPW][KNN_MAXCAND];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int sg = lane / LW, sl = lane % LW;
    const int slot = wave * RPW + sg;
    const uint64_t rmask = LW == 64 ? ~0ull : (0xFFFFFFFFull << (32 * sg));  // this row's lanes
    const uint32_t total = KM * splits;  // <= HH * LW
    unsigned long long local = 0;
    for (uint32_t i = (blockIdx.x * RS_WAVES + wave) * RPW + sg; i < n; i += gridDim.x * RS_WAVES * RPW) {
        const uint64_t base = (uint64_t)i * total;
        float sv[HH];
        uint32_t jv[HH];
#pragma unroll
        for (int h = 0; h < HH; h++) {
            const uint32_t e = sl + LW * h;
            jv[h] = e < total ? cand_idx[base + e] : KNONE;
            sv[h] = jv[h] != KNONE ? cand_s[base + e] : INFINITY;
            sS[slot][e] = sv[h];
        }
        // a split whose list is full (its last entry is a point) may hide closer points; the
        // smallest last value over those splits bounds every point not in the lists. A list that
        // is not full hides only points at or above the row's thresholds (pruned search), which
        // are strictly farther than k points of the lists
        float last = INFINITY;
        bool full = false;
        if ((uint32_t)sl < splits) {
            const uint32_t jl = cand_idx[base + (uint64_t)sl * KM + KM - 1];
            full = jl != KNONE;
            if (full) last = cand_s[base + (uint64_t)sl * KM + KM - 1];
        }
        const bool any_full = (__ballot(full) & rmask) != 0ull;
#pragma unroll
        for (int o = LW / 2; o > 0; o >>= 1) last = fminf(last, __shfl_xor(last, o, LW));
        const float amin = last;
        uint32_t cnt = 0;
#pragma unroll
        for (int h = 0; h < HH; h++) cnt += (uint32_t)__popcll(__ballot(jv[h] != KNONE) & rmask);
        __threadfence_block();

        // S_k: the approximate value of rank k - 1 (ties by list position)
        const double eps = eps_coef * sqrt(nx64[i]) + eps_const;
        double cut = INFINITY;
        if (cnt > k) {
            // each split's list is sorted (ascending values, empty entries last), so a candidate's
            // rank by (value, list position) is its position in its own list plus, per other list,
            // the entries before it: values at or below it in an earlier list, below it in a later
            // one (a binary search of KM entries, KM a power of two)
            uint32_t rank[HH];
#pragma unroll
            for (int h = 0; h < HH; h++) {
                rank[h] = 0u;
                const uint32_t e = (uint32_t)sl + LW * h;
                if (e >= total) continue;
                const uint32_t sp = e / KM;
                const float x = sv[h];
                uint32_t r = e - sp * KM;
                for (uint32_t s2 = 0; s2 < splits; s2++) {
                    if (s2 == sp) continue;
                    const float *L = &sS[slot][s2 * KM];
                    const bool le = s2 < sp;
                    uint32_t lo = 0;
                    for (uint32_t step = KM >> 1; step > 0; step >>= 1) {
                        const float y = L[lo + step - 1];
                        if (le ? y <= x : y < x) lo += step;
                    }
                    const float y = L[lo];
                    if (le ? y <= x : y < x) lo++;
                    r += lo;
                }
                rank[h] = r;
            }
            float sk = INFINITY;
#pragma unroll
            for (int h = 0; h < HH; h++) {
                const uint64_t m = __ballot(jv[h] != KNONE && rank[h] == k - 1) & rmask;
                if (m) sk = __shfl(sv[h], (int)__ffsll((unsigned long long)m) - 1);
            }
            cut = (double)sk + 2.0 * eps;
        }

        // exact distances of the candidates that can still be among the k nearest, packed to the
        // front of the row's LDS arrays (the list order is no longer needed)
        double dd[HH];
        bool pass[HH];
#pragma unroll
        for (int h = 0; h < HH; h++) {
            pass[h] = jv[h] != KNONE && (double)sv[h] <= cut;
            dd[h] = pass[h] ? exact_d2(X, d, i, jv[h]) : INFINITY;
        }
        uint64_t pm[HH];
        uint32_t P = 0, Pb[HH];
#pragma unroll
        for (int h = 0; h < HH; h++) {
            pm[h] = __ballot(pass[h]) & rmask;
            Pb[h] = P;
            P += (uint32_t)__popcll(pm[h]);
        }
        local += sl == 0 ? P : 0;
#pragma unroll
        for (int h = 0; h < HH; h++) {
            if (pass[h]) {
                const uint32_t at = Pb[h] + __builtin_amdgcn_mbcnt_hi((uint32_t)(pm[h] >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)pm[h], 0u));
                sD[slot][at] = dd[h];
                sI[slot][at] = jv[h];
            }
        }
        __threadfence_block();

        // rank by (d2, index) among the rescored candidates; the first k are the answer
        double kth = INFINITY;
#pragma unroll
        for (int h = 0; h < HH; h++) {
            uint32_t r = 0;
            if (pass[h]) {
                for (uint32_t e = 0; e < P; e++) {
                    const double v = sD[slot][e];
                    r += (v < dd[h] || (v == dd[h] && sI[slot][e] < jv[h])) ? 1u : 0u;
                }
                if (r < k) {
                    idx_out[(uint64_t)i * k + r] = jv[h];
                    d2_out[(uint64_t)i * k + r] = dd[h];
                }
            }
            const uint64_t m = __ballot(pass[h] && r == k - 1) & rmask;
            if (m) kth = __shfl(dd[h], (int)__ffsll((unsigned long long)m) - 1);
        }
        for (uint32_t q = (uint32_t)sl; q < k; q += LW)
            if (q >= P) {
                idx_out[(uint64_t)i * k + q] = KNONE;
                d2_out[(uint64_t)i * k + q] = INFINITY;
            }
        // certified when no full split can hide a point at or below the k-th distance
        bool ok = true;
        if (any_full) ok = cnt >= k && (nx64[i] + (double)amin - eps > kth);
        if ((!ok || fallback_all) && sl == 0) {
            const unsigned int fslot = atomicAdd(n_fallback, 1u);
            fallback[fslot] = i;
        }
        __threadfence_block();  // this row's LDS reads are done before the next row's writes
    }
    if (sl == 0 && local) atomicAdd(n_rescored, local);
}

// Exact FP64 scan for the rows the certificate rejected: workgroup (r, w) (128 threads) scans
// the w-th of W equal ranges of points for fallback row r and writes its k nearest, sorted by
// (d2, index), to part[(r W + w) k ..]; k_knn_fallback_merge then merges the W lists of a row.
// Ranges spread one row over many workgroups: a single row scanned by one workgroup waited on
// ~n/128 dependent distance computations per thread.
constexpr int FB_THREADS = 128;
constexpr int FB_MAXK = 32;

__global__ void __launch_bounds__(FB_THREADS) k_knn_fallback(const float *X, uint32_t n, uint32_t d, uint32_t k,
                                                             const uint32_t *rows, uint32_t *pidx, double *pd2) {
    __shared__ double sd[FB_THREADS * FB_MAXK];
    __shared__ uint32_t si[FB_THREADS * FB_MAXK];
    __shared__ uint32_t head[FB_THREADS];
    const uint32_t i = rows[blockIdx.x];
    const uint32_t W = gridDim.y, w = blockIdx.y;
    const uint32_t chunk = (n + W - 1) / W;
    const uint32_t lo = min(n, w * chunk), hi = min(n, lo + chunk);
    const int t = threadIdx.x;
    double *md = sd + t * FB_MAXK;
    uint32_t *mi = si + t * FB_MAXK;
    uint32_t cnt = 0;
    for (uint32_t j = lo + t; j < hi; j += FB_THREADS) {
        if (j == i) continue;
        double dd = exact_d2(X, d, i, j);
        if (cnt == k && !(dd < md[k - 1])) continue;  // j increases: ties keep the lower index
        uint32_t p = cnt < k ? cnt : k - 1;
        while (p > 0 && md[p - 1] > dd) {
            md[p] = md[p - 1];
            mi[p] = mi[p - 1];
            p--;
        }
        md[p] = dd;
        mi[p] = j;
        if (cnt < k) cnt++;
    }
    for (uint32_t q = cnt; q < k; q++) {
        md[q] = INFINITY;
        mi[q] = KNONE;
    }
    head[t] = 0;
    __syncthreads();
    if (t == 0) {
        const uint64_t o = ((uint64_t)blockIdx.x * W + w) * k;
        for (uint32_t out = 0; out < k; out++) {
            int best = -1;
            double bdv = INFINITY;
            uint32_t bidx = KNONE;
            for (int q = 0; q < FB_THREADS; q++) {
                uint32_t h = head[q];
                if (h >= k) continue;
                double v = sd[q * FB_MAXK + h];
                uint32_t ix = si[q * FB_MAXK + h];
                if (ix == KNONE) continue;
                if (v < bdv || (v == bdv && ix < bidx)) {
                    bdv = v;
                    bidx = ix;
                    best = q;
                }
            }
            if (best >= 0) head[best]++;
            pidx[o + out] = bidx;
            pd2[o + out] = bdv;
        }
    }
}

// The k nearest of fallback row r from its W partial lists (each sorted by (d2, index), KNONE
// past its end): a W-way merge by (d2, index), one wave per row. Lane l holds the heads of lists
// l, l + 64, l + 128 and l + 192 in registers; per output the wave reduces the smallest head and
// only the list it came from loads its next entry (the points of different lists differ, so
// the winning head is unique).
constexpr uint32_t FB_MAXW = 256;
__device__ __forceinline__ bool fb_less(double a, uint32_t ai, double b, uint32_t bi) {
    return a < b || (a == b && ai < bi);
}
__global__ void __launch_bounds__(64) k_knn_fallback_merge(uint32_t k, uint32_t W, const uint32_t *rows,
                                                           const uint32_t *pidx, const double *pd2, uint32_t *idx_out,
                                                           double *d2_out) {
    constexpr int HQ = FB_MAXW / 64;
    const uint32_t i = rows[blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * W * k;
    const uint32_t lane = threadIdx.x;
    double hv[HQ];
    uint32_t hi[HQ], hp[HQ];
    auto load = [&](int q) {
        const uint32_t w = lane + 64u * (uint32_t)q;
        const uint32_t ix = (w < W && hp[q] < k) ? pidx[base + (uint64_t)w * k + hp[q]] : KNONE;
        hi[q] = ix;
        hv[q] = ix != KNONE ? pd2[base + (uint64_t)w * k + hp[q]] : INFINITY;
    };
#pragma unroll
    for (int q = 0; q < HQ; q++) {
        hp[q] = 0;
        load(q);
    }
    for (uint32_t out = 0; out < k; out++) {
        double bv = hv[0];
        uint32_t bi = hi[0];
#pragma unroll
        for (int q = 1; q < HQ; q++)
            if (fb_less(hv[q], hi[q], bv, bi)) {
                bv = hv[q];
                bi = hi[q];
            }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(bv, o);
            const uint32_t oi = __shfl_xor(bi, o);
            if (fb_less(ov, oi, bv, bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            idx_out[(uint64_t)i * k + out] = bi;
            d2_out[(uint64_t)i * k + out] = bi != KNONE ? bv : INFINITY;
        }
        if (bi == KNONE) continue;  // every list is exhausted (wave-uniform)
#pragma unroll
        for (int q = 0; q < HQ; q++)
            if (hi[q] == bi && hv[q] == bv) {
                hp[q]++;
                load(q);
            }
    }
}

// ---- pruned search: spatial grouping and block bounds -----------------------------------------
// The points are reordered so that nearby points are consecutive (k-means groups), every block
// of consecutive points gets a bounding sphere, and a row block R skips the column blocks B that
// provably hold none of its rows' k nearest points: R holds at least k + 1 points, so every row
// x of R has k other points within 2 rR (rR: R's radius), while every y in B is at least
// |cR - cB| - rR - rB away; B is skipped when |cR - cB| - rR - rB > 2 rR. A skipped point is
// therefore strictly farther than the k-th nearest and cannot tie with it: the result is the
// brute-force result. Only the amount of skipped work depends on the grouping.
constexpr int KMC = 64;  // centroids per LDS chunk of the assignment

__global__ void k_fill_f64_knn(double *p, uint64_t n, double x) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = x;
}

__global__ void k_fill_u32_knn(uint32_t *p, uint32_t n, uint32_t x) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = x;
}

__global__ void k_knn_iota(uint32_t *p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

// Every stride-th point of X (n_out of them), as rows of d values.
__global__ void k_gather_rows(const float *X, uint32_t d, uint32_t stride, uint32_t n_out, float *out) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < (uint64_t)n_out * d;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = e / d, t = e % d;
        out[e] = X[i * stride * d + t];
    }
}

// Initial centroids: C points spread evenly over the input order; f32 squared norms.
__global__ void k_kmeans_init(const float *X, uint32_t n, uint32_t d, uint32_t C, float *cen) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < (uint64_t)C * d;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = (uint32_t)(e / d), t = (uint32_t)(e % d);
        const uint64_t src = (uint64_t)c * n / C;
        cen[e] = X[src * d + t];
    }
}

__global__ void k_kmeans_norms(const float *cen, uint32_t d, uint32_t C, float *cnorm) {
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (uint32_t t = 0; t < d; t++) s = fmaf(cen[(uint64_t)c * d + t], cen[(uint64_t)c * d + t], s);
        cnorm[c] = s;
    }
}

// Centroids as the matrix cores' B rows for the assignment below (d + 2 <= 32): split-BF16
// coordinates (hi in cb_h, lo in cb_l) and the f32 squared norm split into slots d (hi) and
// d + 1 (lo) of cb_h; rows C .. Cpad - 1 are padding whose norm is +inf (never nearest).
__global__ void k_kmeans_pack(const float *cen, uint32_t d, uint32_t C, uint32_t Cpad, unsigned short *cb_h,
                              unsigned short *cb_l) {
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < Cpad; c += gridDim.x * blockDim.x) {
        float q = 0.0f;
        for (uint32_t t = 0; t < 32; t++) {
            unsigned short h = 0, l = 0;
            if (c < C && t < d) {
                const float v = cen[(uint64_t)c * d + t];
                q = fmaf(v, v, q);
                h = bf16_bits_rn(v);
                l = bf16_bits_rn(v - bf16_bits_to_f32(h));
            }
            cb_h[(uint64_t)c * 32 + t] = h;
            cb_l[(uint64_t)c * 32 + t] = l;
        }
        if (c < C) {
            const unsigned short qh = bf16_bits_rn(q);
            cb_h[(uint64_t)c * 32 + d] = qh;
            cb_h[(uint64_t)c * 32 + d + 1] = bf16_bits_rn(q - bf16_bits_to_f32(qh));
        } else {
            cb_h[(uint64_t)c * 32 + d] = 0x7F80u;  // +inf
        }
    }
}

// Nearest centroid of every point (d + 2 <= 32) by |c|^2 - 2 x.c on the matrix cores: the point
// as split-BF16 (-2 x, 1, 1) against the centroid (c, |c|^2 hi, |c|^2 lo), three products per
// 16 x 16 block (x_h c_h + x_h c_l + x_l c_h), FP32 accumulation; ties to the lower centroid.
// The assignment only groups points for the pruned search (any assignment gives the same kNN
// result; the groups' spheres are computed from the points each group holds), so the products'
// rounding may pick a near-tied centroid. A wave holds 64 points as four row tiles; 64
// centroids at a time pass through LDS; each lane keeps, for its column slot, the best value and
// centroid of its 16 (row tile, row) pairs, reduced across the 16 column slots at the end.
__global__ void __launch_bounds__(256) k_kmeans_assign_mfma(const float *X, uint32_t n, uint32_t d,
                                                            const unsigned short *cb_h, const unsigned short *cb_l,
                                                            uint32_t Cpad, uint32_t *assign) {
    constexpr int STR = 40;  // staged centroid stride (bf16 elements)
    __shared__ __attribute__((aligned(16))) unsigned short sBh[64 * STR], sBl[64 * STR];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, quarter = lane >> 4;
    for (uint64_t r0 = (uint64_t)blockIdx.x * 256; r0 < n; r0 += (uint64_t)gridDim.x * 256) {  // workgroup-uniform
        const uint64_t wrow0 = r0 + (uint64_t)wave * 64;
        bf16x8 ah[4], al[4];
#pragma unroll
        for (int rt = 0; rt < 4; rt++) {
            const uint64_t arow = wrow0 + rt * 16 + (lane & 15);
            union {
                bf16x8 v;
                unsigned short u[8];
            } h, l;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const uint32_t t = 8 * quarter + e;
                float xv = 0.0f;
                if (arow < n && t < d) xv = X[arow * d + t];
                if (t < d) {
                    const unsigned short xh = bf16_bits_rn(xv);
                    const unsigned short xl = bf16_bits_rn(xv - bf16_bits_to_f32(xh));
                    h.u[e] = bf16_bits_rn(-2.0f * bf16_bits_to_f32(xh));  // exact: scaling by -2
                    l.u[e] = bf16_bits_rn(-2.0f * bf16_bits_to_f32(xl));
                } else {
                    h.u[e] = t < d + 2 ? (unsigned short)0x3F80u : (unsigned short)0;  // 1.0 at the norm slots
                    l.u[e] = 0;
                }
            }
            ah[rt] = h.v;
            al[rt] = l.v;
        }
        float bv[4][4];
        uint32_t bi[4][4];
#pragma unroll
        for (int rt = 0; rt < 4; rt++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                bv[rt][i] = INFINITY;
                bi[rt][i] = 0xFFFFFFFFu;
            }
        for (uint32_t c0 = 0; c0 < Cpad; c0 += 64) {
            __syncthreads();
            {
                // 64 centroids x 32 values x 2 arrays = 512 chunks of 8 values, two per thread
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const int ch = tid + 256 * q;
                    const int arr = ch >> 8, r = (ch >> 2) & 63, part = ch & 3;
                    const unsigned short *src = (arr ? cb_l : cb_h) + (uint64_t)(c0 + r) * 32 + part * 8;
                    unsigned short *dst = (arr ? sBl : sBh) + r * STR + part * 8;
                    *reinterpret_cast<uint4 *>(dst) = *reinterpret_cast<const uint4 *>(src);
                }
            }
            __syncthreads();
#pragma unroll
            for (int ct = 0; ct < 4; ct++) {
                const int off = (ct * 16 + (lane & 15)) * STR + 8 * quarter;
                const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(sBh + off);
                const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(sBl + off);
                const uint32_t col = c0 + ct * 16 + (lane & 15);
#pragma unroll
                for (int rt = 0; rt < 4; rt++) {
                    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt], bh, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt], bl, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rt], bh, acc, 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        if (acc[i] < bv[rt][i]) {  // columns arrive in increasing order: ties keep the lower
                            bv[rt][i] = acc[i];
                            bi[rt][i] = col;
                        }
                }
            }
        }
        // per row: the best over the 16 column slots (lanes with the same quarter), ties to the
        // lower centroid
#pragma unroll
        for (int rt = 0; rt < 4; rt++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                float v = bv[rt][i];
                uint32_t c = bi[rt][i];
#pragma unroll
                for (int off = 1; off < 16; off <<= 1) {
                    const float ov = __shfl_xor(v, off, 16);
                    const uint32_t oc = (uint32_t)__shfl_xor((int)c, off, 16);
                    if (ov < v || (ov == v && oc < c)) {
                        v = ov;
                        c = oc;
                    }
                }
                const uint64_t row = wrow0 + rt * 16 + 4 * quarter + i;
                if ((lane & 15) == 0 && row < n) assign[row] = c;
            }
    }
}

// Nearest centroid of every point by |c|^2 - 2 x.c (f32; ties to the lower centroid). The
// centroids pass through LDS in chunks of KMC, transposed so a thread reads one value per FMA.
__global__ void __launch_bounds__(256) k_kmeans_assign(const float *X, uint32_t n, uint32_t d, const float *cen,
                                                       const float *cnorm, uint32_t C, uint32_t *assign) {
    __shared__ float sc[KMC * 128];
    __shared__ float sn[KMC];
    for (uint32_t i0 = blockIdx.x * 256u; i0 < n; i0 += gridDim.x * 256u) {  // workgroup-uniform
        const uint32_t i = i0 + threadIdx.x;
        const bool ok = i < n;
        float best = INFINITY;
        uint32_t bid = 0;
        for (uint32_t c0 = 0; c0 < C; c0 += KMC) {
            const uint32_t cc = min((uint32_t)KMC, C - c0);
            __syncthreads();
            for (uint32_t e = threadIdx.x; e < cc * d; e += 256) {
                const uint32_t c = e / d, t = e % d;
                sc[t * KMC + c] = cen[(uint64_t)(c0 + c) * d + t];
            }
            if (threadIdx.x < cc) sn[threadIdx.x] = cnorm[c0 + threadIdx.x];
            __syncthreads();
            if (ok) {
                float acc[KMC];
#pragma unroll
                for (int c = 0; c < KMC; c++) acc[c] = 0.0f;
                const float *x = X + (uint64_t)i * d;
                for (uint32_t t = 0; t < d; t++) {
                    const float xv = x[t];
#pragma unroll
                    for (int c = 0; c < KMC; c++) acc[c] = fmaf(xv, sc[t * KMC + c], acc[c]);
                }
#pragma unroll
                for (int c = 0; c < KMC; c++) {
                    const float dist = sn[c] - 2.0f * acc[c];
                    if ((uint32_t)c < cc && dist < best) {
                        best = dist;
                        bid = c0 + (uint32_t)c;
                    }
                }
            }
        }
        if (ok) assign[i] = bid;
    }
}

// The same assignment with the point's coordinates in registers (DMAX >= d), loaded once instead
// of once per centroid chunk; coordinates t >= d and absent centroids of the last chunk are
// zero in LDS, and a zero term leaves the FMA chain unchanged, so the sums equal the kernel above.
template <int DMAX>
__global__ void __launch_bounds__(256) k_kmeans_assign_r(const float *X, uint32_t n, uint32_t d, const float *cen,
                                                         const float *cnorm, uint32_t C, uint32_t *assign) {
    __shared__ float sc[KMC * DMAX];
    __shared__ float sn[KMC];
    for (uint32_t i0 = blockIdx.x * 256u; i0 < n; i0 += gridDim.x * 256u) {  // workgroup-uniform
        const uint32_t i = i0 + threadIdx.x;
        const bool ok = i < n;
        float xr[DMAX];
#pragma unroll
        for (int t = 0; t < DMAX; t++) xr[t] = (ok && (uint32_t)t < d) ? X[(uint64_t)i * d + t] : 0.0f;
        float best = INFINITY;
        uint32_t bid = 0;
        for (uint32_t c0 = 0; c0 < C; c0 += KMC) {
            const uint32_t cc = min((uint32_t)KMC, C - c0);
            __syncthreads();
            for (uint32_t e = threadIdx.x; e < (uint32_t)KMC * DMAX; e += 256) {
                const uint32_t c = e / DMAX, t = e % DMAX;
                sc[t * KMC + c] = (c < cc && t < d) ? cen[(uint64_t)(c0 + c) * d + t] : 0.0f;
            }
            if (threadIdx.x < KMC) sn[threadIdx.x] = threadIdx.x < cc ? cnorm[c0 + threadIdx.x] : 0.0f;
            __syncthreads();
            float acc[KMC];
#pragma unroll
            for (int c = 0; c < KMC; c++) acc[c] = 0.0f;
#pragma unroll
            for (int t = 0; t < DMAX; t++) {
#pragma unroll
                for (int c = 0; c < KMC; c++) acc[c] = fmaf(xr[t], sc[t * KMC + c], acc[c]);
            }
#pragma unroll
            for (int c = 0; c < KMC; c++) {
                const float dist = sn[c] - 2.0f * acc[c];
                if ((uint32_t)c < cc && dist < best) {
                    best = dist;
                    bid = c0 + (uint32_t)c;
                }
            }
        }
        if (ok) assign[i] = bid;
    }
}

// start[g] = first position of the sorted keys holding a key >= g (g = 0..G).
__global__ void k_key_starts_u32(const uint32_t *keys, uint32_t n, uint32_t G, uint32_t *start) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = n;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (keys[mid] < g) lo = mid + 1;
            else hi = mid;
        }
        start[g] = lo;
    }
}

__device__ __forceinline__ double block_sum_f64(double v, double *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

__device__ __forceinline__ double block_max_f64(double v, double *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Centroid of every group: the mean of its members (positions start[g] .. start[g + 1] of perm);
// an empty group keeps its centroid. One workgroup per group: thread (slice, dimension) adds the
// dimension over every slices-th member, then the slices' sums are added in LDS (the order only
// moves the grouping, never a result).
__global__ void __launch_bounds__(256) k_kmeans_update(const float *X, uint32_t d, const uint32_t *perm,
                                                       const uint32_t *start, uint32_t C, float *cen) {
    __shared__ float part[256];
    __shared__ uint32_t sidx[256];
    const uint32_t slices = 256 / d;  // d <= 128
    const uint32_t sl = threadIdx.x / d, t = threadIdx.x % d;
    for (uint32_t g = blockIdx.x; g < C; g += gridDim.x) {
        const uint32_t b = start[g], e = start[g + 1];
        if (b == e) continue;  // workgroup-uniform
        // slice sl sums the group's positions q = sl, sl + slices, sl + 2 slices, ... in order,
        // the even-numbered ones into acc and the odd-numbered ones into acc2; the point indices
        // pass through LDS 256 at a time, so eight coordinate loads are in flight per batch
        float acc = 0.0f, acc2 = 0.0f;
        for (uint32_t c0 = 0; c0 < e - b; c0 += 256) {  // workgroup-uniform
            const uint32_t cl = min(256u, e - b - c0);
            __syncthreads();
            if (threadIdx.x < cl) sidx[threadIdx.x] = perm[b + c0 + threadIdx.x];
            __syncthreads();
            if (sl < slices) {
                uint32_t p = c0 + (sl + slices - c0 % slices) % slices;  // first position of the slice
                uint32_t q = (p - sl) / slices;
                for (; p < c0 + cl; p += 8 * slices, q += 8) {
                    float v[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const uint32_t pp = p + (uint32_t)u * slices;
                        v[u] = pp < c0 + cl ? X[(uint64_t)sidx[pp - c0] * d + t] : 0.0f;
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        if (p + (uint32_t)u * slices >= c0 + cl) break;
                        if ((q + (uint32_t)u) & 1u) acc2 += v[u];
                        else acc += v[u];
                    }
                }
            }
        }
        acc += acc2;
        __syncthreads();
        part[threadIdx.x] = acc;
        __syncthreads();
        if (threadIdx.x < d) {
            float sum = 0.0f;
            for (uint32_t q = 0; q < slices; q++) sum += part[q * d + threadIdx.x];
            cen[(uint64_t)g * d + threadIdx.x] = sum / (float)(e - b);
        }
    }
}

// Bounding sphere of every block of bsz (<= 256) consecutive layout positions (n of them,
// perm[j] == KNONE for padding): centre (mean of the block's points, f64) and radius (the
// largest distance to it, rounded up so that it bounds every point's exact distance); cnt =
// points in the block.
__global__ void __launch_bounds__(256) k_block_spheres(const float *X, const uint32_t *perm, uint32_t n, uint32_t d,
                                                       uint32_t bsz, uint32_t nblocks, double *cen, double *rad,
                                                       uint32_t *cnt) {
    // one wave per block: the block's point ids go to LDS, lane t sums coordinate t over the
    // points in block order (one coalesced row read per point), then each lane takes the
    // squared distances of its points (lane, lane + 64, ...) and a butterfly takes their maximum
    __shared__ uint32_t ssrc[4][256];
    __shared__ double sc[4][128];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t b = blockIdx.x * 4u + (uint32_t)wave; b < nblocks; b += gridDim.x * 4u) {  // wave-uniform
        uint32_t creal = 0;
        for (uint32_t p = (uint32_t)lane; p < bsz; p += 64) {
            const uint64_t j = (uint64_t)b * bsz + p;
            const uint32_t src = j < n ? perm[j] : KNONE;  // n: layout rows here
            ssrc[wave][p] = src;
            creal += src != KNONE ? 1u : 0u;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) creal += __shfl_xor(creal, o);
        const double c = (double)creal;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t t = (uint32_t)lane; t < d; t += 64) {
            double v = 0.0;
#pragma unroll 8
            for (uint32_t p = 0; p < bsz; p++) {
                const uint32_t src = ssrc[wave][p];
                if (src != KNONE) v += (double)X[(uint64_t)src * d + t];
            }
            sc[wave][t] = c > 0.0 ? v / c : 0.0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double r2 = 0.0;
        for (uint32_t p = (uint32_t)lane; p < bsz; p += 64) {
            const uint32_t src = ssrc[wave][p];
            if (src == KNONE) continue;
            double q = 0.0;
            for (uint32_t t = 0; t < d; t++) {
                const double df = (double)X[(uint64_t)src * d + t] - sc[wave][t];
                q += df * df;
            }
            r2 = fmax(r2, q);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o));
        if (lane == 0) {
            rad[b] = c > 0.0 ? sqrt(r2) * (1.0 + 1e-12) + 1e-300 : 0.0;
            cnt[b] = creal;
        }
        for (uint32_t t = (uint32_t)lane; t < d; t += 64) cen[(uint64_t)b * d + t] = sc[wave][t];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");  // ssrc and sc are rewritten by the next block
    }
}

// Group of every row block: groups occupy whole row blocks of the layout, group g the positions
// [poff[g], poff[g + 1]), so row block R belongs to the largest g with poff[g] <= 256 R (groups
// without points have empty ranges and are passed over by the search).
__global__ void k_rb_groups(const uint32_t *poff, uint32_t C, uint32_t nrb, uint32_t *rbg) {
    for (uint32_t R = blockIdx.x * blockDim.x + threadIdx.x; R < nrb; R += gridDim.x * blockDim.x) {
        const uint32_t p = R * (uint32_t)KNN_BM;
        uint32_t lo = 0, hi = C;  // poff[lo] <= p < poff[hi] (poff[C] = layout rows > p)
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (poff[mid] <= p) lo = mid;
            else hi = mid;
        }
        rbg[R] = lo;
    }
}

// Core of every group: its points run from the centre outwards (the layout's order), and its
// first ceil(0.9 size / 64) blocks of 64 positions are its core, ending at cend[g]; the rest
// (the farthest tenth, which holds the outliers that set the group's radius) is its tail.
__global__ void k_group_core(const uint32_t *start, const uint32_t *poff, uint32_t C, uint32_t *cend) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < C; g += gridDim.x * blockDim.x) {
        const uint32_t size = start[g + 1] - start[g];
        const uint32_t ncore = min((9u * size + 639u) / 640u, (size + 63u) / 64u);
        cend[g] = poff[g] + 64u * ncore;
    }
}

// Squared radii of every group about its centre cen_g: r2bits[g] = the bits of the largest
// |y - cen_g|^2 over the group's points, r2core[g] over its core's, summed in f64 from the f32
// values (relative error below (d + 2) 2^-53; the readers inflate the radius by 1e-12).
// Non-negative doubles order like their bit patterns, so the maximum is an integer atomic. A
// wave covers 64 consecutive positions of one row block, hence of one group and of its core or
// its tail: one atomic per wave and radius.
__global__ void __launch_bounds__(256) k_group_radii(const float *X, const uint32_t *layout, uint32_t rows, uint32_t d,
                                                     const float *cen, const uint32_t *rbg, const uint32_t *cend,
                                                     unsigned long long *r2bits, unsigned long long *r2core) {
    for (uint32_t p0 = blockIdx.x * 256u; p0 < rows; p0 += gridDim.x * 256u) {  // rows: whole row blocks
        const uint32_t p = p0 + threadIdx.x;
        const uint32_t g = rbg[p / (uint32_t)KNN_BM];
        const uint32_t y = layout[p];
        double r2 = -1.0;
        if (y != KNONE) {
            r2 = 0.0;
            for (uint32_t t = 0; t < d; t++) {
                const double df = (double)X[(uint64_t)y * d + t] - (double)cen[(uint64_t)g * d + t];
                r2 += df * df;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o));
        if ((threadIdx.x & 63) == 0 && r2 >= 0.0) {
            const unsigned long long b = (unsigned long long)__double_as_longlong(r2);
            atomicMax(&r2bits[g], b);
            if (p < cend[g]) atomicMax(&r2core[g], b);
        }
    }
}

// Per-group data of the group-level pruning: centres transposed (cenT[t * C + g]) and their f32
// norms, and the radii rounded up: r64 = sqrt(r2) (1 + 1e-12) (f64), rf = r64 rounded up to f32,
// rcf the same for the core.
__global__ void k_group_prep(const float *cen, uint32_t C, uint32_t d, const unsigned long long *r2bits,
                             const unsigned long long *r2core, float *cenT, float *cn, double *r64, float *rf,
                             float *rcf) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < C; g += gridDim.x * blockDim.x) {
        float q = 0.0f;
        for (uint32_t t = 0; t < d; t++) {
            const float c = cen[(uint64_t)g * d + t];
            cenT[(uint64_t)t * C + g] = c;
            q = fmaf(c, c, q);
        }
        cn[g] = sqrtf(q);
        // a group without points keeps radius 0; its layout range, hence its block list, is empty
        const double r = sqrt(__longlong_as_double((long long)r2bits[g])) * (1.0 + 1e-12);
        r64[g] = r;
        rf[g] = f32_up(r);
        rcf[g] = f32_up(sqrt(__longlong_as_double((long long)r2core[g])) * (1.0 + 1e-12));
    }
}

// Groups each row block needs, as two bit masks of mwords words per row block: gmask[R][0] the
// groups needed at all, gmask[R][1] those whose core is needed (a subset: the core's radius is
// at most the group's). Every point y of group g has |x - y| >= |x - cen_g| - r_g, so g holds
// none of row x's k nearest points when |x - cen_g| - r_g exceeds u_x, the row's bound on its
// k-th nearest distance (sqrt(seed_d2)), and its core none when |x - cen_g| - rc_g does; the
// block needs what some row needs. The bound is per row: a row block's bounding sphere
// adds the block's radius to the group's, which in high dimensions is most of the distance
// between neighbouring clusters. The block's sphere (cR, rR, uR >= every u_x) first passes over
// the groups no row can reach (|cR - cen_g| - rR - r_g > uR, the test of the column-block lists
// with its f32 error terms), and the per-row test runs on the rest. Per-row distances are
// summed in f32 from the f32 coordinates and centres, a relative error below (DMAX + 3) 2^-24 on
// the square; r_g and u_x are rounded up to f32 and their sum inflated by 1e-6, and the test
// takes 2e-5 relatively off the distance. A row without a bound (+inf) needs every group. One
// workgroup per row block, one thread per row with its coordinates in registers; 64 centres at a
// time pass through LDS.
constexpr uint32_t KNN_MAX_GROUPS = 2048;
template <int DMAX>
__global__ void __launch_bounds__(256) k_row_groups(const float *X, const uint32_t *layout, uint32_t d, uint32_t nrb,
                                                    uint32_t C, const float *cen, const float *cenT, const float *cn,
                                                    const double *r64, const float *rf, const float *rcf,
                                                    const double *cR, const double *rR, const double *uR,
                                                    const double *seed_d2, uint32_t mwords, uint32_t *gmask) {
    __shared__ float sc[64 * DMAX];
    __shared__ float sr[64], src[64];
    __shared__ float scr[DMAX];
    __shared__ uint32_t sm[2 * (KNN_MAX_GROUPS / 32)];
    __shared__ uint32_t sg[KNN_MAX_GROUPS];
    __shared__ uint32_t s_ng;
    __shared__ float s_nr;
    const int lane = threadIdx.x & 63;
    for (uint32_t R = blockIdx.x; R < nrb; R += gridDim.x) {
        const uint64_t j = (uint64_t)R * KNN_BM + threadIdx.x;
        const uint32_t y = layout[j];
        const bool real = y != KNONE;
        const double sd = real ? seed_d2[j] : 0.0;
        float xr[DMAX];
#pragma unroll
        for (int t = 0; t < DMAX; t++) xr[t] = (real && (uint32_t)t < d) ? X[(uint64_t)y * d + t] : 0.0f;
        const float ux = real ? f32_up(sqrt(sd)) : 0.0f;
        for (uint32_t w = threadIdx.x; w < 2 * mwords; w += 256) sm[w] = 0u;
        for (uint32_t t = threadIdx.x; t < DMAX; t += 256) scr[t] = t < d ? (float)cR[(uint64_t)R * d + t] : 0.0f;
        if (threadIdx.x == 0) s_ng = 0u;
        const bool all = __syncthreads_or(real && !(sd < INFINITY)) != 0;
        const double ur = uR[R];
        if (all) {
            for (uint32_t w = threadIdx.x; w < mwords; w += 256)
                sm[w] = sm[mwords + w] = (w + 1) * 32 <= C ? 0xFFFFFFFFu : ((1u << (C - w * 32)) - 1u);
        } else if (ur >= 0.0) {
            if (threadIdx.x == 0) {
                float q = 0.0f;
                for (uint32_t t = 0; t < d; t++) q = fmaf(scr[t], scr[t], q);
                s_nr = sqrtf(q);
            }
            __syncthreads();
            const double lim = rR[R] + ur;
            const double nrn = (double)s_nr;
            for (uint32_t g = threadIdx.x; g < C; g += 256) {
                float s2 = 0.0f;
                for (uint32_t t = 0; t < d; t++) {
                    const float df = scr[t] - cenT[(uint64_t)t * C + g];
                    s2 = fmaf(df, df, s2);
                }
                const bool skip = (double)sqrtf(s2) * (1.0 - 1e-5) - 1e-6 * (nrn + (double)cn[g]) > (lim + r64[g]) * (1.0 + 1e-12);
                if (!skip) sg[atomicAdd(&s_ng, 1u)] = g;
            }
            __syncthreads();
            const uint32_t ng = s_ng;
            for (uint32_t g0 = 0; g0 < ng; g0 += 64) {
                const uint32_t gn = min(64u, ng - g0);
                __syncthreads();
                for (uint32_t e = threadIdx.x; e < 64 * DMAX; e += 256) {
                    const uint32_t c = e / DMAX, t = e % DMAX;
                    sc[e] = (c < gn && t < d) ? cen[(uint64_t)sg[g0 + c] * d + t] : 0.0f;
                }
                if (threadIdx.x < gn) {
                    sr[threadIdx.x] = rf[sg[g0 + threadIdx.x]];
                    src[threadIdx.x] = rcf[sg[g0 + threadIdx.x]];
                }
                __syncthreads();
                for (uint32_t c = 0; c < gn; c++) {
                    float s2 = 0.0f;
#pragma unroll
                    for (int t = 0; t < DMAX; t++) {
                        const float df = xr[t] - sc[c * DMAX + t];
                        s2 = fmaf(df, df, s2);
                    }
                    const float dl = sqrtf(s2) * (1.0f - 2e-5f);
                    const bool need = real && !(dl > (sr[c] + ux) * (1.0f + 1e-6f));
                    const bool need_core = real && !(dl > (src[c] + ux) * (1.0f + 1e-6f));
                    const bool any = __ballot(need) != 0ull, any_core = __ballot(need_core) != 0ull;
                    if (lane == 0 && any) {
                        const uint32_t g = sg[g0 + c];
                        atomicOr(&sm[g >> 5], 1u << (g & 31u));
                        if (any_core) atomicOr(&sm[mwords + (g >> 5)], 1u << (g & 31u));
                    }
                }
            }
        }
        __syncthreads();
        for (uint32_t w = threadIdx.x; w < 2 * mwords; w += 256) gmask[(uint64_t)R * 2 * mwords + w] = sm[w];
        __syncthreads();
    }
}

// Column blocks each row block needs: the blocks holding points of the groups in its masks
// (k_row_groups; a group's core blocks when its core is needed, its tail blocks when the group
// is), its own group first (its nearest points tighten the tile's thresholds early), then the
// others in increasing order, each group's blocks in layout order. A group's points fill the
// start of its layout range, so a block holds points when its first position does. Row blocks
// without points (uR < 0) need nothing. One wave per row block. Count pass (blist null):
// counts[R]; fill pass: blist[boff[R] + rank] = B.
__global__ void __launch_bounds__(256) k_group_lists(const double *uR, const uint32_t *layout, uint32_t nrb,
                                                     const uint32_t *poff, const uint32_t *cend, const uint32_t *rbg,
                                                     uint32_t mwords, const uint32_t *gmask, uint32_t tn,
                                                     uint32_t *counts, const uint32_t *boff, uint32_t *blist) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t R = blockIdx.x * 4u + wave; R < nrb; R += gridDim.x * 4u) {  // wave-uniform
        uint32_t total = 0;
        if (uR[R] >= 0.0) {
            auto range = [&](uint32_t b0, uint32_t b1) {
                for (uint32_t B0 = b0; B0 < b1; B0 += 64) {
                    const uint32_t B = B0 + (uint32_t)lane;
                    const bool need = B < b1 && layout[(uint64_t)B * tn] != KNONE;
                    const uint64_t m = __ballot(need);
                    if (blist && need)
                        blist[boff[R] + total +
                              __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = B;
                    total += (uint32_t)__popcll(m);
                }
            };
            const uint32_t *mall = gmask + (uint64_t)R * 2 * mwords, *mcore = mall + mwords;
            auto group = [&](uint32_t g) {
                if ((mcore[g >> 5] >> (g & 31u)) & 1u) range(poff[g] / tn, cend[g] / tn);
                range(cend[g] / tn, poff[g + 1] / tn);
            };
            const uint32_t own = rbg[R];
            if ((mall[own >> 5] >> (own & 31u)) & 1u) group(own);
            for (uint32_t w = 0; w < mwords; w++) {
                uint32_t word = mall[w];
                if (w == (own >> 5)) word &= ~(1u << (own & 31u));
                while (word) {
                    const uint32_t bit = (uint32_t)__builtin_ctz(word);
                    word &= word - 1u;
                    group(w * 32u + bit);
                }
            }
        }
        if (!blist && lane == 0) counts[R] = total;
    }
}

// Column-block spheres for the per-row list filter, in f32: centre cB32 = cB rounded to f32
// (moving it by at most 2^-24 |cB|), its squared norm (f32 sum) and norm, and the radius rounded up.
__global__ void k_block_prep(const double *cB, const double *rB, uint32_t ncb, uint32_t d, float *cB32, float *rBf,
                             float *nbn, float *nb2) {
    for (uint32_t B = blockIdx.x * blockDim.x + threadIdx.x; B < ncb; B += gridDim.x * blockDim.x) {
        float q = 0.0f;
        for (uint32_t t = 0; t < d; t++) {
            const float c = (float)cB[(uint64_t)B * d + t];
            cB32[(uint64_t)B * d + t] = c;
            q = fmaf(c, c, q);
        }
        nb2[B] = q;
        nbn[B] = sqrtf(q);
        rBf[B] = f32_up(rB[B]);
    }
}

// Per-row filter of every row block's list of column blocks, in place: block B stays when some
// row x of the block has |x - cB| - rB <= u_x (u_x: the row's bound on its k-th nearest
// distance). The squared distance to cB32 is evaluated as |x|^2 + |cB32|^2 - 2 x.cB32 in f32
// (one FMA per coordinate); its error is below (3 DMAX + 8) 2^-24 (|x|^2 + |cB32|^2) (the two
// norms and the dot product, |x.c| <= (|x|^2 + |c|^2) / 2, and the final additions), which is
// taken off before the square root, and the root's rounding and 1e-6 |cB32| (rounding of the
// centre) off after; rB and u_x are rounded up to f32 and their sum inflated by 1e-6. A row
// block with a row without a bound keeps its list. The list of row block R is blist[lbeg[R] .. lbeg[R] +
// lcnt[R]); the kept blocks move to its front in order, and lkept[R] gets their count. One
// workgroup per row block, one thread per row; 64 blocks at a time pass through LDS, read
// before any of them is written back (a block is written at or before its own position).
template <int DMAX>
__global__ void __launch_bounds__(256) k_list_filter(const float *X, const uint32_t *layout, uint32_t d, uint32_t nrb,
                                                     const double *seed_d2, const float *cB32, const float *rBf,
                                                     const float *nbn, const float *nb2, const uint32_t *lbeg,
                                                     const uint32_t *lcnt, uint32_t *blist, uint32_t *lkept) {
    __shared__ float sc[64 * DMAX];
    __shared__ float sr[64], sn[64], sq[64];
    __shared__ uint32_t sB[64];
    __shared__ uint32_t sf[4][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t R = blockIdx.x; R < nrb; R += gridDim.x) {
        const uint64_t j = (uint64_t)R * KNN_BM + threadIdx.x;
        const uint32_t y = layout[j];
        const bool real = y != KNONE;
        const double sd = real ? seed_d2[j] : 0.0;
        float xr[DMAX];
#pragma unroll
        for (int t = 0; t < DMAX; t++) xr[t] = (real && (uint32_t)t < d) ? X[(uint64_t)y * d + t] : 0.0f;
        const float ux = real ? f32_up(sqrt(sd)) : 0.0f;
        float xn2 = 0.0f;
#pragma unroll
        for (int t = 0; t < DMAX; t++) xn2 = fmaf(xr[t], xr[t], xn2);
        constexpr float ecoef = (float)(3 * DMAX + 8) / 16777216.0f;
        const uint32_t L = lcnt[R], base = lbeg[R];
        const bool all = __syncthreads_or(real && !(sd < INFINITY)) != 0;
        if (all) {
            if (threadIdx.x == 0) lkept[R] = L;
            continue;  // workgroup-uniform
        }
        uint32_t kept = 0;  // wave 0's count (the only wave that writes)
        for (uint32_t c0 = 0; c0 < L; c0 += 64) {
            const uint32_t cn = min(64u, L - c0);
            __syncthreads();
            if (threadIdx.x < cn) {
                const uint32_t B = blist[base + c0 + threadIdx.x];
                sB[threadIdx.x] = B;
                sr[threadIdx.x] = rBf[B];
                sn[threadIdx.x] = nbn[B];
                sq[threadIdx.x] = nb2[B];
            }
            __syncthreads();
            for (uint32_t e = threadIdx.x; e < 64 * DMAX; e += 256) {
                const uint32_t c = e / DMAX, t = e % DMAX;
                sc[e] = (c < cn && t < d) ? cB32[(uint64_t)sB[c] * d + t] : 0.0f;
            }
            __syncthreads();
            uint64_t keep = 0;  // bit c: some row of this wave needs block c (wave-uniform)
            for (uint32_t c = 0; c < cn; c++) {
                float dot = 0.0f;
#pragma unroll
                for (int t = 0; t < DMAX; t++) dot = fmaf(xr[t], sc[c * DMAX + t], dot);
                const float nn = xn2 + sq[c];
                const float lo2 = fmaf(-2.0f, dot, nn) - ecoef * nn;
                const bool need =
                    real && !(sqrtf(fmaxf(lo2, 0.0f)) * (1.0f - 1e-6f) - 1e-6f * sn[c] > (sr[c] + ux) * (1.0f + 1e-6f));
                keep |= (uint64_t)(__ballot(need) != 0ull) << c;
            }
            if (lane == 0) {
                sf[wave][0] = (uint32_t)keep;
                sf[wave][1] = (uint32_t)(keep >> 32);
            }
            __syncthreads();
            if (wave == 0) {
                const uint32_t lo = sf[0][0] | sf[1][0] | sf[2][0] | sf[3][0];
                const uint32_t hi = sf[0][1] | sf[1][1] | sf[2][1] | sf[3][1];
                const bool k = (uint32_t)lane < cn && (((lane < 32 ? lo >> lane : hi >> (lane - 32)) & 1u) != 0u);
                const uint64_t m = __ballot(k);
                if (k)
                    blist[base + kept +
                          __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = sB[lane];
                kept += (uint32_t)__popcll(m);
            }
        }
        if (threadIdx.x == 0) lkept[R] = kept;
        __syncthreads();
    }
}

// The column blocks' centres as the matrix cores' B rows (d + 2 <= 32): split-BF16 cB32 (hi in
// cbh, lo in cbl) with the f32 squared norm nb2 split into slots d (hi) and d + 1 (lo) of cbh.
__global__ void k_block_pack(const float *cB32, const float *nb2, uint32_t ncb, uint32_t d, unsigned short *cbh,
                             unsigned short *cbl) {
    for (uint32_t B = blockIdx.x * blockDim.x + threadIdx.x; B < ncb; B += gridDim.x * blockDim.x) {
        for (uint32_t t = 0; t < 32; t++) {
            unsigned short h = 0, l = 0;
            if (t < d) {
                const float v = cB32[(uint64_t)B * d + t];
                h = bf16_bits_rn(v);
                l = bf16_bits_rn(v - bf16_bits_to_f32(h));
            }
            cbh[(uint64_t)B * 32 + t] = h;
            cbl[(uint64_t)B * 32 + t] = l;
        }
        const float q = nb2[B];
        const unsigned short qh = bf16_bits_rn(q);
        cbh[(uint64_t)B * 32 + d] = qh;
        cbh[(uint64_t)B * 32 + d + 1] = bf16_bits_rn(q - bf16_bits_to_f32(qh));
    }
}

// The per-row list filter of k_list_filter (same decision rule) for d + 2 <= 32, with the
// squared distances on the matrix cores: s = |cB32|^2 - 2 x.cB32 as split-BF16 products of the
// row (-2 x, 1, 1) and the centre (cB32, |cB32|^2 hi, lo), FP32 accumulation. Its error against
// the exact value is below e1 |x| |c| + e2 |c|^2 (the split truncation and the accumulation of
// the 3 x 32 + 2 products, the f32 norm and its split; the same terms as the tile's bound, per
// row and centre instead of with the largest norm). The test is k_list_filter's without the
// root: block B is dropped for row x when |x - cB32| > R = rB + u_x + 2^-24 |cB| (the centre's
// rounding), i.e. when s > T with T an f32 upper bound of R^2 + err - |x|^2 (R rounded up with a
// 1e-6 relative margin, each positive term inflated and the subtracted |x|^2 lower bound
// deflated by three units of rounding, which covers the two f32 additions forming T). A padding
// row (|x|^2 bound +inf) needs no block. A wave holds 64 rows as four row tiles; 64 listed
// blocks at a time are staged in LDS, read before any of them is written back.
__global__ void __launch_bounds__(256) k_list_filter_mfma(const float *X, const uint32_t *layout, uint32_t d,
                                                          uint32_t nrb, const double *seed_d2, const unsigned short *cbh,
                                                          const unsigned short *cbl, const float *rBf, const float *nbn,
                                                          const uint32_t *lbeg, const uint32_t *lcnt, uint32_t *blist,
                                                          uint32_t *lkept, float e1, float e2) {
    constexpr int STR = 40;
    __shared__ __attribute__((aligned(16))) unsigned short sBh[64 * STR], sBl[64 * STR];
    __shared__ float sr[64], sn[64];
    __shared__ uint32_t sB[64];
    __shared__ float s_lox[256], s_ex[256], s_ux[256];  // per row: |x|^2 lower bound, e1 |x|, u_x
    __shared__ uint32_t sf[4][2];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, quarter = lane >> 4;
    constexpr float u24 = 1.0f / 16777216.0f;
    for (uint32_t R = blockIdx.x; R < nrb; R += gridDim.x) {  // workgroup-uniform
        // this thread's row: its norm and bound; a padding row gets u_x = -inf, so that it
        // needs no block (x > -inf holds for the finite left side of the test)
        bool norow;
        {
            const uint64_t j = (uint64_t)R * KNN_BM + tid;
            const uint32_t y = layout[j];
            const bool real = y != KNONE;
            float xn2 = 0.0f;
            for (uint32_t t = 0; real && t < d; t++) {
                const float v = X[(uint64_t)y * d + t];
                xn2 = fmaf(v, v, xn2);
            }
            const double sd = real ? seed_d2[j] : 0.0;
            norow = real && !(sd < INFINITY);
            // the f32 sum of d squares is within 1.01 d 2^-24 |x|^2 of |x|^2; the factor's own
            // rounding and the product's take at most 2 more units
            // (then three more units off: the subtracted term of T)
            s_lox[tid] = real ? xn2 * (1.0f - (1.01f * (float)d + 6.0f) * u24) * (1.0f - 3.0f * u24) : INFINITY;
            // e1 |x| with the product's and the error fma's rounding covered
            s_ex[tid] = real ? e1 * sqrtf(xn2) * (1.0f + 4.0f * u24) * (1.0f + 4.0f * u24) : 0.0f;
            s_ux[tid] = real ? f32_up(sqrt(sd)) * (1.0f + 1e-6f) : 0.0f;
        }
        const uint32_t L = lcnt[R], base = lbeg[R];
        if (__syncthreads_or(norow)) {
            if (tid == 0) lkept[R] = L;
            continue;  // workgroup-uniform: a row without a bound keeps the whole list
        }
        // A fragments of the wave's four row tiles
        bf16x8 ah[4], al[4];
#pragma unroll
        for (int rt = 0; rt < 4; rt++) {
            const uint64_t j = (uint64_t)R * KNN_BM + wave * 64 + rt * 16 + (lane & 15);
            const uint32_t y = layout[j];
            union {
                bf16x8 v;
                unsigned short u[8];
            } h, l;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const uint32_t t = 8 * quarter + e;
                if (t < d) {
                    const float xv = y != KNONE ? X[(uint64_t)y * d + t] : 0.0f;
                    const unsigned short xh = bf16_bits_rn(xv);
                    const unsigned short xl = bf16_bits_rn(xv - bf16_bits_to_f32(xh));
                    h.u[e] = bf16_bits_rn(-2.0f * bf16_bits_to_f32(xh));  // exact: scaling by -2
                    l.u[e] = bf16_bits_rn(-2.0f * bf16_bits_to_f32(xl));
                } else {
                    h.u[e] = t < d + 2 ? (unsigned short)0x3F80u : (unsigned short)0;  // 1.0 at the norm slots
                    l.u[e] = 0;
                }
            }
            ah[rt] = h.v;
            al[rt] = l.v;
        }
        // this lane's 16 rows (row tile rt, value i: row 16 rt + 4 quarter + i of the wave)
        float lox[4][4], ex[4][4], ux[4][4];
#pragma unroll
        for (int rt = 0; rt < 4; rt++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = wave * 64 + rt * 16 + 4 * quarter + i;
                lox[rt][i] = s_lox[r];
                ex[rt][i] = s_ex[r];
                ux[rt][i] = s_ux[r];
            }
        uint32_t kept = 0;  // wave 0's count (the only wave that writes)
        for (uint32_t c0 = 0; c0 < L; c0 += 64) {
            const uint32_t cn = min(64u, L - c0);
            __syncthreads();
            if (tid < 64) {
                const uint32_t B = tid < (int)cn ? blist[base + c0 + tid] : 0u;
                sB[tid] = B;
                sr[tid] = tid < (int)cn ? rBf[B] : INFINITY;
                sn[tid] = tid < (int)cn ? nbn[B] : 0.0f;
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int ch = tid + 256 * q;
                const int arr = ch >> 8, r = (ch >> 2) & 63, part = ch & 3;
                const unsigned short *src = (arr ? cbl : cbh) + (uint64_t)sB[r] * 32 + part * 8;
                unsigned short *dst = (arr ? sBl : sBh) + r * STR + part * 8;
                *reinterpret_cast<uint4 *>(dst) = *reinterpret_cast<const uint4 *>(src);
            }
            __syncthreads();
            uint64_t keep = 0;  // bit c: some row of this wave needs block c (wave-uniform)
#pragma unroll
            for (int ct = 0; ct < 4; ct++) {
                const int col = ct * 16 + (lane & 15);
                const int off = col * STR + 8 * quarter;
                const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(sBh + off);
                const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(sBl + off);
                const float cn_ = sn[col] * (1.0f + 4.0f * u24);  // |cB32|, rounded up
                const float e2c = e2 * cn_ * cn_ * (1.0f + 4.0f * u24) * (1.0f + 4.0f * u24);
                // rB and the centre's rounding, with the 1e-6 margin (u_x carries its own)
                const float rcp = sr[col] * (1.0f + 1e-6f) + 1e-6f * cn_;
                bool need = false;
#pragma unroll
                for (int rt = 0; rt < 4; rt++) {
                    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt], bh, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt], bl, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rt], bh, acc, 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const float Rx = rcp + ux[rt][i];
                        const float R2 = Rx * (Rx * (1.0f + 5.0f * u24));  // >= R^2 (1 + 3 units)
                        const float err = fmaf(ex[rt][i], cn_, e2c);
                        const float T = (R2 + err) - lox[rt][i];
                        need |= !(acc[i] > T);
                    }
                }
                // the column's four quarters (lanes col, col + 16, col + 32, col + 48)
                const uint64_t m = __ballot(need);
                const uint32_t m16 = (uint32_t)(m | (m >> 16) | (m >> 32) | (m >> 48)) & 0xFFFFu;
                keep |= (uint64_t)m16 << (16 * ct);
            }
            if (lane == 0) {
                sf[wave][0] = (uint32_t)keep;
                sf[wave][1] = (uint32_t)(keep >> 32);
            }
            __syncthreads();
            if (wave == 0) {
                const uint32_t lo = sf[0][0] | sf[1][0] | sf[2][0] | sf[3][0];
                const uint32_t hi = sf[0][1] | sf[1][1] | sf[2][1] | sf[3][1];
                const bool k = (uint32_t)lane < cn && (((lane < 32 ? lo >> lane : hi >> (lane - 32)) & 1u) != 0u);
                const uint64_t mk = __ballot(k);
                if (k)
                    blist[base + kept +
                          __builtin_amdgcn_mbcnt_hi((uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u))] = sB[lane];
                kept += (uint32_t)__popcll(mk);
            }
        }
        if (tid == 0) lkept[R] = kept;
        __syncthreads();
    }
}

// Sort keys of the longest-list-first order: the complement of the list length, and the row block.
__global__ void k_lpt_keys(const uint32_t *bcnt, uint32_t nrb, uint32_t *keys, uint32_t *ids) {
    for (uint32_t R = blockIdx.x * blockDim.x + threadIdx.x; R < nrb; R += gridDim.x * blockDim.x) {
        keys[R] = ~bcnt[R];
        ids[R] = R;
    }
}

// Seed of every layout row's threshold in the tile's s = |y|^2 - 2 x.y space: s_exact of a point
// within the row's k-th nearest distance is at most seed_d2 - |x|^2, and its approximation at most
// eps above, so starting the threshold at seed_d2 - |x|^2 + 2.02 eps (rounded up) records every
// k-nearest point, while a value at or above it belongs to a point strictly farther than the
// row's k-th nearest (no candidate a certificate needs is lost).
__global__ void k_seed_thresholds(const double *seed_d2, const float *ny, uint32_t rows, double eps_coef,
                                  double eps_const, float *seed_s) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < rows; j += gridDim.x * blockDim.x) {
        const double sd = seed_d2[j];
        const double y = (double)ny[j];
        float out = INFINITY;
        if (sd < INFINITY && y < INFINITY) {
            const double v = sd - y + 2.02 * (eps_coef * sqrt(y) + eps_const);
            out = (float)v;
            if ((double)out < v) out = nextafterf(out, INFINITY);
        }
        seed_s[j] = out;
    }
}

// Twice the error bound of every layout row's approximate values (2.02 eps, rounded up; +inf for
// padding rows): the tile's k-th-value threshold adds it to the row's k-th listed value.
__global__ void k_thr_init(const float *seed_s, uint32_t rows, uint32_t *gthr) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < rows; j += gridDim.x * blockDim.x)
        gthr[j] = f32_ord(seed_s ? seed_s[j] : INFINITY);
}

__global__ void k_knn_margins(const float *ny, uint32_t rows, double eps_coef, double eps_const, float *margin) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < rows; j += gridDim.x * blockDim.x) {
        const double y = (double)ny[j];
        float out = INFINITY;
        if (y < INFINITY) {
            const double v = 2.02 * (eps_coef * sqrt(y) + eps_const);
            out = (float)v;
            if ((double)out < v) out = nextafterf(out, INFINITY);
        }
        margin[j] = out;
    }
}

// Candidate lists of layout row j (per_row entries; rows layout rows) -> the point's own row,
// columns mapped through perm, so the rescore works on the points in their own order and index
// ties are broken on the points' own indices. Padding rows are dropped.
__global__ void k_knn_remap(const uint32_t *perm, uint32_t rows, uint32_t per_row, const uint32_t *cidx,
                            const float *cs, uint32_t *oidx, float *os) {
    const uint64_t total = (uint64_t)rows * per_row;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t j = (uint32_t)(e / per_row), q = (uint32_t)(e % per_row);
        const uint32_t src = perm[j];
        if (src == KNONE) continue;  // padding row
        const uint32_t c = cidx[e];
        const uint64_t o = (uint64_t)src * per_row + q;
        oidx[o] = c < rows ? perm[c] : KNONE;
        os[o] = cs[e];
    }
}

// uR of every row block: an upper bound, over its points, of the distance to the point's k-th
// nearest other point within a window of two row blocks (the block and the one before it, or
// after it for the first block), so every point of the block has k points within uR; +inf when
// the window holds k other points or fewer, -1 when the block holds none. Distances are summed in f32 from
// the stored f32 coordinates: every term is non-negative, so the sum's relative error is below
// (d + 2) 2^-24 and uR inflates the largest k-th squared distance by 1e-4 before the square
// root. One workgroup per row block, one thread per row with its coordinates in registers; the
// block's points pass through LDS in chunks of 64. seed_d2[row] gets the row's own bound on its
// k-th nearest squared distance (+inf when it has none).
// KK: entries of the per-row insertion network (KK >= k; k is at most FB_MAXK).
static_assert(FB_MAXK <= 32, "the widest insertion network holds 32 entries");
template <int DMAX, int KK>
__global__ void __launch_bounds__(256) k_block_kth(const float *X, const uint32_t *perm, uint32_t d, uint32_t k,
                                                   uint32_t nrb, double *uR, double *seed_d2) {
    __shared__ float sp[64 * DMAX];
    __shared__ uint32_t sv[64];
    __shared__ double red[4];
    for (uint32_t R = blockIdx.x; R < nrb; R += gridDim.x) {
        const uint32_t my = perm[(uint64_t)R * 256 + threadIdx.x];
        const bool real = my != KNONE;
        const double cnt = block_sum_f64(real ? 1.0 : 0.0, red);
        float xr[DMAX];
#pragma unroll
        for (int t = 0; t < DMAX; t++) xr[t] = (real && (uint32_t)t < d) ? X[(uint64_t)my * d + t] : 0.0f;
        float best[KK];
#pragma unroll
        for (int q = 0; q < KK; q++) best[q] = INFINITY;
        float kcur = INFINITY;  // best[k - 1]
        // window positions: this block's and the neighbour block's (when there is one)
        const uint64_t own = (uint64_t)R * 256;
        const uint64_t nb = nrb < 2 ? own : (R > 0 ? own - 256 : own + 256);
        const uint32_t wlen = nrb < 2 ? 256u : 512u;
        for (uint32_t c0 = 0; c0 < wlen; c0 += 64) {
            const uint64_t base = c0 < 256 ? own + c0 : nb + (c0 - 256);
            __syncthreads();
            for (uint32_t e = threadIdx.x; e < 64 * DMAX; e += 256) {
                const uint32_t c = e / DMAX, t = e % DMAX;
                const uint32_t src = perm[base + c];
                sp[e] = (src != KNONE && t < d) ? X[(uint64_t)src * d + t] : 0.0f;
            }
            if (threadIdx.x < 64) sv[threadIdx.x] = perm[base + threadIdx.x];
            __syncthreads();
            for (uint32_t c = 0; c < 64; c++) {
                float s2 = 0.0f;
#pragma unroll
                for (int t = 0; t < DMAX; t++) {
                    const float df = xr[t] - sp[c * DMAX + t];
                    s2 = fmaf(df, df, s2);
                }
                const bool cand = real && sv[c] != KNONE && !(c0 < 256 && c0 + c == threadIdx.x) && s2 < kcur;
                if (__any(cand)) {
                    float v = cand ? s2 : INFINITY;
#pragma unroll
                    for (int q = 0; q < KK; q++) {
                        const bool sw = v < best[q];
                        const float tmp = best[q];
                        best[q] = sw ? v : tmp;
                        v = sw ? tmp : v;
                    }
#pragma unroll
                    for (int q = 0; q < KK; q++)
                        if ((uint32_t)q == k - 1) kcur = best[q];
                }
            }
        }
        // the row's own bound on its k-th nearest squared distance (+inf: none)
        seed_d2[(uint64_t)R * 256 + threadIdx.x] = (real && kcur < INFINITY) ? (double)kcur * (1.0 + 1e-4) : INFINITY;
        const double m = block_max_f64(real ? (double)kcur : 0.0, red);
        if (threadIdx.x == 0) {
            if (cnt < 1.0) uR[R] = -1.0;
            else if (!(m < INFINITY)) uR[R] = INFINITY;  // some row has k other points or fewer in the window
            else uR[R] = sqrt(m * (1.0 + 1e-4)) + 1e-30;
        }
    }
}

// Grouped layout: group g (sorted positions start[g] .. start[g + 1] of perm) occupies layout
// positions poff[g] .. poff[g] + size, padded to whole row blocks; other positions are KNONE.
// The group of sorted position p is sorted_keys[p] >> shift.
__global__ void k_layout_fill(const uint32_t *sorted_keys, uint32_t shift, const uint32_t *perm, const uint32_t *start,
                              const uint32_t *poff, uint32_t n, uint32_t *layout) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const uint32_t g = sorted_keys[p] >> shift;
        layout[poff[g] + (p - start[g])] = perm[p];
    }
}

// Distance of every point (sorted position p: point perm[p] of group keys[p]) to its group's
// centre, f32, and their maximum (bits of a non-negative float order as unsigned integers).
__global__ void k_group_dist(const float *X, uint32_t d, const float *cen, const uint32_t *keys, const uint32_t *perm,
                             uint32_t n, float *dist, unsigned int *dmax) {
    float m = 0.0f;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const uint32_t y = perm[p], g = keys[p];
        float q = 0.0f;
        for (uint32_t t = 0; t < d; t++) {
            const float df = X[(uint64_t)y * d + t] - cen[(uint64_t)g * d + t];
            q = fmaf(df, df, q);
        }
        const float r = sqrtf(q);
        dist[p] = r;
        m = fmaxf(m, r);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(dmax, __float_as_uint(m));
}

// keys[p] = group << qbits | the distance quantised to qbits bits (of the largest): sorted, the
// points of a group run from its centre outwards.
__global__ void k_dist_keys(uint32_t *keys, const float *dist, const unsigned int *dmax, uint32_t n, uint32_t qbits) {
    const float m = __uint_as_float(*dmax);
    const float top = (float)((1u << qbits) - 1u);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const float f = m > 0.0f ? fminf(dist[p] / m * top, top) : 0.0f;
        keys[p] = (keys[p] << qbits) | (uint32_t)f;
    }
}

__global__ void k_padded_sizes(const uint32_t *start, uint32_t C, uint32_t *psz) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g <= C; g += gridDim.x * blockDim.x)
        psz[g] = g < C ? (start[g + 1] - start[g] + 255u) / 256u * 256u : 0u;
}

static bool knn_prune_enabled(uint32_t n) {
    const char *e = std::getenv("SCGPU_KNN_PRUNE");
    if (e && e[0] == '0') return false;
    if (e && e[0] == '1') return n >= 2 * (uint32_t)KNN_BM;
    return n >= 16384;
}

// Layout of the pruned search: points grouped by k-means (C centroids, Lloyd iterations on a
// sample, then one assignment of every point), within a group in increasing distance to its
// centre (quantised; ties in the points' own order), every group starting on a row block; returns the layout's length (a multiple of 256), layout[j] = point or KNONE, and the
// groups: gC of them, centres gcen (gC x d), layout offsets gpoff (gC + 1, whole row blocks),
// sizes gstart[g + 1] - gstart[g].
static uint32_t knn_group_order(const float *d_X, uint32_t n, uint32_t d, DevBuf<uint32_t> &layout,
                                DevBuf<float> &gcen, DevBuf<uint32_t> &gpoff, DevBuf<uint32_t> &gstart, uint32_t &gC,
                                hipStream_t s) {
    const uint32_t C = std::min<uint32_t>(n, std::max<uint32_t>(16u, std::min<uint32_t>(2048u, n / 512u)));
    constexpr int ITERS = 2;
    // Lloyd iterations on a sample (every stride-th point, about 256 per centre), then one
    // assignment of all points to the centres
    const uint32_t stride = std::max<uint32_t>(1u, n / std::min<uint32_t>(n, C * 256u));
    const uint32_t ns = (n + stride - 1) / stride;
    DevBuf<float> Xs;
    const float *Xp = d_X;
    if (stride > 1) {
        Xs.resize((size_t)ns * d);
        hipLaunchKernelGGL(k_gather_rows, dim3(grid_for((uint64_t)ns * d, 256, 8192)), dim3(256), 0, s, d_X, d, stride,
                           ns, Xs.get());
        KERNEL_OK();
        Xp = Xs.get();
    }
    DevBuf<float> cen((size_t)C * d), cnorm(C);
    DevBuf<uint32_t> start(C + 1);
    hipLaunchKernelGGL(k_kmeans_init, dim3(grid_for((uint64_t)C * d, 256, 4096)), dim3(256), 0, s, Xp, ns, d, C,
                       cen.get());
    KERNEL_OK();
    const int kb = std::max(1, bits_for(C));
    // distance bits of the final sort key: at most 20, so the quantiser's top value (2^qbits - 1)
    // and every value below it are exact in f32 and the quantised distance never reaches the
    // group bits
    const uint32_t qbits = std::min<uint32_t>(20u, 32u - (uint32_t)kb);
    // nearest centre of every point of P (np points) -> keys, then the points sorted by centre
    // (perm, stable) and each centre's first position (start)
    auto assign_sort = [&](const float *P, uint32_t np, DevBuf<uint32_t> &keys, DevBuf<uint32_t> &perm) {
        hipLaunchKernelGGL(k_kmeans_norms, dim3(grid_for(C, 256, 1024)), dim3(256), 0, s, cen.get(), d, C, cnorm.get());
        KERNEL_OK();
        const dim3 ag(grid_for(np, 256, 8192));
        if (d + 2 <= 32) {
            const uint32_t Cpad = (C + 63) / 64 * 64;
            DevBuf<unsigned short> cbh((size_t)Cpad * 32), cbl((size_t)Cpad * 32);
            hipLaunchKernelGGL(k_kmeans_pack, dim3(grid_for(Cpad, 256, 1024)), dim3(256), 0, s, cen.get(), d, C, Cpad,
                               cbh.get(), cbl.get());
            KERNEL_OK();
            hipLaunchKernelGGL(k_kmeans_assign_mfma, dim3(grid_for(np, 256, 4096)), dim3(256), 0, s, P, np, d,
                               cbh.get(), cbl.get(), Cpad, keys.get());
        } else if (d <= 32)
            hipLaunchKernelGGL(k_kmeans_assign_r<32>, ag, dim3(256), 0, s, P, np, d, cen.get(), cnorm.get(), C, keys.get());
        else if (d <= 56)
            hipLaunchKernelGGL(k_kmeans_assign_r<56>, ag, dim3(256), 0, s, P, np, d, cen.get(), cnorm.get(), C, keys.get());
        else if (d <= 64)
            hipLaunchKernelGGL(k_kmeans_assign_r<64>, ag, dim3(256), 0, s, P, np, d, cen.get(), cnorm.get(), C, keys.get());
        else
            hipLaunchKernelGGL(k_kmeans_assign, ag, dim3(256), 0, s, P, np, d, cen.get(), cnorm.get(), C, keys.get());
        KERNEL_OK();
        hipLaunchKernelGGL(k_knn_iota, dim3(grid_for(np, 256, 4096)), dim3(256), 0, s, perm.get(), np);
        KERNEL_OK();
        radix_sort_pairs(keys, perm, np, kb, s);
        hipLaunchKernelGGL(k_key_starts_u32, dim3(grid_for(C + 1, 256, 1024)), dim3(256), 0, s, keys.get(), np, C,
                           start.get());
        KERNEL_OK();
    };
    {
        DevBuf<uint32_t> skeys(ns), sperm(ns);
        for (int it = 0; it < ITERS; it++) {
            assign_sort(Xp, ns, skeys, sperm);
            hipLaunchKernelGGL(k_kmeans_update, dim3(std::min<uint32_t>(C, 8192)), dim3(256), 0, s, Xp, d, sperm.get(),
                               start.get(), C, cen.get());
            KERNEL_OK();
        }
    }
    DevBuf<uint32_t> keys(n), perm(n);
    assign_sort(d_X, n, keys, perm);
    {
        // within each group, points in increasing distance to the centre: a group's outliers
        // share its last column blocks, so the spheres of its other blocks stay tight
        DevBuf<float> gdist(n);
        DevBuf<unsigned int> dmax(1);
        HIP_OK(hipMemsetAsync(dmax.get(), 0, sizeof(unsigned int), s));
        hipLaunchKernelGGL(k_group_dist, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, d_X, d, cen.get(), keys.get(),
                           perm.get(), n, gdist.get(), dmax.get());
        KERNEL_OK();
        hipLaunchKernelGGL(k_dist_keys, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, keys.get(), gdist.get(),
                           dmax.get(), n, qbits);
        KERNEL_OK();
        radix_sort_pairs(keys, perm, n, kb + (int)qbits, s);
    }
    DevBuf<uint32_t> psz(C + 1), poff(C + 1);
    hipLaunchKernelGGL(k_padded_sizes, dim3(grid_for(C + 1, 256, 1024)), dim3(256), 0, s, start.get(), C, psz.get());
    KERNEL_OK();
    exclusive_scan_u32(psz.get(), poff.get(), C + 1, s);
    uint32_t rows = 0;
    to_host(&rows, poff.get() + C, 1, s);
    sync_stream(s);
    layout.resize(rows);
    hipLaunchKernelGGL(k_fill_u32_knn, dim3(grid_for(rows, 256, 4096)), dim3(256), 0, s, layout.get(), rows, KNONE);
    KERNEL_OK();
    hipLaunchKernelGGL(k_layout_fill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, keys.get(), qbits, perm.get(),
                       start.get(), poff.get(), n, layout.get());
    KERNEL_OK();
    // the centres of the final assignment and the groups' layout offsets (the group-level pruning)
    gcen = std::move(cen);
    gpoff = std::move(poff);
    gstart = std::move(start);
    gC = C;
    return rows;
}

template <bool BF16, int NKC, int KM>
static void launch_tile_km(dim3 grid, const unsigned short *Xh, const unsigned short *Xl, const float *Xf,
                           uint32_t d, uint32_t n, uint32_t npad, uint32_t cps, const uint32_t *blist,
                           const uint32_t *boff, const uint32_t *bcnt, const uint32_t *rbo, const float *seed_s,
                           const float *margin_s, uint32_t kk, uint32_t *gthr, uint32_t *cidx, float *cs,
                           hipStream_t s) {
    constexpr size_t lds = knn_tile_lds_bytes<BF16, NKC>();
    if (lds > 64 * 1024) throw ScError(SC_INVALID_ARGUMENT, "kNN tile does not fit in LDS for this d");
    const uint32_t row_blocks = grid.x, splits = grid.y;
    const dim3 g1(8 * ((row_blocks * splits + 7) / 8));  // one dimension; the kernel maps it XCD-aware
    hipLaunchKernelGGL((k_knn_tile<BF16, NKC, KM>), g1, dim3(256), lds, s, Xh, Xl, Xf, d, n, npad, cps, row_blocks, splits, blist,
                       boff, bcnt, rbo, seed_s, margin_s, kk, gthr, cidx, cs);
    KERNEL_OK();
}

template <bool BF16, int NKC>
static void launch_tile(dim3 grid, const unsigned short *Xh, const unsigned short *Xl, const float *Xf,
                        uint32_t d, uint32_t n, uint32_t npad, uint32_t KM, uint32_t cps, const uint32_t *blist,
                        const uint32_t *boff, const uint32_t *bcnt, const uint32_t *rbo, const float *seed_s,
                        const float *margin_s, uint32_t kk, uint32_t *gthr, uint32_t *cidx, float *cs, hipStream_t s) {
    switch (KM) {
        case 16: launch_tile_km<BF16, NKC, 16>(grid, Xh, Xl, Xf, d, n, npad, cps, blist, boff, bcnt, rbo, seed_s, margin_s, kk, gthr, cidx, cs, s); break;
        case 32: launch_tile_km<BF16, NKC, 32>(grid, Xh, Xl, Xf, d, n, npad, cps, blist, boff, bcnt, rbo, seed_s, margin_s, kk, gthr, cidx, cs, s); break;
        default: launch_tile_km<BF16, NKC, 64>(grid, Xh, Xl, Xf, d, n, npad, cps, blist, boff, bcnt, rbo, seed_s, margin_s, kk, gthr, cidx, cs, s); break;
    }
}

void knn_device(const float *d_X, uint32_t n, uint32_t d, uint32_t k, int precision, uint32_t *d_idx,
                double *d_d2, sc_knn_stats *st, hipStream_t s) {
    if (k == 0 || k > FB_MAXK) throw ScError(SC_INVALID_ARGUMENT, "k must be in [1, 32]");
    if (d == 0 || d > 126) throw ScError(SC_INVALID_ARGUMENT, "d must be in [1, 126]");
    if (n < 2) throw ScError(SC_INVALID_ARGUMENT, "need at least 2 points");
    const bool bf16 = precision == 0;
    // coordinates, then the two slots of the folded squared norm (k_knn_prep), padded to 32
    const uint32_t nkc = (d + 2 + 31) / 32;
    const uint32_t Kp = 32 * nkc;
    const uint32_t npad = (n + KNN_BM - 1) / KNN_BM * KNN_BM;
    // candidates kept per row and split: the smallest of 16, 32, 64 above k (the certificate
    // needs one candidate beyond the k-th)
    uint32_t K2 = 16;
    while (K2 < k + 1 && K2 < (uint32_t)KNN_MAXK2) K2 *= 2;
    hipEvent_t e0, e1, e2, e3;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventCreate(&e2));
    HIP_OK(hipEventCreate(&e3));
    HIP_OK(hipEventRecord(e0, s));
    // pruned search: the grouped layout (rows: its length, whole row blocks per group)
    bool prune = knn_prune_enabled(n);
    DevBuf<uint32_t> layout;
    DevBuf<float> gcen;
    DevBuf<uint32_t> gpoff, gstart;
    uint32_t gC = 0;
    uint32_t rows = prune ? knn_group_order(d_X, n, d, layout, gcen, gpoff, gstart, gC, s) : npad;
    // the lists of column blocks are addressed with 32-bit offsets: with more (row block, column
    // block) pairs than that the search runs unpruned
    if (prune && (uint64_t)(rows / KNN_BM) * (rows / (nkc <= 2 ? 64u : 32u)) >= 0xFFFFFFFFull) {
        prune = false;
        rows = npad;
    }
    // column splits: the count whose last wave of workgroups is fullest (time ~ waves / splits),
    // with a small charge per split for the extra candidates it sends to the rescore
    const uint32_t row_blocks = rows / KNN_BM;
    int dev = 0, cus = 256;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t slots = (uint32_t)cus * ((K2 == 16 && nkc <= 2) ? 2u : 1u);  // workgroups resident at once
    // at least two splits where the list allows it: a single list holds the k + 1 nearest
    // points only, and the certificate then needs the (k + 1)-th to stand clear of the k-th by the
    // error bound, which fails for most rows of dense clusters (every such row costs an FP64 scan)
    const uint32_t min_splits = 2 * K2 <= (uint32_t)KNN_MAXCAND ? 2u : 1u;
    uint32_t splits = min_splits;
    double best = 1e30;
    for (uint32_t sp = min_splits; sp <= 8 && sp * K2 <= (uint32_t)KNN_MAXCAND; sp++) {
        const double waves = (double)((row_blocks * sp + slots - 1) / slots);
        const double cost = waves / sp + 0.02 * sp;
        if (cost < best) {
            best = cost;
            splits = sp;
        }
    }
    // the online tuner measures the alternatives when the grid is small enough for the last
    // wave of workgroups to matter (the result is exact for every split count)
    std::vector<uint32_t> cand{splits};
    if (prune) {
        // the pruned lists vary in length between row blocks: more splits balance the work, fewer
        // send fewer first hits and candidates (the splits share their thresholds); with two the
        // certificate can run short of margin (rows fall back to the FP64 scan), which the tuner
        // sees, since it times the whole search
        cand.clear();
        for (uint32_t sp : {4u, 2u, 3u})
            if (sp * K2 <= (uint32_t)KNN_MAXCAND) cand.push_back(sp);
        if (cand.empty()) cand.push_back(min_splits);
    } else if (row_blocks < 4 * slots) {
        for (uint32_t sp = std::max(min_splits, splits > 2 ? splits - 2 : 1u); sp <= splits + 2 && sp <= 8 && sp * K2 <= (uint32_t)KNN_MAXCAND; sp++)
            if (sp != splits) cand.push_back(sp);
    }
    const std::string tkey = "knn_tile/rb2^" + std::to_string(bits_for(row_blocks)) + "/kc" + std::to_string(nkc) +
                             "/km" + std::to_string(K2) + (bf16 ? "/bf16" : "/fp32") + (prune ? "/pruned" : "");
    // each candidate is measured twice (the minimum counts): a process's first build also pays
    // for loading the kernels, which would bias the comparison against the first candidate
    const TuneChoice tc = tune_enabled() && cand.size() > 1 ? tune_pick(tkey, (int)cand.size(), 2) : TuneChoice{0, false};
    splits = cand[tc.index];
    if (const char *e = std::getenv("SCGPU_KNN_SPLITS")) {  // fixed split count (measurements)
        const uint32_t f = (uint32_t)std::atoi(e);
        if (f >= 1 && f <= 8 && f * K2 <= (uint32_t)KNN_MAXCAND) splits = f;
    }
    uint32_t cps = (rows + splits - 1) / splits;
    cps = (cps + KNN_TN - 1) / KNN_TN * KNN_TN;
    if (!prune) splits = (rows + cps - 1) / cps;
    DevBuf<unsigned short> Xh(bf16 ? (size_t)rows * Kp : 0), Xl(bf16 ? (size_t)rows * Kp : 0);
    DevBuf<float> Xf(bf16 ? 0 : (size_t)rows * Kp), ny(rows);
    DevBuf<double> nx64(npad);
    DevBuf<unsigned int> ymax(1);
    HIP_OK(hipMemsetAsync(ymax.get(), 0, sizeof(unsigned int), s));
    // pruned search: block spheres, in-block k-th distance bounds, each row block's list of
    // column blocks
    const uint32_t TN = nkc <= 2 ? 64u : 32u;  // KnnTileShape<NKC>::TN
    DevBuf<uint32_t> blist, boff, bcnt, rb_order;
    DevBuf<double> seed_d2;
    uint64_t list_total = 0;
    if (prune) {
        seed_d2.resize(rows);
        const uint32_t ncb = rows / TN;
        DevBuf<double> cR((size_t)row_blocks * d), rR(row_blocks), uR(row_blocks);
        DevBuf<uint32_t> nR(row_blocks), cnt(row_blocks + 1);
        hipLaunchKernelGGL(k_block_spheres, dim3(std::min<uint32_t>((row_blocks + 3) / 4, 16384)), dim3(256), 0, s, d_X,
                           layout.get(), rows, d, (uint32_t)KNN_BM, row_blocks, cR.get(), rR.get(), nR.get());
        KERNEL_OK();
        // the coordinates held per row (DMAX >= d) and the insertion network (KK >= k) are sized
        // to the search; the first k entries of the network are the same for any KK >= k
        const dim3 kg(std::min<uint32_t>(row_blocks, 16384));
        auto kth = [&](auto kern) {
            hipLaunchKernelGGL(kern, kg, dim3(256), 0, s, d_X, layout.get(), d, k, row_blocks, uR.get(), seed_d2.get());
        };
        const bool k16 = k <= 16;
        if (d <= 32 && k <= 8) kth(k_block_kth<32, 8>);
        else if (d <= 32) kth(k16 ? k_block_kth<32, 16> : k_block_kth<32, 32>);
        else if (d <= 56) kth(k16 ? k_block_kth<56, 16> : k_block_kth<56, 32>);
        else if (d <= 64) kth(k16 ? k_block_kth<64, 16> : k_block_kth<64, 32>);
        else kth(k16 ? k_block_kth<128, 16> : k_block_kth<128, 32>);
        KERNEL_OK();
        hipLaunchKernelGGL(k_fill_u32_knn, dim3(1), dim3(64), 0, s, cnt.get() + row_blocks, 1u, 0u);
        KERNEL_OK();
        // groups each row block needs (per-row bounds against the groups' spheres), then its list
        // of column blocks within those groups
        DevBuf<uint32_t> rbg(row_blocks);
        hipLaunchKernelGGL(k_rb_groups, dim3(grid_for(row_blocks, 256, 4096)), dim3(256), 0, s, gpoff.get(), gC,
                           row_blocks, rbg.get());
        KERNEL_OK();
        DevBuf<uint32_t> cend(gC);
        hipLaunchKernelGGL(k_group_core, dim3(grid_for(gC, 256, 64)), dim3(256), 0, s, gstart.get(), gpoff.get(), gC,
                           cend.get());
        KERNEL_OK();
        DevBuf<unsigned long long> r2bits(gC), r2core(gC);
        HIP_OK(hipMemsetAsync(r2bits.get(), 0, sizeof(unsigned long long) * gC, s));
        HIP_OK(hipMemsetAsync(r2core.get(), 0, sizeof(unsigned long long) * gC, s));
        hipLaunchKernelGGL(k_group_radii, dim3(std::min<uint32_t>(row_blocks, 16384)), dim3(256), 0, s, d_X,
                           layout.get(), rows, d, gcen.get(), rbg.get(), cend.get(), r2bits.get(), r2core.get());
        KERNEL_OK();
        if (gC == 0 || gC > KNN_MAX_GROUPS) throw ScError(SC_INVALID_ARGUMENT, "kNN grouping: group count out of range");
        const uint32_t mwords = (gC + 31) / 32;
        DevBuf<float> cenT((size_t)gC * d), gcn(gC), grf(gC), grc(gC);
        DevBuf<double> gr64(gC);
        hipLaunchKernelGGL(k_group_prep, dim3(grid_for(gC, 256, 64)), dim3(256), 0, s, gcen.get(), gC, d, r2bits.get(),
                           r2core.get(), cenT.get(), gcn.get(), gr64.get(), grf.get(), grc.get());
        KERNEL_OK();
        DevBuf<uint32_t> gmask((size_t)row_blocks * 2 * mwords);
        auto rowg = [&](auto kern) {
            hipLaunchKernelGGL(kern, kg, dim3(256), 0, s, d_X, layout.get(), d, row_blocks, gC, gcen.get(), cenT.get(),
                               gcn.get(), gr64.get(), grf.get(), grc.get(), cR.get(), rR.get(), uR.get(), seed_d2.get(),
                               mwords, gmask.get());
        };
        if (d <= 32) rowg(k_row_groups<32>);
        else if (d <= 64) rowg(k_row_groups<64>);
        else rowg(k_row_groups<128>);
        KERNEL_OK();
        const dim3 ngrid(std::min<uint32_t>((row_blocks + 3) / 4, 16384));
        hipLaunchKernelGGL(k_group_lists, ngrid, dim3(256), 0, s, uR.get(), layout.get(), row_blocks, gpoff.get(),
                           cend.get(), rbg.get(), mwords, gmask.get(), TN, cnt.get(),
                           (const uint32_t *)nullptr, (uint32_t *)nullptr);
        KERNEL_OK();
        boff.resize(row_blocks + 1);
        exclusive_scan_u32(cnt.get(), boff.get(), row_blocks + 1, s);
        uint32_t lt = 0;
        to_host(&lt, boff.get() + row_blocks, 1, s);
        sync_stream(s);
        blist.resize(std::max<uint64_t>(lt, 1));
        hipLaunchKernelGGL(k_group_lists, ngrid, dim3(256), 0, s, uR.get(), layout.get(), row_blocks, gpoff.get(),
                           cend.get(), rbg.get(), mwords, gmask.get(), TN, (uint32_t *)nullptr,
                           (const uint32_t *)boff.get(), blist.get());
        KERNEL_OK();
        // per-row filter of the lists against the column blocks' spheres (in place; bcnt: kept)
        DevBuf<double> cB((size_t)ncb * d), rB(ncb);
        DevBuf<uint32_t> nB(ncb);
        hipLaunchKernelGGL(k_block_spheres, dim3(std::min<uint32_t>((ncb + 3) / 4, 16384)), dim3(256), 0, s, d_X, layout.get(),
                           rows, d, TN, ncb, cB.get(), rB.get(), nB.get());
        KERNEL_OK();
        DevBuf<float> cB32((size_t)ncb * d), rBf(ncb), nbn(ncb), nb2(ncb);
        hipLaunchKernelGGL(k_block_prep, dim3(grid_for(ncb, 256, 4096)), dim3(256), 0, s, cB.get(), rB.get(), ncb, d,
                           cB32.get(), rBf.get(), nbn.get(), nb2.get());
        KERNEL_OK();
        bcnt.resize(row_blocks + 1);
        hipLaunchKernelGGL(k_fill_u32_knn, dim3(1), dim3(64), 0, s, bcnt.get() + row_blocks, 1u, 0u);
        KERNEL_OK();
        auto lfilt = [&](auto kern) {
            hipLaunchKernelGGL(kern, kg, dim3(256), 0, s, d_X, layout.get(), d, row_blocks, seed_d2.get(), cB32.get(),
                               rBf.get(), nbn.get(), nb2.get(), boff.get(), cnt.get(), blist.get(), bcnt.get());
        };
        if (d + 2 <= 32) {
            // the matrix-core filter: its error coefficients follow the tile's bound (see there)
            const double u24 = std::ldexp(1.0, -24), u16 = std::ldexp(1.0, -16);
            const double acc_c = 3.03 * (3.0 * 32 + 2.0) * u24, split_c = 3.01 * u16;
            const float e1 = (float)(2.0 * 2.0 * (split_c + acc_c) * 1.001);
            const float e2 = (float)(2.0 * (1.01 * d * u24 + u16 + acc_c) * 1.001);
            DevBuf<unsigned short> cbh((size_t)ncb * 32), cbl((size_t)ncb * 32);
            hipLaunchKernelGGL(k_block_pack, dim3(grid_for(ncb, 256, 4096)), dim3(256), 0, s, cB32.get(), nb2.get(),
                               ncb, d, cbh.get(), cbl.get());
            KERNEL_OK();
            hipLaunchKernelGGL(k_list_filter_mfma, kg, dim3(256), 0, s, d_X, layout.get(), d, row_blocks,
                               seed_d2.get(), cbh.get(), cbl.get(), rBf.get(), nbn.get(), boff.get(), cnt.get(),
                               blist.get(), bcnt.get(), e1, e2);
        } else if (d <= 32) lfilt(k_list_filter<32>);
        else if (d <= 64) lfilt(k_list_filter<64>);
        else lfilt(k_list_filter<128>);
        KERNEL_OK();
        // entries kept over all row blocks (the pairs the tile computes)
        exclusive_scan_u32(bcnt.get(), cnt.get(), row_blocks + 1, s);
        to_host(&lt, cnt.get() + row_blocks, 1, s);
        // row blocks by decreasing list length (ties in row-block order) for the tile's grid
        DevBuf<uint32_t> lkeys(row_blocks);
        rb_order.resize(row_blocks);
        hipLaunchKernelGGL(k_lpt_keys, dim3(grid_for(row_blocks, 256, 4096)), dim3(256), 0, s, bcnt.get(), row_blocks,
                           lkeys.get(), rb_order.get());
        KERNEL_OK();
        radix_sort_pairs(lkeys, rb_order, row_blocks, 32, s);
        sync_stream(s);
        list_total = lt;
    }
    const uint32_t *pm = prune ? layout.get() : nullptr;
    const uint32_t *bl = prune ? blist.get() : nullptr;
    const uint32_t *bo = prune ? boff.get() : nullptr;
    const uint32_t *bc = prune ? bcnt.get() : nullptr;
    const uint32_t *ro = prune ? rb_order.get() : nullptr;
    const uint64_t total = (uint64_t)rows * Kp;
    {
        // waves per workgroup: as many 64-row LDS tiles as fit in 64 KB (4 up to d = 63)
        const uint32_t nw = std::max<uint32_t>(1u, std::min<uint32_t>(4u, 65536u / (256u * (d + 1))));
        const size_t lds = (size_t)nw * 64 * (d + 1) * sizeof(float);
        hipLaunchKernelGGL(k_knn_norms, dim3(grid_for(rows, 64 * nw, 8192)), dim3(64 * nw), lds, s, d_X, pm, n, d,
                           rows, ny.get(), nx64.get(), ymax.get());
        KERNEL_OK();
    }
    hipLaunchKernelGGL(k_knn_prep, dim3(grid_for(total, 256, 16384)), dim3(256), 0, s, d_X, pm, (const float *)ny.get(),
                       n, d, Kp, total, bf16 ? Xh.get() : nullptr, bf16 ? Xl.get() : nullptr, bf16 ? nullptr : Xf.get());
    KERNEL_OK();
    // the error bound of the approximate values (needed by the tile's seeded thresholds and the
    // rescore): |s_approx - s| <= norm error + 2 |dot error|, doubled for safety; see the header
    unsigned int ymax_bits = 0;
    to_host(&ymax_bits, ymax.get(), 1, s);
    sync_stream(s);
    float ymaxf;
    std::memcpy(&ymaxf, &ymax_bits, 4);
    double Y = (double)ymaxf;
    // The tile's value is one FP32-accumulated sum over the Kp slots of the products (-2 x_t)(y_t)
    // (bf16: the three split products) and 1 * (|y|^2 hi) + 1 * (|y|^2 lo). Its error against the
    // exact s = |y|^2 - 2 x.y: the split truncation of the dot part (3 * 2^-16 |x||y|, doubled by
    // the factor 2), the f32 norm (d * 2^-24 |y|^2), the bf16 hi + lo split of the norm (2^-16
    // |y|^2) and the accumulation (3 Kp + 2 terms of magnitude at most 2 |x||y| or |y|^2, taken
    // at three times 2^-24 per term, whatever order and rounding the matrix core's accumulation
    // uses), all with |y| <= Y; the whole bound is doubled for safety.
    const double u24 = std::ldexp(1.0, -24), u16 = std::ldexp(1.0, -16);
    const double terms = bf16 ? 3.0 * Kp + 2.0 : (double)Kp + 1.0;
    const double acc_coef = 3.03 * terms * u24;  // 3x the round-to-nearest bound per addition
    const double split_coef = bf16 ? 3.01 * u16 : 0.0;
    double eps_coef = 2.0 * (2.0 * (split_coef + acc_coef) * Y);
    double eps_const = 2.0 * ((1.01 * d * u24 + (bf16 ? u16 : 0.0) + acc_coef) * Y * Y);
    DevBuf<float> seed_s;
    if (prune) {
        seed_s.resize(rows);
        hipLaunchKernelGGL(k_seed_thresholds, dim3(grid_for(rows, 256, 4096)), dim3(256), 0, s, seed_d2.get(),
                           ny.get(), rows, eps_coef, eps_const, seed_s.get());
        KERNEL_OK();
    }
    const float *sd = prune ? seed_s.get() : nullptr;
    DevBuf<float> margin_s(rows);
    hipLaunchKernelGGL(k_knn_margins, dim3(grid_for(rows, 256, 4096)), dim3(256), 0, s, ny.get(), rows, eps_coef,
                       eps_const, margin_s.get());
    KERNEL_OK();
    const float *mg = margin_s.get();
    // the rows' thresholds shared by their splits (starting at the seeds)
    DevBuf<uint32_t> gthr(rows);
    hipLaunchKernelGGL(k_thr_init, dim3(grid_for(rows, 256, 4096)), dim3(256), 0, s, sd, rows, gthr.get());
    KERNEL_OK();
    uint32_t *const gt = splits > 1 ? gthr.get() : nullptr;
    // candidate lists per layout row (the pruned layout has padding rows)
    const uint32_t crow = prune ? rows : n;
    DevBuf<uint32_t> cidx((size_t)crow * splits * K2);
    DevBuf<float> cs((size_t)crow * splits * K2);
    HIP_OK(hipEventRecord(e1, s));
    dim3 grid(row_blocks, splits);
    const unsigned short *xh = Xh.get(), *xl = Xl.get();
    const float *xf = Xf.get();
    switch (nkc * 2 + (bf16 ? 0 : 1)) {
        case 2: launch_tile<true, 1>(grid, xh, xl, xf, d, crow, rows, K2, cps, bl, bo, bc, ro, sd, mg, k, gt, cidx.get(), cs.get(), s); break;
        case 3: launch_tile<false, 1>(grid, xh, xl, xf, d, crow, rows, K2, cps, bl, bo, bc, ro, sd, mg, k, gt, cidx.get(), cs.get(), s); break;
        case 4: launch_tile<true, 2>(grid, xh, xl, xf, d, crow, rows, K2, cps, bl, bo, bc, ro, sd, mg, k, gt, cidx.get(), cs.get(), s); break;
        case 5: launch_tile<false, 2>(grid, xh, xl, xf, d, crow, rows, K2, cps, bl, bo, bc, ro, sd, mg, k, gt, cidx.get(), cs.get(), s); break;
        case 6: launch_tile<true, 3>(grid, xh, xl, xf, d, crow, rows, K2, cps, bl, bo, bc, ro, sd, mg, k, gt, cidx.get(), cs.get(), s); break;
        case 7: launch_tile<false, 3>(grid, xh, xl, xf, d, crow, rows, K2, cps, bl, bo, bc, ro, sd, mg, k, gt, cidx.get(), cs.get(), s); break;
        case 8: launch_tile<true, 4>(grid, xh, xl, xf, d, crow, rows, K2, cps, bl, bo, bc, ro, sd, mg, k, gt, cidx.get(), cs.get(), s); break;
        default: launch_tile<false, 4>(grid, xh, xl, xf, d, crow, rows, K2, cps, bl, bo, bc, ro, sd, mg, k, gt, cidx.get(), cs.get(), s); break;
    }
    HIP_OK(hipEventRecord(e2, s));
    if (prune) {
        // candidates back to the points' own rows and indices
        DevBuf<uint32_t> cidx2((size_t)n * splits * K2);
        DevBuf<float> cs2((size_t)n * splits * K2);
        hipLaunchKernelGGL(k_knn_remap, dim3(grid_for((uint64_t)rows * splits * K2, 256, 16384)), dim3(256), 0, s, pm,
                           rows, splits * K2, cidx.get(), cs.get(), cidx2.get(), cs2.get());
        KERNEL_OK();
        std::swap(cidx, cidx2);
        std::swap(cs, cs2);
    }
    DevBuf<uint32_t> fb(n);
    DevBuf<unsigned int> nfb(1);
    DevBuf<unsigned long long> nres(1);
    HIP_OK(hipMemsetAsync(nfb.get(), 0, sizeof(unsigned int), s));
    HIP_OK(hipMemsetAsync(nres.get(), 0, sizeof(unsigned long long), s));
    // SCGPU_KNN_FALLBACK_ALL=1 sends every row through the exact FP64 scan as well (a test of
    // that path: its rows then replace the certified ones)
    const char *fa = std::getenv("SCGPU_KNN_FALLBACK_ALL");
    const int fallback_all = fa && fa[0] == '1' ? 1 : 0;
    // a few long-lived workgroups per CU walk the rows: launching one short workgroup per few rows
    // left most of the machine idle waiting for workgroup creation
    // two rows per wave when a row's candidates fit in half a wave
    if (K2 * splits <= 32)
        hipLaunchKernelGGL(k_knn_rescore<32>, dim3(std::min<uint32_t>((n + 2 * RS_WAVES - 1) / (2 * RS_WAVES), 2048)),
                           dim3(64 * RS_WAVES), 0, s, d_X, n, d, k, K2, splits, cidx.get(), cs.get(), nx64.get(),
                           eps_coef, eps_const, d_idx, d_d2, fallback_all, fb.get(), nfb.get(), nres.get());
    else
        hipLaunchKernelGGL(k_knn_rescore<64>, dim3(std::min<uint32_t>((n + RS_WAVES - 1) / RS_WAVES, 2048)),
                           dim3(64 * RS_WAVES), 0, s, d_X, n, d, k, K2, splits, cidx.get(), cs.get(), nx64.get(),
                           eps_coef, eps_const, d_idx, d_d2, fallback_all, fb.get(), nfb.get(), nres.get());
    KERNEL_OK();
    unsigned int nf = 0;
    unsigned long long nr = 0;
    to_host(&nf, nfb.get(), 1, s);
    to_host(&nr, nres.get(), 1, s);
    sync_stream(s);
    if (nf) {
        // about 2048 workgroups in all: up to 256 ranges per row, one range per row above 2048 rows
        const uint32_t W = std::max<uint32_t>(1u, std::min<uint32_t>(FB_MAXW, 2048u / nf));
        DevBuf<uint32_t> pidx((size_t)nf * W * k);
        DevBuf<double> pd2((size_t)nf * W * k);
        hipLaunchKernelGGL(k_knn_fallback, dim3(nf, W), dim3(FB_THREADS), 0, s, d_X, n, d, k, fb.get(), pidx.get(),
                           pd2.get());
        KERNEL_OK();
        hipLaunchKernelGGL(k_knn_fallback_merge, dim3(nf), dim3(64), 0, s, k, W, fb.get(), pidx.get(), pd2.get(),
                           d_idx, d_d2);
        KERNEL_OK();
    }
    HIP_OK(hipEventRecord(e3, s));
    sync_stream(s);
    float t_tile = 0.0f, t_all = 0.0f;
    HIP_OK(hipEventElapsedTime(&t_tile, e1, e2));
    HIP_OK(hipEventElapsedTime(&t_all, e0, e3));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    (void)hipEventDestroy(e3);
    // the whole search is timed: a split count that starves the certificate pays in fallbacks
    if (tune_enabled() && cand.size() > 1) tune_record(tkey, tc.index, t_all);
    if (st) {
        st->n = n;
        st->candidates_rescored = nr;
        st->rows_fallback = nf;
        st->tile_kernel_ms = t_tile;
        st->total_ms = t_all;
        // pairs the tile computed (padding included); the brute force computes npad^2
        st->pairs_computed = prune ? (double)list_total * TN * KNN_BM : (double)npad * (double)npad;
        st->flops = 2.0 * st->pairs_computed * (double)d;
    }
}

// ---- Gaussian connectivity (neighborhood/mod.rs:104-168) --------------------------------------
__global__ void k_sigma(const double *d2, uint32_t n, uint32_t k, int mode, double *sigma) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (mode == 0) {
            double dk = d2[(uint64_t)i * k + k - 1];
            sigma[i] = isfinite(dk) ? sqrt(dk) : 1.0;
        } else {
            // upper median of the nonzero squared distances (self at 0 is filtered, :108-121)
            double v[FB_MAXK];
            uint32_t c = 0;
            for (uint32_t t = 0; t < k; t++) {
                double x = d2[(uint64_t)i * k + t];
                if (x > 0.0 && isfinite(x)) v[c++] = x;
            }
            if (c == 0) {
                sigma[i] = 1.0;
                continue;
            }
            for (uint32_t a = 1; a < c; a++) {
                double x = v[a];
                int b = (int)a - 1;
                while (b >= 0 && v[b] > x) {
                    v[b + 1] = v[b];
                    b--;
                }
                v[b + 1] = x;
            }
            sigma[i] = sqrt(v[c / 2]);
        }
    }
}

__device__ __forceinline__ float gauss_w(double si, double sj, double d2) {
    double si2 = si * si, sj2 = sj * sj;
    double den = si2 + sj2;
    if (!(den > 0.0)) return 0.0f;
    double num = 2.0 * si * sj;
    return (float)(sqrt(num / den) * exp(-d2 / den));
}

// Emits (row << nb | col, w) entries, nb = bits_for(n) (so the key has 2 nb bits and the
// all-ones sentinel `sent` is above every real key). mode 0: both directions of the unordered
// pair, weight from the canonical (min, max) order; duplicates from the two endpoints' lists
// collapse later. mode 1: only pairs i <= j from row i's list with w > 1e-14, both directions
// (:131-168).
__global__ void k_gauss_emit(const uint32_t *idx, const double *d2, const double *sigma, uint32_t n, uint32_t k,
                             int mode, uint32_t nb, uint64_t sent, uint64_t *keys, float *vals) {
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < (uint64_t)n * k;
         q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t i = (uint32_t)(q / k);
        uint32_t j = idx[q];
        uint64_t o = 2 * q;
        keys[o] = sent;
        keys[o + 1] = sent;
        vals[o] = 0.0f;
        vals[o + 1] = 0.0f;
        if (j == KNONE) continue;
        if (mode == 0) {
            uint32_t a = i < j ? i : j, b = i < j ? j : i;
            float w = gauss_w(sigma[a], sigma[b], d2[q]);
            if (!(w > 0.0f)) continue;
            keys[o] = ((uint64_t)a << nb) | b;
            vals[o] = w;
            keys[o + 1] = ((uint64_t)b << nb) | a;
            vals[o + 1] = w;
        } else {
            if (!(i <= j)) continue;
            float w = gauss_w(sigma[i], sigma[j], d2[q]);
            if (!(w > 1e-14f)) continue;
            keys[o] = ((uint64_t)i << nb) | j;
            vals[o] = w;
            if (i != j) {
                keys[o + 1] = ((uint64_t)j << nb) | i;
                vals[o + 1] = w;
            }
        }
    }
}

__global__ void k_dedup_heads(const uint64_t *keys, uint64_t m, uint64_t sent, uint32_t *head) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x)
        head[e] = (keys[e] != sent && (e == 0 || keys[e] != keys[e - 1])) ? 1u : 0u;
}

// symmetrize by max: equal keys carry equal weights here, the max keeps that explicit
__global__ void k_dedup_emit(const uint64_t *keys, const float *vals, const uint32_t *head, const uint32_t *pos,
                             uint64_t m, uint32_t nb, uint32_t *nbrs, float *wts, unsigned long long *row_count) {
    const int lane = threadIdx.x & 63;
    for (uint64_t e0 = blockIdx.x * (uint64_t)blockDim.x; e0 < m; e0 += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t e = e0 + threadIdx.x;  // the loop bound is uniform: every lane reaches the ballots
        const bool in = e < m;
        const bool h = in && head[e] != 0u;
        const uint64_t key = in ? keys[e] : ~0ull;
        if (h) {
            float w = vals[e];
            for (uint64_t f = e + 1; f < m && keys[f] == key; f++) w = fmaxf(w, vals[f]);
            const uint32_t o = pos[e];
            nbrs[o] = (uint32_t)(key & ((1ull << nb) - 1));
            wts[o] = w;
        }
        // row counts: the keys are sorted, so a wave's entries of one row are consecutive lanes;
        // the first lane of each such run adds the run's heads with one atomic
        const uint64_t row = key >> nb;
        const uint64_t prow = __shfl_up(row, 1);
        const uint64_t starts = __ballot(lane == 0 || row != prow);
        const uint64_t heads = __ballot(h);
        if (in && ((starts >> lane) & 1ull)) {
            const uint64_t above = lane == 63 ? 0ull : (starts >> (lane + 1)) << (lane + 1);  // later run starts
            const int end = above ? __ffsll((unsigned long long)above) - 1 : 64;             // this run's end
            const uint64_t run = (end == 64 ? ~0ull : ((1ull << end) - 1ull)) & ~((1ull << lane) - 1ull);
            const uint32_t c = (uint32_t)__popcll(heads & run);
            if (c) atomicAdd(&row_count[row], (unsigned long long)c);
        }
    }
}

std::unique_ptr<sc_graph> gaussian_graph_device(const uint32_t *d_idx, const double *d_d2, uint32_t n, uint32_t k,
                                                int mode, hipStream_t s) {
    DevBuf<double> sigma(n);
    hipLaunchKernelGGL(k_sigma, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, d_d2, n, k, mode, sigma.get());
    KERNEL_OK();
    uint64_t m = 2ull * n * k;
    DevBuf<uint64_t> keys(m);
    DevBuf<float> vals(m);
    // (row, col) keys of 2 nb bits: the sort needs ceil(2 nb / 8) digit passes, not 8
    const uint32_t nb = (uint32_t)std::max(1, bits_for(n));
    const uint64_t sent = 2 * nb >= 64 ? ~0ull : (1ull << (2 * nb)) - 1;
    hipLaunchKernelGGL(k_gauss_emit, dim3(grid_for((uint64_t)n * k, 256, 8192)), dim3(256), 0, s, d_idx, d_d2,
                       sigma.get(), n, k, mode, nb, sent, keys.get(), vals.get());
    KERNEL_OK();
    radix_sort_pairs(keys, vals, m, (int)(2 * nb), s);
    DevBuf<uint32_t> head(m), pos(m);
    hipLaunchKernelGGL(k_dedup_heads, dim3(grid_for(m, 256, 8192)), dim3(256), 0, s, keys.get(), m, sent, head.get());
    KERNEL_OK();
    exclusive_scan_u32(head.get(), pos.get(), m, s);
    uint32_t lp = 0, lh = 0;
    to_host(&lp, pos.get() + m - 1, 1, s);
    to_host(&lh, head.get() + m - 1, 1, s);
    sync_stream(s);
    uint64_t nnz = (uint64_t)lp + lh;
    auto g = std::make_unique<sc_graph>();
    g->n = n;
    g->nnz = nnz;
    g->nbrs.resize(nnz);
    g->wts.resize(nnz);
    DevBuf<unsigned long long> rc(n + 1);
    HIP_OK(hipMemsetAsync(rc.get(), 0, (n + 1) * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_dedup_emit, dim3(grid_for(m, 256, 8192)), dim3(256), 0, s, keys.get(), vals.get(), head.get(),
                       pos.get(), m, nb, g->nbrs.get(), g->wts.get(), rc.get());
    KERNEL_OK();
    g->ptrs.resize(n + 1);
    exclusive_scan_u64(reinterpret_cast<uint64_t *>(rc.get()), g->ptrs.get(), n + 1, s);
    g->node_w.resize(n);
    hipLaunchKernelGGL(k_fill_f64_knn, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, g->node_w.get(), (uint64_t)n, 1.0);
    KERNEL_OK();
    graph_finish_sorted(*g, s);
    sync_stream(s);
    return g;
}

}  // namespace scgpu

using namespace scgpu;

static int knn_precision(const sc_knn_opts *o) { return o ? o->precision : 0; }
static int knn_mode(const sc_knn_opts *o) { return o ? o->mode : 0; }

extern "C" {

sc_status sc_knn(const float *data, uint64_t n, uint32_t d, uint32_t k, const sc_knn_opts *opts, uint32_t *idx_out,
                 double *d2_out, sc_knn_stats *stats, sc_error *err) {
    try {
        if (!data) throw ScError(SC_INVALID_ARGUMENT, "null data");
        check_node_count(n);
        hipStream_t s = thread_stream();
        DevBuf<float> X((size_t)n * d);
        to_device(X.get(), data, (size_t)n * d, s);
        DevBuf<uint32_t> idx((size_t)n * k);
        DevBuf<double> d2((size_t)n * k);
        knn_device(X.get(), (uint32_t)n, d, k, knn_precision(opts), idx.get(), d2.get(), stats, s);
        if (idx_out) to_host(idx_out, idx.get(), (size_t)n * k, s);
        if (d2_out) to_host(d2_out, d2.get(), (size_t)n * k, s);
        sync_stream(s);
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

sc_status sc_knn_graph(const float *data, uint64_t n, uint32_t d, uint32_t k, const sc_knn_opts *opts, sc_graph **out,
                       sc_knn_stats *stats, sc_error *err) {
    try {
        if (!data || !out) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        check_node_count(n);
        check_entry_count(2ull * n * k);  // the Gaussian assembly's entries, before any search
        *out = nullptr;
        hipStream_t s = thread_stream();
        DevBuf<float> X((size_t)n * d);
        to_device(X.get(), data, (size_t)n * d, s);
        DevBuf<uint32_t> idx((size_t)n * k);
        DevBuf<double> d2((size_t)n * k);
        knn_device(X.get(), (uint32_t)n, d, k, knn_precision(opts), idx.get(), d2.get(), stats, s);
        std::unique_ptr<sc_graph> g = gaussian_graph_device(idx.get(), d2.get(), (uint32_t)n, k, knn_mode(opts), s);
        *out = g.release();
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

sc_status sc_knn_graph_device(const float *d_data, uint64_t n, uint32_t d, uint32_t k, const sc_knn_opts *opts,
                              sc_graph **out, sc_knn_stats *stats, sc_error *err) {
    try {
        if (!d_data || !out) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        check_node_count(n);
        check_entry_count(2ull * n * k);  // the Gaussian assembly's entries, before any search
        *out = nullptr;
        hipStream_t s = thread_stream();
        DevBuf<uint32_t> idx((size_t)n * k);
        DevBuf<double> d2((size_t)n * k);
        knn_device(d_data, (uint32_t)n, d, k, knn_precision(opts), idx.get(), d2.get(), stats, s);
        std::unique_ptr<sc_graph> g = gaussian_graph_device(idx.get(), d2.get(), (uint32_t)n, k, knn_mode(opts), s);
        *out = g.release();
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

sc_status sc_knn_gaussian_csr(const uint32_t *idx, const double *d2, uint64_t n, uint32_t k, int32_t mode,
                              uint64_t *node_ptrs, uint32_t *neighbors, float *weights, uint64_t *nnz_out,
                              sc_error *err) {
    try {
        if (!idx || !d2 || !node_ptrs || !neighbors || !weights || !nnz_out)
            throw ScError(SC_INVALID_ARGUMENT, "null argument");
        if (k == 0 || k > 32) throw ScError(SC_INVALID_ARGUMENT, "k must be in [1, 32]");
        check_node_count(n);
        check_entry_count(2ull * n * k);
        for (uint64_t q = 0; q < n * k; q++)
            if (idx[q] != KNONE && (uint64_t)idx[q] >= n) {
                ScError e(SC_NODE_INDEX_OUT_OF_RANGE, "neighbour index " + std::to_string(idx[q]) +
                                                          " is out of range for " + std::to_string(n) + " points");
                e.payload.node = idx[q];
                e.payload.n_nodes = n;
                throw e;
            }
        hipStream_t s = thread_stream();
        DevBuf<uint32_t> di((size_t)n * k);
        DevBuf<double> dd((size_t)n * k);
        to_device(di.get(), idx, (size_t)n * k, s);
        to_device(dd.get(), d2, (size_t)n * k, s);
        std::unique_ptr<sc_graph> g = gaussian_graph_device(di.get(), dd.get(), (uint32_t)n, k, mode, s);
        to_host(node_ptrs, g->ptrs.get(), n + 1, s);
        to_host(neighbors, g->nbrs.get(), g->nnz, s);
        to_host(weights, g->wts.get(), g->nnz, s);
        sync_stream(s);
        *nnz_out = g->nnz;
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

}  // extern "C"
