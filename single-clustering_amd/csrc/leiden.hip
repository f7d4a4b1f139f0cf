// leiden.hip -- the Leiden core on gfx950: local moving, refinement, aggregation, driver.
//
// Restates leiden/{local_move.rs:39-122, refine.rs:29-153, mod.rs:168-289} and
// network/csr_network.rs:516-608 (aggregate) for a GPU:
//
//  * Local moving runs synchronous sub-sweeps. Each iteration splits the active vertices into P
//    pseudo-random subsets; within a subset every vertex picks its best community from one
//    snapshot (decide), then all moves are applied (apply) and neighbours outside the new
//    community are re-activated (activate) -- the reference's queue re-push rule
//    (local_move.rs:114-118). A singleton may join another singleton only if the target id is
//    lower, which removes swap oscillation. Iterations stop when nothing moves.
//  * Neighbour-community weights are accumulated per vertex in an LDS open-addressing hash
//    (global memory for very high degree), vertex groups of G lanes sized by degree bucket.
//    Weights and community totals are int64 fixed point: atomic accumulation is exact and order
//    independent, so a fixed seed gives bit-identical labels.
//  * Delta-quality is the reference's formula in f64 (objective.rs:91-106, :158-169), ties go
//    to the lower community id, acceptance needs gain > 1e-12 * max(m, 1) (local_move.rs:22-28).
//  * Refinement (only singletons move, moves stay inside the working community,
//    refine.rs:9-11) runs in rounds. A community that is targeted in a round does not move in
//    that round (so every join attaches to a stable community and connectivity is kept), except
//    a mutual pair, where the higher id joins the lower when nobody else targets it. The
//    exp(gain/theta) draw of refine.rs:110-153 is taken with the Gumbel-max trick, which has the
//    same distribution and does not depend on candidate order.
//  * Aggregation sums every coarse edge in the reference's order (members ascending, rows
//    ascending) via a stable radix sort by (group, neighbour group), so coarse weights are
//    bit-identical to the CPU path.
#include "common.hpp"
#include "mlp.hpp"
#include "tune.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <memory>
#include <random>
#include <vector>
#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <string>

namespace scgpu {

double quality_device(const sc_graph &g, const uint64_t *d_labels, int mode, double gamma, hipStream_t s);
double quality_singletons_device(const sc_graph &g, int mode, double gamma, hipStream_t s);

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t EMPTY_KEY = 0xFFFFFFFFu;
constexpr int MOVE_SUBSWEEPS = 2;
constexpr int MAX_MOVE_ITERS = 100;
constexpr int MAX_REFINE_ROUNDS = 64;
// multi-launch and graph-replayed phases: iterations launched (or replayed) per host check
constexpr int ITER_BATCH = 4;

thread_local sc_leiden_stats t_stats;

struct DecideArgs {
    // level graph
    uint32_t n;
    const uint64_t *ptrs;
    const uint32_t *nbrs;
    const float *wts;
    const double *strength;
    const double *node_w;
    const double *self_loop;
    // vertices of this bucket
    const uint32_t *list;
    uint32_t list_len;
    // partition state (working partition for local moving, refined one for refinement)
    const uint32_t *memb;
    const uint32_t *constraint;
    const long long *Kfx;
    const long long *Wfx;
    const uint32_t *csize;
    uint8_t *active;
    uint32_t *target;
    uint32_t *tround;
    uint32_t *movers;
    unsigned int *n_movers;
    uint32_t *targeted_cnt;
    // local moving without a weight cap (null otherwise): joiner admission (join_admitted).
    // jsum / jmin hold two sets of jstride entries, used by alternating steps
    long long *jsum;
    uint32_t *jmin;
    double *mgain;
    uint64_t jstride;
    // objective
    int objective;
    double gamma;
    double four_m;
    double threshold;
    int sK, sW, sw;
    int has_max;
    double max_w;
    // randomness / scheduling
    uint64_t seed;
    uint32_t P;
    uint32_t sub;
    uint32_t round;
    double theta;
    // graph-replayed iterations: the iteration index is read from device memory (null: seed
    // and round above are used); local moving derives its seed from iter_seed_base the way the
    // persistent sweep does
    const uint32_t *iter_dev;
    uint64_t iter_seed_base;
    // multi-launch steps: this step's deciders of the bucket, compacted by k_compact (vertices;
    // list positions for the global-table bucket), their count in device memory, and the phase's
    // stop flag (set once it has converged: every later launch of the phase returns at once)
    const uint32_t *sel;
    const unsigned int *sel_len;
    const unsigned int *stop;
    // global hash tables for the highest-degree bucket
    uint32_t *gkeys;
    long long *gvals;
    const uint64_t *goff;
    // counters
    unsigned long long *stats;  // [vertices, edges, candidates]
    // batched runs (null otherwise): job of every vertex and the jobs' parameters
    const uint32_t *vjob;
    const struct JobParams *jobs;
    // learned filter support: the level index, and (trace runs) where local-moving decisions
    // are recorded as (features, moved) samples
    uint32_t level;
    float *trace_x;
    uint8_t *trace_y;
    uint64_t *trace_key;  // (level << 56 | iteration << 32 | vertex): the samples' canonical order
    unsigned long long *trace_n;
    uint64_t trace_cap;
    // per-vertex filter features kept by local moving when a filter or a trace is on (else
    // null): the last decision's gain margin and own-community weight fraction (-1 before the
    // vertex's first decision at this level), and neighbour moves since that decision
    float *f_margin;
    float *f_own;
    uint32_t *f_nmv;
};

// First list position of every degree bucket (and the total), passed by value to kernels.
struct BucketStarts {
    uint32_t s[NUM_BUCKETS + 1];
};

// Parameters of one job of a batched run (one graph at one resolution, sharing a disjoint-union
// graph with the other jobs). A job's vertices are contiguous at every level, starting at lo.
struct JobParams {
    double gamma, four_m, threshold, theta;
    int sK, sW, sw;
    uint32_t lo;      // first vertex of the job at the current level
    uint32_t frozen;  // the job has stopped (its level loop or its passes ended): it never moves
    uint32_t pad;
};

// Per-vertex view of the objective parameters: the graph's own, or its job's in a batched run.
struct VParams {
    double gamma, four_m, threshold, theta;
    int sK, sW, sw;
    uint32_t lo;
    bool frozen;
};

// BATCH = false: a run without a job table, whose parameters are the same for every vertex (the
// persistent sweeps are compiled for it separately, so they stay scalar)
template <bool BATCH = true>
__device__ __forceinline__ VParams vparams(const DecideArgs &a, uint32_t v) {
    if (!BATCH || !a.vjob) return VParams{a.gamma, a.four_m, a.threshold, a.theta, a.sK, a.sW, a.sw, 0u, false};
    const JobParams &p = a.jobs[a.vjob[v]];
    return VParams{p.gamma, p.four_m, p.threshold, p.theta, p.sK, p.sW, p.sw, p.lo, p.frozen != 0u};
}

__device__ __forceinline__ double delta_insert(int objective, const VParams &p, double w_c, double loop, double k,
                                               double na, double Kc, double Wc) {
    if (objective == 0) {
        if (p.four_m == 0.0) return 0.0;
        return (w_c + loop) - p.gamma * k * (2.0 * Kc + k) / p.four_m;
    }
    return (w_c + loop) - p.gamma * na * (2.0 * Wc + na - 1.0) / 2.0;
}

// Joiner admission (local moving without a weight cap). Every decision of a sub-sweep reads one
// snapshot, so the joiners of one community each assume it unchanged; if they all land, its
// strength (Rb) or node weight (CPM) grows by theirs. The sequential reference sees each earlier
// joiner land (local_move.rs:58-119), so a flood of joiners that each gain on the snapshot but
// lose together cannot happen there. A mover whose gain stays above the threshold when every
// other joiner of its target lands moves; of the remaining joiners of a community only the lowest
// vertex id moves (its snapshot gain holds if no one else lands), the others stay active and
// decide again. The joiners' sums are fixed point (exact, order independent) and the lowest id is
// an atomicMin, so admission is deterministic.
__device__ __forceinline__ long long join_amount(const DecideArgs &a, const VParams &vp, double k, double na) {
    return a.objective == 0 ? to_fx(k, vp.sK) : to_fx(na, vp.sW);
}

__device__ __forceinline__ bool join_admitted(const DecideArgs &a, const VParams &vp, uint32_t v, uint32_t t,
                                              uint32_t set) {
    const uint64_t idx = (uint64_t)set * a.jstride + t;
    if (a.jmin[idx] == v) return true;
    const double k = a.strength[v], na = a.node_w[v];
    const long long others = a.jsum[idx] - join_amount(a, vp, k, na);
    double extra;
    if (a.objective == 0)
        extra = vp.four_m == 0.0 ? 0.0 : vp.gamma * k * (2.0 * from_fx(others, vp.sK)) / vp.four_m;
    else
        extra = vp.gamma * na * from_fx(others, vp.sW);
    return a.mgain[v] - extra > vp.threshold;
}

// Features of an active vertex for the learned filter, from the state every decision reads
// and the vertex's last decision at this level: degree, own-community weight fraction and
// best-gain margin of the last decision, neighbour moves since then, whether it has decided at
// this level yet, iteration, level and own community size (each scaled to about [0, 1]).
__device__ __forceinline__ void vertex_features(const DecideArgs &a, const VParams &vp, const uint32_t *mstep,
                                                uint32_t v, uint32_t it, float (&x)[MLP_NF]) {
    (void)vp;
    (void)mstep;
    const uint32_t own = a.memb[v];
    const double deg = (double)(a.ptrs[v + 1] - a.ptrs[v]);
    const float fo = a.f_own[v];
    x[0] = (float)(log2(1.0 + deg) / 16.0);
    x[1] = fo >= 0.0f ? fo : 0.0f;
    x[2] = a.f_margin[v];
    x[3] = (float)(log2(1.0 + (double)a.f_nmv[v]) / 8.0);
    x[4] = fo >= 0.0f ? 1.0f : 0.0f;
    x[5] = (float)min(it, 15u) / 15.0f;
    x[6] = (float)min(a.level, 15u) / 15.0f;
    x[7] = (float)(log2(1.0 + (double)a.csize[own]) / 16.0);
}

__device__ __forceinline__ uint32_t subset_of(uint64_t seed, uint32_t v, uint32_t P) {
    return (uint32_t)((mix64(seed ^ ((uint64_t)v * 0x9E3779B97F4A7C15ull)) >> 33) % P);
}

__device__ __forceinline__ double gumbel(uint64_t seed, uint32_t round, uint32_t v, uint32_t c) {
    uint64_t h = mix64(seed ^ mix64(((uint64_t)v << 32) ^ c) ^ ((uint64_t)round * 0xD1B54A32D192ED03ull));
    double u = u01_open(h);
    return -log(-log(u));
}

template <typename KeyPtr, typename ValPtr>
__device__ __forceinline__ void table_insert(KeyPtr keys, ValPtr vals, uint32_t mask, uint32_t c, long long w) {
    uint32_t h = hash32(c) & mask;
    while (true) {
        uint32_t prev = atomicCAS(&keys[h], EMPTY_KEY, c);
        if (prev == EMPTY_KEY || prev == c) {
            atomicAdd((unsigned long long *)&vals[h], (unsigned long long)w);
            return;
        }
        h = (h + 1) & mask;
    }
}

// argmax over (score, id): larger score wins, ties to the smaller id
__device__ __forceinline__ void better(double &bs, uint32_t &bc, double s, uint32_t c) {
    if (s > bs || (s == bs && c < bc)) { bs = s; bc = c; }
}

template <int G>
__device__ __forceinline__ void group_reduce_wave(double &bs, uint32_t &bc, long long &wown) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        double os = __shfl_xor(bs, off, G);
        uint32_t oc = __shfl_xor(bc, off, G);
        long long ow = __shfl_xor(wown, off, G);
        better(bs, bc, os, oc);
        wown += ow;
    }
}

// Ordering point between the phases of a decide unit (clear table / insert / scan). A group of
// G <= 64 lanes lies inside one wave and owns its LDS table, and a wave's LDS operations are
// performed in order, so only the compiler must not move them across this point: the waves of a
// workgroup then run their units independently. Wider groups and global tables synchronize the
// workgroup.
template <int G, bool GTAB>
__device__ __forceinline__ void group_sync() {
    if constexpr (GTAB || G > 64) {
        __syncthreads();
        if (GTAB) __threadfence_block();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Per-step values of a decide pass (the multi-launch path takes them from DecideArgs).
struct StepParams {
    uint64_t seed;
    uint32_t sub;
    uint32_t round;
    uint32_t step;
    unsigned int *n_movers;  // mover count: global, or the workgroup's LDS counter with mv_list
    uint32_t *mstep;         // step stamp of accepted movers (persistent local moving), else null
    uint32_t *mv_list;       // persistent sweeps: the workgroup's own mover list (else a.movers)
    uint32_t *mv_tgt;        // persistent local moving: the movers' targets, beside mv_list (else null)
};

struct DecideStats {
    unsigned long long v = 0, e = 0, c = 0;
};

// One unit of a decide pass: NG = BS / G vertices of one degree bucket, G lanes each. blist is
// the bucket's vertex list; sel, when given, holds this step's vertices (their positions within
// blist for the global-table bucket), which are then already known to be active and in the
// step's subset. base is the index (into sel, or blist) of the unit's first vertex. Every thread
// of the block calls this.
// WAVE: the unit is one wave's (64 / G vertices, G <= 64): groups are numbered within the wave
// and the tables lie in the wave's own part of s_keys / s_vals (WAVE_SLOTS entries), so the
// waves of a workgroup can run units of different buckets at the same time.
constexpr int WAVE_SLOTS = 1024;

template <int BS, int G, int T, bool REFINE, bool GTAB, bool BATCH = true, bool WAVE = false>
__device__ __forceinline__ void decide_unit(const DecideArgs &a, const StepParams &sp, const uint32_t *blist,
                                            const uint32_t *sel, uint32_t len, uint64_t base, uint32_t *s_keys,
                                            long long *s_vals, double *r_s, uint32_t *r_c, long long *r_w,
                                            DecideStats &ds) {
    static_assert(!WAVE || (G <= 64 && !GTAB && (64 / G) * T <= WAVE_SLOTS), "wave units need a wave's table");
    const int lane = threadIdx.x % G;
    const int grp = WAVE ? (int)((threadIdx.x & 63) / G) : (int)(threadIdx.x / G);
    if (WAVE) {
        s_keys += (threadIdx.x >> 6) * WAVE_SLOTS;
        s_vals += (threadIdx.x >> 6) * WAVE_SLOTS;
    }
    const uint64_t idx = base + grp;
    bool valid = idx < len;
    // sel holds vertices, or list positions for the global-table bucket (it needs goff[pos])
    uint32_t pos = valid ? (sel ? sel[idx] : (uint32_t)idx) : 0u;
    uint32_t v = valid ? ((sel && !GTAB) ? pos : blist[pos]) : 0u;
    const VParams vp = vparams<BATCH>(a, v);
    if (valid && !sel) {
        valid = a.active[v] != 0 && !vp.frozen && subset_of(sp.seed, v - vp.lo, a.P) == sp.sub;
        if (REFINE) valid = valid && a.memb[v] == v && a.csize[v] == 1;
    }
    uint32_t *keys;
    long long *vals;
    uint32_t tsize;
    if (GTAB) {
        uint64_t o = valid ? a.goff[pos] : 0;
        tsize = valid ? (uint32_t)(a.goff[pos + 1] - o) : 0;
        keys = a.gkeys + o;
        vals = a.gvals + o;
    } else {
        keys = s_keys + grp * T;
        vals = s_vals + grp * T;
        tsize = T;
    }
    for (uint32_t s = lane; s < tsize; s += G) {
        keys[s] = EMPTY_KEY;
        vals[s] = 0;
    }
    group_sync<G, GTAB>();

    uint64_t beg = 0, end = 0;
    uint32_t own = 0, cv = 0;
    if (valid) {
        beg = a.ptrs[v];
        end = a.ptrs[v + 1];
        own = a.memb[v];
        if (REFINE) cv = a.constraint[v];
    }
    const uint32_t mask = tsize - 1;
    // EB entries per lane in flight: their neighbour ids, then their memberships, are gathered
    // with independent loads before the table inserts (the inserts' order does not matter:
    // fixed-point sums are exact and candidates are ranked by gain, then id)
    constexpr int EB = 4;
    for (uint64_t e0 = beg + lane; e0 < end; e0 += (uint64_t)EB * G) {
        uint32_t u[EB];
        float w[EB];
#pragma unroll
        for (int j = 0; j < EB; j++) {
            const uint64_t e = e0 + (uint64_t)j * G;
            const bool in = e < end;
            u[j] = in ? a.nbrs[e] : v;  // v: no entry (self-loops are the partition's, partition.rs:430)
            w[j] = in ? a.wts[e] : 0.0f;
        }
        uint32_t c[EB];
#pragma unroll
        for (int j = 0; j < EB; j++) {
            bool ok = u[j] != v;
            c[j] = ok ? a.memb[u[j]] : EMPTY_KEY;
            if (REFINE && ok && a.constraint[u[j]] != cv) c[j] = EMPTY_KEY;
        }
#pragma unroll
        for (int j = 0; j < EB; j++)
            if (c[j] != EMPTY_KEY) table_insert(keys, vals, mask, c[j], to_fx((double)w[j], vp.sw));
    }
    group_sync<G, GTAB>();

    double bs = -INFINITY;
    uint32_t bc = NONE;
    long long wown = 0;
    double k = 0.0, na = 0.0, loop = 0.0, rbase = 0.0;
    if (valid) {
        k = a.strength[v];
        na = a.node_w[v];
        loop = a.self_loop[v];
        // refinement: the node leaves its singleton, which becomes empty (refine.rs:71-74)
        if (REFINE) rbase = delta_insert(a.objective, vp, 0.0, loop, k, na, 0.0, 0.0);
        // SB candidates per lane at a time: the lane walks its table slots (stride G) in LDS
        // until it holds SB occupied ones (its own community's weight is taken on the way),
        // then their community totals are gathered with independent loads (an unused entry
        // reads the vertex's own community, a valid index) before any is scored. Candidates
        // are few next to the slots, so most lanes gather once. Every candidate is scored
        // exactly as one at a time, and the argmax does not depend on the order.
        // (A table in global memory, for the highest degrees, is read SB slots at a time with
        // independent loads instead: walking it slot by slot would chain its loads.)
        constexpr int SB = 4;
        uint32_t sl = lane;
        while (true) {
            uint32_t cs[SB];
            long long wv[SB], Wv[SB], Kv[SB];
            if (GTAB) {
#pragma unroll
                for (int j = 0; j < SB; j++) {
                    const uint32_t at = sl + (uint32_t)j * G;
                    cs[j] = at < tsize ? keys[at] : EMPTY_KEY;
                    wv[j] = cs[j] != EMPTY_KEY ? vals[at] : 0;
                }
                sl += (uint32_t)SB * G;
#pragma unroll
                for (int j = 0; j < SB; j++)
                    if (cs[j] == own) {
                        wown = wv[j];
                        cs[j] = EMPTY_KEY;
                    }
                if (sl - (uint32_t)SB * G >= tsize) break;
            } else {
#pragma unroll
                for (int j = 0; j < SB; j++) {
                    cs[j] = EMPTY_KEY;
                    wv[j] = 0;
                    while (sl < tsize) {
                        const uint32_t key = keys[sl];
                        const uint32_t at = sl;
                        sl += G;
                        if (key == EMPTY_KEY) continue;
                        if (key == own) {
                            wown = vals[at];
                            continue;
                        }
                        cs[j] = key;
                        wv[j] = vals[at];
                        break;
                    }
                }
                if (cs[0] == EMPTY_KEY) break;
            }
#pragma unroll
            for (int j = 0; j < SB; j++) {
                const uint32_t ci = cs[j] != EMPTY_KEY ? cs[j] : own;
                // Rb reads community strengths only, CPM node weights only; the cap reads weights
                Wv[j] = (a.objective != 0 || a.has_max) ? a.Wfx[ci] : 0;
                Kv[j] = a.objective == 0 ? a.Kfx[ci] : 0;
            }
#pragma unroll
            for (int j = 0; j < SB; j++) {
                const uint32_t c = cs[j];
                if (c == EMPTY_KEY) continue;
                ds.c++;
                double Wc = from_fx(Wv[j], vp.sW);
                if (a.has_max && !(Wc + na <= a.max_w)) continue;
                double Kc = a.objective == 0 ? from_fx(Kv[j], vp.sK) : 0.0;
                double gain = delta_insert(a.objective, vp, from_fx(wv[j], vp.sw), loop, k, na, Kc, Wc);
                if (REFINE) {
                    double rel = gain - rbase;
                    if (!(rel >= 0.0)) continue;  // candidates that do not lose quality (refine.rs:91-95)
                    double score =
                        vp.theta > 0.0 ? rel / vp.theta + gumbel(sp.seed, sp.round, v - vp.lo, c - vp.lo) : rel;
                    better(bs, bc, score, c);
                } else {
                    better(bs, bc, gain, c);
                }
            }
            if ((!GTAB && cs[SB - 1] == EMPTY_KEY) || sl >= tsize) break;
        }
    }
    // group-wide reduction
    if (G <= 64) {
        group_reduce_wave<(G <= 64 ? G : 64)>(bs, bc, wown);
    } else {
        static_assert(G <= 64 || G == BS, "groups wider than a wave span the whole block");
        group_reduce_wave<64>(bs, bc, wown);
        const int wave = threadIdx.x / 64;
        if ((threadIdx.x & 63) == 0) {
            r_s[wave] = bs;
            r_c[wave] = bc;
            r_w[wave] = wown;
        }
        __syncthreads();
        if (lane == 0) {
            bs = r_s[0];
            bc = r_c[0];
            wown = r_w[0];
            for (int w = 1; w < BS / 64; w++) {
                better(bs, bc, r_s[w], r_c[w]);
                wown += r_w[w];
            }
        }
    }
    bool push = false;
    if (valid && lane == 0) {
        a.active[v] = 0;
        ds.v++;
        ds.e += end - beg;
        if (!REFINE) {
            // everything the rest reads, with independent loads: the own community's totals, the
            // sizes that name the empty community, and the sizes of the singleton-swap rule
            const uint32_t bc0 = bc;
            const long long Kofx = a.objective == 0 ? a.Kfx[own] : 0;
            const long long Wofx = (a.objective != 0 || a.has_max) ? a.Wfx[own] : 0;
            const uint32_t sz_v = a.csize[v], sz_nv = a.csize[a.n + v], sz_own = a.csize[own];
            const uint32_t sz_bc0 = a.csize[bc0 != NONE ? bc0 : own];
            long long kfx = a.objective == 0 ? to_fx(k, vp.sK) : 0;
            long long wfxv = to_fx(na, vp.sW);
            double Kown = a.objective == 0 ? from_fx(Kofx - kfx, vp.sK) : 0.0;
            double Wown = (a.objective != 0 || a.has_max) ? from_fx(Wofx - wfxv, vp.sW) : 0.0;
            double base_q = delta_insert(a.objective, vp, from_fx(wown, vp.sw), loop, k, na, Kown, Wown);
            // the empty community (local_move.rs:93-100): this vertex's own id when no one holds
            // it, else its private id n + v (community ids span [0, 2n), so a vertex whose own
            // id names a community -- its own or another -- can still leave to an empty one)
            const uint32_t e = sz_v == 0 ? v : (sz_nv == 0 ? a.n + v : NONE);
            if (e != NONE && (!a.has_max || na <= a.max_w)) {
                double ge = delta_insert(a.objective, vp, 0.0, loop, k, na, 0.0, 0.0);
                better(bs, bc, ge, e);
            }
            bool accept = bc != NONE && bc != own && bs - base_q > vp.threshold;
            // the empty community, when it won, has size 0
            const uint32_t sz_bc = bc == bc0 ? sz_bc0 : 0u;
            if (accept && sz_own == 1 && sz_bc == 1 && bc > own) accept = false;
            if (a.trace_x) {
                const unsigned long long slot = atomicAdd(a.trace_n, 1ull);
                if (slot < a.trace_cap) {
                    float x[MLP_NF];
                    vertex_features(a, vp, sp.mstep, v, sp.round, x);
#pragma unroll
                    for (int f = 0; f < MLP_NF; f++) a.trace_x[slot * MLP_NF + f] = x[f];
                    a.trace_y[slot] = accept ? 1 : 0;
                    a.trace_key[slot] = ((uint64_t)a.level << 56) | ((uint64_t)min(sp.round, 0xFFFFFFu) << 32) | v;
                }
            }
            if (a.f_margin) {
                // this decision becomes the vertex's "last decision" for its next features
                const double mg = bc != NONE ? (bs - base_q) / fmax(k, 1e-300) : -1e300;
                const float m = (float)fmin(fmax(mg, -1e30), 1e30);
                a.f_margin[v] = m / (1.0f + fabsf(m));
                a.f_own[v] = k > 0.0 ? (float)fmin(fmax(from_fx(wown, vp.sw) / k, 0.0), 1.0) : 0.0f;
                a.f_nmv[v] = 0u;
            }
            if (accept) {
                a.target[v] = bc;
                if (sp.mstep) sp.mstep[v] = sp.step;
                if (a.jsum) {
                    a.mgain[v] = bs - base_q;
                    if (!sp.mv_tgt) {  // persistent sweeps sum a workgroup's joiners first
                        const uint64_t idx = (uint64_t)(sp.step & 1u) * a.jstride + bc;
                        atomicAdd((unsigned long long *)&a.jsum[idx], (unsigned long long)join_amount(a, vp, k, na));
                        atomicMin(&a.jmin[idx], v);
                    }
                }
                push = true;
            }
        } else if (bc != NONE) {
            a.target[v] = bc;
            a.tround[v] = sp.round;
            atomicAdd(&a.targeted_cnt[bc], 1u);
            push = true;
        }
    }
    // append the wave's movers with one counter update
    const uint64_t pm = __ballot(push);
    if (pm) {
        const int leader = __ffsll((unsigned long long)pm) - 1;
        unsigned int slot0 = 0;
        if ((int)(threadIdx.x & 63) == leader) slot0 = atomicAdd(sp.n_movers, (unsigned int)__popcll(pm));
        slot0 = __shfl(slot0, leader);
        if (push) {
            const unsigned int slot = slot0 + __builtin_amdgcn_mbcnt_hi((uint32_t)(pm >> 32),
                                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u));
            (sp.mv_list ? sp.mv_list : a.movers)[slot] = v;
            if (sp.mv_tgt) sp.mv_tgt[slot] = bc;
        }
    }
    group_sync<G, GTAB>();
}

// Refinement's first step, in which every vertex is still a singleton (memb[u] = u, size 1, totals
// its own): every neighbour in the working community is a candidate of its own with the entry's
// weight (a row holds each neighbour once), so no table is built and no membership is gathered.
// Candidate set, scores and tie order are decide_unit's, so the decisions are the same. sel holds
// the step's vertices (list positions for the global-table bucket).
template <int BS, int G, bool GTAB, bool BATCH>
__device__ __forceinline__ void decide_unit_single(const DecideArgs &a, const StepParams &sp, const uint32_t *blist,
                                                   const uint32_t *sel, uint32_t len, uint64_t base, double *r_s,
                                                   uint32_t *r_c, long long *r_w, DecideStats &ds,
                                                   const uint4 *rpk) {
    const int lane = threadIdx.x % G;
    const int grp = threadIdx.x / G;
    const uint64_t idx = base + grp;
    const bool valid = idx < len;
    const uint32_t pos = valid ? sel[idx] : 0u;
    const uint32_t v = valid ? (GTAB ? blist[pos] : pos) : 0u;
    const VParams vp = vparams<BATCH>(a, v);
    double bs = -INFINITY;
    uint32_t bc = NONE;
    long long wown = 0;
    uint64_t beg = 0, end = 0;
    if (valid) {
        beg = a.ptrs[v];
        end = a.ptrs[v + 1];
        const uint32_t cv = a.constraint[v];
        const double k = a.strength[v], na = a.node_w[v], loop = a.self_loop[v];
        // the node leaves its singleton, which becomes empty (refine.rs:71-74)
        const double rbase = delta_insert(a.objective, vp, 0.0, loop, k, na, 0.0, 0.0);
        constexpr int EB = 4;
        for (uint64_t e0 = beg + lane; e0 < end; e0 += (uint64_t)EB * G) {
            uint32_t u[EB];
            float w[EB];
#pragma unroll
            for (int j = 0; j < EB; j++) {
                const uint64_t e = e0 + (uint64_t)j * G;
                const bool in = e < end;
                u[j] = in ? a.nbrs[e] : v;  // v: no entry (self-loops are the partition's)
                w[j] = in ? a.wts[e] : 0.0f;
            }
            bool ok[EB];
            long long Wv[EB], Kv[EB];
            if (rpk) {
                // a neighbour's working community and singleton total in one load
#pragma unroll
                for (int j = 0; j < EB; j++) {
                    const uint4 q = rpk[u[j]];
                    const long long val = (long long)(((unsigned long long)q.y << 32) | q.x);
                    ok[j] = u[j] != v && q.z == cv;
                    Kv[j] = a.objective == 0 ? val : 0;
                    Wv[j] = a.objective != 0 ? val : 0;
                }
            } else {
#pragma unroll
                for (int j = 0; j < EB; j++) ok[j] = u[j] != v && a.constraint[u[j]] == cv;
#pragma unroll
                for (int j = 0; j < EB; j++) {
                    const uint32_t ci = ok[j] ? u[j] : v;
                    Wv[j] = (a.objective != 0 || a.has_max) ? a.Wfx[ci] : 0;
                    Kv[j] = a.objective == 0 ? a.Kfx[ci] : 0;
                }
            }
#pragma unroll
            for (int j = 0; j < EB; j++) {
                if (!ok[j]) continue;
                ds.c++;
                const double Wc = from_fx(Wv[j], vp.sW);
                if (a.has_max && !(Wc + na <= a.max_w)) continue;
                const double Kc = a.objective == 0 ? from_fx(Kv[j], vp.sK) : 0.0;
                const double gain =
                    delta_insert(a.objective, vp, from_fx(to_fx((double)w[j], vp.sw), vp.sw), loop, k, na, Kc, Wc);
                const double rel = gain - rbase;
                if (!(rel >= 0.0)) continue;  // candidates that do not lose quality (refine.rs:91-95)
                const double score =
                    vp.theta > 0.0 ? rel / vp.theta + gumbel(sp.seed, sp.round, v - vp.lo, u[j] - vp.lo) : rel;
                better(bs, bc, score, u[j]);
            }
        }
    }
    if (G <= 64) {
        group_reduce_wave<(G <= 64 ? G : 64)>(bs, bc, wown);
    } else {
        group_reduce_wave<64>(bs, bc, wown);
        const int wave = threadIdx.x / 64;
        if ((threadIdx.x & 63) == 0) {
            r_s[wave] = bs;
            r_c[wave] = bc;
        }
        __syncthreads();
        if (lane == 0) {
            bs = r_s[0];
            bc = r_c[0];
            for (int w = 1; w < BS / 64; w++) better(bs, bc, r_s[w], r_c[w]);
        }
    }
    bool push = false;
    if (valid && lane == 0) {
        a.active[v] = 0;
        ds.v++;
        ds.e += end - beg;
        if (bc != NONE) {
            a.target[v] = bc;
            a.tround[v] = sp.round;
            atomicAdd(&a.targeted_cnt[bc], 1u);
            push = true;
        }
    }
    const uint64_t pm = __ballot(push);
    if (pm) {
        const int leader = __ffsll((unsigned long long)pm) - 1;
        unsigned int slot0 = 0;
        if ((int)(threadIdx.x & 63) == leader) slot0 = atomicAdd(sp.n_movers, (unsigned int)__popcll(pm));
        slot0 = __shfl(slot0, leader);
        if (push) {
            const unsigned int slot = slot0 + __builtin_amdgcn_mbcnt_hi((uint32_t)(pm >> 32),
                                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u));
            (sp.mv_list ? sp.mv_list : a.movers)[slot] = v;
        }
    }
    if (G > 64) __syncthreads();  // r_s is rewritten by the next unit
}

__device__ __forceinline__ void flush_decide_stats(const DecideStats &ds, unsigned long long *s_stats,
                                                   unsigned long long *stats) {
    if (ds.v) {
        atomicAdd(&s_stats[0], ds.v);
        atomicAdd(&s_stats[1], ds.e);
    }
    if (ds.c) atomicAdd(&s_stats[2], ds.c);
    __syncthreads();
    if (threadIdx.x < 3 && s_stats[threadIdx.x]) atomicAdd(&stats[threadIdx.x], s_stats[threadIdx.x]);
}

// Multi-launch decide: one launch per degree bucket, grid-stride over the bucket's vertices.
template <int BS, int G, int T, bool REFINE, bool GTAB>
__global__ void __launch_bounds__(BS) k_decide(DecideArgs a) {
    constexpr int NG = BS / G;
    constexpr int TS = GTAB ? 1 : NG * T;
    __shared__ uint32_t s_keys[TS];
    __shared__ long long s_vals[TS];
    __shared__ double r_s[BS / 64];
    __shared__ uint32_t r_c[BS / 64];
    __shared__ long long r_w[BS / 64];
    __shared__ unsigned long long s_stats[3];
    if (threadIdx.x < 3) s_stats[threadIdx.x] = 0;
    uint64_t seed = a.seed;
    uint32_t round = a.round;
    if (a.iter_dev) {
        round = *a.iter_dev;
        if (!REFINE) seed = mix64(a.iter_seed_base ^ (0x5851F42D4C957F2Dull * (uint64_t)(round + 1)));
    }
    StepParams sp{seed, a.sub, round, 0u, a.n_movers, nullptr, nullptr};
    DecideStats ds;
    if (a.stop && *a.stop) return;  // workgroup-uniform
    const uint64_t len = a.sel_len ? (uint64_t)*a.sel_len : (uint64_t)a.list_len;
    for (uint64_t base = (uint64_t)blockIdx.x * NG; base < len; base += (uint64_t)gridDim.x * NG)
        decide_unit<BS, G, T, REFINE, GTAB>(a, sp, a.list, a.sel, len, base, s_keys, s_vals, r_s, r_c, r_w, ds);
    flush_decide_stats(ds, s_stats, a.stats);
}

// The vertices of every degree bucket that decide in this step of the multi-launch path: active,
// in the step's subset and not frozen (refinement: still singletons), appended to the bucket's
// part of clist (vertices; list positions for the global-table bucket) with one global atomic per
// bucket and workgroup; their flags are cleared (they leave the active set by deciding).
template <bool REFINE>
__global__ void __launch_bounds__(256) k_compact(DecideArgs a, const uint32_t *lists, BucketStarts bs,
                                                 uint32_t *clist, unsigned int *ccount) {
    constexpr int U = 4;  // positions per thread and pass, their loads in flight together
    __shared__ unsigned int sc[NUM_BUCKETS], sb[NUM_BUCKETS];
    if (a.stop && *a.stop) return;
    uint64_t seed = a.seed;
    if (a.iter_dev && !REFINE) seed = mix64(a.iter_seed_base ^ (0x5851F42D4C957F2Dull * (uint64_t)(*a.iter_dev + 1)));
    const uint32_t total = bs.s[NUM_BUCKETS];
    const int lane = threadIdx.x & 63;
    for (uint32_t i0 = blockIdx.x * (256u * U); i0 < total; i0 += gridDim.x * (256u * U)) {  // workgroup-uniform
        if (threadIdx.x < NUM_BUCKETS) sc[threadIdx.x] = 0;
        __syncthreads();
        uint32_t ip[U], v[U], local[U];
        int kb[U];
        bool take[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            ip[q] = i0 + (uint32_t)q * 256u + threadIdx.x;
            v[q] = ip[q] < total ? lists[ip[q]] : 0u;
        }
        uint8_t f[U];
#pragma unroll
        for (int q = 0; q < U; q++) f[q] = ip[q] < total ? a.active[v[q]] : (uint8_t)0;
#pragma unroll
        for (int q = 0; q < U; q++) {
            kb[q] = 0;
#pragma unroll
            for (int k = 1; k < NUM_BUCKETS; k++) kb[q] += ip[q] >= bs.s[k] ? 1 : 0;
            take[q] = false;
            if (f[q] == 1) {
                const VParams vp = vparams<true>(a, v[q]);
                take[q] = !vp.frozen && subset_of(seed, v[q] - vp.lo, a.P) == a.sub;
            }
        }
        if (REFINE) {
            uint32_t m[U], cs[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                m[q] = take[q] ? a.memb[v[q]] : 0u;
                cs[q] = take[q] ? a.csize[v[q]] : 0u;
            }
#pragma unroll
            for (int q = 0; q < U; q++) take[q] = take[q] && m[q] == v[q] && cs[q] == 1u;
        }
        // a wave's takes per bucket reserve their places with one LDS atomic
#pragma unroll
        for (int q = 0; q < U; q++) {
            local[q] = 0;
#pragma unroll
            for (int k = 0; k < NUM_BUCKETS; k++) {
                const uint64_t mk = __ballot(take[q] && kb[q] == k);
                if (!mk) continue;
                const int leader = __ffsll((unsigned long long)mk) - 1;
                unsigned int basei = 0;
                if (lane == leader) basei = atomicAdd(&sc[k], (unsigned int)__popcll(mk));
                basei = __shfl(basei, leader);
                if (take[q] && kb[q] == k)
                    local[q] = basei + __builtin_amdgcn_mbcnt_hi((uint32_t)(mk >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u));
            }
        }
        __syncthreads();
        if (threadIdx.x < NUM_BUCKETS && sc[threadIdx.x]) sb[threadIdx.x] = atomicAdd(&ccount[threadIdx.x], sc[threadIdx.x]);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < U; q++)
            if (take[q]) {
                clist[bs.s[kb[q]] + sb[kb[q]] + local[q]] = kb[q] == GTAB_BUCKET ? ip[q] - bs.s[kb[q]] : v[q];
                a.active[v[q]] = 0;
            }
        __syncthreads();
    }
}

// Start of a multi-launch step: the compacted counts and the step's mover counters start at 0.
__global__ void k_step_begin(const unsigned int *stop, unsigned int *ccount, unsigned int *zero, uint32_t nzero) {
    if (*stop) return;
    if (threadIdx.x < NUM_BUCKETS) ccount[threadIdx.x] = 0;
    if (threadIdx.x < nzero) zero[threadIdx.x] = 0;
}

// End of an iteration (local moving: both sub-sweeps; refinement: one round). moved[0 .. nmoved)
// count what it moved; none (or the iteration limit) sets the stop flag. ctl: [0] stop,
// [1] iterations run, [2] their moves. The graph-replayed path's iteration index advances.
__global__ void k_iter_end(unsigned int *ctl, const unsigned int *moved, uint32_t nmoved, uint32_t max_iters,
                           uint32_t *iter_dev) {
    if (ctl[0]) return;
    unsigned int m = 0;
    for (uint32_t i = 0; i < nmoved; i++) m += moved[i];
    const unsigned int it = ++ctl[1];
    ctl[2] += m;
    if (m == 0 || it >= max_iters) ctl[0] = 1u;
    if (iter_dev) *iter_dev += 1u;
}

// Refinement: the targeted counts this round's movers set go back to 0 for the next round.
__global__ void k_clear_targeted(const unsigned int *stop, const uint32_t *movers, const unsigned int *count,
                                 const uint32_t *target, uint32_t *targeted_cnt) {
    if (*stop) return;
    const unsigned int cnt = *count;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x)
        targeted_cnt[target[movers[i]]] = 0;
}

// Fill that does nothing once the phase has stopped.
__global__ void k_fill_u32_live(const unsigned int *stop, uint32_t *p, uint64_t n, uint32_t x) {
    if (*stop) return;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = x;
}

struct ApplyArgs {
    const uint32_t *movers;
    const unsigned int *n_movers;
    uint32_t *memb;
    const uint32_t *target;
    const uint32_t *tround;
    const uint32_t *targeted_cnt;
    const uint32_t *csize_snapshot;  // unused placeholder for symmetry with decide
    long long *Kfx;
    long long *Wfx;
    uint32_t *csize;
    const double *strength;
    const double *node_w;
    int sK, sW;
    uint32_t round;
    uint8_t *active;
    uint32_t *moved;
    unsigned int *n_moved;
    const uint32_t *claim;  // with a community-weight cap: lowest mover id per target, else null
    const uint32_t *iter_dev;  // graph-replayed rounds: the round index in device memory, else null
    uint8_t *admitted;      // local moving without a cap: 1 if the mover was admitted (join_admitted)
    const unsigned int *stop;  // the phase's stop flag (every launch returns at once once set)
    uint32_t *mstep;        // refinement: the round in which a vertex moved (NONE before)
};

// Local moving: a decided mover moves if its target admits it (the lowest claim with a weight
// cap, join_admitted without; set 0 of the joiner sums, which k_clear_joins resets afterwards).
__global__ void k_apply_move(ApplyArgs a, DecideArgs d) {
    if (a.stop && *a.stop) return;
    unsigned int cnt = *a.n_movers;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
        const uint64_t v = a.movers[i];
        const uint64_t t = a.target[v];
        if (a.claim && a.claim[t] != (uint32_t)v) {
            a.active[v] = 1;  // another vertex claimed the capped community first: retry later
            continue;
        }
        if (d.jsum) {
            const bool go = join_admitted(d, vparams<true>(d, (uint32_t)v), (uint32_t)v, (uint32_t)t, 0u);
            a.admitted[v] = go ? 1 : 0;
            if (!go) {
                a.active[v] = 1;  // the target is crowded this sub-sweep: decide again later
                continue;
            }
        }
        const uint64_t old = a.memb[v];
        a.memb[v] = (uint32_t)t;
        long long kf = to_fx(a.strength[v], a.sK), wf = to_fx(a.node_w[v], a.sW);
        atomicAdd((unsigned long long *)&a.Kfx[t], (unsigned long long)kf);
        atomicAdd((unsigned long long *)&a.Kfx[old], (unsigned long long)(-kf));
        atomicAdd((unsigned long long *)&a.Wfx[t], (unsigned long long)wf);
        atomicAdd((unsigned long long *)&a.Wfx[old], (unsigned long long)(-wf));
        atomicAdd(&a.csize[t], 1u);
        atomicSub(&a.csize[old], 1u);
    }
}

// Refinement: a targeted singleton stays and absorbs its joiners, except the higher id of a
// mutual pair that nobody else targets.
// Indices are widened to 64 bits where they are loaded and the condition is evaluated without
// short-circuit branches: every operand is in bounds (t is a community id < n), and a
// branchy form of this test was compiled into address arithmetic that, on the tc == 0 path,
// took the high half of a 64-bit index from a stale register.
__global__ void k_apply_refine(ApplyArgs a) {
    if (a.stop && *a.stop) return;
    unsigned int cnt = *a.n_movers;
    const uint32_t round = a.iter_dev ? *a.iter_dev : a.round;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
        const uint64_t v = a.movers[i];
        const uint64_t t = a.target[v];
        const uint32_t tc = a.targeted_cnt[v];
        const uint32_t tr = a.tround[t];
        const uint64_t tt = a.target[t];
        const uint32_t cl = a.claim ? a.claim[t] : (uint32_t)v;
        const bool go = ((tc == 0u) | ((tc == 1u) & (t < v) & (tr == round) & (tt == v))) & (cl == (uint32_t)v);
        if (!go) {
            a.active[v] = 1;  // blocked: look again next round
            continue;
        }
        a.memb[v] = t;
        if (a.mstep) a.mstep[v] = round;
        long long kf = to_fx(a.strength[v], a.sK), wf = to_fx(a.node_w[v], a.sW);
        atomicAdd((unsigned long long *)&a.Kfx[t], (unsigned long long)kf);
        atomicAdd((unsigned long long *)&a.Kfx[v], (unsigned long long)(-kf));
        atomicAdd((unsigned long long *)&a.Wfx[t], (unsigned long long)wf);
        atomicAdd((unsigned long long *)&a.Wfx[v], (unsigned long long)(-wf));
        atomicAdd(&a.csize[t], 1u);
        atomicSub(&a.csize[v], 1u);
        unsigned int slot = atomicAdd(a.n_moved, 1u);
        a.moved[slot] = v;
    }
}

// With a community-weight cap, each target community admits one joiner per sub-sweep (the
// lowest vertex id), so the snapshot check W[c] + w_v <= cap stays true after the moves.
__global__ void k_claim(const uint32_t *movers, const unsigned int *count, const uint32_t *target, uint32_t *claim,
                        const unsigned int *stop = nullptr) {
    if (stop && *stop) return;
    unsigned int cnt = *count;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
        const uint64_t v = movers[i];
        const uint64_t t = target[v];
        atomicMin(&claim[t], (uint32_t)v);
    }
}

// Resets the joiner sums and lowest ids of this sub-sweep's targets (set 0) for the next one.
__global__ void k_clear_joins(const uint32_t *movers, const unsigned int *count, const uint32_t *target,
                              long long *jsum, uint32_t *jmin, const unsigned int *stop = nullptr) {
    if (stop && *stop) return;
    unsigned int cnt = *count;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
        const uint32_t t = target[movers[i]];
        jsum[t] = 0;
        jmin[t] = 0xFFFFFFFFu;
    }
}

// Re-activation after moves. Local moving: neighbours not in the mover's new community
// (local_move.rs:114-118). Refinement: neighbours inside the same working community. With a
// weight cap (claim != null) a local mover whose claim lost did not move and activates nothing.
template <int G, bool REFINE>
__global__ void __launch_bounds__(256) k_activate(const uint32_t *list, const unsigned int *count,
                                                  const uint64_t *ptrs, const uint32_t *nbrs,
                                                  const uint32_t *memb, const uint32_t *constraint,
                                                  uint8_t *active, const uint32_t *target = nullptr,
                                                  const uint32_t *claim = nullptr,
                                                  const uint8_t *admitted = nullptr,
                                                  const unsigned int *stop = nullptr) {
    constexpr int NG = 256 / G;
    if (stop && *stop) return;
    unsigned int cnt = *count;
    const int lane = threadIdx.x % G;
    for (unsigned int i = blockIdx.x * NG + threadIdx.x / G; i < cnt; i += gridDim.x * NG) {
        uint32_t v = list[i];
        if (claim && claim[target[v]] != v) continue;
        if (admitted && !admitted[v]) continue;  // not admitted: it did not move
        uint32_t mv = REFINE ? constraint[v] : memb[v];
        for (uint64_t e = ptrs[v] + lane; e < ptrs[v + 1]; e += G) {
            uint32_t u = nbrs[e];
            if (u == v) continue;
            if (REFINE ? (constraint[u] == mv) : (memb[u] != mv)) active[u] = 1;
        }
    }
}

// Refinement's re-activation, from the singletons' side (as the persistent sweep does it): an
// inactive vertex that is still a singleton becomes active when a neighbour in its working
// community moved in this round (stamped by k_apply_refine). One thread per vertex.
__global__ void __launch_bounds__(256) k_refine_reactivate(uint32_t n, const uint64_t *ptrs, const uint32_t *nbrs,
                                                           const uint32_t *memb, const uint32_t *csize,
                                                           const uint32_t *constraint, const uint32_t *mstep,
                                                           uint8_t *active, const uint32_t *iter_dev, uint32_t round,
                                                           const unsigned int *stop) {
    if (stop && *stop) return;
    const uint32_t r = iter_dev ? *iter_dev : round;
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
        if (active[u] || memb[u] != (uint32_t)u || csize[u] != 1u) continue;
        const uint32_t cu = constraint[u];
        const uint64_t e1 = ptrs[u + 1];
        for (uint64_t e = ptrs[u]; e < e1; e++) {
            const uint32_t w = nbrs[e];
            if (w != (uint32_t)u && mstep[w] == r && constraint[w] == cu) {
                active[u] = 1;
                break;
            }
        }
    }
}

// Community totals of a membership. A level that starts from earlier communities has few, large
// ones, whose same-address atomics would serialize: the workgroup first sums its vertices per
// community in an LDS table (integer sums, so any order gives the same totals) and adds each
// community's sums with one global atomic; a community without a slot within eight probes is
// added directly.
__global__ void __launch_bounds__(256) k_init_totals(const uint32_t *memb, const double *strength,
                                                     const double *node_w, uint32_t n, int sK, int sW, long long *Kfx,
                                                     long long *Wfx, uint32_t *csize) {
    constexpr uint32_t PT = 1024;
    __shared__ uint32_t tk[PT];
    __shared__ long long tK[PT], tW[PT];
    __shared__ uint32_t tN[PT];
    for (uint32_t i = threadIdx.x; i < PT; i += blockDim.x) {
        tk[i] = EMPTY_KEY;
        tK[i] = 0;
        tW[i] = 0;
        tN[i] = 0;
    }
    __syncthreads();
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint32_t c = memb[v];
        const long long kf = to_fx(strength[v], sK), wf = to_fx(node_w[v], sW);
        uint32_t h = hash32(c) & (PT - 1);
        bool done = false;
        for (int p = 0; p < 8 && !done; p++) {
            const uint32_t prev = atomicCAS(&tk[h], EMPTY_KEY, c);
            if (prev == EMPTY_KEY || prev == c) {
                atomicAdd((unsigned long long *)&tK[h], (unsigned long long)kf);
                atomicAdd((unsigned long long *)&tW[h], (unsigned long long)wf);
                atomicAdd(&tN[h], 1u);
                done = true;
            }
            h = (h + 1) & (PT - 1);
        }
        if (!done) {
            atomicAdd((unsigned long long *)&Kfx[c], (unsigned long long)kf);
            atomicAdd((unsigned long long *)&Wfx[c], (unsigned long long)wf);
            atomicAdd(&csize[c], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < PT; i += blockDim.x) {
        const uint32_t c = tk[i];
        if (c == EMPTY_KEY) continue;
        atomicAdd((unsigned long long *)&Kfx[c], (unsigned long long)tK[i]);
        atomicAdd((unsigned long long *)&Wfx[c], (unsigned long long)tW[i]);
        atomicAdd(&csize[c], tN[i]);
    }
}

__global__ void k_iota(uint32_t *p, uint32_t n) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) p[v] = v;
}

__global__ void k_fill_u8(uint8_t *p, uint64_t n, uint8_t x) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = x;
}

// Fills and counters of the multi-launch iterations are kernels (not memset nodes), so that an
// iteration can be captured into a graph and replayed.
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t x) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = x;
}

__global__ void k_iter_advance(uint32_t *it) { *it += 1u; }

static void fill_u32(uint32_t *p, uint64_t n, uint32_t x, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, n, x);
    KERNEL_OK();
}

// One iteration's launch sequence captured once (per level) and replayed per iteration.
// An executable graph kept across the levels of one run: a level's capture updates it in place
// (hipGraphExecUpdate) when its launch sequence has the same shape, so a run instantiates once
// per phase kind instead of once per level and phase.
struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    GraphSlot() = default;
    GraphSlot(const GraphSlot &) = delete;
    GraphSlot &operator=(const GraphSlot &) = delete;
    ~GraphSlot() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

struct IterGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool owned = true;  // exec belongs to a GraphSlot otherwise
    IterGraph() = default;
    IterGraph(const IterGraph &) = delete;
    IterGraph &operator=(const IterGraph &) = delete;
    template <typename F>
    void capture(hipStream_t s, F &&record, GraphSlot *slot = nullptr) {
        HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            record();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIP_OK(hipStreamEndCapture(s, &graph));
        if (slot) {
            owned = false;
            if (slot->exec) {
                hipGraphNode_t bad = nullptr;
                hipGraphExecUpdateResult res = hipGraphExecUpdateError;
                if (hipGraphExecUpdate(slot->exec, graph, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
                    exec = slot->exec;
                    return;
                }
                (void)hipGetLastError();  // a refused update is not an error of the run
                (void)hipGraphExecDestroy(slot->exec);
                slot->exec = nullptr;
            }
            HIP_OK(hipGraphInstantiate(&slot->exec, graph, nullptr, nullptr, 0));
            exec = slot->exec;
            return;
        }
        HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    void launch(hipStream_t s) { HIP_OK(hipGraphLaunch(exec, s)); }
    ~IterGraph() {
        if (exec && owned) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

// ---- degree buckets (bucket_of: common.hpp) ---------------------------------------------
__global__ void k_bucket_count(const uint64_t *ptrs, uint32_t n, BucketBounds bb, uint32_t *counts) {
    __shared__ uint32_t sc[NUM_BUCKETS];
    if (threadIdx.x < NUM_BUCKETS) sc[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        atomicAdd(&sc[bucket_of(ptrs[v + 1] - ptrs[v], bb)], 1u);
    __syncthreads();
    if (threadIdx.x < NUM_BUCKETS && sc[threadIdx.x]) atomicAdd(&counts[threadIdx.x], sc[threadIdx.x]);
}


// Block-local counting, one global reservation per bucket per block, then placement.
__global__ void k_bucket_fill(const uint64_t *ptrs, uint32_t n, BucketBounds bb, BucketStarts starts, uint32_t *cursor,
                              uint32_t *lists) {
    __shared__ uint32_t cnt[NUM_BUCKETS], base[NUM_BUCKETS];
    for (uint32_t v0 = blockIdx.x * blockDim.x; v0 < n; v0 += gridDim.x * blockDim.x) {
        if (threadIdx.x < NUM_BUCKETS) cnt[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t v = v0 + threadIdx.x;
        int b = 0;
        uint32_t local = 0;
        if (v < n) {
            b = bucket_of(ptrs[v + 1] - ptrs[v], bb);
            local = atomicAdd(&cnt[b], 1u);
        }
        __syncthreads();
        if (threadIdx.x < NUM_BUCKETS && cnt[threadIdx.x])
            base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], cnt[threadIdx.x]);
        __syncthreads();
        if (v < n) lists[starts.s[b] + base[b] + local] = v;
        __syncthreads();
    }
}

__global__ void k_gtab_sizes(const uint32_t *list, uint32_t len, const uint64_t *ptrs, uint64_t *sizes) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= len; i += gridDim.x * blockDim.x) {
        if (i == len) { sizes[i] = 0; continue; }
        uint32_t v = list[i];
        uint64_t d = ptrs[v + 1] - ptrs[v];
        uint64_t t = 1;
        while (t < 2 * d) t <<= 1;
        sizes[i] = t;
    }
}

struct Buckets {
    DevBuf<uint32_t> lists;
    uint32_t start[NUM_BUCKETS + 1] = {0};
    DevBuf<uint64_t> goff;
    DevBuf<uint32_t> gkeys;
    DevBuf<long long> gvals;
};

static bool default_bounds(const BucketBounds &bb) {
    for (int k = 0; k < NUM_BUCKETS - 1; k++)
        if (bb.hi[k] != DEFAULT_BUCKET_BOUNDS.hi[k]) return false;
    return true;
}

static void build_buckets(const sc_graph &L, Buckets &b, const BucketBounds &bb, hipStream_t s) {
    uint32_t n = (uint32_t)L.n;
    for (int k = 0; k < NUM_BUCKETS - 1; k++)
        if (bb.hi[k] > DEFAULT_BUCKET_BOUNDS.hi[k]) throw ScError(SC_INVALID_ARGUMENT, "degree class bound above its capacity");
    DevBuf<uint32_t> cnt(2 * NUM_BUCKETS + 2);
    HIP_OK(hipMemsetAsync(cnt.get(), 0, (2 * NUM_BUCKETS + 2) * sizeof(uint32_t), s));
    uint32_t h[NUM_BUCKETS];
    if (L.has_bucket_counts && default_bounds(bb)) {
        // counted while the graph was finished (read back with its other aggregates)
        for (int i = 0; i < NUM_BUCKETS; i++) h[i] = L.bucket_counts[i];
    } else {
        hipLaunchKernelGGL(k_bucket_count, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, L.ptrs.get(), n, bb, cnt.get());
        KERNEL_OK();
        to_host(h, cnt.get(), NUM_BUCKETS, s);
        sync_stream(s);
    }
    b.start[0] = 0;
    for (int i = 0; i < NUM_BUCKETS; i++) b.start[i + 1] = b.start[i] + h[i];
    BucketStarts starts;
    for (int i = 0; i <= NUM_BUCKETS; i++) starts.s[i] = b.start[i];
    b.lists.resize(n);
    hipLaunchKernelGGL(k_bucket_fill, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, L.ptrs.get(), n, bb, starts,
                       cnt.get() + NUM_BUCKETS, b.lists.get());
    KERNEL_OK();
    uint32_t nb6 = h[GTAB_BUCKET];
    if (nb6) {
        b.goff.resize(nb6 + 1);
        DevBuf<uint64_t> sizes(nb6 + 1);
        hipLaunchKernelGGL(k_gtab_sizes, dim3(grid_for(nb6 + 1, 256, 1024)), dim3(256), 0, s,
                           b.lists.get() + b.start[GTAB_BUCKET], nb6, L.ptrs.get(), sizes.get());
        KERNEL_OK();
        exclusive_scan_u64(sizes.get(), b.goff.get(), nb6 + 1, s);
        uint64_t total = 0;
        to_host(&total, b.goff.get() + nb6, 1, s);
        sync_stream(s);
        b.gkeys.resize(total);
        b.gvals.resize(total);
    }
}

// ---- persistent level sweeps -------------------------------------------------------------------
// One cooperative launch runs a whole local-moving (or refinement) phase of a level: every
// sub-sweep is compact -> decide -> apply (+ activate), separated by grid barriers, and the
// convergence test reads device counters, so nothing returns to the host between sub-sweeps.
constexpr int SWEEP_BS = 256;
constexpr int SWEEP_TAB = 4096;  // LDS hash entries per workgroup (the largest bucket layout)
constexpr int SWEEP_ACT_G = 16;  // lanes per mover when re-activating neighbours
// A barrier wait longer than this (ticks of the 100 MHz real-time clock) is reported as an error
// instead of spinning on; every workgroup then leaves the kernel.
constexpr long long SWEEP_TIMEOUT_TICKS = 400000000ll;

struct BarrierCtl;

struct SweepArgs {
    DecideArgs d;
    const uint32_t *lists;  // vertices grouped by degree bucket
    uint32_t bstart[NUM_BUCKETS + 1];
    uint32_t *work;         // compacted positions of one step, same bucket layout as lists
    unsigned int *wcnt;     // [max_steps][NUM_BUCKETS] compacted counts
    unsigned int *mcnt;     // [max_steps] deciders that chose to move
    unsigned int *ccnt;     // [max_steps] vertices that moved (refinement)
    uint32_t *mstep;        // step stamp of accepted movers
    // refinement's first step (every vertex a singleton): per vertex its own total of the
    // objective's kind (strength for Rb, node weight for CPM; fixed point) and its working
    // community, packed so one 16-byte load brings both; null under a weight cap
    uint4 *rpk;
    uint32_t mv_stride;     // each workgroup's mover list: d.movers + blockIdx.x * mv_stride
    uint32_t *mv_tgt;       // local moving with joiner admission: the movers' targets, same layout
    uint32_t *claim;        // lowest mover id per capped target, or null without a weight cap
    BarrierCtl *bc;         // grid barrier words (zeroed before the launch)
    unsigned int *ctl;      // [0] steps run, [2] iterations, [3] moves
    // writable views of d.memb, d.Kfx, d.Wfx, d.csize
    uint32_t *memb;
    long long *Kfx, *Wfx;
    uint32_t *csize;
    uint32_t max_iters;
    uint32_t max_steps;
    uint64_t level_seed;
    long long *trace;       // optional: real-time stamps after each barrier (workgroup 0)
    uint32_t trace_cap;
    // learned filter (local moving): packed BF16 weights (null = plain), skip when the logit is
    // below filter_thr. A skipped vertex leaves the active set like one that decided to stay,
    // with a mark; when the filtered iterations converge, the judge activates every vertex
    // (filter_judge 1) or the marked ones (2) and plain iterations run to convergence.
    // fstats: [0] skip events, [1] vertices the judge iteration decided, [2] of them moved
    const MlpPacked *filter_pk;
    float filter_thr;
    int filter_judge;             // 0 none, 1 full plain sweep, 2 the marked vertices only
    uint8_t *f_mark;              // 1 while a vertex's last evaluation was a skip
    unsigned long long *fstats;   // also [3] marked vertices at the judge, [4] of them moved
};

__device__ __forceinline__ unsigned int ld_acquire_ctr(const unsigned int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide barrier for co-resident workgroups (cooperative launch). Each XCD has its own L2:
// every wave first waits until its own stores (and atomics) have reached L2, then the last
// workgroup of each XCD to arrive writes that L2 back (one agent-scope release per XCD) before
// arriving at the root. After the release every workgroup performs its own agent-scope acquire
// (an invalidate issued by one workgroup does not make the others' loads fresh). A wait longer
// than the timeout sets the error flag and every workgroup returns false.
struct BarrierCtl {
    // separate 128-byte lines for the words that many workgroups touch
    unsigned int root, pad0[31];
    unsigned int release, pad1[31];
    unsigned int err, pad2[31];
    unsigned int joined, pad3[31];
    unsigned int xarrive[8][32];
    unsigned int xblocks[8];
};

struct BarrierState {
    unsigned int xcc, nb, nx, epoch;
};

__device__ __forceinline__ unsigned int xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 7u; }

__device__ __forceinline__ bool spin_until(const unsigned int *p, unsigned int v, unsigned int *err, long long t0) {
    while (ld_acquire_ctr(p) < v) {
        if (ld_acquire_ctr(err) != 0u) return false;
        if (wall_clock64() - t0 > SWEEP_TIMEOUT_TICKS) {
            atomicExch(err, 1u);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return true;
}

// Registers the workgroup with its XCD; all workgroups must call it once before any barrier.
__device__ __forceinline__ bool barrier_join(BarrierCtl *bc, BarrierState &bs, int *s_flag) {
    __shared__ unsigned int s_x, s_nb, s_nx;
    if (threadIdx.x == 0) {
        const unsigned int x = xcc_id();
        atomicAdd(&bc->xblocks[x], 1u);
        atomicAdd(&bc->joined, 1u);
        int bad = spin_until(&bc->joined, gridDim.x, &bc->err, wall_clock64()) ? 0 : 1;
        unsigned int nx = 0;
        for (int k = 0; k < 8; k++) nx += ld_acquire_ctr(&bc->xblocks[k]) ? 1u : 0u;
        s_x = x;
        s_nb = ld_acquire_ctr(&bc->xblocks[x]);
        s_nx = nx;
        *s_flag = bad;
    }
    __syncthreads();
    bs.xcc = s_x;
    bs.nb = s_nb;
    bs.nx = s_nx;
    bs.epoch = 0;
    return *s_flag == 0;
}

__device__ __forceinline__ bool grid_barrier(BarrierCtl *bc, BarrierState &bs, int *s_flag,
                                             long long *trace = nullptr, uint32_t trace_cap = 0) {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores and atomics are in L2
    __syncthreads();
    bs.epoch++;
    if (threadIdx.x == 0) {
        const unsigned int e = bs.epoch, x = bs.xcc;
        const long long t0 = wall_clock64();
        int bad = 0;
        const unsigned int old = atomicAdd(&bc->xarrive[x][0], 1u);
        const bool leader = old + 1 == bs.nb * e;
        if (leader) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            const unsigned int r = atomicAdd(&bc->root, 1u);
            if (r + 1 == bs.nx * e) atomicExch(&bc->release, e);
        }
        if (!spin_until(&bc->release, e, &bc->err, t0)) bad = 1;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __builtin_amdgcn_s_waitcnt(0x0F70);  // the invalidate has completed before the block reads
        *s_flag = bad;
        if (trace && blockIdx.x == 0 && e - 1 < trace_cap) trace[e - 1] = wall_clock64();
    }
    __syncthreads();
    return *s_flag == 0;
}

__device__ __forceinline__ int bucket_of_pos(const SweepArgs &A, uint32_t i) {
    int b = 0;
#pragma unroll
    for (int k = 1; k < NUM_BUCKETS; k++) b += i >= A.bstart[k] ? 1 : 0;
    return b;
}

// Positions (within their bucket) of the vertices that decide in this step, wave-aggregated
// appends into work[bstart[b] + ...].
// Workgroup b's share of bucket k's list positions, [lo, hi): every bucket is split evenly over
// the workgroups, so each compacts and decides its own share (no global counters, no barrier
// between compaction and decide).
__device__ __forceinline__ void bucket_share(const SweepArgs &A, int k, uint32_t &lo, uint32_t &hi) {
    const uint64_t L = (uint64_t)(A.bstart[k + 1] - A.bstart[k]);
    lo = A.bstart[k] + (uint32_t)(L * blockIdx.x / gridDim.x);
    hi = A.bstart[k] + (uint32_t)(L * (blockIdx.x + 1) / gridDim.x);
}

// The vertices of this workgroup's share that decide in this step (list positions, relative to
// their bucket, for the global-table bucket), stored at the start of the share's own range of
// work[]; s_wc[k] gets their count.
template <bool REFINE, bool FILTER, bool BATCH>
__device__ __forceinline__ void sweep_compact(const SweepArgs &A, uint64_t seed, uint32_t sub, uint32_t it,
                                              bool filter_on, unsigned int *s_wc) {
    const DecideArgs &a = A.d;
    const int lane = threadIdx.x & 63;
    unsigned long long skipped = 0;
    // the workgroup's shares of all buckets are walked as one range, so a thread's loads for
    // positions of different buckets are in flight together; appends stay per bucket
    uint32_t lo[NUM_BUCKETS], cum[NUM_BUCKETS + 1];
    cum[0] = 0;
#pragma unroll
    for (int k = 0; k < NUM_BUCKETS; k++) {
        uint32_t hi;
        bucket_share(A, k, lo[k], hi);
        cum[k + 1] = cum[k] + (hi - lo[k]);
    }
    const uint32_t total = cum[NUM_BUCKETS];
    // U positions per thread and pass: their list entries, then their vertices' flags (and, in
    // refinement, memberships and sizes) are loaded with independent loads before any is used
    constexpr int U = 4;
    for (uint32_t j0 = 0; j0 < total; j0 += SWEEP_BS * U) {  // workgroup-uniform bounds
        int kb[U];
        uint32_t ip[U], v[U];
        bool in[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const uint32_t j = j0 + (uint32_t)q * SWEEP_BS + threadIdx.x;
            // bucket of range position j: the last k with cum[k] <= j (empty buckets are passed over)
            kb[q] = 0;
            ip[q] = 0;
#pragma unroll
            for (int k = 0; k < NUM_BUCKETS; k++)
                if (j >= cum[k]) {
                    kb[q] = k;
                    ip[q] = lo[k] + (j - cum[k]);
                }
            in[q] = j < total;
            v[q] = in[q] ? A.lists[ip[q]] : 0u;
        }
        uint8_t f[U];
#pragma unroll
        for (int q = 0; q < U; q++) f[q] = in[q] ? a.active[v[q]] : (uint8_t)0;
        bool take[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            take[q] = false;
            if (f[q] == 1) {
                const VParams vp = vparams<BATCH>(a, v[q]);
                take[q] = !vp.frozen && subset_of(seed, v[q] - vp.lo, a.P) == sub;
            }
        }
        if (REFINE) {
            uint32_t m[U], cs[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                m[q] = take[q] ? a.memb[v[q]] : 0u;
                cs[q] = take[q] ? a.csize[v[q]] : 0u;
            }
#pragma unroll
            for (int q = 0; q < U; q++) take[q] = take[q] && m[q] == v[q] && cs[q] == 1u;
        }
#pragma unroll
        for (int q = 0; q < U; q++) {
            bool tq = take[q];
            if (FILTER && filter_on) {
                // learned filter: a vertex predicted to stay is not decided now; every lane of
                // the wave reaches the wave-collective inference
                float x[MLP_NF] = {};
                if (tq) vertex_features(a, vparams<BATCH>(a, v[q]), A.mstep, v[q], it, x);
                if (__any(tq)) {
                    const float z = mlp_forward_wave(A.filter_pk, x);
                    if (tq && z < A.filter_thr) {
                        // it leaves the active set like a vertex that decided to stay (a
                        // neighbour's move re-activates it), marked as skipped
                        tq = false;
                        skipped++;
                        a.active[v[q]] = A.filter_judge == 2 ? 2 : 0;
                        A.f_mark[v[q]] = 1;
                    } else if (tq) {
                        A.f_mark[v[q]] = 0;
                    }
                }
            }
            // one wave-aggregated append per bucket the wave's positions touch
#pragma unroll
            for (int k = 0; k < NUM_BUCKETS; k++) {
                const uint64_t mk = __ballot(tq && kb[q] == k);
                if (!mk) continue;
                const int leader = __ffsll((unsigned long long)mk) - 1;
                unsigned int basei = 0;
                if (lane == leader) basei = atomicAdd(&s_wc[k], (unsigned int)__popcll(mk));
                basei = __shfl(basei, leader);
                if (tq && kb[q] == k) {
                    const uint32_t below =
                        __builtin_amdgcn_mbcnt_hi((uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u));
                    // the vertex itself; the global-table bucket needs its list position
                    A.work[lo[k] + basei + below] = k == GTAB_BUCKET ? ip[q] - A.bstart[k] : v[q];
                }
            }
        }
    }
    if (FILTER && filter_on) {
        skipped = block_sum(skipped);
        if (threadIdx.x == 0 && skipped) atomicAdd(&A.fstats[0], skipped);
    }
    __threadfence_block();  // the work[] entries are read by the workgroup's other waves
    __syncthreads();
}

template <int B>
struct BucketShape;
// Narrow groups keep many vertices in flight per workgroup (each lane walks a few edges).
template <> struct BucketShape<0> { static constexpr int G = 4, T = 16; };
template <> struct BucketShape<1> { static constexpr int G = 4, T = 32; };
template <> struct BucketShape<2> { static constexpr int G = 8, T = 64; };
template <> struct BucketShape<3> { static constexpr int G = 16, T = 256; };
template <> struct BucketShape<4> { static constexpr int G = 64, T = 1024; };  // a wave per vertex
template <> struct BucketShape<5> { static constexpr int G = 256, T = 2048; };
template <> struct BucketShape<6> { static constexpr int G = 256, T = 4096; };
template <> struct BucketShape<7> { static constexpr int G = 256, T = 1; };

template <bool REFINE, bool BATCH>
__device__ __forceinline__ void sweep_decide(const SweepArgs &A, const StepParams &sp, const unsigned int *s_wc,
                                             uint32_t *s_keys, long long *s_vals, double *r_s, uint32_t *r_c,
                                             long long *r_w, DecideStats &ds) {
    static_assert((SWEEP_BS / 64) * WAVE_SLOTS <= SWEEP_TAB, "wave tables do not fit the sweep's LDS");
    // buckets 0-4 (groups of at most a wave): wave units of every bucket in one list, dealt to the
    // waves in turn, each wave with its own tables, so a step whose few deciders are spread over
    // the buckets does not run the buckets one after another
    if constexpr (!REFINE) {
        const uint32_t wave = threadIdx.x >> 6;
        uint32_t us[6];
        us[0] = 0;
#define SC_UNITS(BK) us[BK + 1] = us[BK] + (s_wc[BK] + (64 / BucketShape<BK>::G) - 1) / (64 / BucketShape<BK>::G);
        SC_UNITS(0)
        SC_UNITS(1)
        SC_UNITS(2)
        SC_UNITS(3)
        SC_UNITS(4)
#undef SC_UNITS
#define SC_WAVE_UNIT(BK)                                                                                      \
    {                                                                                                         \
        constexpr int G = BucketShape<BK>::G, T = BucketShape<BK>::T;                                         \
        uint32_t lo, hi;                                                                                      \
        bucket_share(A, BK, lo, hi);                                                                          \
        decide_unit<SWEEP_BS, G, T, REFINE, false, BATCH, true>(A.d, sp, A.lists + A.bstart[BK], A.work + lo, \
                                                               s_wc[BK], (uint64_t)(u - us[BK]) * (64 / G),   \
                                                               s_keys, s_vals, r_s, r_c, r_w, ds);           \
    }
        for (uint32_t u = wave; u < us[5]; u += SWEEP_BS / 64) {  // wave-uniform
            if (u < us[1]) SC_WAVE_UNIT(0)
            else if (u < us[2]) SC_WAVE_UNIT(1)
            else if (u < us[3]) SC_WAVE_UNIT(2)
            else if (u < us[4]) SC_WAVE_UNIT(3)
            else SC_WAVE_UNIT(4)
        }
#undef SC_WAVE_UNIT
    }
    // buckets 5-7 (a workgroup per vertex; refinement: every bucket): bucket by bucket after
    // every wave has left the previous one (the tables span the workgroup's LDS)
    bool first = REFINE;
#define SC_DECIDE_BUCKET(BK)                                                                          \
    {                                                                                                 \
        constexpr int G = BucketShape<BK>::G, T = BucketShape<BK>::T, NG = SWEEP_BS / G;               \
        static_assert(BK == GTAB_BUCKET || NG * T <= SWEEP_TAB, "bucket table does not fit");         \
        const uint32_t len = s_wc[BK];                                                                \
        if (len) {                                                                                    \
            if (!first) __syncthreads();                                                              \
            first = false;                                                                            \
            uint32_t lo, hi;                                                                          \
            bucket_share(A, BK, lo, hi);                                                              \
            for (uint64_t base = 0; base < len; base += NG)                                           \
                decide_unit<SWEEP_BS, G, T, REFINE, BK == GTAB_BUCKET, BATCH>(A.d, sp, A.lists + A.bstart[BK],  \
                                                                       A.work + lo,                   \
                                                             len, base, s_keys, s_vals, r_s, r_c, r_w, ds); \
        }                                                                                             \
    }
    if constexpr (REFINE) {
        SC_DECIDE_BUCKET(0)
        SC_DECIDE_BUCKET(1)
        SC_DECIDE_BUCKET(2)
        SC_DECIDE_BUCKET(3)
        SC_DECIDE_BUCKET(4)
    }
    SC_DECIDE_BUCKET(5)
    SC_DECIDE_BUCKET(6)
    SC_DECIDE_BUCKET(7)
#undef SC_DECIDE_BUCKET
}

// sweep_decide for refinement's first step (every vertex a singleton): decide_unit_single over
// the workgroup's share of each bucket, in the buckets' group shapes.
template <bool BATCH>
__device__ __forceinline__ void sweep_decide_single(const SweepArgs &A, const StepParams &sp, const unsigned int *s_wc,
                                                    double *r_s, uint32_t *r_c, long long *r_w, DecideStats &ds) {
#define SC_DECIDE_SINGLE(BK)                                                                                 \
    {                                                                                                        \
        constexpr int G = BucketShape<BK>::G, NG = SWEEP_BS / G;                                             \
        const uint32_t len = s_wc[BK];                                                                       \
        if (len) {                                                                                           \
            uint32_t lo, hi;                                                                                 \
            bucket_share(A, BK, lo, hi);                                                                     \
            for (uint64_t base = 0; base < len; base += NG)                                                  \
                decide_unit_single<SWEEP_BS, G, BK == GTAB_BUCKET, BATCH>(A.d, sp, A.lists + A.bstart[BK],    \
                                                                          A.work + lo, len, base, r_s, r_c,   \
                                                                          r_w, ds, A.rpk);                   \
        }                                                                                                    \
    }
    SC_DECIDE_SINGLE(0)
    SC_DECIDE_SINGLE(1)
    SC_DECIDE_SINGLE(2)
    SC_DECIDE_SINGLE(3)
    SC_DECIDE_SINGLE(4)
    SC_DECIDE_SINGLE(5)
    SC_DECIDE_SINGLE(6)
    SC_DECIDE_SINGLE(7)
#undef SC_DECIDE_SINGLE
}

// A workgroup's mover totals, summed per community in LDS before they go to global memory:
// late steps send many movers into the same few large communities, and same-address device
// atomics serialize. AGG_SLOTS keys with K, W and size sums (integers, so the totals do not
// depend on the order); a community that finds no slot within a few probes is updated directly.
constexpr uint32_t AGG_SLOTS = 1024;

struct AggTable {
    uint32_t *key;
    long long *K, *W, *N;
};

__device__ __forceinline__ void agg_clear(const AggTable &t, long long n_init = 0) {
    for (uint32_t i = threadIdx.x; i < AGG_SLOTS; i += blockDim.x) {
        t.key[i] = EMPTY_KEY;
        t.K[i] = 0;
        t.W[i] = 0;
        t.N[i] = n_init;
    }
}

__device__ __forceinline__ void agg_add(const SweepArgs &A, const AggTable &t, uint32_t c, long long kf, long long wf,
                                        int dn) {
    uint32_t h = hash32(c) & (AGG_SLOTS - 1);
    for (int p = 0; p < 8; p++) {
        const uint32_t prev = atomicCAS(&t.key[h], EMPTY_KEY, c);
        if (prev == EMPTY_KEY || prev == c) {
            if (kf) atomicAdd((unsigned long long *)&t.K[h], (unsigned long long)kf);
            if (wf) atomicAdd((unsigned long long *)&t.W[h], (unsigned long long)wf);
            atomicAdd((unsigned long long *)&t.N[h], (unsigned long long)(long long)dn);
            return;
        }
        h = (h + 1) & (AGG_SLOTS - 1);
    }
    if (kf) atomicAdd((unsigned long long *)&A.Kfx[c], (unsigned long long)kf);
    if (wf) atomicAdd((unsigned long long *)&A.Wfx[c], (unsigned long long)wf);
    atomicAdd(&A.csize[c], (unsigned int)dn);
}

// v moves from community `from` to `to`: its strength for Rb, node weight for CPM or under a
// weight cap (the only readers of either), size always
template <bool BATCH>
__device__ __forceinline__ void agg_move(const SweepArgs &A, const AggTable &t, uint64_t v, uint32_t from,
                                         uint32_t to) {
    const VParams vp = vparams<BATCH>(A.d, (uint32_t)v);
    const long long kf = A.d.objective == 0 ? to_fx(A.d.strength[v], vp.sK) : 0;
    const long long wf = (A.d.objective != 0 || A.d.has_max) ? to_fx(A.d.node_w[v], vp.sW) : 0;
    agg_add(A, t, to, kf, wf, 1);
    agg_add(A, t, from, -kf, -wf, -1);
}

// agg_move without the table: the totals of both communities by device atomics
template <bool BATCH>
__device__ __forceinline__ void direct_move(const SweepArgs &A, uint64_t v, uint32_t from, uint32_t to) {
    const VParams vp = vparams<BATCH>(A.d, (uint32_t)v);
    if (A.d.objective == 0) {
        const long long kf = to_fx(A.d.strength[v], vp.sK);
        atomicAdd((unsigned long long *)&A.Kfx[to], (unsigned long long)kf);
        atomicAdd((unsigned long long *)&A.Kfx[from], (unsigned long long)-kf);
    }
    if (A.d.objective != 0 || A.d.has_max) {
        const long long wf = to_fx(A.d.node_w[v], vp.sW);
        atomicAdd((unsigned long long *)&A.Wfx[to], (unsigned long long)wf);
        atomicAdd((unsigned long long *)&A.Wfx[from], (unsigned long long)-wf);
    }
    atomicAdd(&A.csize[to], 1u);
    atomicSub(&A.csize[from], 1u);
}

__device__ __forceinline__ void agg_flush(const SweepArgs &A, const AggTable &t) {
    for (uint32_t i = threadIdx.x; i < AGG_SLOTS; i += blockDim.x) {
        const uint32_t c = t.key[i];
        if (c == EMPTY_KEY) continue;
        if (t.K[i]) atomicAdd((unsigned long long *)&A.Kfx[c], (unsigned long long)t.K[i]);
        if (t.W[i]) atomicAdd((unsigned long long *)&A.Wfx[c], (unsigned long long)t.W[i]);
        if (t.N[i]) atomicAdd(&A.csize[c], (unsigned int)t.N[i]);
    }
}

__global__ void k_stamp_clock(long long *out) { *out = wall_clock64(); }

// After a sweep: its iterations, moves, decide statistics and barrier-timeout flag join the run's
// counters, and the barrier words, result words and statistics are zeroed for the next sweep (no
// memset between sweeps).
__global__ void k_sweep_collect(unsigned int *ctl, unsigned int *bc_words, uint32_t n_bc_words, uint32_t err_word,
                                unsigned long long *run, unsigned long long *run_err, unsigned long long *stat,
                                unsigned long long *run_stat) {
    if (threadIdx.x >= 64 && threadIdx.x < 67) {
        run_stat[threadIdx.x - 64] += stat[threadIdx.x - 64];
        stat[threadIdx.x - 64] = 0;
    }
    if (threadIdx.x == 0) {
        run[0] += ctl[2];
        run[1] += ctl[3];
        if (bc_words[err_word]) *run_err = 1ull;
    }
    __syncthreads();
    if (threadIdx.x < 4) ctl[threadIdx.x] = 0;
    for (uint32_t i = threadIdx.x; i < n_bc_words; i += blockDim.x) bc_words[i] = 0;
}

template <bool REFINE, bool FILTER, bool BATCH>
__global__ void __launch_bounds__(SWEEP_BS) __attribute__((amdgpu_waves_per_eu(3))) k_sweep(SweepArgs A) {
    __shared__ uint32_t s_keys[SWEEP_TAB];
    __shared__ long long s_vals[SWEEP_TAB];
    __shared__ double r_s[SWEEP_BS / 64];
    __shared__ uint32_t r_c[SWEEP_BS / 64];
    __shared__ long long r_w[SWEEP_BS / 64];
    __shared__ unsigned long long s_stats[3];
    __shared__ int s_flag;
    __shared__ unsigned int s_wc[NUM_BUCKETS];  // this step's compacted vertices per bucket
    __shared__ unsigned int s_mc;               // this step's movers of the workgroup
    const DecideArgs &a = A.d;
    const uint32_t n = a.n;
    const uint64_t gtid = (uint64_t)blockIdx.x * SWEEP_BS + threadIdx.x;
    const uint64_t gsize = (uint64_t)gridDim.x * SWEEP_BS;
    const bool capped = A.claim != nullptr;
    if (threadIdx.x < 3) s_stats[threadIdx.x] = 0;
    DecideStats ds;
    BarrierState bst;
    if (!barrier_join(A.bc, bst, &s_flag)) return;

    // prologue: state of the level (refinement starts from singletons)
    for (uint64_t v = gtid; v < n; v += gsize) {
        A.Kfx[v] = 0;
        A.Wfx[v] = 0;
        A.csize[v] = 0;
        a.active[v] = 1;
        if (REFINE) {
            A.memb[v] = (uint32_t)v;
            A.mstep[v] = NONE;
            if (A.rpk) {
                const VParams vp = vparams<BATCH>(a, (uint32_t)v);
                const long long val = a.objective == 0 ? to_fx(a.strength[v], vp.sK) : to_fx(a.node_w[v], vp.sW);
                A.rpk[v] = make_uint4((uint32_t)(unsigned long long)val, (uint32_t)((unsigned long long)val >> 32),
                                      a.constraint[v], 0u);
            }
            a.tround[v] = NONE;
            a.targeted_cnt[v] = 0;
        } else {
            // the vertex's private empty community n + v starts empty
            A.Kfx[n + v] = 0;
            A.Wfx[n + v] = 0;
            A.csize[n + v] = 0;
            if (capped) A.claim[n + v] = NONE;
            if (a.jsum) {
#pragma unroll
                for (uint64_t set = 0; set < 2; set++) {
                    a.jsum[set * a.jstride + v] = 0;
                    a.jsum[set * a.jstride + n + v] = 0;
                    a.jmin[set * a.jstride + v] = NONE;
                    a.jmin[set * a.jstride + n + v] = NONE;
                }
            }
            A.mstep[v] = NONE;
            if (a.f_margin) {
                a.f_margin[v] = 0.0f;
                a.f_own[v] = -1.0f;
                a.f_nmv[v] = 0u;
            }
            if (FILTER) A.f_mark[v] = 0;
        }
        if (capped) A.claim[v] = NONE;
    }
    for (uint64_t i = gtid; i < (uint64_t)A.max_steps * (NUM_BUCKETS + 2); i += gsize) {
        if (i < (uint64_t)A.max_steps * NUM_BUCKETS) A.wcnt[i] = 0;
        else if (i < (uint64_t)A.max_steps * (NUM_BUCKETS + 1)) A.mcnt[i - (uint64_t)A.max_steps * NUM_BUCKETS] = 0;
        else A.ccnt[i - (uint64_t)A.max_steps * (NUM_BUCKETS + 1)] = 0;
    }
    if (!grid_barrier(A.bc, bst, &s_flag, A.trace, A.trace_cap)) return;
    if (REFINE) {
        // singletons: every community gets exactly one vertex, no two atomics share an address
        for (uint64_t v = gtid; v < n; v += gsize) {
            const VParams vp = vparams<BATCH>(a, (uint32_t)v);
            atomicAdd((unsigned long long *)&A.Kfx[v], (unsigned long long)to_fx(a.strength[v], vp.sK));
            atomicAdd((unsigned long long *)&A.Wfx[v], (unsigned long long)to_fx(a.node_w[v], vp.sW));
            atomicAdd(&A.csize[v], 1u);
        }
    } else {
        // a level that starts from earlier communities has few, large ones: the workgroup first
        // sums its vertices per community in an LDS table (integer sums, so any order gives the
        // same totals), then adds each community's sums with one global atomic per array; a
        // vertex whose community finds no slot within a few probes is added directly
        constexpr uint32_t PT = 1024;
        static_assert(3 * PT <= SWEEP_TAB, "prologue table does not fit the sweep's LDS");
        static_assert(3 * AGG_SLOTS <= SWEEP_TAB, "mover table does not fit the sweep's LDS");
        uint32_t *tk = s_keys;
        long long *tK = s_vals, *tW = s_vals + PT, *tN = s_vals + 2 * PT;
        for (uint32_t i = threadIdx.x; i < PT; i += SWEEP_BS) {
            tk[i] = EMPTY_KEY;
            tK[i] = 0;
            tW[i] = 0;
            tN[i] = 0;
        }
        __syncthreads();
        for (uint64_t v = gtid; v < n; v += gsize) {
            const uint32_t c = a.memb[v];
            const VParams vp = vparams<BATCH>(a, (uint32_t)v);
            const long long kf = to_fx(a.strength[v], vp.sK), wf = to_fx(a.node_w[v], vp.sW);
            uint32_t h = hash32(c) & (PT - 1);
            bool done = false;
            for (int p = 0; p < 8 && !done; p++) {
                const uint32_t prev = atomicCAS(&tk[h], EMPTY_KEY, c);
                if (prev == EMPTY_KEY || prev == c) {
                    atomicAdd((unsigned long long *)&tK[h], (unsigned long long)kf);
                    atomicAdd((unsigned long long *)&tW[h], (unsigned long long)wf);
                    atomicAdd((unsigned long long *)&tN[h], 1ull);
                    done = true;
                }
                h = (h + 1) & (PT - 1);
            }
            if (!done) {
                atomicAdd((unsigned long long *)&A.Kfx[c], (unsigned long long)kf);
                atomicAdd((unsigned long long *)&A.Wfx[c], (unsigned long long)wf);
                atomicAdd(&A.csize[c], 1u);
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < PT; i += SWEEP_BS) {
            const uint32_t c = tk[i];
            if (c == EMPTY_KEY) continue;
            atomicAdd((unsigned long long *)&A.Kfx[c], (unsigned long long)tK[i]);
            atomicAdd((unsigned long long *)&A.Wfx[c], (unsigned long long)tW[i]);
            atomicAdd(&A.csize[c], (uint32_t)tN[i]);
        }
    }
    if (!grid_barrier(A.bc, bst, &s_flag, A.trace, A.trace_cap)) return;

    const int lane = threadIdx.x % SWEEP_ACT_G;
    const uint32_t gslot = threadIdx.x / SWEEP_ACT_G;
    constexpr uint32_t NGA = SWEEP_BS / SWEEP_ACT_G;
    uint32_t *const mv_list = a.movers + (uint64_t)blockIdx.x * A.mv_stride;  // this workgroup's movers
    uint32_t *const mv_tgt = A.mv_tgt ? A.mv_tgt + (uint64_t)blockIdx.x * A.mv_stride : nullptr;
    if (threadIdx.x == 0) s_mc = 0;
    uint32_t step = 0, it = 0;
    bool ok = true;
    // with a learned filter, local moving runs filtered to convergence; with a judge, plain
    // iterations follow (the first plain iteration's moves are the filter's errors)
    bool filter_on = FILTER;
    uint32_t judge_it = 0xFFFFFFFFu;
    unsigned long long total_moves = 0, judge_dec = 0, judge_mv = 0;  // thread 0's tallies
    for (; it < A.max_iters && ok; it++) {
        const uint64_t seed = REFINE ? a.seed : mix64(A.level_seed ^ (0x5851F42D4C957F2Dull * (uint64_t)(it + 1)));
        const uint32_t P = REFINE ? 1u : a.P;
        for (uint32_t sub = 0; sub < P && ok; sub++, step++) {
            // joiner admission alternates two sets of sums by step: the set the previous step used
            // is reset here from this workgroup's movers of that step; this step writes the other
            // set, and the next step's decide follows the barriers of this one
            if (!REFINE && mv_tgt && step > 0) {
                const unsigned int pm = s_mc;
                const uint64_t off = (uint64_t)((step - 1) & 1u) * a.jstride;
                for (uint32_t i = threadIdx.x; i < pm; i += SWEEP_BS) {
                    const uint32_t t = mv_tgt[i];
                    a.jsum[off + t] = 0;
                    a.jmin[off + t] = NONE;
                }
            }
            // clear the claims and join counts this workgroup's movers of the previous step
            // left behind; every workgroup finishes that before anyone decides again
            if (step > 0 && (capped || REFINE)) {
                const unsigned int pm = s_mc;
                for (uint32_t i = threadIdx.x; i < pm; i += SWEEP_BS) {
                    const uint32_t t = a.target[mv_list[i]];
                    if (capped) A.claim[t] = NONE;
                    if (REFINE) a.targeted_cnt[t] = 0;
                }
                if (!(ok = grid_barrier(A.bc, bst, &s_flag, A.trace, A.trace_cap))) break;
            }
            __syncthreads();
            if (threadIdx.x < NUM_BUCKETS) s_wc[threadIdx.x] = 0;
            if (threadIdx.x == 0) s_mc = 0;
            __syncthreads();
            const long long t_dec = A.trace ? wall_clock64() : 0;
            // compact and decide the workgroup's own share of every bucket
            sweep_compact<REFINE, FILTER, BATCH>(A, seed, sub, it, filter_on, s_wc);
            StepParams sp{seed, sub, it, step, &s_mc, REFINE ? nullptr : A.mstep, mv_list, REFINE ? nullptr : mv_tgt};
            if (REFINE && step == 0) sweep_decide_single<BATCH>(A, sp, s_wc, r_s, r_c, r_w, ds);
            else sweep_decide<REFINE, BATCH>(A, sp, s_wc, s_keys, s_vals, r_s, r_c, r_w, ds);
            __syncthreads();
            const unsigned int nm = s_mc;
            if (!REFINE && mv_tgt) {
                // joiner admission sums: the workgroup's joiners of each target are summed (and
                // their lowest id taken) in LDS, then added with one atomic per target and array
                const AggTable agg{s_keys, s_vals, s_vals + AGG_SLOTS, s_vals + 2 * AGG_SLOTS};
                agg_clear(agg, (long long)NONE);  // N: the lowest joiner id
                __syncthreads();
                const uint64_t off = (uint64_t)(step & 1u) * a.jstride;
                for (uint32_t i = threadIdx.x; i < nm; i += SWEEP_BS) {
                    const uint32_t v = mv_list[i], t = mv_tgt[i];
                    const VParams vp = vparams<BATCH>(a, v);
                    const long long amt = join_amount(a, vp, a.strength[v], a.node_w[v]);
                    uint32_t h = hash32(t) & (AGG_SLOTS - 1);
                    bool done = false;
                    for (int p = 0; p < 8 && !done; p++) {
                        const uint32_t prev = atomicCAS(&agg.key[h], EMPTY_KEY, t);
                        if (prev == EMPTY_KEY || prev == t) {
                            atomicAdd((unsigned long long *)&agg.K[h], (unsigned long long)amt);
                            atomicMin((unsigned long long *)&agg.N[h], (unsigned long long)v);
                            done = true;
                        }
                        h = (h + 1) & (AGG_SLOTS - 1);
                    }
                    if (!done) {
                        atomicAdd((unsigned long long *)&a.jsum[off + t], (unsigned long long)amt);
                        atomicMin(&a.jmin[off + t], v);
                    }
                }
                __syncthreads();
                for (uint32_t i = threadIdx.x; i < AGG_SLOTS; i += SWEEP_BS) {
                    const uint32_t t = agg.key[i];
                    if (t == EMPTY_KEY) continue;
                    atomicAdd((unsigned long long *)&a.jsum[off + t], (unsigned long long)agg.K[i]);
                    atomicMin(&a.jmin[off + t], (uint32_t)agg.N[i]);
                }
            }
            if (threadIdx.x == 0) {
                if (A.trace && step < A.max_steps) {
                    // traced runs: deciders and movers of every step, and the largest and summed
                    // compact + decide time of the workgroups (clock ticks)
                    unsigned int dec = 0;
                    for (int b = 0; b < NUM_BUCKETS; b++) dec += s_wc[b];
                    if (dec) atomicAdd(&A.wcnt[(uint64_t)step * NUM_BUCKETS + 0], dec);
                    if (nm) atomicAdd(&A.wcnt[(uint64_t)step * NUM_BUCKETS + 1], nm);
                    const unsigned int dt = (unsigned int)min(wall_clock64() - t_dec, 0x7FFFFFFFll);
                    atomicMax(&A.wcnt[(uint64_t)step * NUM_BUCKETS + 2], dt);
                    atomicAdd(&A.wcnt[(uint64_t)step * NUM_BUCKETS + 3], dt);
                }
                if (!REFINE) {
                    if (nm) A.mcnt[step] = 1u;  // someone chose to move: not converged
                    total_moves += nm;
                }
                if (!REFINE && it == judge_it) {
                    unsigned long long dec = 0;
                    for (int b = 0; b < NUM_BUCKETS; b++) dec += s_wc[b];
                    judge_dec += dec;
                    judge_mv += nm;
                }
            }
            if (!(ok = grid_barrier(A.bc, bst, &s_flag, A.trace, A.trace_cap))) break;

            if (capped) {
                for (uint32_t i = threadIdx.x; i < nm; i += SWEEP_BS) {
                    const uint32_t v = mv_list[i];
                    atomicMin(&A.claim[a.target[v]], v);
                }
                if (!(ok = grid_barrier(A.bc, bst, &s_flag, A.trace, A.trace_cap))) break;
            }
            // apply this workgroup's movers, one lane each; refinement then re-activates in the
            // same phase (it reads the working communities, which do not change here), local
            // moving after all moves landed (below), like the multi-launch path
            auto admitted = [&](uint64_t v, uint64_t t) -> bool {
                if (REFINE) {
                    const uint32_t tc = a.targeted_cnt[v];
                    const uint32_t tr = a.tround[t];
                    const uint64_t tt = a.target[t];
                    const uint32_t cl = capped ? A.claim[t] : (uint32_t)v;
                    return ((tc == 0u) | ((tc == 1u) & (t < v) & (tr == it) & (tt == v))) & (cl == (uint32_t)v);
                }
                if (capped) return A.claim[t] == (uint32_t)v;
                return !a.jsum || join_admitted(a, vparams<BATCH>(a, (uint32_t)v), (uint32_t)v, (uint32_t)t, step & 1u);
            };
            unsigned int applied = 0;
            // more movers than table slots (early steps, whose targets are mostly distinct): the
            // table would only add probes, so the totals take direct atomics (integer sums: the
            // same totals either way)
            const bool direct = nm > 2 * AGG_SLOTS;
            const AggTable agg{s_keys, s_vals, s_vals + AGG_SLOTS, s_vals + 2 * AGG_SLOTS};
            if (!direct) agg_clear(agg);  // the decide tables are free: every wave has left the decide phase
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < nm; i += SWEEP_BS) {
                const uint64_t v = mv_list[i];
                const uint64_t t = a.target[v];
                if (!admitted(v, t)) {
                    a.active[v] = 1;  // blocked: look again next step
                } else {
                    const uint64_t from = REFINE ? v : a.memb[v];
                    A.memb[v] = (uint32_t)t;
                    if (REFINE) A.mstep[v] = step;  // read by the re-activation scan below
                    if (direct) direct_move<BATCH>(A, v, (uint32_t)from, (uint32_t)t);
                    else agg_move<BATCH>(A, agg, v, (uint32_t)from, (uint32_t)t);
                    applied++;
                    if (FILTER && it == judge_it && A.f_mark[v]) atomicAdd(&A.fstats[4], 1ull);
                }
            }
            __syncthreads();
            if (!direct) agg_flush(A, agg);
            if (REFINE) {
                const unsigned long long moved_here = block_sum((unsigned long long)applied);
                if (threadIdx.x == 0) {
                    if (moved_here) A.ccnt[step] = 1u;  // someone moved: not converged
                    total_moves += moved_here;
                }
            }
            if (!(ok = grid_barrier(A.bc, bst, &s_flag, A.trace, A.trace_cap))) break;
            if (REFINE && ld_acquire_ctr(A.ccnt + step) != 0u) {
                // refinement re-activates the singletons that have a neighbour in their working
                // community that moved in this step (refine.rs: only singletons move); gathered
                // from the singletons' side: after the first steps few singletons are left, and a
                // vertex that is not a singleton never becomes one again. Each lane takes one
                // vertex; the step's next barrier orders these flags before compaction reads them.
                for (uint64_t u = gtid; u < n; u += gsize) {
                    if (a.active[u] || A.memb[u] != (uint32_t)u || A.csize[u] != 1u) continue;
                    const uint32_t cu = a.constraint[u];
                    const uint64_t e1 = a.ptrs[u + 1];
                    for (uint64_t e = a.ptrs[u]; e < e1; e++) {
                        const uint32_t w = a.nbrs[e];
                        if (w != (uint32_t)u && A.mstep[w] == step && a.constraint[w] == cu) {
                            a.active[u] = 1;
                            break;
                        }
                    }
                }
            }
            if (!REFINE) {
                // local moving: after all moves landed, a mover that moved re-activates its
                // neighbours outside its new community (local_move.rs:114-118)
                for (uint32_t i = gslot; i < nm; i += NGA) {
                    const uint64_t v = mv_list[i];
                    const uint32_t t = a.target[v];
                    // a mover that landed is in its target now; one that was not admitted (or lost
                    // its claim) is still in its old community, which differs from its target
                    if (a.memb[v] != t) continue;
                    for (uint64_t e = a.ptrs[v] + lane; e < a.ptrs[v + 1]; e += SWEEP_ACT_G) {
                        const uint32_t u = a.nbrs[e];
                        if (u == v) continue;
                        if (a.f_nmv) atomicAdd(&a.f_nmv[u], 1u);
                        if (a.memb[u] != t) a.active[u] = 1;
                    }
                }
                if (!(ok = grid_barrier(A.bc, bst, &s_flag, A.trace, A.trace_cap))) break;
            }
        }
        if (!ok) break;
        // convergence, read identically by every workgroup after the barrier: the step flags
        // say whether anyone chose to move (local moving) or moved (refinement)
        unsigned int moved = 0;
        if (REFINE) {
            moved = ld_acquire_ctr(A.ccnt + step - 1);
        } else {
            for (uint32_t s2 = step - P; s2 < step; s2++) moved |= ld_acquire_ctr(A.mcnt + s2);
        }
        if (!REFINE && it == judge_it && threadIdx.x == 0) {
            atomicAdd(&A.fstats[1], judge_dec);
            atomicAdd(&A.fstats[2], judge_mv);
        }
        if (moved == 0) {
            if (filter_on && A.filter_judge) {
                // the judge: plain iterations from every vertex (a full sweep) or from the
                // vertices whose last evaluation was a skip, run to convergence
                filter_on = false;
                judge_it = it + 1;
                unsigned long long marked = 0;
                for (uint64_t v = gtid; v < n; v += gsize) {
                    marked += A.f_mark[v];
                    if (A.filter_judge == 1 || a.active[v] == 2) a.active[v] = 1;
                }
                marked = block_sum(marked);
                if (threadIdx.x == 0 && marked) atomicAdd(&A.fstats[3], marked);
                if (!(ok = grid_barrier(A.bc, bst, &s_flag, A.trace, A.trace_cap))) break;
                continue;
            }
            it++;
            break;
        }
    }
    flush_decide_stats(ds, s_stats, a.stats);
    if (threadIdx.x == 0 && total_moves) atomicAdd(&A.ctl[3], (unsigned int)total_moves);
    if (gtid == 0) {
        A.ctl[0] = step;
        A.ctl[2] = it;
    }
}

// ---- profiling ---------------------------------------------------------------------------
struct KernelTimer {
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    size_t used = 0;
    double total_ms = 0.0;
    uint64_t launches = 0;
    bool on = false;
    hipEvent_t next() {
        if (used == pool.size()) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(hipStream_t s, hipEvent_t &a) {
        if (!on) return;
        a = next();
        HIP_OK(hipEventRecord(a, s));
    }
    void end(hipStream_t s, hipEvent_t a) {
        if (!on) return;
        hipEvent_t b = next();
        HIP_OK(hipEventRecord(b, s));
        pending.emplace_back(a, b);
        launches++;
    }
    // call after a stream synchronize
    void harvest() {
        for (auto &p : pending) {
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, p.first, p.second));
            total_ms += ms;
        }
        pending.clear();
        used = 0;
    }
    ~KernelTimer() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// ---- level state -----------------------------------------------------------------------------
struct LevelState {
    DevBuf<long long> Kfx, Wfx;
    DevBuf<uint32_t> csize;
    DevBuf<uint8_t> active;
    DevBuf<uint32_t> target, tround, movers, moved, targeted, claim;
    DevBuf<unsigned int> counters;           // [0] movers, [1] moved
    DevBuf<unsigned long long> stat;         // vertices, edges, candidates
    // learned filter / trace runs only: decision features and skip marks (see DecideArgs)
    DevBuf<float> f_margin, f_own;
    DevBuf<uint32_t> f_nmv;
    DevBuf<uint8_t> f_mark;
    // persistent sweeps: compacted step lists, mover stamps, step counters + barrier + control
    DevBuf<uint32_t> work, mstep;
    DevBuf<uint4> rpk;
    // joiner admission (local moving without a cap): sums and lowest ids per target (two sets),
    // the movers' snapshot gain margins, their targets beside the persistent mover lists, and
    // (multi-launch) whether each mover was admitted
    DevBuf<long long> jsum;
    DevBuf<uint32_t> jmin, mv_tgt;
    DevBuf<double> mgain;
    DevBuf<uint8_t> admitted;
    DevBuf<unsigned int> sweep;
    DevBuf<unsigned char> sweep_ctl;
};

struct BatchState;

// Counters a run accumulates on the device and reads back once, at its end (no readback per
// phase): per phase (0 local moving, 1 refinement) the decide statistics [vertices, edges,
// candidates] and the sweep results [iterations, moves]; then a sweep barrier-timeout flag.
constexpr int RUN_STAT = 0, RUN_SWEEP = 6, RUN_ERR = 10, RUN_COUNTERS = 11;

struct EventPool {
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    hipEvent_t next() {
        if (used == pool.size()) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    ~EventPool() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// A tuner measurement whose events are read when the run ends.
struct TuneSample {
    std::string key;
    int index;
    hipEvent_t t0, t1;
};

struct RunCtx {
    const sc_leiden_config *cfg;
    hipStream_t s;
    KernelTimer timer;
    sc_leiden_stats st{};
    double move_bytes = 0.0;
    DevBuf<unsigned long long> counters;                 // RUN_COUNTERS device counters
    EventPool events;                                    // phase marks and tuner samples
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> phase_marks;
    std::vector<TuneSample> tune_samples;
    // batched run (null otherwise): the jobs and, for the current level, every vertex's job
    BatchState *batch = nullptr;
    const uint32_t *vjob = nullptr;
    const JobParams *jobs = nullptr;
    // learned filter / trace collection (device pointers; null when off) and the level index
    const struct FilterOpts *fopt = nullptr;
    uint32_t level = 0;
    // graph-replayed sweeps: one executable graph per phase kind, updated level by level
    GraphSlot move_graph, refine_graph;
};

// Learned-filter inference or trace collection for local moving (device pointers).
struct FilterOpts {
    const MlpPacked *pk = nullptr;        // packed BF16 weights; null: no filtering
    int judge = 1;                        // 0 none, 1 full plain sweep, 2 skipped vertices only
    float logit_thr = 0.0f;               // skip when the logit is below this
    unsigned long long *fstats = nullptr; // [skip events, judged, judge moves]
    float *trace_x = nullptr;             // trace runs: samples (features, moved, order key)
    uint8_t *trace_y = nullptr;
    uint64_t *trace_key = nullptr;
    unsigned long long *trace_n = nullptr;
    uint64_t trace_cap = 0;
};

template <int BS, int G, int T, bool REFINE, bool GTAB>
static void launch_decide(const DecideArgs &a, hipStream_t s) {
    constexpr int NG = BS / G;
    // over a compacted list (count on the device) the grid covers the bucket but stays within
    // about two waves of workgroups, so a small step's launch is cheap
    unsigned grid = grid_for(a.list_len, NG, a.sel_len ? 2048u : 8192u);
    hipLaunchKernelGGL((k_decide<BS, G, T, REFINE, GTAB>), dim3(grid), dim3(BS), 0, s, a);
    KERNEL_OK();
}

template <bool REFINE>
static void decide_all(DecideArgs a, const Buckets &b, hipStream_t s, KernelTimer *timer,
                       const uint32_t *clist = nullptr, const unsigned int *ccount = nullptr,
                       bool every_bucket = false) {
    // every_bucket: empty buckets launch too (one workgroup that finds nothing), so a captured
    // step has the same launch sequence on every level
    for (int bk = 0; bk < NUM_BUCKETS; bk++) {
        uint32_t len = b.start[bk + 1] - b.start[bk];
        if (!len && !every_bucket) continue;
        a.list = b.lists.get() + b.start[bk];
        a.list_len = len;
        if (clist) {
            a.sel = clist + b.start[bk];
            a.sel_len = ccount + bk;
        }
        hipEvent_t ev = nullptr;
        if (timer) timer->begin(s, ev);
        switch (bk) {
            // the persistent sweep's group shapes (narrow groups keep many vertices in flight)
            case 0: launch_decide<256, BucketShape<0>::G, BucketShape<0>::T, REFINE, false>(a, s); break;
            case 1: launch_decide<256, BucketShape<1>::G, BucketShape<1>::T, REFINE, false>(a, s); break;
            case 2: launch_decide<256, BucketShape<2>::G, BucketShape<2>::T, REFINE, false>(a, s); break;
            case 3: launch_decide<256, BucketShape<3>::G, BucketShape<3>::T, REFINE, false>(a, s); break;
            case 4: launch_decide<256, BucketShape<4>::G, BucketShape<4>::T, REFINE, false>(a, s); break;
            case 5: launch_decide<256, BucketShape<5>::G, BucketShape<5>::T, REFINE, false>(a, s); break;
            case 6: launch_decide<256, BucketShape<6>::G, BucketShape<6>::T, REFINE, false>(a, s); break;
            default:
                a.gkeys = b.gkeys.get();
                a.gvals = b.gvals.get();
                a.goff = b.goff.get();
                launch_decide<256, 256, 1, REFINE, true>(a, s);
                break;
        }
        if (timer) timer->end(s, ev);
    }
}

static DecideArgs base_args(const sc_graph &L, LevelState &ls, const sc_leiden_config &cfg, int sK, int sW,
                            int sw) {
    DecideArgs a{};
    a.n = (uint32_t)L.n;
    a.ptrs = L.ptrs.get();
    a.nbrs = L.nbrs.get();
    a.wts = L.wts.get();
    a.strength = L.strength.get();
    a.node_w = L.node_w.get();
    a.self_loop = L.self_loop.get();
    a.Kfx = ls.Kfx.get();
    a.Wfx = ls.Wfx.get();
    a.csize = ls.csize.get();
    a.active = ls.active.get();
    a.target = ls.target.get();
    a.tround = ls.tround.get();
    a.movers = ls.movers.get();
    a.n_movers = ls.counters.get();
    a.targeted_cnt = ls.targeted.get();
    a.objective = cfg.objective;
    a.gamma = cfg.resolution;
    a.four_m = 4.0 * L.total_weight;
    a.threshold = 1e-12 * std::max(std::fabs(L.total_weight), 1.0);
    a.sK = sK;
    a.sW = sW;
    a.sw = sw;
    a.has_max = cfg.has_max_community_weight;
    a.max_w = cfg.max_community_weight;
    a.P = 1;
    a.stats = ls.stat.get();
    return a;
}

// Community-indexed arrays span [0, 2n): ids below n are vertex ids (a vertex's own id is its
// first empty community), n + v is vertex v's private empty community.
static void alloc_level(LevelState &ls, uint32_t n, hipStream_t s) {
    ls.Kfx.resize(2 * (size_t)n);
    ls.Wfx.resize(2 * (size_t)n);
    ls.csize.resize(2 * (size_t)n);
    ls.active.resize(n);
    ls.target.resize(n);
    ls.tround.resize(n);
    ls.movers.resize(n);
    ls.moved.resize(n);
    ls.targeted.resize(n);
    ls.claim.resize(2 * (size_t)n);
    ls.counters.resize(4);
    ls.stat.resize(4);
    HIP_OK(hipMemsetAsync(ls.stat.get(), 0, 4 * sizeof(unsigned long long), s));
    ls.work.resize(n);
    ls.mstep.resize(n);
    ls.rpk.resize(n);
    ls.sweep.resize((size_t)MAX_MOVE_ITERS * MOVE_SUBSWEEPS * (NUM_BUCKETS + 2) + 8);
    ls.sweep_ctl.resize(sizeof(BarrierCtl));
    // sweep counters and barrier words start at zero; each sweep leaves them zero again
    HIP_OK(hipMemsetAsync(ls.sweep.get(), 0, ls.sweep.bytes(), s));
    HIP_OK(hipMemsetAsync(ls.sweep_ctl.get(), 0, ls.sweep_ctl.bytes(), s));
}

// Sweep mode: 1 = one cooperative launch per phase (default), 0 = one launch per sub-sweep
// kernel, 2 = mode 0's launches of one iteration captured into a graph once per level and
// replayed per iteration. SCGPU_SWEEP_MODE=0/1/2 in the environment selects the mode at
// start-up (SCGPU_PERSISTENT=0 selects 0).
static std::atomic<int> g_sweep_mode{-1};

static int sweep_mode() {
    int m = g_sweep_mode.load(std::memory_order_relaxed);
    if (m < 0) {
        const char *e = std::getenv("SCGPU_SWEEP_MODE");
        const char *p = std::getenv("SCGPU_PERSISTENT");
        m = (e && e[0] >= '0' && e[0] <= '2' && !e[1]) ? e[0] - '0' : ((p && p[0] == '0') ? 0 : 1);
        g_sweep_mode.store(m, std::memory_order_relaxed);
    }
    return m;
}

static bool persistent_sweeps() { return sweep_mode() == 1; }
static bool graph_sweeps() { return sweep_mode() == 2; }

// Workgroups of a sweep launch: all must be resident at once (cooperative launch); small
// levels use fewer, which makes each grid barrier cheaper.
// Workgroups of a sweep launch that can be resident at once.
template <bool REFINE, bool FILTER = false, bool BATCH = false>
static uint32_t sweep_resident() {
    static thread_local int resident[64] = {0};
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    int &r = resident[dev & 63];
    if (!r) {
        int per = 0, cus = 0;
        HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_sweep<REFINE, FILTER, BATCH>, SWEEP_BS, 0));
        HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        r = std::max(1, per) * std::max(1, cus);
    }
    return (uint32_t)r;
}

template <bool REFINE, bool FILTER = false, bool BATCH = false>
static uint32_t sweep_grid(uint32_t n, uint64_t nnz) {
    const uint32_t r = sweep_resident<REFINE, FILTER, BATCH>();
    static const long long forced = [] {
        const char *e = std::getenv("SCGPU_SWEEP_BLOCKS");
        return e ? std::atoll(e) : 0ll;
    }();
    const uint64_t by_work = std::max<uint64_t>(((uint64_t)n + 63) / 64, nnz / 2048 + 1);
    const uint32_t want = forced > 0 ? (uint32_t)forced : (uint32_t)std::min<uint64_t>(std::max<uint64_t>(8, by_work), 1u << 20);
    return std::min<uint32_t>((uint32_t)r, want);
}

static bool sweep_trace_on() {
    static const bool on = [] {
        const char *e = std::getenv("SCGPU_SWEEP_TRACE");
        return e && e[0] == '1';
    }();
    return on;
}

// One cooperative launch for the whole phase; returns {steps, iterations, moves}.
template <bool REFINE>
static void run_sweep(const DecideArgs &d, const Buckets &b, LevelState &ls, RunCtx &ctx, uint64_t level_seed,
                      uint32_t *memb_w, uint64_t nnz) {
    hipStream_t s = ctx.s;
    const uint32_t n = d.n;
    SweepArgs A{};
    A.d = d;
    A.lists = b.lists.get();
    for (int i = 0; i <= NUM_BUCKETS; i++) A.bstart[i] = b.start[i];
    A.work = ls.work.get();
    A.max_iters = REFINE ? MAX_REFINE_ROUNDS : MAX_MOVE_ITERS;
    A.max_steps = REFINE ? MAX_REFINE_ROUNDS : MAX_MOVE_ITERS * MOVE_SUBSWEEPS;
    unsigned int *base = ls.sweep.get();
    A.wcnt = base;
    A.mcnt = base + (size_t)A.max_steps * NUM_BUCKETS;
    A.ccnt = A.mcnt + A.max_steps;
    A.ctl = A.ccnt + A.max_steps;
    A.bc = reinterpret_cast<BarrierCtl *>(ls.sweep_ctl.get());
    A.mstep = ls.mstep.get();
    A.rpk = ctx.cfg->has_max_community_weight ? nullptr : ls.rpk.get();
    A.claim = ctx.cfg->has_max_community_weight ? ls.claim.get() : nullptr;
    A.level_seed = level_seed;
    A.memb = memb_w;
    const bool filtered = !REFINE && ctx.fopt && ctx.fopt->pk;
    if (filtered) {
        A.filter_pk = ctx.fopt->pk;
        A.filter_thr = ctx.fopt->logit_thr;
        A.filter_judge = ctx.fopt->judge;
        A.f_mark = ls.f_mark.get();
        A.fstats = ctx.fopt->fstats;
    }
    A.Kfx = ls.Kfx.get();
    A.Wfx = ls.Wfx.get();
    A.csize = ls.csize.get();
    // the barrier and result words are zero here: alloc_level clears them, and each sweep's
    // k_sweep_collect clears them for the next one
    // workgroup count: the heuristic first, then a few alternatives the online tuner measures
    // (any count gives identical labels, so exploring never changes a result)
    // a run without a job table uses the sweep compiled for uniform parameters
    const bool batch = A.d.vjob != nullptr;
    const uint32_t heur = filtered ? sweep_grid<false, true>(n, nnz)
                                   : (batch ? sweep_grid<REFINE, false, true>(n, nnz) : sweep_grid<REFINE>(n, nnz));
    std::vector<uint32_t> grids{heur};
    {
        const uint32_t res = filtered ? sweep_resident<false, true>()
                                      : (batch ? sweep_resident<REFINE, false, true>() : sweep_resident<REFINE>());
        // candidates stay within 4x of the heuristic: on a small level every count measures
        // about the same (launch-bound), but each step's barrier and compaction grow with it
        const uint32_t lo = std::max<uint32_t>(8u, heur / 4);
        const uint64_t hi = std::min<uint64_t>(res, 4ull * heur);
        for (uint32_t c : {heur / 2, heur * 2, res})
            if (c >= lo && c <= hi && std::find(grids.begin(), grids.end(), c) == grids.end()) grids.push_back(c);
    }
    const std::string tkey = std::string(REFINE ? "sweep/refine" : (filtered ? "sweep/move-filtered" : "sweep/move")) +
                             (batch ? "-batch" : "") + "/n2^" + std::to_string(bits_for(n)) + "/e2^" +
                             std::to_string(bits_for(nnz));
    const TuneChoice tc = tune_enabled() ? tune_pick(tkey, (int)grids.size()) : TuneChoice{0, false};
    const uint32_t grid = grids[tc.index];
    // each workgroup keeps its movers in its own slice of the mover array: its shares of the
    // buckets hold at most ceil(n / grid) + NUM_BUCKETS vertices
    const uint64_t stride = ((uint64_t)n + grid - 1) / grid + NUM_BUCKETS;
    if (ls.movers.size() < stride * grid) ls.movers.resize(stride * grid);
    A.d.movers = ls.movers.get();
    A.mv_stride = (uint32_t)stride;
    if (!REFINE && A.d.jsum) {
        if (ls.mv_tgt.size() < stride * grid) ls.mv_tgt.resize(stride * grid);
        A.mv_tgt = ls.mv_tgt.get();
    }
    constexpr uint32_t TRACE_CAP = 2048;
    DevBuf<long long> trace;
    if (sweep_trace_on()) {
        trace.resize(TRACE_CAP + 1);
        A.trace = trace.get() + 1;
        A.trace_cap = TRACE_CAP;
        hipLaunchKernelGGL(k_stamp_clock, dim3(1), dim3(1), 0, s, trace.get());
        KERNEL_OK();
    }
    void *args[] = {&A};
    hipEvent_t ev = nullptr;
    // a tuner measurement is taken only while the shape is explored; its events are read when
    // the run ends, so nothing here waits for the device
    const bool sample = tune_enabled() && tc.exploring;
    TuneSample ts{};
    if (sample) {
        ts = TuneSample{tkey, tc.index, ctx.events.next(), ctx.events.next()};
        HIP_OK(hipEventRecord(ts.t0, s));
    }
    if (ctx.timer.on && !REFINE) ctx.timer.begin(s, ev);
    const void *kern = filtered ? (const void *)k_sweep<false, true, false>
                                : (batch ? (const void *)k_sweep<REFINE, false, true> : (const void *)k_sweep<REFINE, false, false>);
    HIP_OK(hipLaunchCooperativeKernel(kern, dim3(grid), dim3(SWEEP_BS), args, 0, s));
    if (ctx.timer.on && !REFINE) ctx.timer.end(s, ev);
    if (sample) {
        HIP_OK(hipEventRecord(ts.t1, s));
        ctx.tune_samples.push_back(ts);
    }
    unsigned int ctl[4] = {0, 0, 0, 0};
    if (sweep_trace_on()) to_host(ctl, A.ctl, 4, s);  // before the collect kernel clears them
    // iterations, moves and the barrier-timeout flag accumulate on the device
    hipLaunchKernelGGL(k_sweep_collect, dim3(1), dim3(256), 0, s, A.ctl, reinterpret_cast<unsigned int *>(A.bc),
                       (uint32_t)(sizeof(BarrierCtl) / sizeof(unsigned int)),
                       (uint32_t)(offsetof(BarrierCtl, err) / sizeof(unsigned int)),
                       ctx.counters.get() + RUN_SWEEP + (REFINE ? 2 : 0), ctx.counters.get() + RUN_ERR, ls.stat.get(),
                       ctx.counters.get() + RUN_STAT + (REFINE ? 3 : 0));
    KERNEL_OK();
    if (sweep_trace_on()) {
        sync_stream(s);
        // barriers: 2 prologue, then per step: decide, apply; a cleanup first (refinement, cap);
        // with a cap also claims; local moving also activation. Phase times in us.
        const bool cp = A.claim != nullptr;
        const uint32_t per = 2u + (cp ? 1u : 0u) + (!REFINE ? 1u : 0u);
        const uint32_t cleanups = (REFINE || cp) && ctl[0] ? ctl[0] - 1u : 0u;
        const uint32_t nb = std::min<uint32_t>(TRACE_CAP, 2 + ctl[0] * per + cleanups);
        std::vector<long long> h(nb + 1);
        to_host(h.data(), trace.get(), nb + 1, s);
        const uint32_t ns = std::min<uint32_t>(ctl[0], A.max_steps);
        std::vector<unsigned int> wc((size_t)ns * NUM_BUCKETS + 1);
        if (ns) to_host(wc.data(), A.wcnt, (size_t)ns * NUM_BUCKETS, s);
        sync_stream(s);
        std::string line = std::string(REFINE ? "[sweep refine]" : "[sweep move]") + " n=" + std::to_string(n) +
                           " grid=" + std::to_string(grid) + " steps=" + std::to_string(ctl[0]) + " buckets=";
        for (int k = 0; k < NUM_BUCKETS; k++)
            line += std::to_string(b.start[k + 1] - b.start[k]) + (k + 1 < NUM_BUCKETS ? "/" : "");
        line += " us:";
        for (uint32_t i = 1; i <= nb; i++) {
            char buf[32];
            std::snprintf(buf, sizeof(buf), " %.1f", (double)(h[i] - h[i - 1]) / 100.0);
            line += buf;
        }
        line += " deciders/movers:";
        for (uint32_t i = 0; i < ns; i++)
            line += " " + std::to_string(wc[(size_t)i * NUM_BUCKETS]) + "/" + std::to_string(wc[(size_t)i * NUM_BUCKETS + 1]);
        line += " decide us max/mean over workgroups:";
        for (uint32_t i = 0; i < ns; i++) {
            char buf[48];
            std::snprintf(buf, sizeof(buf), " %.1f/%.1f", wc[(size_t)i * NUM_BUCKETS + 2] / 100.0,
                          wc[(size_t)i * NUM_BUCKETS + 3] / 100.0 / grid);
            line += buf;
        }
        std::fprintf(stderr, "%s\n", line.c_str());
    }
}

static void init_totals(const sc_graph &L, const uint32_t *memb, LevelState &ls, int sK, int sW, hipStream_t s) {
    uint32_t n = (uint32_t)L.n;
    HIP_OK(hipMemsetAsync(ls.Kfx.get(), 0, 2 * (size_t)n * sizeof(long long), s));
    HIP_OK(hipMemsetAsync(ls.Wfx.get(), 0, 2 * (size_t)n * sizeof(long long), s));
    HIP_OK(hipMemsetAsync(ls.csize.get(), 0, 2 * (size_t)n * sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_init_totals, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, memb, L.strength.get(),
                       L.node_w.get(), n, sK, sW, ls.Kfx.get(), ls.Wfx.get(), ls.csize.get());
    KERNEL_OK();
}

__global__ void k_stats_collect(unsigned long long *stat, unsigned long long *run) {
    if (threadIdx.x < 3) {
        run[threadIdx.x] += stat[threadIdx.x];
        stat[threadIdx.x] = 0;
    }
}

// The phase's decide statistics join the run's device counters (read when the run ends).
static void harvest_stats(LevelState &ls, RunCtx &ctx, bool move_phase) {
    hipLaunchKernelGGL(k_stats_collect, dim3(1), dim3(64), 0, ctx.s, ls.stat.get(),
                       ctx.counters.get() + RUN_STAT + (move_phase ? 0 : 3));
    KERNEL_OK();
}

// Run-end readback of the device counters, phase marks, tuner samples and kernel timers
// (after the stream has been synchronized).
static void finish_run_stats(RunCtx &ctx) {
    unsigned long long h[RUN_COUNTERS];
    to_host(h, ctx.counters.get(), RUN_COUNTERS, ctx.s);
    sync_stream(ctx.s);
    ctx.st.move_iterations += h[RUN_SWEEP + 0];
    ctx.st.moves += h[RUN_SWEEP + 1];
    ctx.st.refine_rounds += h[RUN_SWEEP + 2];
    for (int ph = 0; ph < 2; ph++) {
        const unsigned long long *c = h + RUN_STAT + 3 * ph;
        ctx.st.vertex_visits += c[0];
        ctx.st.edge_visits += c[1];
        // minimum unique bytes: neighbour id + weight + membership per entry (refinement also
        // reads the neighbour's working community: 4 more), 32 B per vertex, one 8-byte
        // community total per candidate scored, and per move two communities' two totals
        // (32 B; refinement's moves are counted by the persistent sweeps only)
        double bytes = (ph == 0 ? 12.0 : 16.0) * (double)c[1] + 32.0 * (double)c[0] + 8.0 * (double)c[2];
        bytes += 32.0 * (double)(ph == 0 ? ctx.st.moves : h[RUN_SWEEP + 3]);
        ctx.st.algorithmic_bytes += bytes;
        if (ph == 0) ctx.move_bytes += bytes;
    }
    for (auto &pm : ctx.phase_marks) {
        float ms = 0.0f;
        HIP_OK(hipEventElapsedTime(&ms, pm.second.first, pm.second.second));
        ctx.st.phase_ms[pm.first] += ms;
    }
    ctx.phase_marks.clear();
    for (auto &t : ctx.tune_samples) {
        float ms = 0.0f;
        HIP_OK(hipEventElapsedTime(&ms, t.t0, t.t1));
        tune_record(t.key, t.index, ms);
    }
    ctx.tune_samples.clear();
    if (ctx.timer.on) ctx.timer.harvest();
    if (h[RUN_ERR]) throw ScError(SC_DEVICE_ERROR, "Leiden sweep: a grid barrier timed out");
}

// Local moving on one level (local_move.rs:39-122), synchronous sub-sweeps.
static void local_move(const sc_graph &L, uint32_t *memb, LevelState &ls, const Buckets &b, RunCtx &ctx,
                       uint64_t level_seed, int sK, int sW, int sw) {
    uint32_t n = (uint32_t)L.n;
    hipStream_t s = ctx.s;
    DecideArgs a = base_args(L, ls, *ctx.cfg, sK, sW, sw);
    a.vjob = ctx.vjob;
    a.jobs = ctx.jobs;
    a.level = ctx.level;
    if (ctx.fopt && ctx.fopt->trace_x) {
        a.trace_x = ctx.fopt->trace_x;
        a.trace_y = ctx.fopt->trace_y;
        a.trace_key = ctx.fopt->trace_key;
        a.trace_n = ctx.fopt->trace_n;
        a.trace_cap = ctx.fopt->trace_cap;
    }
    if (ctx.fopt) {
        // filter features (kept by the sweep; the arrays hold the largest level's size)
        if (ls.f_margin.size() < n) {
            ls.f_margin.resize(n);
            ls.f_own.resize(n);
            ls.f_nmv.resize(n);
            ls.f_mark.resize(n);
        }
        a.f_margin = ls.f_margin.get();
        a.f_own = ls.f_own.get();
        a.f_nmv = ls.f_nmv.get();
    }
    a.memb = memb;
    a.P = MOVE_SUBSWEEPS;
    const bool capped = ctx.cfg->has_max_community_weight != 0;
    if (!capped) {
        // joiner admission: two sets of 2n sums / lowest ids (alternating persistent steps; the
        // multi-launch iterations use set 0 and clear it after every sub-sweep)
        if (ls.jsum.size() < 4 * (size_t)n) {
            ls.jsum.resize(4 * (size_t)n);
            ls.jmin.resize(4 * (size_t)n);
        }
        if (ls.mgain.size() < n) ls.mgain.resize(n);
        a.jsum = ls.jsum.get();
        a.jmin = ls.jmin.get();
        a.mgain = ls.mgain.get();
        a.jstride = 2 * (uint64_t)n;
    }
    if (b.start[GTAB_BUCKET] < b.start[GTAB_BUCKET + 1]) {
        a.gkeys = b.gkeys.get();
        a.gvals = b.gvals.get();
        a.goff = b.goff.get();
    }
    if (persistent_sweeps() || ctx.vjob || ctx.fopt) {
        run_sweep<false>(a, b, ls, ctx, level_seed, memb, L.nnz);  // its collect kernel takes the statistics
        return;
    }
    init_totals(L, memb, ls, sK, sW, s);
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, ls.active.get(), (uint64_t)n, (uint8_t)1);
    KERNEL_OK();
    if (!capped) {
        HIP_OK(hipMemsetAsync(ls.jsum.get(), 0, 2 * (size_t)n * sizeof(long long), s));
        HIP_OK(hipMemsetAsync(ls.jmin.get(), 0xFF, 2 * (size_t)n * sizeof(uint32_t), s));
        if (ls.admitted.size() < n) ls.admitted.resize(n);
    }
    // ctl: [0] stop, [1] iterations, [2] moves, [3] unused, [4 ..] movers per sub-sweep, then the
    // compacted counts per bucket
    constexpr uint32_t CTL = 4 + MOVE_SUBSWEEPS + NUM_BUCKETS;
    DevBuf<unsigned int> ctl(CTL);
    HIP_OK(hipMemsetAsync(ctl.get(), 0, CTL * sizeof(unsigned int), s));
    unsigned int *const mv_ctr = ctl.get() + 4, *const ccount = ctl.get() + 4 + MOVE_SUBSWEEPS;
    DevBuf<uint32_t> clist(n);
    BucketStarts bst;
    for (int i = 0; i <= NUM_BUCKETS; i++) bst.s[i] = b.start[i];
    a.stop = ctl.get();
    ApplyArgs ap{};
    ap.movers = ls.movers.get();
    ap.memb = memb;
    ap.target = ls.target.get();
    ap.Kfx = ls.Kfx.get();
    ap.Wfx = ls.Wfx.get();
    ap.csize = ls.csize.get();
    ap.strength = L.strength.get();
    ap.node_w = L.node_w.get();
    ap.sK = sK;
    ap.sW = sW;
    ap.active = ls.active.get();
    ap.claim = capped ? ls.claim.get() : nullptr;
    ap.admitted = capped ? nullptr : ls.admitted.get();
    ap.stop = ctl.get();
    const bool use_graph = graph_sweeps();
    DevBuf<uint32_t> iter_dev;
    if (use_graph) {
        iter_dev.resize(1);
        fill_u32(iter_dev.get(), 1, 0u, s);
        a.iter_dev = iter_dev.get();
        a.iter_seed_base = level_seed;
    }
    // one iteration: both sub-sweeps' compaction / decide / claim / apply / activate, then the
    // convergence test on the device (sync-free, so it can be captured); once the phase has
    // stopped every launch returns at once. With a graph the iteration index and seed come
    // from iter_dev.
    auto iteration = [&](int it) {
        if (!use_graph) a.seed = mix64(level_seed ^ (0x5851F42D4C957F2Dull * (uint64_t)(it + 1)));
        for (uint32_t sub = 0; sub < (uint32_t)MOVE_SUBSWEEPS; sub++) {
            a.sub = sub;
            a.n_movers = mv_ctr + sub;
            ap.n_movers = mv_ctr + sub;
            hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, s, ctl.get(), ccount, mv_ctr + sub, 1u);
            KERNEL_OK();
            hipLaunchKernelGGL(k_compact<false>, dim3(grid_for(n, 1024, 4096)), dim3(256), 0, s, a, b.lists.get(), bst,
                               clist.get(), ccount);
            KERNEL_OK();
            decide_all<false>(a, b, s, (ctx.timer.on && !use_graph) ? &ctx.timer : nullptr, clist.get(), ccount,
                              use_graph);
            if (capped) {
                hipLaunchKernelGGL(k_fill_u32_live, dim3(grid_for(2 * (uint64_t)n, 256, 4096)), dim3(256), 0, s,
                                   ctl.get(), ls.claim.get(), 2 * (uint64_t)n, 0xFFFFFFFFu);
                KERNEL_OK();
                hipLaunchKernelGGL(k_claim, dim3(1024), dim3(256), 0, s, ls.movers.get(), mv_ctr + sub,
                                   ls.target.get(), ls.claim.get(), ctl.get());
                KERNEL_OK();
            }
            hipLaunchKernelGGL(k_apply_move, dim3(1024), dim3(256), 0, s, ap, a);
            KERNEL_OK();
            hipLaunchKernelGGL((k_activate<16, false>), dim3(1024), dim3(256), 0, s, ls.movers.get(), mv_ctr + sub,
                               L.ptrs.get(), L.nbrs.get(), (const uint32_t *)memb, (const uint32_t *)nullptr,
                               ls.active.get(), (const uint32_t *)ls.target.get(),
                               (const uint32_t *)(capped ? ls.claim.get() : nullptr),
                               (const uint8_t *)(capped ? nullptr : ls.admitted.get()), (const unsigned int *)ctl.get());
            KERNEL_OK();
            if (!capped) {
                hipLaunchKernelGGL(k_clear_joins, dim3(1024), dim3(256), 0, s, ls.movers.get(), mv_ctr + sub,
                                   ls.target.get(), ls.jsum.get(), ls.jmin.get(), (const unsigned int *)ctl.get());
                KERNEL_OK();
            }
        }
        hipLaunchKernelGGL(k_iter_end, dim3(1), dim3(1), 0, s, ctl.get(), (const unsigned int *)mv_ctr,
                           (uint32_t)MOVE_SUBSWEEPS, (uint32_t)MAX_MOVE_ITERS, use_graph ? iter_dev.get() : nullptr);
        KERNEL_OK();
    };
    // ITER_BATCH iterations per replay (or per host check): the host reads the stop flag once per
    // batch, the iterations after convergence return at once
    IterGraph ig;
    if (use_graph)
        ig.capture(
            s,
            [&] {
                for (int k = 0; k < ITER_BATCH; k++) iteration(0);
            },
            &ctx.move_graph);
    for (int it = 0; it < MAX_MOVE_ITERS; it += ITER_BATCH) {
        if (use_graph)
            ig.launch(s);
        else
            for (int k = 0; k < ITER_BATCH; k++) iteration(it + k);
        unsigned int stop = 0;
        to_host(&stop, ctl.get(), 1, s);
        sync_stream(s);
        if (ctx.timer.on) ctx.timer.harvest();
        if (stop) break;
    }
    unsigned int h[3];
    to_host(h, ctl.get(), 3, s);
    sync_stream(s);
    ctx.st.move_iterations += h[1];
    ctx.st.moves += h[2];
    harvest_stats(ls, ctx, true);
}

// Refinement on one level (refine.rs:29-103); returns the refined membership in rmemb.
static void refine(const sc_graph &L, const uint32_t *memb, uint32_t *rmemb, LevelState &ls, const Buckets &b,
                   RunCtx &ctx, uint64_t level_seed, int sK, int sW, int sw) {
    uint32_t n = (uint32_t)L.n;
    hipStream_t s = ctx.s;
    DecideArgs a = base_args(L, ls, *ctx.cfg, sK, sW, sw);
    a.vjob = ctx.vjob;
    a.jobs = ctx.jobs;
    a.memb = rmemb;
    a.constraint = memb;
    a.P = 1;
    a.sub = 0;
    double mean_w = L.edge_count > 0 ? L.total_weight / (double)L.edge_count : 1.0;
    a.theta = ctx.cfg->refine_randomness * mean_w;  // refine.rs:44-50
    a.seed = mix64(level_seed ^ 0xA0761D6478BD642Full);
    if (b.start[GTAB_BUCKET] < b.start[GTAB_BUCKET + 1]) {
        a.gkeys = b.gkeys.get();
        a.gvals = b.gvals.get();
        a.goff = b.goff.get();
    }
    if (persistent_sweeps() || ctx.vjob) {
        run_sweep<true>(a, b, ls, ctx, level_seed, rmemb, L.nnz);  // its collect kernel takes the statistics
        return;
    }
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, rmemb, n);
    KERNEL_OK();
    init_totals(L, rmemb, ls, sK, sW, s);
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, ls.active.get(), (uint64_t)n, (uint8_t)1);
    KERNEL_OK();
    HIP_OK(hipMemsetAsync(ls.tround.get(), 0xFF, n * sizeof(uint32_t), s));
    HIP_OK(hipMemsetAsync(ls.mstep.get(), 0xFF, n * sizeof(uint32_t), s));  // no vertex has moved
    fill_u32(ls.targeted.get(), n, 0u, s);  // each round then clears what its movers set
    // ctl: [0] stop, [1] rounds, [2] moves, [3] unused, then the compacted counts per bucket
    constexpr uint32_t CTL = 4 + NUM_BUCKETS;
    DevBuf<unsigned int> ctl(CTL);
    HIP_OK(hipMemsetAsync(ctl.get(), 0, CTL * sizeof(unsigned int), s));
    unsigned int *const ccount = ctl.get() + 4;
    DevBuf<uint32_t> clist(n);
    BucketStarts bst;
    for (int i = 0; i <= NUM_BUCKETS; i++) bst.s[i] = b.start[i];
    a.stop = ctl.get();
    ApplyArgs ap{};
    ap.movers = ls.movers.get();
    ap.n_movers = ls.counters.get();
    ap.memb = rmemb;
    ap.target = ls.target.get();
    ap.tround = ls.tround.get();
    ap.targeted_cnt = ls.targeted.get();
    ap.Kfx = ls.Kfx.get();
    ap.Wfx = ls.Wfx.get();
    ap.csize = ls.csize.get();
    ap.strength = L.strength.get();
    ap.node_w = L.node_w.get();
    ap.sK = sK;
    ap.sW = sW;
    ap.active = ls.active.get();
    ap.moved = ls.moved.get();
    ap.n_moved = ls.counters.get() + 1;
    const bool capped = ctx.cfg->has_max_community_weight != 0;
    ap.claim = capped ? ls.claim.get() : nullptr;
    ap.stop = ctl.get();
    ap.mstep = ls.mstep.get();
    const bool use_graph = graph_sweeps();
    DevBuf<uint32_t> iter_dev;
    if (use_graph) {
        iter_dev.resize(1);
        fill_u32(iter_dev.get(), 1, 0u, s);
        a.iter_dev = iter_dev.get();
        ap.iter_dev = iter_dev.get();
    }
    auto round_launches = [&](uint32_t r) {
        if (!use_graph) {
            a.round = r;
            ap.round = r;
        }
        hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, s, ctl.get(), ccount, ls.counters.get(), 2u);
        KERNEL_OK();
        hipLaunchKernelGGL(k_compact<true>, dim3(grid_for(n, 1024, 4096)), dim3(256), 0, s, a, b.lists.get(), bst,
                           clist.get(), ccount);
        KERNEL_OK();
        decide_all<true>(a, b, s, nullptr, clist.get(), ccount, use_graph);
        if (capped) {
            hipLaunchKernelGGL(k_fill_u32_live, dim3(grid_for(2 * (uint64_t)n, 256, 4096)), dim3(256), 0, s,
                               ctl.get(), ls.claim.get(), 2 * (uint64_t)n, 0xFFFFFFFFu);
            KERNEL_OK();
            hipLaunchKernelGGL(k_claim, dim3(1024), dim3(256), 0, s, ls.movers.get(), ls.counters.get(),
                               ls.target.get(), ls.claim.get(), ctl.get());
            KERNEL_OK();
        }
        hipLaunchKernelGGL(k_apply_refine, dim3(1024), dim3(256), 0, s, ap);
        KERNEL_OK();
        hipLaunchKernelGGL(k_refine_reactivate, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, n, L.ptrs.get(),
                           L.nbrs.get(), (const uint32_t *)rmemb, (const uint32_t *)ls.csize.get(), memb,
                           (const uint32_t *)ls.mstep.get(), ls.active.get(),
                           use_graph ? (const uint32_t *)iter_dev.get() : nullptr, ap.round,
                           (const unsigned int *)ctl.get());
        KERNEL_OK();
        hipLaunchKernelGGL(k_clear_targeted, dim3(1024), dim3(256), 0, s, (const unsigned int *)ctl.get(),
                           ls.movers.get(), (const unsigned int *)ls.counters.get(), ls.target.get(),
                           ls.targeted.get());
        KERNEL_OK();
        hipLaunchKernelGGL(k_iter_end, dim3(1), dim3(1), 0, s, ctl.get(), (const unsigned int *)ls.counters.get() + 1,
                           1u, (uint32_t)MAX_REFINE_ROUNDS, use_graph ? iter_dev.get() : nullptr);
        KERNEL_OK();
    };
    IterGraph ig;
    if (use_graph)
        ig.capture(
            s,
            [&] {
                for (int k = 0; k < ITER_BATCH; k++) round_launches(0);
            },
            &ctx.refine_graph);
    for (uint32_t r = 0; r < (uint32_t)MAX_REFINE_ROUNDS; r += ITER_BATCH) {
        if (use_graph)
            ig.launch(s);
        else
            for (int k = 0; k < ITER_BATCH; k++) round_launches(r + k);
        unsigned int stop = 0;
        to_host(&stop, ctl.get(), 1, s);
        sync_stream(s);
        if (stop) break;
    }
    unsigned int h[3];
    to_host(h, ctl.get(), 3, s);
    sync_stream(s);
    ctx.st.refine_rounds += h[1];
    harvest_stats(ls, ctx, false);
}

// ---- renumbering -----------------------------------------------------------------------------
__global__ void k_used_flags(const uint32_t *ids, uint32_t n, uint32_t *flags) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        flags[ids[v]] = 1u;
}

__global__ void k_remap(const uint32_t *ids, uint32_t n, const uint32_t *newid, uint32_t *out) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        out[v] = newid[ids[v]];
}

// Ascending compaction of the ids in use (Partition::renumber, partition.rs:265-296).
// ids are < id_space. Writes compacted ids to out (may alias ids) and returns their count, or
// 0 without waiting for the device when the caller does not need it.
static uint32_t renumber(const uint32_t *ids, uint32_t n, uint32_t id_space, uint32_t *out, hipStream_t s,
                         bool need_count = true) {
    DevBuf<uint32_t> flags(id_space + 1), newid(id_space + 1);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for((uint64_t)id_space + 1, 256, 4096)), dim3(256), 0, s, flags.get(),
                       (uint64_t)id_space + 1, 0u);
    KERNEL_OK();
    hipLaunchKernelGGL(k_used_flags, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, ids, n, flags.get());
    KERNEL_OK();
    exclusive_scan_u32(flags.get(), newid.get(), id_space + 1, s);
    hipLaunchKernelGGL(k_remap, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, ids, n, newid.get(), out);
    KERNEL_OK();
    if (!need_count) return 0;
    uint32_t count = 0;
    to_host(&count, newid.get() + id_space, 1, s);
    sync_stream(s);
    return count;
}

// ---- aggregation (csr_network.rs:516-608) ----------------------------------------------------
__global__ void k_pairs_for_sort(const uint32_t *groups, uint32_t n, uint32_t *keys, uint32_t *vals) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        keys[v] = groups[v];
        vals[v] = v;
    }
}

__global__ void k_group_starts(const uint32_t *sorted_groups, uint32_t n, uint32_t ng, uint64_t *starts) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) {
        if (i == n) { starts[ng] = n; continue; }
        if (i == 0 || sorted_groups[i] != sorted_groups[i - 1]) starts[sorted_groups[i]] = i;
    }
}

// Coarse node weight and summed self-loop weight of each group, members in ascending order.
// Per group, in ascending member order (csr_network.rs:520-523): node weight and self-loop
// sums. 16 lanes per group; the lanes' tree sum is the sequential sum whenever the f64_quantum
// bound holds, otherwise lane 0 sums sequentially.
__global__ void __launch_bounds__(256) k_group_sums(const uint64_t *starts, const uint32_t *members, uint32_t ng,
                                                    const double *node_w, const double *self_loop, double *cnw,
                                                    double *cloops) {
    constexpr int G = 16, NG = 256 / G;
    const int lane = threadIdx.x % G;
    for (uint32_t g0 = blockIdx.x * NG; g0 < ng; g0 += gridDim.x * NG) {
        const uint32_t g = g0 + threadIdx.x / G;
        uint64_t b = 0, e = 0;
        if (g < ng) {
            b = starts[g];
            e = starts[g + 1];
        }
        double a = 0.0, l = 0.0, aa = 0.0, la = 0.0;
        int qa = 4096, ql = 4096;
        for (uint64_t j = b + lane; j < e; j += G) {
            const uint32_t v = members[j];
            const double x = node_w[v], y = self_loop[v];
            a += x;
            l += y;
            aa += fabs(x);
            la += fabs(y);
            qa = min(qa, f64_quantum(x));
            ql = min(ql, f64_quantum(y));
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            a += __shfl_xor(a, off, G);
            l += __shfl_xor(l, off, G);
            aa += __shfl_xor(aa, off, G);
            la += __shfl_xor(la, off, G);
            qa = min(qa, __shfl_xor(qa, off, G));
            ql = min(ql, __shfl_xor(ql, off, G));
        }
        if (g < ng && lane == 0) {
            if (!(aa < ldexp(1.0, min(53 + qa, 1000)))) {
                a = 0.0;
                for (uint64_t j = b; j < e; j++) a += node_w[members[j]];
            }
            if (!(la < ldexp(1.0, min(53 + ql, 1000)))) {
                l = 0.0;
                for (uint64_t j = b; j < e; j++) l += self_loop[members[j]];
            }
            cnw[g] = a;
            cloops[g] = l;
        }
    }
}

__global__ void k_member_degrees(const uint32_t *members, uint32_t n, const uint64_t *ptrs, uint64_t *deg) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j <= n; j += gridDim.x * blockDim.x) {
        if (j == n) { deg[j] = 0; continue; }
        uint32_t v = members[j];
        deg[j] = ptrs[v + 1] - ptrs[v];
    }
}

// Fine entries in member order; key (g, h) for h >= g, a sentinel above every key otherwise.
template <int G>
__global__ void __launch_bounds__(256) k_emit_fine(const uint32_t *members, const uint32_t *sorted_groups, uint32_t n,
                                                   const uint64_t *moff, const uint64_t *ptrs, const uint32_t *nbrs,
                                                   const float *wts, const uint32_t *groups, int hbits,
                                                   uint64_t sentinel, uint64_t *keys, float *vals) {
    constexpr int NG = 256 / G;
    const int lane = threadIdx.x % G;
    for (uint32_t j = blockIdx.x * NG + threadIdx.x / G; j < n; j += gridDim.x * NG) {
        uint32_t v = members[j];
        uint64_t g = sorted_groups[j];
        uint64_t o = moff[j], lo = ptrs[v], hi = ptrs[v + 1];
        for (uint64_t e = lo + lane; e < hi; e += G) {
            uint64_t h = groups[nbrs[e]];
            uint64_t out = o + (e - lo);
            keys[out] = h >= g ? (g << hbits) | h : sentinel;
            vals[out] = wts[e];
        }
    }
}

// Run heads of the sorted fine keys; the sentinel keys (sorted last) start no run.
__global__ void k_key_heads(const uint64_t *keys, uint64_t n, uint64_t sentinel, uint32_t *head) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        head[i] = (keys[i] != sentinel && (i == 0 || keys[i] != keys[i - 1])) ? 1u : 0u;
}

// Coarse pair sums. The reference adds each pair's fine weights sequentially in f64 in the
// sorted (member, row) order (csr_network.rs:548-564). Long segments use a wave-parallel tree,
// which equals the sequential chain whenever the f64_quantum bound holds, and keep the
// sequential chain otherwise. Short segments are summed sequentially by one thread.
constexpr uint32_t PAIR_SHORT = 64;

__device__ __forceinline__ void emit_pair(uint64_t key, double acc, uint32_t r, int hbits, const double *cloops,
                                          uint32_t *pa, uint32_t *pb, float *pw, uint32_t *keep) {
    const uint64_t hmask = (1ull << hbits) - 1;
    uint32_t g = (uint32_t)(key >> hbits), h = (uint32_t)(key & hmask);
    pa[r] = g;
    pb[r] = h;
    keep[r] = acc > 0.0 ? (g == h ? 1u : 2u) : 0u;
    pw[r] = (g == h) ? (float)(0.5 * (acc + cloops[g])) : (float)acc;
}

// Segment (pair) starts; start[np] is the end of the valid keys: the first sentinel, or n.
__global__ void k_seg_starts(const uint64_t *keys, uint64_t sentinel, const uint32_t *head, const uint32_t *run,
                             uint64_t n, uint32_t np, uint64_t *start) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i == n) {
            if (n == 0 || keys[n - 1] != sentinel) start[np] = n;
        } else if (keys[i] == sentinel) {
            if (i == 0 || keys[i - 1] != sentinel) start[np] = i;
        } else if (head[i]) {
            start[run[i]] = i;
        }
    }
}

__global__ void k_pair_sums_short(const uint64_t *keys, const float *vals, const uint64_t *start, uint32_t np,
                                  int hbits, const double *cloops, uint32_t *pa, uint32_t *pb, float *pw,
                                  uint32_t *keep, uint32_t *long_list, unsigned int *n_long) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < np; r += gridDim.x * blockDim.x) {
        const uint64_t b = start[r], e = start[r + 1];
        if (e - b > PAIR_SHORT) {
            long_list[atomicAdd(n_long, 1u)] = r;
            continue;
        }
        double acc = 0.0;
        for (uint64_t j = b; j < e; j++) acc += (double)vals[j];
        emit_pair(keys[b], acc, r, hbits, cloops, pa, pb, pw, keep);
    }
}

// One wave per long segment.
__global__ void __launch_bounds__(256) k_pair_sums_long(const uint64_t *keys, const float *vals, const uint64_t *start,
                                                        const uint32_t *long_list, const unsigned int *n_long,
                                                        int hbits, const double *cloops, uint32_t *pa, uint32_t *pb,
                                                        float *pw, uint32_t *keep) {
    const unsigned int cnt = *n_long;
    const int lane = threadIdx.x & 63;
    for (unsigned int q = blockIdx.x * 4 + (threadIdx.x >> 6); q < cnt; q += gridDim.x * 4) {
        const uint32_t r = long_list[q];
        const uint64_t b = start[r], e = start[r + 1];
        double acc = 0.0, sabs = 0.0;
        int qz = 4096;
        for (uint64_t j = b + lane; j < e; j += 64) {
            const double w = (double)vals[j];
            acc += w;
            sabs += fabs(w);
            qz = min(qz, f64_quantum(w));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            acc += __shfl_xor(acc, off, 64);
            sabs += __shfl_xor(sabs, off, 64);
            qz = min(qz, __shfl_xor(qz, off, 64));
        }
        const bool exact = sabs < ldexp(1.0, min(53 + qz, 1000));
        if (lane == 0) {
            if (!exact) {
                acc = 0.0;
                for (uint64_t j = b; j < e; j++) acc += (double)vals[j];
            }
            emit_pair(keys[b], acc, r, hbits, cloops, pa, pb, pw, keep);
        }
    }
}

__global__ void k_pair_emit(const uint32_t *pa, const uint32_t *pb, const float *pw, const uint32_t *keep,
                            const uint32_t *off, uint32_t np, int cbits, uint64_t *keys, float *vals) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < np; r += gridDim.x * blockDim.x) {
        uint32_t k = keep[r];
        if (!k) continue;
        uint32_t o = off[r];
        uint64_t a = pa[r], b = pb[r];
        keys[o] = (a << cbits) | b;
        vals[o] = pw[r];
        if (k == 2) {
            keys[o + 1] = (b << cbits) | a;
            vals[o + 1] = pw[r];
        }
    }
}

__global__ void k_csr_from_keys(const uint64_t *keys, const float *vals, uint64_t n, int cbits, uint32_t *nbrs,
                                float *wts) {
    const uint64_t cmask = (1ull << cbits) - 1;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k = keys[i];
        nbrs[i] = (uint32_t)(k & cmask);
        wts[i] = vals[i];
    }
}

// Row offsets of key-sorted entries: ptrs[r] = first index whose row (key >> cbits) is >= r.
__global__ void k_ptrs_from_keys(const uint64_t *keys, uint64_t n, int cbits, uint32_t rows, uint64_t *ptrs) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r <= rows; r += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = n;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if ((keys[mid] >> cbits) < r) lo = mid + 1;
            else hi = mid;
        }
        ptrs[r] = lo;
    }
}

// 64-bit-key variant (any size): fine entries sorted by (g, h), coarse entries by (row, column).
static std::unique_ptr<sc_graph> aggregate_device_wide(const sc_graph &L, const uint32_t *groups, uint32_t ng,
                                                       hipStream_t s) {
    uint32_t n = (uint32_t)L.n;
    auto out = std::make_unique<sc_graph>();
    out->n = ng;
    // members grouped by group, ascending vertex within a group (stable sort)
    DevBuf<uint32_t> sg(n), members(n);
    hipLaunchKernelGGL(k_pairs_for_sort, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, groups, n, sg.get(), members.get());
    KERNEL_OK();
    radix_sort_pairs(sg, members, n, std::max(1, bits_for(ng)), s);
    DevBuf<uint64_t> starts(ng + 1);
    hipLaunchKernelGGL(k_group_starts, dim3(grid_for(n + 1, 256, 4096)), dim3(256), 0, s, sg.get(), n, ng, starts.get());
    KERNEL_OK();
    out->node_w.resize(ng);
    DevBuf<double> cloops(ng);
    hipLaunchKernelGGL(k_group_sums, dim3(grid_for(ng, 16, 4096)), dim3(256), 0, s, starts.get(), members.get(), ng,
                       L.node_w.get(), L.self_loop.get(), out->node_w.get(), cloops.get());
    KERNEL_OK();
    // fine entries in member order
    DevBuf<uint64_t> deg(n + 1), moff(n + 1);
    hipLaunchKernelGGL(k_member_degrees, dim3(grid_for(n + 1, 256, 4096)), dim3(256), 0, s, members.get(), n,
                       L.ptrs.get(), deg.get());
    KERNEL_OK();
    exclusive_scan_u64(deg.get(), moff.get(), n + 1, s);
    uint64_t nnz = L.nnz;
    const int hbits = std::max(1, bits_for(ng));
    const uint64_t sentinel = (uint64_t)ng << hbits;
    const int kbits = std::min(64, hbits + bits_for(ng) + 1);
    DevBuf<uint64_t> keys(nnz);
    DevBuf<float> vals(nnz);
    if (nnz) {
        hipLaunchKernelGGL((k_emit_fine<16>), dim3(grid_for(n, 16, 2048)), dim3(256), 0, s, members.get(), sg.get(), n,
                           moff.get(), L.ptrs.get(), L.nbrs.get(), L.wts.get(), groups, hbits, sentinel, keys.get(),
                           vals.get());
        KERNEL_OK();
        radix_sort_pairs(keys, vals, nnz, kbits, s);
    }
    // runs of equal keys over all nnz entries (sentinels start none), so the number of valid
    // keys is never needed on the host: one wait gives the pair count
    DevBuf<uint32_t> head(nnz + 1), run(nnz + 1);
    uint32_t np = 0;
    if (nnz) {
        hipLaunchKernelGGL(k_key_heads, dim3(grid_for(nnz, 256, 8192)), dim3(256), 0, s, keys.get(), nnz, sentinel,
                           head.get());
        KERNEL_OK();
        exclusive_scan_u32(head.get(), run.get(), nnz, s);
        uint32_t lr = 0, lh = 0;
        to_host(&lr, run.get() + nnz - 1, 1, s);
        to_host(&lh, head.get() + nnz - 1, 1, s);
        sync_stream(s);
        np = lr + lh;
    }
    DevBuf<uint32_t> pa(np + 1), pb(np + 1), keep(np + 1), off(np + 1);
    DevBuf<float> pw(np + 1);
    uint64_t total_entries = 0;
    if (np) {
        DevBuf<uint64_t> seg(np + 1);
        DevBuf<uint32_t> long_list(np);
        DevBuf<unsigned int> n_long(1);
        HIP_OK(hipMemsetAsync(n_long.get(), 0, sizeof(unsigned int), s));
        hipLaunchKernelGGL(k_seg_starts, dim3(grid_for(nnz + 1, 256, 8192)), dim3(256), 0, s, keys.get(), sentinel,
                           head.get(), run.get(), nnz, np, seg.get());
        KERNEL_OK();
        hipLaunchKernelGGL(k_pair_sums_short, dim3(grid_for(np, 256, 8192)), dim3(256), 0, s, keys.get(), vals.get(),
                           seg.get(), np, hbits, cloops.get(), pa.get(), pb.get(), pw.get(), keep.get(),
                           long_list.get(), n_long.get());
        KERNEL_OK();
        hipLaunchKernelGGL(k_pair_sums_long, dim3(1024), dim3(256), 0, s, keys.get(), vals.get(), seg.get(),
                           long_list.get(), n_long.get(), hbits, cloops.get(), pa.get(), pb.get(), pw.get(),
                           keep.get());
        KERNEL_OK();
        exclusive_scan_u32(keep.get(), off.get(), np, s);
        uint32_t lo = 0, lk = 0;
        to_host(&lo, off.get() + np - 1, 1, s);
        to_host(&lk, keep.get() + np - 1, 1, s);
        sync_stream(s);
        total_entries = (uint64_t)lo + lk;
    }
    const int cbits = std::max(1, bits_for(ng));
    DevBuf<uint64_t> ckeys(total_entries);
    DevBuf<float> cvals(total_entries);
    if (np) {
        hipLaunchKernelGGL(k_pair_emit, dim3(grid_for(np, 256, 4096)), dim3(256), 0, s, pa.get(), pb.get(), pw.get(),
                           keep.get(), off.get(), np, cbits, ckeys.get(), cvals.get());
        KERNEL_OK();
        radix_sort_pairs(ckeys, cvals, total_entries, std::min(64, 2 * cbits), s);
    }
    out->nnz = total_entries;
    out->nbrs.resize(total_entries);
    out->wts.resize(total_entries);
    if (total_entries) {
        hipLaunchKernelGGL(k_csr_from_keys, dim3(grid_for(total_entries, 256, 8192)), dim3(256), 0, s, ckeys.get(),
                           cvals.get(), total_entries, cbits, out->nbrs.get(), out->wts.get());
        KERNEL_OK();
    }
    out->ptrs.resize(ng + 1);
    hipLaunchKernelGGL(k_ptrs_from_keys, dim3(grid_for((uint64_t)ng + 1, 256, 4096)), dim3(256), 0, s, ckeys.get(),
                       total_entries, cbits, ng, out->ptrs.get());
    KERNEL_OK();
    graph_finish_sorted(*out, s);
    return out;
}

// ---- aggregation with 32-bit keys --------------------------------------------------------------
// The reference walks the groups in order and, per group g, adds the weights of its members'
// entries (members ascending, rows ascending) towards groups h >= g (csr_network.rs:548-564).
// Here every fine entry t toward a group h >= g (member order, rows in order) is keyed by h and
// carries (g, w) packed in 64 bits; a stable sort by h leaves each pair's entries contiguous
// and in t order, with the pairs ordered by (h, g). So the pair list directly
// gives each row r its entries (g, r), g <= r, ascending in g; a second stable sort of the
// pairs with g < h by g (carrying (h, w)) gives each row its entries (r, h), h > r, ascending in
// h. Both sorts use 32-bit keys of bits_for(ng) bits.

__device__ __forceinline__ uint64_t pack_gw(uint32_t g, float w) { return ((uint64_t)g << 32) | __float_as_uint(w); }
__device__ __forceinline__ uint32_t gw_group(uint64_t x) { return (uint32_t)(x >> 32); }
__device__ __forceinline__ float gw_weight(uint64_t x) { return __uint_as_float((uint32_t)x); }

// The entries aggregation keeps are those toward groups at or above the row's own: per stored
// entry its neighbour's group (hb[e], entry order), per row the number it keeps (ucnt[v]). G
// lanes per row.
template <int G>
__global__ void __launch_bounds__(256) k_upper_counts(const uint64_t *ptrs, const uint32_t *nbrs, const uint32_t *groups,
                                                      uint32_t n, uint32_t *hb, uint32_t *ucnt) {
    constexpr int NG = 256 / G;
    constexpr int EB = 4;  // entries per lane with their loads in flight together
    const int lane = threadIdx.x % G;
    for (uint32_t v = blockIdx.x * NG + threadIdx.x / G; v < n; v += gridDim.x * NG) {
        const uint32_t g = groups[v];
        const uint64_t lo = ptrs[v], hi = ptrs[v + 1];
        uint32_t c = 0;
        for (uint64_t e0 = lo + lane; e0 < hi; e0 += (uint64_t)EB * G) {
            uint32_t u[EB];
#pragma unroll
            for (int q = 0; q < EB; q++) {
                const uint64_t e = e0 + (uint64_t)q * G;
                u[q] = e < hi ? nbrs[e] : 0u;
            }
            uint32_t h[EB];
#pragma unroll
            for (int q = 0; q < EB; q++) h[q] = e0 + (uint64_t)q * G < hi ? groups[u[q]] : 0u;
#pragma unroll
            for (int q = 0; q < EB; q++) {
                const uint64_t e = e0 + (uint64_t)q * G;
                if (e < hi) {
                    hb[e] = h[q];
                    c += h[q] >= g ? 1u : 0u;
                }
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) c += __shfl_xor(c, off, G);
        if (lane == 0) ucnt[v] = c;
    }
}

__global__ void k_member_upper(const uint32_t *members, uint32_t n, const uint32_t *ucnt, uint64_t *deg) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j <= n; j += gridDim.x * blockDim.x)
        deg[j] = j == n ? 0u : ucnt[members[j]];
}

// The kept entries of the member rows in member order (rows ascending within a member), packed:
// entry t of the list gets key h and payload (g, w). A group of G lanes walks a row EB * G
// entries at a time; a kept entry's place is the row's running count plus the kept entries
// before it in the batch (compare ballots of the group's lanes).
template <int G>
__global__ void __launch_bounds__(256) k_emit_upper_h(const uint32_t *members, const uint32_t *sorted_groups, uint32_t n,
                                                      const uint64_t *moff, const uint64_t *ptrs, const uint32_t *hb,
                                                      const float *wts, uint32_t *keys, uint64_t *gw) {
    constexpr int NG = 256 / G;
    constexpr int EB = 4;
    const int lane = threadIdx.x % G;
    const int wl = threadIdx.x & 63;
    const int gl0 = wl - lane;  // the group's first lane in the wave
    const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    const uint64_t lower = (1ull << lane) - 1ull;  // the group's lanes before this one
    for (uint32_t j = blockIdx.x * NG + threadIdx.x / G; j < n; j += gridDim.x * NG) {
        const uint32_t v = members[j];
        const uint32_t g = sorted_groups[j];
        const uint64_t lo = ptrs[v], hi = ptrs[v + 1];
        uint64_t t = moff[j];
        for (uint64_t b = lo; b < hi; b += (uint64_t)EB * G) {  // uniform over the group
            uint32_t h[EB];
            float w[EB];
#pragma unroll
            for (int q = 0; q < EB; q++) {
                const uint64_t e = b + (uint64_t)q * G + lane;
                h[q] = e < hi ? hb[e] : 0u;
                w[q] = e < hi ? wts[e] : 0.0f;
            }
#pragma unroll
            for (int q = 0; q < EB; q++) {
                const bool keep = b + (uint64_t)q * G + lane < hi && h[q] >= g;
                const uint64_t gm = (__ballot(keep) >> gl0) & gmask;
                if (keep) {
                    const uint64_t at = t + (uint64_t)__popcll(gm & lower);
                    keys[at] = h[q];
                    gw[at] = pack_gw(g, w[q]);
                }
                t += (uint64_t)__popcll(gm);
            }
        }
    }
}

// A pair starts where the neighbour group or the row group changes (sentinels start none).
__global__ void k_pair_heads_h(const uint32_t *keys, const uint64_t *gw, uint64_t n, uint32_t sent, uint32_t *head) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = keys[i];
        head[i] = (k != sent && (i == 0 || k != keys[i - 1] || gw_group(gw[i]) != gw_group(gw[i - 1]))) ? 1u : 0u;
    }
}

// Pair starts; start[np] is the first sentinel position (or n).
__global__ void k_seg_starts_h(const uint32_t *keys, uint32_t sent, const uint32_t *head, const uint32_t *run,
                               uint64_t n, uint32_t np, uint64_t *start) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i == n) {
            if (n == 0 || keys[n - 1] != sent) start[np] = n;
        } else if (keys[i] == sent) {
            if (i == 0 || keys[i - 1] != sent) start[np] = i;
        } else if (head[i]) {
            start[run[i]] = i;
        }
    }
}

__device__ __forceinline__ void emit_pair_h(uint32_t g, uint32_t h, double acc, uint32_t r, const double *cloops,
                                            uint32_t *pa, uint32_t *pb, float *pw, uint32_t *keep) {
    pa[r] = g;
    pb[r] = h;
    keep[r] = acc > 0.0 ? 1u : 0u;  // non-positive sums are not emitted (csr_network.rs:568-570)
    pw[r] = (g == h) ? (float)(0.5 * (acc + cloops[g])) : (float)acc;
}

// Short pairs: one thread adds the entries in their t order.
__global__ void k_pair_sums_short_h(const uint32_t *keys, const uint64_t *gw, const uint64_t *start, uint32_t np,
                                    const double *cloops, uint32_t *pa, uint32_t *pb, float *pw, uint32_t *keep,
                                    uint32_t *long_list, unsigned int *n_long) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < np; r += gridDim.x * blockDim.x) {
        const uint64_t b = start[r], e = start[r + 1];
        if (e - b > PAIR_SHORT) {
            long_list[atomicAdd(n_long, 1u)] = r;
            continue;
        }
        double acc = 0.0;
        for (uint64_t j = b; j < e; j++) acc += (double)gw_weight(gw[j]);
        emit_pair_h(gw_group(gw[b]), keys[b], acc, r, cloops, pa, pb, pw, keep);
    }
}

// Long pairs: one workgroup each; the tree sum equals the sequential chain whenever the
// f64_quantum bound holds, otherwise thread 0 adds sequentially.
__global__ void __launch_bounds__(256) k_pair_sums_long_h(const uint32_t *keys, const uint64_t *gw,
                                                          const uint64_t *start, const uint32_t *long_list,
                                                          const unsigned int *n_long, const double *cloops,
                                                          uint32_t *pa, uint32_t *pb, float *pw, uint32_t *keep) {
    __shared__ double s_acc[4], s_abs[4];
    __shared__ int s_q[4];
    const unsigned int cnt = *n_long;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (unsigned int q = blockIdx.x; q < cnt; q += gridDim.x) {
        const uint32_t r = long_list[q];
        const uint64_t b = start[r], e = start[r + 1];
        double acc = 0.0, sabs = 0.0;
        int qz = 4096;
        // four independent loads in flight per thread
        uint64_t j = b + threadIdx.x;
        for (; j + 768 < e; j += 1024) {
            const double w0 = (double)gw_weight(gw[j]), w1 = (double)gw_weight(gw[j + 256]);
            const double w2 = (double)gw_weight(gw[j + 512]), w3 = (double)gw_weight(gw[j + 768]);
            acc += (w0 + w1) + (w2 + w3);
            sabs += (fabs(w0) + fabs(w1)) + (fabs(w2) + fabs(w3));
            qz = min(qz, min(min(f64_quantum(w0), f64_quantum(w1)), min(f64_quantum(w2), f64_quantum(w3))));
        }
        for (; j < e; j += 256) {
            const double w = (double)gw_weight(gw[j]);
            acc += w;
            sabs += fabs(w);
            qz = min(qz, f64_quantum(w));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            acc += __shfl_xor(acc, off, 64);
            sabs += __shfl_xor(sabs, off, 64);
            qz = min(qz, __shfl_xor(qz, off, 64));
        }
        if (lane == 0) {
            s_acc[wave] = acc;
            s_abs[wave] = sabs;
            s_q[wave] = qz;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            acc = (s_acc[0] + s_acc[1]) + (s_acc[2] + s_acc[3]);
            sabs = (s_abs[0] + s_abs[1]) + (s_abs[2] + s_abs[3]);
            qz = min(min(s_q[0], s_q[1]), min(s_q[2], s_q[3]));
            if (!(sabs < ldexp(1.0, min(53 + qz, 1000)))) {
                acc = 0.0;
                for (uint64_t i = b; i < e; i++) acc += (double)gw_weight(gw[i]);
            }
            emit_pair_h(gw_group(gw[b]), keys[b], acc, r, cloops, pa, pb, pw, keep);
        }
        __syncthreads();
    }
}

// Key of the second sort: the row group of a kept pair with g < h (an upper entry of row g),
// the sentinel otherwise; the payload carries (h, w).
__global__ void k_upper_keys(const uint32_t *pa, const uint32_t *pb, const float *pw, const uint32_t *keep,
                             uint32_t np, uint32_t sent, uint32_t *key2, uint64_t *hw) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < np; r += gridDim.x * blockDim.x) {
        key2[r] = (keep[r] && pa[r] != pb[r]) ? pa[r] : sent;
        hw[r] = pack_gw(pb[r], pw[r]);
    }
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Row r: its part A (kept pairs (g, r), g <= r) is the pair list range with pb == r, its part B
// (pairs (r, h), h > r) the range of the second sort with key2 == r.
__global__ void k_row_parts(const uint32_t *pb, const uint32_t *kscan, const uint32_t *key2, uint32_t np,
                            uint32_t ng, uint32_t *a0, uint32_t *b0, uint32_t *cnt_a, uint64_t *row_len) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r <= ng; r += gridDim.x * blockDim.x) {
        if (r == ng) {
            row_len[ng] = 0;
            continue;
        }
        const uint32_t x0 = lower_bound_u32(pb, np, r), x1 = lower_bound_u32(pb, np, r + 1);
        const uint32_t y0 = lower_bound_u32(key2, np, r), y1 = lower_bound_u32(key2, np, r + 1);
        const uint32_t ca = kscan[x1] - kscan[x0];
        a0[r] = x0;
        b0[r] = y0;
        cnt_a[r] = ca;
        row_len[r] = (uint64_t)ca + (y1 - y0);
    }
}

__global__ void k_fill_rows(const uint32_t *pa, const uint32_t *pb, const float *pw, const uint32_t *keep,
                            const uint32_t *kscan, const uint32_t *key2, const uint64_t *hw, uint32_t np,
                            uint32_t sent, const uint32_t *a0, const uint32_t *b0, const uint32_t *cnt_a,
                            const uint64_t *ptrs, uint32_t *nbrs, float *wts) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
        if (keep[i]) {
            // entry (g, r) of row r = pb[i], at its rank among the row's kept part-A pairs
            const uint32_t r = pb[i];
            const uint64_t o = ptrs[r] + (kscan[i] - kscan[a0[r]]);
            nbrs[o] = pa[i];
            wts[o] = pw[i];
        }
        const uint32_t g = key2[i];
        if (g != sent) {
            // entry (g, h) of row g, after the row's part A
            const uint64_t x = hw[i];
            const uint64_t o = ptrs[g] + cnt_a[g] + (i - b0[g]);
            nbrs[o] = gw_group(x);
            wts[o] = gw_weight(x);
        }
    }
}

std::unique_ptr<sc_graph> aggregate_device(const sc_graph &L, const uint32_t *groups, uint32_t ng, hipStream_t s) {
    const uint32_t n = (uint32_t)L.n;
    const uint64_t nnz = L.nnz;
    if (nnz >= 0xFFFFFFFFull || ng >= 0x7FFFFFFFu) return aggregate_device_wide(L, groups, ng, s);
    auto out = std::make_unique<sc_graph>();
    out->n = ng;
    // members grouped by group, ascending vertex within a group (stable sort)
    DevBuf<uint32_t> sg(n), members(n);
    hipLaunchKernelGGL(k_pairs_for_sort, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, groups, n, sg.get(), members.get());
    KERNEL_OK();
    radix_sort_pairs(sg, members, n, std::max(1, bits_for(ng)), s);
    DevBuf<uint64_t> starts(ng + 1);
    hipLaunchKernelGGL(k_group_starts, dim3(grid_for(n + 1, 256, 4096)), dim3(256), 0, s, sg.get(), n, ng, starts.get());
    KERNEL_OK();
    out->node_w.resize(ng);
    DevBuf<double> cloops(ng);
    hipLaunchKernelGGL(k_group_sums, dim3(grid_for(ng, 16, 4096)), dim3(256), 0, s, starts.get(), members.get(), ng,
                       L.node_w.get(), L.self_loop.get(), out->node_w.get(), cloops.get());
    KERNEL_OK();
    const uint32_t sent = ng;
    const int kbits = std::max(1, bits_for(ng));
    // only the entries toward groups at or above the row's own enter the sort (the others are
    // the mirror images the pair list restores): their neighbour groups and counts first, then
    // the packed list in member order
    uint64_t nk = 0;
    DevBuf<uint64_t> moff(n + 1);
    DevBuf<uint32_t> keys, head, run;
    DevBuf<uint64_t> gw;
    uint32_t np = 0;
    if (nnz) {
        // lanes per row from the level's mean degree (about four entries per lane)
        const double mean_deg = n ? (double)nnz / (double)n : 0.0;
        const int eg = mean_deg <= 16.0 ? 4 : (mean_deg <= 32.0 ? 8 : (mean_deg <= 64.0 ? 16 : 32));
        DevBuf<uint32_t> hb(nnz), ucnt(n);
        auto count = [&](auto kern, int grp) {
            hipLaunchKernelGGL(kern, dim3(grid_for(n, 256 / grp, 4096)), dim3(256), 0, s, L.ptrs.get(), L.nbrs.get(),
                               groups, n, hb.get(), ucnt.get());
            KERNEL_OK();
        };
        if (eg == 4) count(k_upper_counts<4>, 4);
        else if (eg == 8) count(k_upper_counts<8>, 8);
        else if (eg == 16) count(k_upper_counts<16>, 16);
        else count(k_upper_counts<32>, 32);
        DevBuf<uint64_t> deg(n + 1);
        hipLaunchKernelGGL(k_member_upper, dim3(grid_for(n + 1, 256, 4096)), dim3(256), 0, s, members.get(), n,
                           ucnt.get(), deg.get());
        KERNEL_OK();
        exclusive_scan_u64(deg.get(), moff.get(), n + 1, s);
        to_host(&nk, moff.get() + n, 1, s);
        sync_stream(s);
        if (nk) {
            keys.resize(nk);
            gw.resize(nk);
            head.resize(nk + 1);
            run.resize(nk + 1);
            auto emit = [&](auto kern, int grp) {
                hipLaunchKernelGGL(kern, dim3(grid_for(n, 256 / grp, 4096)), dim3(256), 0, s, members.get(), sg.get(),
                                   n, moff.get(), L.ptrs.get(), hb.get(), L.wts.get(), keys.get(), gw.get());
                KERNEL_OK();
            };
            if (eg == 4) emit(k_emit_upper_h<4>, 4);
            else if (eg == 8) emit(k_emit_upper_h<8>, 8);
            else if (eg == 16) emit(k_emit_upper_h<16>, 16);
            else emit(k_emit_upper_h<32>, 32);
            radix_sort_pairs(keys, gw, nk, kbits, s);
            hipLaunchKernelGGL(k_pair_heads_h, dim3(grid_for(nk, 256, 8192)), dim3(256), 0, s, keys.get(), gw.get(),
                               nk, sent, head.get());
            KERNEL_OK();
            exclusive_scan_u32(head.get(), run.get(), nk, s);
            uint32_t lr = 0, lh = 0;
            to_host(&lr, run.get() + nk - 1, 1, s);
            to_host(&lh, head.get() + nk - 1, 1, s);
            sync_stream(s);
            np = lr + lh;
        }
    }
    DevBuf<uint32_t> pa(np + 1), pb(np + 1), keep(np + 1), kscan(np + 1), key2(np + 1);
    DevBuf<uint64_t> hw(np + 1);
    DevBuf<float> pw(np + 1);
    DevBuf<uint32_t> a0(ng + 1), b0(ng + 1), cnt_a(ng + 1);
    DevBuf<uint64_t> row_len(ng + 1);
    out->ptrs.resize((uint64_t)ng + 1);
    if (np) {
        DevBuf<uint64_t> seg(np + 1);
        DevBuf<uint32_t> long_list(np);
        DevBuf<unsigned int> n_long(1);
        HIP_OK(hipMemsetAsync(n_long.get(), 0, sizeof(unsigned int), s));
        hipLaunchKernelGGL(k_seg_starts_h, dim3(grid_for(nk + 1, 256, 8192)), dim3(256), 0, s, keys.get(), sent,
                           head.get(), run.get(), nk, np, seg.get());
        KERNEL_OK();
        hipLaunchKernelGGL(k_pair_sums_short_h, dim3(grid_for(np, 256, 8192)), dim3(256), 0, s, keys.get(), gw.get(),
                           seg.get(), np, cloops.get(), pa.get(), pb.get(), pw.get(), keep.get(), long_list.get(),
                           n_long.get());
        KERNEL_OK();
        hipLaunchKernelGGL(k_pair_sums_long_h, dim3(2048), dim3(256), 0, s, keys.get(), gw.get(), seg.get(),
                           long_list.get(), n_long.get(), cloops.get(), pa.get(), pb.get(), pw.get(), keep.get());
        KERNEL_OK();
        // kscan[np] = kept pairs in total
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, keep.get() + np, (uint64_t)1, 0u);
        KERNEL_OK();
        exclusive_scan_u32(keep.get(), kscan.get(), (uint64_t)np + 1, s);
        hipLaunchKernelGGL(k_upper_keys, dim3(grid_for(np, 256, 8192)), dim3(256), 0, s, pa.get(), pb.get(), pw.get(),
                           keep.get(), np, sent, key2.get(), hw.get());
        KERNEL_OK();
        radix_sort_pairs(key2, hw, np, kbits, s);
    }
    hipLaunchKernelGGL(k_row_parts, dim3(grid_for((uint64_t)ng + 1, 256, 8192)), dim3(256), 0, s, pb.get(),
                       kscan.get(), key2.get(), np, ng, a0.get(), b0.get(), cnt_a.get(), row_len.get());
    KERNEL_OK();
    exclusive_scan_u64(row_len.get(), out->ptrs.get(), (uint64_t)ng + 1, s);
    // a pair gives at most two entries (its row and, for g < h, the mirrored one), so 2 np
    // bounds the coarse graph without waiting for its count; graph_finish_sorted reads the count
    // with the other row statistics
    out->nbrs.resize(2 * (uint64_t)np);
    out->wts.resize(2 * (uint64_t)np);
    if (np) {
        hipLaunchKernelGGL(k_fill_rows, dim3(grid_for(np, 256, 8192)), dim3(256), 0, s, pa.get(), pb.get(), pw.get(),
                           keep.get(), kscan.get(), key2.get(), hw.get(), np, sent, a0.get(), b0.get(), cnt_a.get(),
                           out->ptrs.get(), out->nbrs.get(), out->wts.get());
        KERNEL_OK();
    }
    graph_finish_sorted(*out, s, true);
    return out;
}

// ---- batched runs: many graphs x resolutions as one disjoint-union graph ----------------------
// Every job (one graph at one resolution) keeps its vertices contiguous at every level; its
// objective parameters, fixed-point scales and the per-level totals they come from are the
// ones a run of that graph alone would use (same reductions, same association order), hashing
// uses job-local vertex ids, and a job stops (is frozen) when its own level loop or its own
// passes would stop. So each job's labels are the labels of the standalone run.
struct BatchState {
    uint32_t J = 0;
    std::vector<double> gamma;       // per job
    std::vector<double> tw0, tnw0;   // level-0 total weight and total node weight per job
    std::vector<uint64_t> voff0;     // level-0 vertex offsets (J + 1)
    std::vector<uint64_t> cnt;       // vertices per job at the current level
    std::vector<uint8_t> pass_done;  // the job's passes have stopped (patience)
    std::vector<uint8_t> level_done; // the job's level loop has stopped in this pass
    DevBuf<uint32_t> vjob0;          // level-0 job of every vertex
    DevBuf<uint64_t> d_voff0;
    DevBuf<JobParams> params;
    bool stopped(uint32_t j) const { return pass_done[j] || level_done[j]; }
    bool all_stopped() const {
        for (uint32_t j = 0; j < J; j++)
            if (!stopped(j)) return false;
        return true;
    }
    bool all_passes_done() const {
        for (uint32_t j = 0; j < J; j++)
            if (!pass_done[j]) return false;
        return true;
    }
};

// Undirected edges (u >= v, loops once) of each job's rows: one workgroup per job, a binary
// search per sorted row.
__global__ void k_job_edge_counts(const uint64_t *ptrs, const uint32_t *nbrs, const uint64_t *vranges,
                                  unsigned long long *out) {
    const uint32_t j = blockIdx.x;
    unsigned long long c = 0;
    for (uint64_t v = vranges[j] + threadIdx.x; v < vranges[j + 1]; v += blockDim.x) {
        uint64_t lo = ptrs[v], hi = ptrs[v + 1];
        const uint64_t end = hi;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (nbrs[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        c += end - lo;
    }
    c = block_sum(c);
    if (threadIdx.x == 0) out[j] = c;
}

// Smallest and largest value of x over each job's range: one workgroup per job.
__global__ void k_job_minmax(const uint32_t *x, const uint64_t *vranges, uint32_t *mn, uint32_t *mx) {
    const uint32_t j = blockIdx.x;
    uint32_t a = 0xFFFFFFFFu, b = 0;
    for (uint64_t v = vranges[j] + threadIdx.x; v < vranges[j + 1]; v += blockDim.x) {
        a = min(a, x[v]);
        b = max(b, x[v]);
    }
    a = 0xFFFFFFFFu - block_max(0xFFFFFFFFu - a);
    b = block_max(b);
    if (threadIdx.x == 0) {
        mn[j] = a;
        mx[j] = b;
    }
}

__global__ void k_scatter_job(const uint32_t *groups, const uint32_t *vjob, uint32_t n, uint32_t *cjob) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        cjob[groups[v]] = vjob[v];
}

__global__ void k_job_rebase(uint32_t *ids, const uint32_t *vjob, uint32_t n, const long long *shift) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        ids[v] = (uint32_t)((long long)ids[v] + shift[vjob[v]]);
}

static std::vector<uint64_t> batch_ranges(const BatchState &B);

// Community ids compacted over the union are contiguous per job; moves them onto the job's
// vertex range (job j's ids become lo_j + local id), so vertex ids and community ids share one
// offset per job as they share one id space in a standalone run (the empty-community candidate
// and the lower-id tie-break compare them).
static void batch_rebase(const BatchState &B, uint32_t *ids, const uint32_t *vjob, uint32_t n, hipStream_t s) {
    const uint32_t J = B.J;
    const std::vector<uint64_t> r = batch_ranges(B);
    DevBuf<uint64_t> d_r(J + 1);
    to_device(d_r.get(), r.data(), J + 1, s);
    DevBuf<uint32_t> mm(2 * (size_t)J);
    hipLaunchKernelGGL(k_job_minmax, dim3(J), dim3(256), 0, s, ids, d_r.get(), mm.get(), mm.get() + J);
    KERNEL_OK();
    std::vector<uint32_t> h(2 * (size_t)J);
    to_host(h.data(), mm.get(), 2 * (size_t)J, s);
    sync_stream(s);
    std::vector<long long> sh(J, 0);
    for (uint32_t j = 0; j < J; j++) sh[j] = B.cnt[j] ? (long long)r[j] - (long long)h[j] : 0;
    DevBuf<long long> d_sh(J);
    to_device(d_sh.get(), sh.data(), J, s);
    hipLaunchKernelGGL(k_job_rebase, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, ids, vjob, n,
                       (const long long *)d_sh.get());
    KERNEL_OK();
    sync_stream(s);  // sh is a host temporary
}

static std::vector<uint64_t> batch_ranges(const BatchState &B) {
    std::vector<uint64_t> r(B.J + 1, 0);
    for (uint32_t j = 0; j < B.J; j++) r[j + 1] = r[j] + B.cnt[j];
    return r;
}

// After local moving of a batched run, community ids span [0, 2N): c < N is a vertex id of its
// job's range [lo, hi), N + v is vertex v's private empty community. Each job's ids are mapped
// in order into [2 lo, 2 hi) (vertex ids to lo + c, private ids to hi + v), so a job's ids are
// contiguous and ordered as in a standalone run of that job ([0, n_j), then n_j + v), and the
// ascending compactions that follow number every job's communities as its standalone run does.
__global__ void k_job_contiguous_ids(uint32_t *memb, uint32_t n, const uint32_t *vjob, const uint64_t *ranges) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint32_t c = memb[v];
        const bool priv = c >= n;
        const uint32_t base = priv ? c - n : c;
        const uint32_t j = vjob[base];
        memb[v] = (uint32_t)(priv ? ranges[j + 1] + base : ranges[j] + base);
    }
}

static void batch_job_contiguous_ids(const BatchState &B, uint32_t *memb, const uint32_t *vjob, uint32_t n,
                                     hipStream_t s) {
    if (n == 0) return;
    const std::vector<uint64_t> r = batch_ranges(B);
    DevBuf<uint64_t> d_r(B.J + 1);
    to_device(d_r.get(), r.data(), B.J + 1, s);
    hipLaunchKernelGGL(k_job_contiguous_ids, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, memb, n, vjob,
                       (const uint64_t *)d_r.get());
    KERNEL_OK();
    sync_stream(s);  // r is a host temporary
}

// Per-level parameters of every job, from the level graph's per-job totals (the same values
// and reductions run_once derives from a graph of that job alone).
static void batch_level_params(BatchState &B, const sc_graph &L, const sc_leiden_config &cfg, hipStream_t s) {
    const uint32_t J = B.J;
    const std::vector<uint64_t> r = batch_ranges(B);
    DevBuf<uint64_t> d_r(J + 1);
    to_device(d_r.get(), r.data(), J + 1, s);
    DevBuf<double> red(3 * (size_t)J);
    DevBuf<unsigned long long> ec(J);
    device_segment_reduce_f64(L.row_total.get(), d_r.get(), J, false, red.get(), s);
    device_segment_reduce_f64(L.node_w.get(), d_r.get(), J, false, red.get() + J, s);
    device_segment_reduce_f64(L.strength.get(), d_r.get(), J, true, red.get() + 2 * J, s);
    hipLaunchKernelGGL(k_job_edge_counts, dim3(J), dim3(256), 0, s, L.ptrs.get(), L.nbrs.get(), d_r.get(), ec.get());
    KERNEL_OK();
    std::vector<double> h(3 * (size_t)J);
    std::vector<unsigned long long> hec(J);
    to_host(h.data(), red.get(), 3 * (size_t)J, s);
    to_host(hec.data(), ec.get(), J, s);
    sync_stream(s);
    std::vector<JobParams> P(J);
    for (uint32_t j = 0; j < J; j++) {
        const double tw = h[j], tnw = h[J + j], ms = B.cnt[j] ? h[2 * J + j] : 0.0;
        JobParams &p = P[j];
        p.gamma = B.gamma[j];
        p.four_m = 4.0 * tw;
        p.threshold = 1e-12 * std::max(std::fabs(tw), 1.0);
        const double mean_w = hec[j] > 0 ? tw / (double)hec[j] : 1.0;  // refine.rs:44-50
        p.theta = cfg.refine_randomness * mean_w;
        p.sK = fixed_scale_for(2.0 * tw);
        p.sW = fixed_scale_for(tnw);
        p.sw = fixed_scale_for(ms);
        p.lo = (uint32_t)r[j];
        p.frozen = (B.stopped(j) || B.cnt[j] == 0) ? 1u : 0u;
        p.pad = 0;
    }
    B.params.resize(J);
    to_device(B.params.get(), P.data(), J, s);
    sync_stream(s);  // P is a host temporary
}

// Groups per job after refinement (ids are ascending and contiguous per job).
static std::vector<uint32_t> batch_group_counts(const BatchState &B, const uint32_t *groups, hipStream_t s) {
    const uint32_t J = B.J;
    const std::vector<uint64_t> r = batch_ranges(B);
    DevBuf<uint64_t> d_r(J + 1);
    to_device(d_r.get(), r.data(), J + 1, s);
    DevBuf<uint32_t> mm(2 * (size_t)J);
    hipLaunchKernelGGL(k_job_minmax, dim3(J), dim3(256), 0, s, groups, d_r.get(), mm.get(), mm.get() + J);
    KERNEL_OK();
    std::vector<uint32_t> h(2 * (size_t)J), out(J, 0);
    to_host(h.data(), mm.get(), 2 * (size_t)J, s);
    sync_stream(s);
    for (uint32_t j = 0; j < J; j++) out[j] = B.cnt[j] ? h[J + j] - h[j] + 1 : 0;
    return out;
}

// ---- driver (leiden/mod.rs:168-289) -----------------------------------------------------------
// Stream time per phase: events mark the phase boundaries in stream order (no host
// synchronization); the intervals are read when the run ends.
struct PhaseClock {
    RunCtx &ctx;
    hipStream_t s;
    hipEvent_t t;
    PhaseClock(RunCtx &c, hipStream_t st) : ctx(c), s(st), t(c.events.next()) { HIP_OK(hipEventRecord(t, s)); }
    void lap(int phase) {
        hipEvent_t now = ctx.events.next();
        HIP_OK(hipEventRecord(now, s));
        ctx.phase_marks.push_back({phase, {t, now}});
        t = now;
    }
};

__global__ void k_scatter_coarse_memb(const uint32_t *groups, const uint32_t *memb, uint32_t n, uint32_t *cm) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        cm[groups[v]] = memb[v];
}

__global__ void k_gather_u32(const uint32_t *map, uint32_t *x, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = map[x[i]];
}

__global__ void k_u32_to_u64(const uint32_t *in, uint64_t *out, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = in[i];
}

// One pass (run_once, mod.rs:214-289). labels (n0, consecutive ids) are read and overwritten.
static uint32_t run_once(const sc_graph &G0, uint32_t *labels, uint64_t pass_seed, RunCtx &ctx) {
    hipStream_t s = ctx.s;
    const sc_leiden_config &cfg = *ctx.cfg;
    uint32_t n0 = (uint32_t)G0.n;
    const sc_graph *level = &G0;
    std::unique_ptr<sc_graph> owned;
    DevBuf<uint32_t> memb(n0), rmemb(n0), node_at(n0), groups(n0);
    HIP_OK(hipMemcpyAsync(memb.get(), labels, n0 * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n0, 256, 4096)), dim3(256), 0, s, node_at.get(), n0);
    KERNEL_OK();
    LevelState ls;
    alloc_level(ls, n0, s);
    BatchState *B = ctx.batch;
    DevBuf<uint32_t> vjob_l, vjob_c;
    if (B) {
        vjob_l.resize(n0);
        vjob_c.resize(n0);
        HIP_OK(hipMemcpyAsync(vjob_l.get(), B->vjob0.get(), n0 * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        for (uint32_t j = 0; j < B->J; j++) {
            B->cnt[j] = B->voff0[j + 1] - B->voff0[j];
            B->level_done[j] = 0;
        }
    }
    for (uint64_t lvl = 0; lvl < cfg.max_levels; lvl++) {
        const sc_graph &L = *level;
        uint32_t n = (uint32_t)L.n;
        ctx.st.levels++;
        int sK = fixed_scale_for(2.0 * L.total_weight);
        int sW = fixed_scale_for(L.total_node_weight);
        int sw = fixed_scale_for(L.max_strength);
        uint64_t level_seed = mix64(pass_seed ^ (0x9E3779B97F4A7C15ull * (lvl + 1)));
        PhaseClock pc(ctx, s);
        ctx.level = (uint32_t)lvl;
        if (B) {
            batch_level_params(*B, L, cfg, s);
            ctx.vjob = vjob_l.get();
            ctx.jobs = B->params.get();
        }
        // degree-class bounds: the defaults, then a few narrower sets the online tuner measures
        // over this level's local moving and refinement (identical decisions under any of them)
        static constexpr BucketBounds BOUNDS[] = {DEFAULT_BUCKET_BOUNDS,
                                                  {{6, 12, 24, 96, 384, 768, 1536}},
                                                  {{8, 16, 32, 64, 512, 1024, 2048}},
                                                  {{4, 12, 32, 128, 512, 1024, 2048}}};
        const std::string bkey = std::string(B ? "buckets-batch" : "buckets") + "/n2^" + std::to_string(bits_for(n)) +
                                 "/e2^" + std::to_string(bits_for(L.nnz)) + (cfg.refine ? "" : "/louvain");
        const TuneChoice btc = tune_enabled() && n >= 4096 ? tune_pick(bkey, 4, 1) : TuneChoice{0, false};
        Buckets b;
        build_buckets(L, b, BOUNDS[btc.index], s);
        TuneSample bts{};
        if (btc.exploring) {
            bts = TuneSample{bkey, btc.index, ctx.events.next(), ctx.events.next()};
            HIP_OK(hipEventRecord(bts.t0, s));
        }
        if (sweep_trace_on())
            std::fprintf(stderr, "[level %llu] n=%u nnz=%llu max_degree=%llu\n", (unsigned long long)lvl, n,
                         (unsigned long long)L.nnz, (unsigned long long)L.max_degree);
        pc.lap(0);
        local_move(L, memb.get(), ls, b, ctx, level_seed, sK, sW, sw);
        // community ids now span [0, 2n); a batched run lays each job's ids out contiguously
        if (B) batch_job_contiguous_ids(*B, memb.get(), vjob_l.get(), n, s);
        pc.lap(1);
        uint32_t ng;
        if (cfg.refine) {
            refine(L, memb.get(), rmemb.get(), ls, b, ctx, level_seed, sK, sW, sw);
            pc.lap(2);
        }
        if (btc.exploring) {
            HIP_OK(hipEventRecord(bts.t1, s));
            ctx.tune_samples.push_back(bts);
        }
        if (cfg.refine) ng = renumber(rmemb.get(), n, n, groups.get(), s);
        else ng = renumber(memb.get(), n, 2 * n, groups.get(), s);
        pc.lap(3);
        std::vector<uint32_t> jng;
        if (B) {
            // each job stops where its own level loop would (mod.rs:263-265)
            jng = batch_group_counts(*B, groups.get(), s);
            for (uint32_t j = 0; j < B->J; j++)
                if (!B->stopped(j) && jng[j] >= B->cnt[j]) B->level_done[j] = 1;
            if (B->all_stopped()) break;
        } else if (ng >= n) {
            break;  // no shrink, nothing left to gain (mod.rs:263-265)
        }
        std::unique_ptr<sc_graph> coarse = aggregate_device(L, groups.get(), ng, s);
        if (B) {
            hipLaunchKernelGGL(k_scatter_job, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, groups.get(),
                               vjob_l.get(), n, vjob_c.get());
            KERNEL_OK();
            std::swap(vjob_l, vjob_c);
            for (uint32_t j = 0; j < B->J; j++) B->cnt[j] = jng[j];
        }
        pc.lap(4);
        // several coarse nodes can share a working community (mod.rs:269-272); compact its ids
        hipLaunchKernelGGL(k_scatter_coarse_memb, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, groups.get(),
                           memb.get(), n, rmemb.get());
        KERNEL_OK();
        renumber(rmemb.get(), ng, 2 * n, memb.get(), s, false);
        if (B) batch_rebase(*B, memb.get(), vjob_l.get(), ng, s);
        hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(n0, 256, 4096)), dim3(256), 0, s, groups.get(), node_at.get(), n0);
        KERNEL_OK();
        pc.lap(3);
        owned = std::move(coarse);
        level = owned.get();
    }
    ctx.vjob = nullptr;
    ctx.jobs = nullptr;
    uint32_t n = (uint32_t)level->n;
    uint32_t nc = renumber(memb.get(), n, 2 * n, rmemb.get(), s);
    // labels[v] = final community of v's node at the last level
    HIP_OK(hipMemcpyAsync(labels, node_at.get(), n0 * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(n0, 256, 4096)), dim3(256), 0, s, rmemb.get(), labels, n0);
    KERNEL_OK();
    sync_stream(s);
    return nc;
}

static double labels_quality(const sc_graph &G0, const uint32_t *labels, const sc_leiden_config &cfg, hipStream_t s) {
    uint32_t n = (uint32_t)G0.n;
    DevBuf<uint64_t> l64(n);
    hipLaunchKernelGGL(k_u32_to_u64, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, labels, l64.get(), n);
    KERNEL_OK();
    return quality_device(G0, l64.get(), cfg.objective == SC_RB ? 1 : 2, cfg.resolution, s);
}

static void validate_config(const sc_leiden_config &c) {
    auto bad = [](const std::string &m) { throw ScError(SC_INVALID_CONFIG, "invalid configuration: " + m); };
    if (c.objective != SC_RB && c.objective != SC_CPM) bad("unknown objective");
    if (!std::isfinite(c.resolution) || c.resolution < 0.0)
        bad("resolution must be finite and non-negative, got " + std::to_string(c.resolution));
    if (c.n_iterations == 0) bad("n_iterations must be at least 1");
    if (c.max_levels == 0) bad("max_levels must be at least 1");
    if (c.patience == 0) bad("patience must be at least 1");
    if (!std::isfinite(c.refine_randomness) || c.refine_randomness < 0.0)
        bad("refine_randomness must be finite and non-negative, got " + std::to_string(c.refine_randomness));
    if (c.has_max_community_weight && (!std::isfinite(c.max_community_weight) || c.max_community_weight <= 0.0))
        bad("max_community_weight must be finite and positive, got " + std::to_string(c.max_community_weight));
}

void quality_jobs_device(const sc_graph &g, const uint64_t *d_labels, int mode, const uint32_t *vjob,
                         const uint64_t *d_vranges, uint32_t J, const double *h_gamma, const double *h_tw,
                         const double *h_tnw, double *h_q, hipStream_t s);

static void batch_quality(const sc_graph &G0, const uint32_t *labels, const sc_leiden_config &cfg, BatchState &B,
                          std::vector<double> &q, hipStream_t s) {
    const uint32_t n = (uint32_t)G0.n;
    DevBuf<uint64_t> l64(n);
    hipLaunchKernelGGL(k_u32_to_u64, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, labels, l64.get(), n);
    KERNEL_OK();
    q.assign(B.J, 0.0);
    quality_jobs_device(G0, l64.get(), cfg.objective == SC_RB ? 1 : 2, B.vjob0.get(), B.d_voff0.get(), B.J,
                        B.gamma.data(), B.tw0.data(), B.tnw0.data(), q.data(), s);
}

void leiden_device(const sc_graph &G0, const sc_leiden_config &cfg, uint32_t *d_labels, uint64_t *n_clusters,
                   hipStream_t s, sc_leiden_stats *stats_out, BatchState *batch = nullptr,
                   const FilterOpts *fopt = nullptr) {
    validate_config(cfg);
    RunCtx ctx;
    ctx.cfg = &cfg;
    ctx.s = s;
    ctx.fopt = fopt;
    ctx.timer.on = profiling_enabled();
    ctx.counters.resize(RUN_COUNTERS);
    HIP_OK(hipMemsetAsync(ctx.counters.get(), 0, RUN_COUNTERS * sizeof(unsigned long long), s));
    hipEvent_t t0, t1;
    HIP_OK(hipEventCreate(&t0));
    HIP_OK(hipEventCreate(&t1));
    HIP_OK(hipEventRecord(t0, s));
    uint32_t n0 = (uint32_t)G0.n;
    uint64_t seed = cfg.seed;
    if (!cfg.has_seed) {
        std::random_device rd;
        seed = ((uint64_t)rd() << 32) ^ rd();
    }
    uint64_t nc = 0;
    if (n0 > 0 && batch) {
        // every job's passes, with its own improvement and patience (mod.rs:180-204)
        ctx.batch = batch;
        BatchState &B = *batch;
        hipLaunchKernelGGL(k_iota, dim3(grid_for(n0, 256, 4096)), dim3(256), 0, s, d_labels, n0);
        KERNEL_OK();
        std::vector<double> q_prev, q;
        std::vector<uint64_t> stagnant(B.J, 0);
        B.pass_done.assign(B.J, 0);
        B.level_done.assign(B.J, 0);
        B.cnt.assign(B.J, 0);
        for (uint32_t j = 0; j < B.J; j++)
            if (B.voff0[j + 1] == B.voff0[j]) B.pass_done[j] = 1;
        PhaseClock pq(ctx, s);
        batch_quality(G0, d_labels, cfg, B, q_prev, s);
        pq.lap(5);
        for (uint64_t pass = 0; pass < cfg.n_iterations && !B.all_passes_done(); pass++) {
            ctx.st.passes++;
            uint64_t pass_seed = mix64(seed + 0x9E3779B97F4A7C15ull * (pass + 1));
            nc = run_once(G0, d_labels, pass_seed, ctx);
            PhaseClock pq2(ctx, s);
            batch_quality(G0, d_labels, cfg, B, q, s);
            // the next pass starts from these labels: put each job's on its vertex range
            for (uint32_t j = 0; j < B.J; j++) B.cnt[j] = B.voff0[j + 1] - B.voff0[j];
            batch_rebase(B, d_labels, B.vjob0.get(), n0, s);
            pq2.lap(5);
            for (uint32_t j = 0; j < B.J; j++) {
                if (B.pass_done[j]) continue;
                const double improvement = q[j] - q_prev[j];
                q_prev[j] = q[j];
                if (improvement <= cfg.tolerance * std::max(std::fabs(B.tw0[j]), 1.0)) {
                    if (++stagnant[j] >= cfg.patience) B.pass_done[j] = 1;
                } else {
                    stagnant[j] = 0;
                }
            }
        }
        ctx.batch = nullptr;
    } else if (n0 > 0) {
        hipLaunchKernelGGL(k_iota, dim3(grid_for(n0, 256, 4096)), dim3(256), 0, s, d_labels, n0);
        KERNEL_OK();
        PhaseClock pq(ctx, s);
        // the starting labels are the singletons
        double q_prev = quality_singletons_device(G0, cfg.objective == SC_RB ? 1 : 2, cfg.resolution, s);
        pq.lap(5);
        uint64_t stagnant = 0;
        double scale = std::max(std::fabs(G0.total_weight), 1.0);
        for (uint64_t pass = 0; pass < cfg.n_iterations; pass++) {
            ctx.st.passes++;
            uint64_t pass_seed = mix64(seed + 0x9E3779B97F4A7C15ull * (pass + 1));
            nc = run_once(G0, d_labels, pass_seed, ctx);
            PhaseClock pq2(ctx, s);
            double q = labels_quality(G0, d_labels, cfg, s);
            pq2.lap(5);
            double improvement = q - q_prev;
            q_prev = q;
            if (improvement <= cfg.tolerance * scale) {
                if (++stagnant >= cfg.patience) break;
            } else {
                stagnant = 0;
            }
        }
    }
    HIP_OK(hipEventRecord(t1, s));
    sync_stream(s);
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    finish_run_stats(ctx);
    ctx.st.total_ms = ms;
    ctx.st.move_kernel_ms = ctx.timer.total_ms;
    ctx.st.move_kernel_launches = ctx.timer.launches;
    ctx.st.move_kernel_bytes = ctx.move_bytes;
    *n_clusters = nc;
    if (stats_out) *stats_out = ctx.st;
}

// One job's source graph and its place in the union.
struct UnionSrc {
    const uint64_t *ptrs;
    const uint32_t *nbrs;
    const float *wts;
    const double *node_w, *strength, *self_loop, *row_total;
    uint64_t n, nnz, voff, eoff;
};

// blockIdx.y = job: node arrays (row offsets shifted, every vertex tagged with its job).
__global__ void k_union_nodes(const UnionSrc *src, uint64_t *ptrs, double *node_w, double *strength,
                              double *self_loop, double *row_total, uint32_t *vjob) {
    const UnionSrc u = src[blockIdx.y];
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < u.n; v += (uint64_t)gridDim.x * blockDim.x) {
        ptrs[u.voff + v] = u.eoff + u.ptrs[v];
        node_w[u.voff + v] = u.node_w[v];
        strength[u.voff + v] = u.strength[v];
        self_loop[u.voff + v] = u.self_loop[v];
        row_total[u.voff + v] = u.row_total[v];
        vjob[u.voff + v] = blockIdx.y;
    }
}

// blockIdx.y = job: adjacency entries with neighbour ids shifted by the job's vertex offset.
__global__ void k_union_edges(const UnionSrc *src, uint32_t *nbrs, float *wts) {
    const UnionSrc u = src[blockIdx.y];
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < u.nnz; e += (uint64_t)gridDim.x * blockDim.x) {
        nbrs[u.eoff + e] = (uint32_t)(u.voff + u.nbrs[e]);
        wts[u.eoff + e] = u.wts[e];
    }
}

}  // namespace scgpu

using namespace scgpu;

extern "C" {

static void labels_to_host(const uint32_t *d, uint64_t n, uint64_t *out, hipStream_t s) {
    DevBuf<uint64_t> l64(n);
    hipLaunchKernelGGL(k_u32_to_u64, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, d, l64.get(), n);
    KERNEL_OK();
    to_host(out, l64.get(), n, s);
    sync_stream(s);
}

sc_status sc_leiden_trace(const sc_graph *g, const sc_leiden_config *cfg, uint64_t cap, float *x_out,
                          uint8_t *y_out, uint64_t *n_out, sc_error *err) {
    try {
        if (!g || !cfg || !n_out || (cap && (!x_out || !y_out))) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        check_node_count(g->n);
        hipStream_t s = thread_stream();
        const uint32_t n = (uint32_t)g->n;
        DevBuf<float> tx(std::max<uint64_t>(cap, 1) * MLP_NF);
        DevBuf<uint8_t> ty(std::max<uint64_t>(cap, 1));
        DevBuf<uint64_t> tk(std::max<uint64_t>(cap, 1));
        DevBuf<unsigned long long> tn(1);
        HIP_OK(hipMemsetAsync(tn.get(), 0, sizeof(unsigned long long), s));
        FilterOpts fo;
        fo.trace_x = tx.get();
        fo.trace_y = ty.get();
        fo.trace_key = tk.get();
        fo.trace_n = tn.get();
        fo.trace_cap = cap;
        DevBuf<uint32_t> labels(n);
        uint64_t nc = 0;
        sc_leiden_stats st{};
        leiden_device(*g, *cfg, labels.get(), &nc, s, &st, nullptr, &fo);
        unsigned long long total = 0;
        to_host(&total, tn.get(), 1, s);
        sync_stream(s);
        const uint64_t m = std::min<uint64_t>(total, cap);
        *n_out = m;
        // samples arrive in atomic-slot order; sorting by (level, iteration, vertex) and then by
        // content (levels repeat across passes) makes the trace, and a model trained on it,
        // reproducible (when the cap is not reached)
        std::vector<float> hx(m * MLP_NF);
        std::vector<uint8_t> hy(m);
        std::vector<uint64_t> hk(m);
        to_host(hx.data(), tx.get(), m * MLP_NF, s);
        to_host(hy.data(), ty.get(), m, s);
        to_host(hk.data(), tk.get(), m, s);
        sync_stream(s);
        std::vector<uint64_t> order(m);
        for (uint64_t i = 0; i < m; i++) order[i] = i;
        std::sort(order.begin(), order.end(), [&](uint64_t p, uint64_t q) {
            if (hk[p] != hk[q]) return hk[p] < hk[q];
            const int c = std::memcmp(hx.data() + p * MLP_NF, hx.data() + q * MLP_NF, MLP_NF * sizeof(float));
            return c != 0 ? c < 0 : hy[p] < hy[q];
        });
        for (uint64_t i = 0; i < m; i++) {
            std::memcpy(x_out + i * MLP_NF, hx.data() + order[i] * MLP_NF, MLP_NF * sizeof(float));
            y_out[i] = hy[order[i]];
        }
        t_stats = st;
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

sc_status sc_leiden_filtered(const sc_graph *g, const sc_leiden_config *cfg, const float *weights, double threshold,
                             int32_t judge, uint64_t *labels_out, uint64_t *n_clusters_out, sc_filter_stats *fst,
                             sc_error *err) {
    try {
        if (!g || !cfg || !weights || !n_clusters_out || (!labels_out && g->n))
            throw ScError(SC_INVALID_ARGUMENT, "null argument");
        if (!(threshold > 0.0 && threshold < 1.0)) throw ScError(SC_INVALID_ARGUMENT, "threshold must be in (0, 1)");
        if (judge < 0 || judge > 2) throw ScError(SC_INVALID_ARGUMENT, "judge must be 0, 1 or 2");
        check_node_count(g->n);
        hipStream_t s = thread_stream();
        const uint32_t n = (uint32_t)g->n;
        DevBuf<float> w(MLP_PARAMS);
        DevBuf<MlpPacked> pk(1);
        DevBuf<unsigned long long> fs(5);
        to_device(w.get(), weights, MLP_PARAMS, s);
        mlp_pack_device(w.get(), pk.get(), s);
        HIP_OK(hipMemsetAsync(fs.get(), 0, 5 * sizeof(unsigned long long), s));
        FilterOpts fo;
        fo.pk = pk.get();
        fo.judge = judge;
        fo.logit_thr = (float)std::log(threshold / (1.0 - threshold));
        fo.fstats = fs.get();
        DevBuf<uint32_t> labels(n);
        sc_leiden_stats st{};
        leiden_device(*g, *cfg, labels.get(), n_clusters_out, s, &st, nullptr, &fo);
        if (n) labels_to_host(labels.get(), n, labels_out, s);
        unsigned long long h[5] = {0, 0, 0, 0, 0};
        to_host(h, fs.get(), 5, s);
        sync_stream(s);
        if (fst) {
            fst->skipped = h[0];
            fst->judged = h[1];
            fst->judge_moves = h[2];
            fst->skipped_at_judge = h[3];
            fst->skipped_moved = h[4];
            fst->skip_error_rate = h[3] ? (double)h[4] / (double)h[3] : 0.0;
        }
        t_stats = st;
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

sc_status sc_leiden_batch(const sc_graph *const *graphs, uint64_t n_graphs, const double *resolutions,
                          uint64_t n_resolutions, const sc_leiden_config *cfg, uint64_t *labels_out,
                          uint64_t *n_clusters_out, sc_error *err) {
    try {
        if (!graphs || !resolutions || !cfg || !n_clusters_out)
            throw ScError(SC_INVALID_ARGUMENT, "null argument");
        if (n_graphs == 0 || n_resolutions == 0) throw ScError(SC_INVALID_ARGUMENT, "empty batch");
        const uint64_t J64 = n_graphs * n_resolutions;
        if (J64 > 65535) throw ScError(SC_INVALID_ARGUMENT, "at most 65535 jobs per batch");
        const uint32_t J = (uint32_t)J64;
        for (uint64_t r = 0; r < n_resolutions; r++) {
            sc_leiden_config c = *cfg;
            c.resolution = resolutions[r];
            validate_config(c);
        }
        uint64_t N = 0, E = 0;
        for (uint64_t g = 0; g < n_graphs; g++) {
            if (!graphs[g]) throw ScError(SC_INVALID_ARGUMENT, "null graph");
            N += graphs[g]->n * n_resolutions;
            E += graphs[g]->nnz * n_resolutions;
        }
        check_node_count(N);
        check_entry_count(E);
        if (N && !labels_out) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        hipStream_t s = thread_stream();
        // the union graph: job j = graph j / R at resolution j % R, vertices and entries in job order
        sc_graph U;
        U.n = N;
        U.nnz = E;
        U.ptrs.resize(N + 1);
        U.nbrs.resize(E);
        U.wts.resize(E);
        U.node_w.resize(N);
        U.strength.resize(N);
        U.self_loop.resize(N);
        U.row_total.resize(N);
        BatchState B;
        B.J = J;
        B.gamma.resize(J);
        B.tw0.resize(J);
        B.tnw0.resize(J);
        B.voff0.assign(J + 1, 0);
        B.vjob0.resize(N);
        std::vector<UnionSrc> src(J);
        uint64_t voff = 0, eoff = 0;
        uint32_t maxdeg = 0;
        bool bucket_known = true;
        for (uint32_t j = 0; j < J; j++) {
            const sc_graph &g = *graphs[j / n_resolutions];
            src[j] = UnionSrc{g.ptrs.get(), g.nbrs.get(), g.wts.get(), g.node_w.get(), g.strength.get(),
                              g.self_loop.get(), g.row_total.get(), g.n, g.nnz, voff, eoff};
            B.gamma[j] = resolutions[j % n_resolutions];
            B.tw0[j] = g.total_weight;
            B.tnw0[j] = g.total_node_weight;
            B.voff0[j] = voff;
            U.total_weight += g.total_weight;
            U.total_node_weight += g.total_node_weight;
            U.edge_count += g.edge_count;
            U.any_self_loops = U.any_self_loops || g.any_self_loops;
            U.max_strength = std::max(U.max_strength, g.max_strength);
            maxdeg = std::max(maxdeg, g.max_degree);
            bucket_known = bucket_known && g.has_bucket_counts;
            for (int k = 0; k < NUM_BUCKETS; k++) U.bucket_counts[k] += g.bucket_counts[k];
            voff += g.n;
            eoff += g.nnz;
        }
        B.voff0[J] = voff;
        U.max_degree = maxdeg;
        U.has_bucket_counts = bucket_known && N <= 0xFFFFFFFFull;
        DevBuf<UnionSrc> d_src(J);
        to_device(d_src.get(), src.data(), J, s);
        B.d_voff0.resize(J + 1);
        to_device(B.d_voff0.get(), B.voff0.data(), J + 1, s);
        hipLaunchKernelGGL(k_union_nodes, dim3(64, J), dim3(256), 0, s, (const UnionSrc *)d_src.get(), U.ptrs.get(),
                           U.node_w.get(), U.strength.get(), U.self_loop.get(), U.row_total.get(), B.vjob0.get());
        KERNEL_OK();
        hipLaunchKernelGGL(k_union_edges, dim3(256, J), dim3(256), 0, s, (const UnionSrc *)d_src.get(), U.nbrs.get(),
                           U.wts.get());
        KERNEL_OK();
        to_device(U.ptrs.get() + N, &E, 1, s);
        sync_stream(s);  // host tables above are temporaries
        sc_leiden_config c = *cfg;
        c.resolution = resolutions[0];
        DevBuf<uint32_t> labels(N);
        uint64_t nc = 0;
        sc_leiden_stats st{};
        leiden_device(U, c, labels.get(), &nc, s, &st, &B);
        // per job: labels are consecutive over the union and contiguous per job
        std::vector<uint32_t> h(N);
        to_host(h.data(), labels.get(), N, s);
        sync_stream(s);
        for (uint32_t j = 0; j < J; j++) {
            const uint64_t lo = B.voff0[j], hi = B.voff0[j + 1];
            uint32_t mn = 0xFFFFFFFFu, mx = 0;
            for (uint64_t v = lo; v < hi; v++) {
                mn = std::min(mn, h[v]);
                mx = std::max(mx, h[v]);
            }
            for (uint64_t v = lo; v < hi; v++) labels_out[v] = (uint64_t)(h[v] - mn);
            n_clusters_out[j] = hi > lo ? (uint64_t)(mx - mn) + 1 : 0;
        }
        t_stats = st;
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

sc_leiden_config sc_leiden_config_default(void) {
    sc_leiden_config c;
    std::memset(&c, 0, sizeof(c));
    c.objective = SC_RB;
    c.resolution = 1.0;
    c.has_seed = 1;
    c.seed = 42;
    c.n_iterations = 2;
    c.max_levels = 64;
    c.tolerance = 1e-9;
    c.refine = 1;
    c.refine_randomness = 0.01;
    c.patience = 2;
    c.has_max_community_weight = 0;
    c.max_community_weight = 0.0;
    return c;
}

sc_status sc_leiden(const sc_graph *g, const sc_leiden_config *cfg, uint64_t *labels_out, uint64_t *n_clusters_out,
                    sc_error *err) {
    try {
        if (!g || !cfg || !n_clusters_out || (!labels_out && g->n)) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        check_node_count(g->n);
        hipStream_t s = thread_stream();
        uint32_t n = (uint32_t)g->n;
        DevBuf<uint32_t> labels(n);
        DevBuf<uint64_t> l64(n);
        sc_leiden_stats st{};
        leiden_device(*g, *cfg, labels.get(), n_clusters_out, s, &st);
        if (n) {
            hipLaunchKernelGGL(k_u32_to_u64, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, labels.get(), l64.get(), n);
            KERNEL_OK();
            to_host(labels_out, l64.get(), n, s);
        }
        sync_stream(s);
        t_stats = st;
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

sc_status sc_set_sweep_mode(int32_t mode) {
    if (mode < 0 || mode > 2) return SC_INVALID_ARGUMENT;
    g_sweep_mode.store(mode, std::memory_order_relaxed);
    return SC_OK;
}

sc_status sc_leiden_last_stats(sc_leiden_stats *out) {
    if (!out) return SC_INVALID_ARGUMENT;
    *out = t_stats;
    return SC_OK;
}

sc_status sc_graph_aggregate(const sc_graph *g, const uint64_t *groups, sc_graph **out, sc_error *err) {
    try {
        if (!g || !out || (!groups && g->n)) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        uint64_t n = g->n;
        uint64_t maxg = 0;
        for (uint64_t v = 0; v < n; v++) maxg = std::max(maxg, groups[v]);
        if (n && maxg >= 0xFFFFFFFFull) throw ScError(SC_INVALID_ARGUMENT, "group ids must be below 2^32 - 1");
        // VectorGrouping::from_assignments normalizes ids to consecutive, ascending
        std::vector<uint32_t> g32(n);
        for (uint64_t v = 0; v < n; v++) g32[v] = (uint32_t)groups[v];
        hipStream_t s = thread_stream();
        DevBuf<uint32_t> d(n), dn(n);
        to_device(d.get(), g32.data(), n, s);
        uint32_t ng = n ? renumber(d.get(), (uint32_t)n, (uint32_t)(maxg + 1), dn.get(), s) : 0;
        std::unique_ptr<sc_graph> c;
        if (n) {
            c = aggregate_device(*g, dn.get(), ng, s);
        } else {
            c = std::make_unique<sc_graph>();
            c->ptrs.resize(1);
            HIP_OK(hipMemsetAsync(c->ptrs.get(), 0, sizeof(uint64_t), s));
        }
        sync_stream(s);
        *out = c.release();
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

sc_status sc_renumber(const uint64_t *labels, uint64_t n, uint64_t *out, uint64_t *n_groups_out, sc_error *err) {
    try {
        if ((!labels || !out) && n) throw ScError(SC_INVALID_ARGUMENT, "null argument");
        check_node_count(n);
        uint64_t maxl = 0;
        for (uint64_t v = 0; v < n; v++) maxl = std::max(maxl, labels[v]);
        if (n && maxl >= 0xFFFFFFFFull) throw ScError(SC_INVALID_ARGUMENT, "labels must be below 2^32 - 1");
        uint32_t ng = 0;
        if (n) {
            std::vector<uint32_t> l32(n);
            for (uint64_t v = 0; v < n; v++) l32[v] = (uint32_t)labels[v];
            hipStream_t s = thread_stream();
            DevBuf<uint32_t> d(n);
            to_device(d.get(), l32.data(), n, s);
            // the device compaction every level of leiden() uses, in place
            ng = renumber(d.get(), (uint32_t)n, (uint32_t)(maxl + 1), d.get(), s);
            to_host(l32.data(), d.get(), n, s);
            sync_stream(s);
            for (uint64_t v = 0; v < n; v++) out[v] = l32[v];
        }
        if (n_groups_out) *n_groups_out = ng;
        return SC_OK;
    } catch (...) {
        return status_from_exception(err);
    }
}

}  // extern "C"
