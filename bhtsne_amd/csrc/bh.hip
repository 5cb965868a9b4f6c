// Barnes-Hut repulsion on the device (src/tsne/arena.rs:137-392, src/tsne/morton.rs:103-120,
// src/tsne/morton/d2.rs:8-47, src/repulsion.rs:95-158), D = 2.
//
//  1. Morton codes of the quantised points (32 bits per axis, u64 words) from the per-axis
//     bounding box, with the reference's f32 quantisation.
//  2. Stable LSD radix sort of (code, index): ties keep index order.
//  3. Breadth-first emission of the compressed quadtree: every round expands the current
//     frontier in node order (children located by searches of the sorted codes, a block scan
//     assigns their slots), the top rounds in one workgroup per job and the subtrees below them in
//     several (k_bh_top, k_bh_sub); then centres of mass bottom-up, round by round, with the
//     reference's f32 summation order.
//  4. Each point's walk in the reference's visit order, a wave walking the union of the walks
//     of 64 consecutive points of the Morton order (k_bh_traverse).  Z is summed in FP64.  The
//     reference pops node ids from a stack (up to 128 at D = 2); the walk keeps a stack of child
//     ranges (first child, children left), packed in a u32: taking the last child left of the top
//     range is the node the reference pops next.  A range is held per internal node on the
//     current path, whose levels increase strictly, so at most 32 are saved.  An internal node's
//     squared half-width is stored beside its child count, so a visit issues its two loads at
//     once.
#include <cfloat>

#include "internal.hpp"

namespace bht {

constexpr uint32_t SENTINEL = 0xffffffffu;
constexpr int BITS = 32;
constexpr int BUILD_THREADS = 1024;
constexpr int TRAV_THREADS = 256;  // four independent waves per walk block
constexpr int MAX_ROUNDS = 64;

__device__ __forceinline__ uint64_t part1by1(uint64_t x) {
    x &= 0x00000000ffffffffull;
    x = (x | (x << 16)) & 0x0000ffff0000ffffull;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    return (x | (x << 1)) & 0x5555555555555555ull;
}

__device__ __forceinline__ uint32_t quantize(float p, float mn, float inv_scale) {
    float scaled = floorf((p - mn) * inv_scale);
    if (scaled > 0.0f) {
        if (scaled >= 18446744073709551616.0f) return 0xffffffffu;
        uint64_t v = (uint64_t)scaled;
        return v < 0xffffffffull ? (uint32_t)v : 0xffffffffu;
    }
    return 0u;
}

__global__ void __launch_bounds__(256) k_bh_codes(const JobDev *jobs, const uint32_t *map) {
    const BlockCtx bc = block_ctx(map, jobs, F_PT);
    const JobDev &J = jobs[bc.job];
    const EpochScalars *scal = J.scal;
    const float *y = job_y(J);
    uint64_t *codes = J.codes;
    uint32_t *idx = J.cidx;
    const int n = J.n;
    const float scale = 4294967296.0f;  // 2^32 as f32
    float mn[2], inv[2];
#pragma unroll
    for (int a = 0; a < 2; a++) {
        mn[a] = scal->amin[a];
        float extent = scal->amax[a] - scal->amin[a];
        inv[a] = extent > 0.0f ? scale / extent : 0.0f;
    }
    for (int i = bc.lb * 256 + threadIdx.x; i < n; i += bc.nb * 256) {
        float2 p = reinterpret_cast<const float2 *>(y)[i];
        uint32_t q0 = quantize(p.x, mn[0], inv[0]);
        uint32_t q1 = quantize(p.y, mn[1], inv[1]);
        codes[i] = part1by1(q0) | (part1by1(q1) << 1);
        idx[i] = (uint32_t)i;
    }
}

// block-wide exclusive scan of one u32 per thread (THREADS threads): a shuffle scan in each
// wave, then every thread adds the totals of the waves before its own
template <int THREADS = BUILD_THREADS>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *buf, uint32_t &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(x, off, 64);
        if (lane >= off) x += t;
    }
    if (lane == 63) buf[w] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < THREADS / 64; k++) {
        const uint32_t c = buf[k];
        pre += k < w ? c : 0u;
        tot += c;
    }
    total = tot;
    __syncthreads();
    return pre + x - v;
}

__device__ __forceinline__ uint32_t lower_bound_digit(const uint64_t *codes, uint32_t lo,
                                                      uint32_t hi, int shift, uint32_t g) {
    // first position in [lo, hi) whose digit at `shift` is >= g (digits are nondecreasing)
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if ((uint32_t)((codes[mid] >> shift) & 3ull) < g) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// First position in [lo, hi) whose digit at `shift` is >= g, or hi (digits nondecreasing there),
// found by one wave: 64 probes per round narrow the range 64-fold, so a range of a million codes
// takes four rounds of one load each instead of twenty dependent loads.  lo and hi stay
// wave-uniform (they follow ballots).
__device__ __forceinline__ uint32_t wave_lower_bound(const uint64_t *codes, uint32_t lo,
                                                     uint32_t hi, int shift, uint32_t g,
                                                     int lane) {
    while (hi - lo > 64) {
        const uint32_t step = (hi - lo + 63) / 64;
        const uint32_t p = lo + (uint32_t)lane * step;
        const bool ge = p >= hi || (uint32_t)((codes[p] >> shift) & 3ull) >= g;
        const uint64_t m = __ballot(ge);
        if (m == 0ull) {  // every probe is below g: the answer is past the last one
            lo = lo + 63u * step + 1u;
            continue;
        }
        const int j = __builtin_ctzll(m);
        if (j == 0) return lo;
        const uint32_t nhi = lo + (uint32_t)j * step;
        lo = lo + (uint32_t)(j - 1) * step + 1u;
        hi = nhi < hi ? nhi : hi;
    }
    const uint32_t p = lo + (uint32_t)lane;
    const bool ge = p >= hi || (uint32_t)((codes[p] >> shift) & 3ull) >= g;
    const uint64_t m = __ballot(ge);
    return m ? lo + (uint32_t)__builtin_ctzll(m) : hi;
}

// The tree is built in three launches, the same nodes and children (in digit order) as the
// reference's breadth-first emission (arena.rs:197-309):
//  1. k_bh_top, one workgroup per job: breadth-first rounds from the root, expanding only nodes
//     of more than SUB_POINTS points; an internal node of at most SUB_POINTS points is deferred
//     (listed, its children not yet emitted).  Rounds of at most 16 nodes search a node's child
//     bounds with a whole wave (wave_lower_bound), larger rounds with a thread per node.  These
//     nodes take indices [0, 2n) in round order.
//  2. k_bh_sub, several workgroups per job: each takes deferred roots and emits the root's
//     subtree breadth-first into a region of its own, [2n + 2s, 2n + 2e) for a root over sorted
//     points [s, e) (a subtree over m points has at most 2m - 1 nodes, and the ranges of the
//     deferred roots are disjoint), then the subtree's centres of mass bottom-up.
//  3. k_bh_top_com, one workgroup per job: the centres of mass of the top nodes, bottom-up.
// The walk reads each node's first child and child count, so the numbering (which differs from
// the reference's below the top) changes no sum; the children of a node are contiguous and in
// digit order as in the reference.
constexpr uint32_t SUB_POINTS = 2048;
constexpr int CNT_TOP = 0, CNT_ROUNDS = 1, CNT_DEFER = 2, CNT_RSTART = 8;  // counts[] layout

// A subtree round rarely holds more than a few hundred nodes: 256 threads per subtree block.  A
// job takes a block per 512 points (the deferred roots hold a few hundred to SUB_POINTS points
// each, so most blocks build one subtree), at most 1024.
constexpr int SUB_THREADS = 256;

int bh_sub_blocks(int n) {
    return (int)std::min<uint64_t>(1024, std::max<uint64_t>(1, ceil_div((uint64_t)(n > 0 ? n : 1),
                                                                        512)));
}

// Children of an internal node over sorted points [s, e) at `shift`: the bounds b[0..4] of the
// digit groups 0..3 (thread-level binary searches).
__device__ __forceinline__ uint32_t child_bounds(const uint64_t *codes, uint32_t s, uint32_t e,
                                                 int shift, uint32_t *b) {
    b[0] = s;
    b[4] = e;
    for (uint32_t g = 1; g < 4; g++) b[g] = lower_bound_digit(codes, s, e, shift, g);
    uint32_t nchild = 0;
    for (int g = 0; g < 4; g++) nchild += b[g + 1] > b[g] ? 1u : 0u;
    return nchild;
}

// Emits the children of node `i` (digit groups b[0..4]) at indices first, first + 1, ... and
// records the node's first child and meta.
__device__ __forceinline__ void emit_children(const BhNodes &t, uint32_t i, uint32_t first,
                                              uint32_t nchild, int level, const float *shw2,
                                              const uint32_t *b) {
    float4 nd = t.node[i];
    nd.w = __uint_as_float(first);
    t.node[i] = nd;
    t.meta[i] = make_uint2(nchild | ((uint32_t)level << 8), __float_as_uint(shw2[level]));
    uint32_t c = first;
    for (int g = 0; g < 4; g++) {
        if (b[g + 1] <= b[g]) continue;
        t.node[c] = make_float4(0.f, 0.f, __uint_as_float(b[g + 1] - b[g]),
                                __uint_as_float(SENTINEL));
        t.meta[c] = make_uint2((uint32_t)BITS << 8, 0u);
        t.range[2 * c] = b[g];
        t.range[2 * c + 1] = b[g + 1];
        c++;
    }
}

// Centre of mass of node i from its members (a leaf) or its children (arena.rs:261-309).
__device__ __forceinline__ void node_com(const BhNodes &t, uint32_t i, const float *y,
                                         const uint32_t *sidx) {
    float4 nd = t.node[i];
    const uint32_t count = __float_as_uint(nd.z);
    const uint32_t fc = __float_as_uint(nd.w);
    float cx = 0.0f, cy = 0.0f;
    if (fc == SENTINEL) {
        for (uint32_t s = t.range[2 * i]; s < t.range[2 * i + 1]; s++) {
            float2 p = reinterpret_cast<const float2 *>(y)[sidx[s]];
            cx += p.x;
            cy += p.y;
        }
    } else {
        const uint32_t cc = t.meta[i].x & 0xffu;
        for (uint32_t ch = fc; ch < fc + cc; ch++) {
            float4 cn = t.node[ch];
            const float w = (float)__float_as_uint(cn.z);
            cx += cn.x * w;
            cy += cn.y * w;
        }
    }
    const float inv = 1.0f / (float)count;
    nd.x = cx * inv;
    nd.y = cy * inv;
    t.node[i] = nd;
}

__global__ void __launch_bounds__(BUILD_THREADS) k_bh_top(const JobDev *jobs,
                                                          const uint32_t *map) {
    __shared__ uint32_t sbuf[BUILD_THREADS / 64];
    __shared__ uint32_t round_start[MAX_ROUNDS + 2];
    __shared__ float shw2[BITS + 1];
    __shared__ uint32_t ndef;
    const BlockCtx bc = block_ctx(map, jobs, F_ONE);
    const JobDev &J = jobs[bc.job];
    const int n = J.n;
    const uint64_t *codes = J.scodes;
    EpochScalars *scal = J.scal;
    const BhNodes t = J.nodes;
    const int tid = threadIdx.x, lane = tid & 63;
    if (tid == 0) {
        // level half-widths from the largest axis extent (arena.rs:167-174)
        float e0 = scal->amax[0] - scal->amin[0];
        float e1 = scal->amax[1] - scal->amin[1];
        float max_extent = fmaxf(fmaxf(0.0f, e0), e1);
        for (int level = 0; level <= BITS; level++) {
            float hw = max_extent / (float)(1ull << (level + 1));
            t.hw2[level] = hw * hw;
            shw2[level] = hw * hw;
        }
        t.node[0] = make_float4(0.f, 0.f, __uint_as_float((uint32_t)n), __uint_as_float(SENTINEL));
        t.meta[0] = make_uint2((uint32_t)BITS << 8, 0u);
        t.range[0] = 0;
        t.range[1] = (uint32_t)n;
        round_start[0] = 0;
        round_start[1] = 1;
        ndef = 0;
    }
    __syncthreads();
    if (n == 0) {
        if (tid == 0) {
            t.counts[CNT_TOP] = 0;
            t.counts[CNT_ROUNDS] = 0;
            t.counts[CNT_DEFER] = 0;
        }
        return;
    }
    // every thread keeps the node count, which each chunk's scan total advances alike
    uint32_t nn = 1;
    int rounds = 1;
    for (int r = 0; r < MAX_ROUNDS; r++) {
        const uint32_t f0 = round_start[r], f1 = round_start[r + 1];
        if (f0 == f1) break;
        if (f1 - f0 <= (uint32_t)(BUILD_THREADS / 64)) {
            // a wave per node
            const uint32_t i = f0 + (uint32_t)(tid >> 6);
            uint32_t nchild = 0, b[5] = {0, 0, 0, 0, 0};
            int level = BITS;
            bool expand = false;
            if (i < f1) {
                const uint32_t s = t.range[2 * i], e = t.range[2 * i + 1];
                if (e - s > 1 && codes[s] != codes[e - 1]) {
                    const uint64_t x = codes[s] ^ codes[e - 1];
                    const int highest = 63 - __clzll((long long)x);
                    level = (BITS - 1) - highest / 2;
                    const int shift = 2 * (BITS - 1 - level);
                    if (e - s > SUB_POINTS) {
                        expand = true;
                        b[0] = s;
                        b[4] = e;
                        for (uint32_t g = 1; g < 4; g++)
                            b[g] = wave_lower_bound(codes, s, e, shift, g, lane);
                        for (int g = 0; g < 4; g++) nchild += b[g + 1] > b[g] ? 1u : 0u;
                    } else if (lane == 0) {
                        t.defer[atomicAdd(&ndef, 1u)] = i;
                    }
                }
            }
            uint32_t total;
            const uint32_t off = block_excl_scan(lane == 0 ? nchild : 0u, sbuf, total);
            if (expand && lane == 0) emit_children(t, i, nn + off, nchild, level, shw2, b);
            nn += total;
        } else {
            for (uint32_t base = f0; base < f1; base += BUILD_THREADS) {
                const uint32_t i = base + tid;
                uint32_t nchild = 0, b[5] = {0, 0, 0, 0, 0};
                int level = BITS;
                bool expand = false;
                if (i < f1) {
                    const uint32_t s = t.range[2 * i], e = t.range[2 * i + 1];
                    if (e - s > 1 && codes[s] != codes[e - 1]) {
                        const uint64_t x = codes[s] ^ codes[e - 1];
                        const int highest = 63 - __clzll((long long)x);
                        level = (BITS - 1) - highest / 2;
                        if (e - s > SUB_POINTS) {
                            expand = true;
                            nchild = child_bounds(codes, s, e, 2 * (BITS - 1 - level), b);
                        } else {
                            t.defer[atomicAdd(&ndef, 1u)] = i;
                        }
                    }
                }
                uint32_t total;
                const uint32_t off = block_excl_scan(nchild, sbuf, total);
                if (expand) emit_children(t, i, nn + off, nchild, level, shw2, b);
                nn += total;
            }
        }
        // the round's nodes and its end are visible to the whole block from here on
        if (tid == 0) round_start[r + 2] = nn;
        rounds = r + 1;
        __syncthreads();
    }
    if (tid == 0) {
        t.counts[CNT_TOP] = nn;
        t.counts[CNT_ROUNDS] = (uint32_t)rounds;
        t.counts[CNT_DEFER] = ndef;
    }
    for (int r = tid; r <= rounds; r += BUILD_THREADS) t.counts[CNT_RSTART + r] = round_start[r];
}

// The subtree of one deferred root is built out of LDS: the root's sorted codes (at most
// SUB_POINTS) and its nodes' point ranges, relative to the root's first point (at most
// 2 SUB_POINTS - 1 nodes), so that the child-bound searches and the range reads of every round
// cost LDS latency rather than a global round trip per probe.  Nodes, meta and ranges still go to
// the job's arrays for the walk and the centres of mass.
__global__ void __launch_bounds__(SUB_THREADS) k_bh_sub(const JobDev *jobs,
                                                          const uint32_t *map) {
    __shared__ uint32_t sbuf[SUB_THREADS / 64];
    __shared__ uint32_t round_start[MAX_ROUNDS + 2];
    __shared__ float shw2[BITS + 1];
    __shared__ uint64_t lcodes[SUB_POINTS];
    __shared__ uint32_t lrange[2 * (2 * SUB_POINTS)];
    const BlockCtx bc = block_ctx(map, jobs, F_BSUB);
    const JobDev &J = jobs[bc.job];
    const float *y = job_y(J);
    const uint32_t n = (uint32_t)J.n;
    const uint64_t *codes = J.scodes;
    const uint32_t *sidx = J.scidx;
    const BhNodes t = J.nodes;
    const int tid = threadIdx.x;
    if (tid <= BITS) shw2[tid] = t.hw2[tid];
    const uint32_t ndef = t.counts[CNT_DEFER];
    __syncthreads();
    for (uint32_t q = (uint32_t)bc.lb; q < ndef; q += (uint32_t)bc.nb) {
        const uint32_t root = t.defer[q];
        const uint32_t s0 = t.range[2 * root];
        const uint32_t m = t.range[2 * root + 1] - s0;  // 2 .. SUB_POINTS (a deferred root)
        const uint32_t base = 2u * n + 2u * s0 - 1u;    // region of local node k >= 1
        auto gid = [&](uint32_t k) { return k == 0 ? root : base + k; };
        for (uint32_t k = tid; k < m; k += SUB_THREADS) lcodes[k] = codes[s0 + k];
        if (tid == 0) {
            round_start[0] = 0;
            round_start[1] = 1;
            lrange[0] = 0;
            lrange[1] = m;
        }
        __syncthreads();
        uint32_t nn = 1;
        int rounds = 1;
        for (int r = 0; r < MAX_ROUNDS; r++) {
            const uint32_t f0 = round_start[r], f1 = round_start[r + 1];
            if (f0 == f1) break;
            for (uint32_t kb = f0; kb < f1; kb += SUB_THREADS) {
                const uint32_t k = kb + tid;
                uint32_t nchild = 0, b[5] = {0, 0, 0, 0, 0};
                int level = BITS;
                bool internal = false;
                if (k < f1) {
                    const uint32_t s = lrange[2 * k], e = lrange[2 * k + 1];
                    if (e - s > 1 && lcodes[s] != lcodes[e - 1]) {
                        internal = true;
                        const uint64_t x = lcodes[s] ^ lcodes[e - 1];
                        const int highest = 63 - __clzll((long long)x);
                        level = (BITS - 1) - highest / 2;
                        nchild = child_bounds(lcodes, s, e, 2 * (BITS - 1 - level), b);
                    }
                }
                uint32_t total;
                const uint32_t off = block_excl_scan<SUB_THREADS>(nchild, sbuf, total);
                if (internal) {
                    // the node (count = its range) with its first child, then the children, in
                    // digit order, at local indices nn + off, ...
                    const uint32_t i = gid(k), kc0 = nn + off;
                    t.node[i] = make_float4(0.f, 0.f, __uint_as_float(b[4] - b[0]),
                                            __uint_as_float(gid(kc0)));
                    t.meta[i] = make_uint2(nchild | ((uint32_t)level << 8),
                                           __float_as_uint(shw2[level]));
                    uint32_t kc = kc0;
                    for (int g = 0; g < 4; g++) {
                        if (b[g + 1] <= b[g]) continue;
                        const uint32_t c = gid(kc);
                        t.node[c] = make_float4(0.f, 0.f, __uint_as_float(b[g + 1] - b[g]),
                                                __uint_as_float(SENTINEL));
                        t.meta[c] = make_uint2((uint32_t)BITS << 8, 0u);
                        t.range[2 * c] = s0 + b[g];
                        t.range[2 * c + 1] = s0 + b[g + 1];
                        lrange[2 * kc] = b[g];
                        lrange[2 * kc + 1] = b[g + 1];
                        kc++;
                    }
                }
                nn += total;
            }
            if (tid == 0) round_start[r + 2] = nn;
            rounds = r + 1;
            __syncthreads();
        }
        for (int r = rounds - 1; r >= 0; r--) {
            const uint32_t f0 = round_start[r], f1 = round_start[r + 1];
            for (uint32_t k = f0 + tid; k < f1; k += SUB_THREADS) node_com(t, gid(k), y, sidx);
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(BUILD_THREADS) k_bh_top_com(const JobDev *jobs,
                                                              const uint32_t *map) {
    const BlockCtx bc = block_ctx(map, jobs, F_ONE);
    const JobDev &J = jobs[bc.job];
    const float *y = job_y(J);
    const uint32_t *sidx = J.scidx;
    const BhNodes t = J.nodes;
    const int rounds = (int)t.counts[CNT_ROUNDS];
    for (int r = rounds - 1; r >= 0; r--) {
        const uint32_t f0 = t.counts[CNT_RSTART + r], f1 = t.counts[CNT_RSTART + r + 1];
        for (uint32_t i = f0 + threadIdx.x; i < f1; i += BUILD_THREADS) {
            // deferred roots (internal, at most SUB_POINTS points) have their centres already
            const float4 nd = t.node[i];
            if (__float_as_uint(nd.w) != SENTINEL && __float_as_uint(nd.z) <= SUB_POINTS) continue;
            node_com(t, i, y, sidx);
        }
        __syncthreads();
    }
}

// The walk.  A wave takes 64 consecutive points of the Morton order and walks the union of their
// walks in the reference's visit order, one node at a time: every lane reads the same node, and a
// lane evaluates it only while it is active there, that is, when it opened every ancestor.  An
// active lane sums the node (a leaf other than its own, or an internal node far enough away) or
// opens it; the wave descends when any lane opens, and the lanes that did not are muted below
// that node until the walk is back at its depth.  Each lane therefore sums exactly the nodes of
// its own walk, in the same order and with the same arithmetic, while the wave issues one scalar
// load per node instead of 64 scattered ones.  The range stack is wave-uniform: entry k (the
// siblings still to visit and their depth) sits in lane k of two registers.  The four waves of a
// block walk independently.
#define BHT_CONST __attribute__((address_space(4)))
typedef float trav_f4 __attribute__((ext_vector_type(4)));
typedef uint32_t trav_u2 __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(TRAV_THREADS) k_bh_traverse(const JobDev *jobs,
                                                              const uint32_t *map) {
    __shared__ double red[TRAV_THREADS / 64];
    const BlockCtx bc = block_ctx(map, jobs, F_TRAV);
    const JobDev &J = jobs[bc.job];
    const float *y = job_y(J);
    const int n = J.n;
    const uint32_t *sidx = J.scidx;
    const BhNodes t = J.nodes;
    // the tree was written by earlier launches and is read-only here: scalar loads
    BHT_CONST const trav_f4 *node = (BHT_CONST const trav_f4 *)(const void *)t.node;
    BHT_CONST const trav_u2 *meta = (BHT_CONST const trav_u2 *)(const void *)t.meta;
    const float theta_sq = J.theta_sq;
    float *frep = J.frep;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double zacc = 0.0;
    for (int s0 = bc.lb * TRAV_THREADS + wid * 64; s0 < n; s0 += bc.nb * TRAV_THREADS) {
        const int s = s0 + lane;
        const bool valid = s < n;
        const uint32_t i = valid ? sidx[s] : 0u;
        const float2 qp = valid ? reinterpret_cast<const float2 *>(y)[i] : make_float2(0.f, 0.f);
        float fx = 0.0f, fy = 0.0f, q_sum = 0.0f;
        int mute = valid ? INT_MAX : -1;  // the lane is active at depth d iff d < mute
        uint32_t first = 0u, left = 1u;   // the range walked (the root), visited last node first
        int dep = 0;                      // its depth
        int below = 0;  // saved ranges (first << 3 | left), in lanes [0, below) of stk_range
        int stk_range = 0, stk_dep = 0;
        while (true) {
            if (left == 0u) {
                if (below == 0) break;
                below--;
                const uint32_t c = (uint32_t)__builtin_amdgcn_readlane(stk_range, below);
                first = c >> 3;
                left = c & 7u;
                dep = __builtin_amdgcn_readlane(stk_dep, below);
                continue;
            }
            const uint32_t id = first + left - 1u;
            left -= 1u;
            const trav_f4 nd = node[id];
            const trav_u2 mt = meta[id];
            const bool act = dep < mute;
            const float dx = qp.x - nd.x, dy = qp.y - nd.y;
            float distance = 0.0f;
            distance += dx * dx;
            distance += dy * dy;
            const uint32_t fc = __float_as_uint(nd.w);
            const bool leaf = fc == SENTINEL;
            const bool open = act && !leaf && __uint_as_float(mt.y) >= theta_sq * distance;
            // a leaf at distance 0 is the query's own
            if (act && !open && !(leaf && distance == 0.0f)) {
                const float inverse = 1.0f / (1.0f + distance);
                float magnitude = (float)__float_as_uint(nd.z) * inverse;
                q_sum += magnitude;
                magnitude = magnitude * inverse;
                fx += magnitude * dx;
                fy += magnitude * dy;
            }
            if (__builtin_amdgcn_ballot_w64(open) != 0ull) {
                // descend; the siblings not yet visited wait with their depth (at most one range
                // per internal level on the path: 32 < 64 lanes)
                if (left > 0u) {
                    if (lane == below) {
                        stk_range = (int)((first << 3) | left);
                        stk_dep = dep;
                    }
                    below++;
                }
                if (act) mute = open ? INT_MAX : dep + 1;
                first = fc;
                left = mt.x & 0xffu;
                dep++;
            }
        }
        if (valid) {
            reinterpret_cast<float2 *>(frep)[i] = make_float2(fx, fy);
            zacc += (double)q_sum;
        }
    }
    const double zb = block_sum_f64(zacc, red);
    if (threadIdx.x == 0) J.dpart[bc.lb] = zb;
}

int traverse_blocks(int n, int dim) {
    // D >= 3 walks one wave per block (bhnd.hip)
    const int threads = dim == 2 ? TRAV_THREADS : 64;
    return (int)std::min<uint64_t>(ceil_div((uint64_t)(n > 0 ? n : 1), threads), 4096);
}

const uint32_t *bh_repulsion(const LaunchSet &L, hipStream_t s) {
    if (L.dim >= 3) return nd_bh_repulsion(L, s);
    hipLaunchKernelGGL(k_bh_codes, dim3(L.nb[F_PT]), dim3(256), 0, s, L.jobs, L.map[F_PT]);
    BHT_LAUNCH_CHECK();
    const bool alt = sort_group<uint64_t>(L, MORTON_KEY_BITS, s);
    hipLaunchKernelGGL(k_bh_top, dim3(L.nb[F_ONE]), dim3(BUILD_THREADS), 0, s, L.jobs,
                       L.map[F_ONE]);
    hipLaunchKernelGGL(k_bh_sub, dim3(L.nb[F_BSUB]), dim3(SUB_THREADS), 0, s, L.jobs,
                       L.map[F_BSUB]);
    hipLaunchKernelGGL(k_bh_top_com, dim3(L.nb[F_ONE]), dim3(BUILD_THREADS), 0, s, L.jobs,
                       L.map[F_ONE]);
    hipLaunchKernelGGL(k_bh_traverse, dim3(L.nb[F_TRAV]), dim3(TRAV_THREADS), 0, s, L.jobs,
                       L.map[F_TRAV]);
    BHT_LAUNCH_CHECK();
    // Z is the plain sum of the per-point q sums (repulsion.rs:155)
    finalize_z(L, F_TRAV, s);
    return alt ? L.vals_alt : L.vals;  // point indices of job 0 in Morton order
}

}  // namespace bht
