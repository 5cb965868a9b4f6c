// Barnes-Hut epochs of embeddings with D = 3..7 coordinates per point: the generic arena
// (src/tsne/arena.rs:137-392), its Z-order codes (src/tsne/morton.rs:1-120, morton/d3.rs ..
// d7.rs), the repulsion step (src/repulsion.rs:95-158), the attraction (src/tsne/arena.rs:398-427),
// the gradient step and the recentring (src/tsne/mod.rs:470-568).
//
//  * Codes: D * BITS(D) bits (BITS = 21 / 16 / 25 / 21 / 18 per axis), one u64 word for D <= 4
//    and two for D >= 5.  The stable LSD radix sort of (code, index) keeps index order among equal
//    codes; a two-word code is sorted by its low word and then, stably, by its high word.
//  * Tree: the compressed 2^D-ary tree in breadth-first order, built by one workgroup per job with
//    the reference's node numbering.  A node's children are the runs of equal D-bit digits of its
//    range at its level, located one after another by binary search.  Centres of mass bottom-up,
//    round by round, in the reference's f32 summation order.
//  * Walk: one lane per point (points in Z-order, so a wave's walks share their paths).  The
//    reference pops node ids from a stack of up to 2304 entries (D = 7); here a lane keeps a stack
//    of child ranges (first child, children left).  Taking the last child left of the top range is
//    the node the reference pops next, so the visit order and every f32 sum are the same.  A range
//    is held per internal node on the current path, which descends strictly in level, so the stack
//    holds at most BITS(D) ranges.  Z is summed in FP64.
//  * Attraction: one thread per row, its edges in CSR order (the reference's per-row sum), over a
//    copy of P's values in CSR order that the first epoch after stop_lying_epoch rescales in place.
//  * Step and recentring: as for D <= 2 (FP64 recentring sums), with D-wide partials.
#include <cfloat>

#include "internal.hpp"

namespace bht {

namespace {

constexpr uint32_t SENTINEL = 0xffffffffu;
constexpr int ND_THREADS = 256;      // point-parallel kernels (F_PT, F_ATT) and the finalisers
constexpr int ND_BUILD_THREADS = 1024;
constexpr int ND_TRAV_THREADS = 64;  // F_TRAV blocks: one wave
constexpr int ND_MAX_ROUNDS = 64;    // tree depth is at most BITS + 1

template <int D>
struct Nd {
    static constexpr int BITS = D == 3 ? 21 : D == 4 ? 16 : D == 5 ? 25 : D == 6 ? 21 : 18;
    static constexpr bool WIDE = D >= 5;              // two-word codes
    static constexpr int S4 = (D + 2 + 3) / 4;        // float4s per node: com[D], count, first
    static constexpr int STACK = BITS + 1;            // child ranges per lane
    static constexpr uint32_t MAX_BUCKET = (1u << BITS) - 1u;
};

// Node i occupies S4 float4s: com[0..D-1], the count and the first child (u32 bits; ~0 = leaf).
template <int D>
__device__ __forceinline__ float *node_ptr(float4 *base, uint32_t i) {
    return reinterpret_cast<float *>(base + (size_t)i * Nd<D>::S4);
}

__device__ __forceinline__ uint32_t quantize_axis(float p, float mn, float inv_scale,
                                                  uint32_t max_bucket) {  // morton.rs:103-120
    const float scaled = floorf((p - mn) * inv_scale);
    if (scaled > 0.0f) {
        if (scaled >= 18446744073709551616.0f) return max_bucket;
        const uint64_t v = (uint64_t)scaled;
        return v < (uint64_t)max_bucket ? (uint32_t)v : max_bucket;
    }
    return 0u;
}

// Z-order interleave: bit b of axis a goes to bit b * D + a of the code (lo: bits 0..63).
template <int D>
__device__ __forceinline__ void encode(const uint32_t *q, uint64_t &lo, uint64_t &hi) {
    lo = 0;
    hi = 0;
#pragma unroll
    for (int b = 0; b < Nd<D>::BITS; b++) {
#pragma unroll
        for (int a = 0; a < D; a++) {
            const int pos = b * D + a;
            const uint64_t bit = (uint64_t)((q[a] >> b) & 1u);
            if (pos < 64) lo |= bit << pos;
            else hi |= bit << (pos - 64);
        }
    }
}

// The sorted codes of a job: the code (D <= 4) or its high word (D >= 5) in w0, the low word in w1.
template <int D>
struct Codes {
    const uint64_t *w0, *w1;
    __device__ __forceinline__ void get(uint32_t s, uint64_t &hi, uint64_t &lo) const {
        if (Nd<D>::WIDE) {
            hi = w0[s];
            lo = w1[s];
        } else {
            hi = 0;
            lo = w0[s];
        }
    }
    __device__ __forceinline__ uint32_t digit(uint32_t s, int shift) const {
        uint64_t hi, lo;
        get(s, hi, lo);
        uint64_t v;
        if (shift >= 64) v = hi >> (shift - 64);
        else if (shift == 0) v = lo;
        else v = (lo >> shift) | (hi << (64 - shift));
        return (uint32_t)(v & ((1ull << D) - 1ull));
    }
    // first position in [lo, hi) whose digit exceeds g (digits are nondecreasing there)
    __device__ __forceinline__ uint32_t upper(uint32_t lo, uint32_t hi, int shift, uint32_t g) const {
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (digit(mid, shift) <= g) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    }
};

template <int D>
__device__ __forceinline__ Codes<D> sorted_codes(const JobDev &J) {
    return Codes<D>{J.scodes, J.slo};
}

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
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
        const uint32_t c = buf[k];
        pre += k < w ? c : 0u;
        tot += c;
    }
    total = tot;
    __syncthreads();
    return pre + x - v;
}

// ---- recentring and bounds (zero_mean, mod.rs:534-568; bounding_box, arena.rs:65-94) ----------

// The mean of the last update's per-block sums (left pending by k_nd_update); advances the job's
// epoch index, which selects its current embedding buffer.
template <int D>
__global__ void __launch_bounds__(ND_THREADS) k_nd_mean(const JobDev *jobs, const uint32_t *map) {
    __shared__ double red[ND_THREADS / 64];
    const BlockCtx bc = block_ctx(map, jobs, F_ONE);
    const JobDev &J = jobs[bc.job];
    EpochScalars *scal = J.scal;
    if (!scal->mean_pending) {
        if (threadIdx.x == 0)
            for (int a = 0; a < D; a++) scal->mean[a] = 0.0f;
        return;
    }
    const int nparts = J.nblk[F_PT];
    float m[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
        double t = 0.0;
        for (int b = threadIdx.x; b < nparts; b += ND_THREADS) t += J.mpart[(size_t)D * b + a];
        t = block_sum_f64(t, red);
        m[a] = (float)(t / (double)J.n);
    }
    if (threadIdx.x == 0) {
        for (int a = 0; a < D; a++) scal->mean[a] = m[a];
        scal->mean_pending = 0;
        scal->epoch = scal->epoch_next;
    }
}

// y -= mean; per-block, per-axis min / max into fpart[2 * D * block + 2 * axis (+1)].
template <int D>
__global__ void __launch_bounds__(ND_THREADS) k_nd_center(const JobDev *jobs, const uint32_t *map) {
    __shared__ float red[2 * D][ND_THREADS / 64];
    const BlockCtx bc = block_ctx(map, jobs, F_PT);
    const JobDev &J = jobs[bc.job];
    float *y = job_y(J);
    float m[D], mn[D], mx[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
        m[a] = J.scal->mean[a];
        mn[a] = INFINITY;
        mx[a] = -INFINITY;
    }
    for (int i = bc.lb * ND_THREADS + threadIdx.x; i < J.n; i += bc.nb * ND_THREADS) {
#pragma unroll
        for (int a = 0; a < D; a++) {
            const float v = y[(size_t)i * D + a] - m[a];
            y[(size_t)i * D + a] = v;
            mn[a] = fminf(mn[a], v);
            mx[a] = fmaxf(mx[a], v);
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < D; a++) {
        mn[a] = wave_min_f32(mn[a]);
        mx[a] = wave_max_f32(mx[a]);
        if (lane == 0) {
            red[2 * a][w] = mn[a];
            red[2 * a + 1][w] = mx[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * D) {
        const int q = threadIdx.x;
        float v = red[q][0];
        for (int k = 1; k < ND_THREADS / 64; k++)
            v = (q & 1) ? fmaxf(v, red[q][k]) : fminf(v, red[q][k]);
        J.fpart[(size_t)2 * D * bc.lb + q] = v;
    }
}

// Combines the per-block bounds into the job's per-axis bounding box.
template <int D>
__global__ void __launch_bounds__(ND_THREADS) k_nd_bounds(const JobDev *jobs, const uint32_t *map) {
    __shared__ float red[2 * D][ND_THREADS / 64];
    const BlockCtx bc = block_ctx(map, jobs, F_ONE);
    const JobDev &J = jobs[bc.job];
    const int nparts = J.nblk[F_PT];
    float mn[D], mx[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
        mn[a] = INFINITY;
        mx[a] = -INFINITY;
    }
    for (int b = threadIdx.x; b < nparts; b += ND_THREADS) {
#pragma unroll
        for (int a = 0; a < D; a++) {
            mn[a] = fminf(mn[a], J.fpart[(size_t)2 * D * b + 2 * a]);
            mx[a] = fmaxf(mx[a], J.fpart[(size_t)2 * D * b + 2 * a + 1]);
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < D; a++) {
        mn[a] = wave_min_f32(mn[a]);
        mx[a] = wave_max_f32(mx[a]);
        if (lane == 0) {
            red[2 * a][w] = mn[a];
            red[2 * a + 1][w] = mx[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * D) {
        const int q = threadIdx.x;
        float v = red[q][0];
        for (int k = 1; k < ND_THREADS / 64; k++)
            v = (q & 1) ? fmaxf(v, red[q][k]) : fminf(v, red[q][k]);
        if (q & 1) J.scal->amax[q >> 1] = v;
        else J.scal->amin[q >> 1] = v;
    }
}

// ---- codes and their sort ----------------------------------------------------------------------

template <int D>
__global__ void __launch_bounds__(ND_THREADS) k_nd_codes(const JobDev *jobs, const uint32_t *map) {
    const BlockCtx bc = block_ctx(map, jobs, F_PT);
    const JobDev &J = jobs[bc.job];
    const float *y = job_y(J);
    const float scale = (float)(1ull << Nd<D>::BITS);
    float mn[D], inv[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
        mn[a] = J.scal->amin[a];
        const float extent = J.scal->amax[a] - J.scal->amin[a];
        inv[a] = extent > 0.0f ? scale / extent : 0.0f;
    }
    for (int i = bc.lb * ND_THREADS + threadIdx.x; i < J.n; i += bc.nb * ND_THREADS) {
        uint32_t q[D];
#pragma unroll
        for (int a = 0; a < D; a++)
            q[a] = quantize_axis(y[(size_t)i * D + a], mn[a], inv[a], Nd<D>::MAX_BUCKET);
        uint64_t lo, hi;
        encode<D>(q, lo, hi);
        J.codes[i] = lo;
        J.cidx[i] = (uint32_t)i;
        if (Nd<D>::WIDE) {
            J.clo[i] = lo;
            J.chi[i] = hi;
        }
    }
}

// Between the two passes of a two-word sort: the high words in the order of the low-word sort
// become the next sort's keys, with the indices as its values.
__global__ void __launch_bounds__(ND_THREADS) k_nd_high_keys(const JobDev *jobs,
                                                             const uint32_t *map) {
    const BlockCtx bc = block_ctx(map, jobs, F_PT);
    const JobDev &J = jobs[bc.job];
    for (int s = bc.lb * ND_THREADS + threadIdx.x; s < J.n; s += bc.nb * ND_THREADS) {
        const uint32_t i = J.scidx[s];
        J.codes[s] = J.chi[i];
        J.cidx[s] = i;
    }
}

// After the high-word sort: the low words in the final order.
__global__ void __launch_bounds__(ND_THREADS) k_nd_low_words(const JobDev *jobs,
                                                             const uint32_t *map) {
    const BlockCtx bc = block_ctx(map, jobs, F_PT);
    const JobDev &J = jobs[bc.job];
    for (int s = bc.lb * ND_THREADS + threadIdx.x; s < J.n; s += bc.nb * ND_THREADS)
        J.slo[s] = J.clo[J.scidx[s]];
}

// ---- tree (arena.rs:137-328) ---------------------------------------------------------------------

template <int D>
__global__ void __launch_bounds__(ND_BUILD_THREADS) k_nd_build(const JobDev *jobs,
                                                               const uint32_t *map) {
    constexpr int BITS = Nd<D>::BITS;
    __shared__ uint32_t sbuf[ND_BUILD_THREADS / 64];
    __shared__ uint32_t round_start[ND_MAX_ROUNDS + 2];
    __shared__ float shw2[BITS + 1];
    const BlockCtx bc = block_ctx(map, jobs, F_ONE);
    const JobDev &J = jobs[bc.job];
    const float *y = job_y(J);
    const int n = J.n;
    const Codes<D> codes = sorted_codes<D>(J);
    const uint32_t *sidx = J.scidx;
    const BhNodes t = J.nodes;
    const int tid = threadIdx.x;
    if (tid == 0) {
        // squared half-width per level from the largest axis extent (arena.rs:167-174)
        float max_extent = 0.0f;
        for (int a = 0; a < D; a++) max_extent = fmaxf(max_extent, J.scal->amax[a] - J.scal->amin[a]);
        for (int level = 0; level <= BITS; level++) {
            const float hw = max_extent / (float)(1ull << (level + 1));
            t.hw2[level] = hw * hw;
            shw2[level] = hw * hw;
        }
        float *r = node_ptr<D>(t.node, 0);
        for (int a = 0; a < D; a++) r[a] = 0.0f;
        r[D] = __uint_as_float((uint32_t)n);
        r[D + 1] = __uint_as_float(SENTINEL);
        t.meta[0] = make_uint2((uint32_t)BITS << 8, 0u);
        t.range[0] = 0;
        t.range[1] = (uint32_t)n;
        round_start[0] = 0;
        round_start[1] = 1;
    }
    __syncthreads();
    if (n == 0) {
        if (tid == 0) t.counts[0] = 0;
        return;
    }
    // breadth-first emission (arena.rs:197-259)
    uint32_t nn = 1;
    int rounds = 1;
    for (int r = 0; r < ND_MAX_ROUNDS; r++) {
        const uint32_t f0 = round_start[r], f1 = round_start[r + 1];
        if (f0 == f1) break;
        for (uint32_t base = f0; base < f1; base += ND_BUILD_THREADS) {
            const uint32_t i = base + tid;
            uint32_t nchild = 0, s = 0, e = 0;
            int level = BITS, shift = 0;
            bool internal = false;
            if (i < f1) {
                s = t.range[2 * i];
                e = t.range[2 * i + 1];
                if (e - s > 1) {
                    uint64_t h0, l0, h1, l1;
                    codes.get(s, h0, l0);
                    codes.get(e - 1, h1, l1);
                    const uint64_t xh = h0 ^ h1, xl = l0 ^ l1;
                    if (xh | xl) {
                        internal = true;
                        const int highest = xh ? 127 - __clzll((long long)xh)
                                               : 63 - __clzll((long long)xl);
                        level = (BITS - 1) - highest / D;
                        shift = D * (BITS - 1 - level);
                        for (uint32_t p = s; p < e; nchild++)
                            p = codes.upper(p + 1, e, shift, codes.digit(p, shift));
                    }
                }
            }
            uint32_t total;
            const uint32_t off = block_excl_scan(nchild, sbuf, total);
            if (internal) {
                const uint32_t first = nn + off;
                node_ptr<D>(t.node, i)[D + 1] = __uint_as_float(first);
                t.meta[i] = make_uint2(nchild | ((uint32_t)level << 8), __float_as_uint(shw2[level]));
                uint32_t c = first;
                for (uint32_t p = s; p < e; c++) {
                    const uint32_t q = codes.upper(p + 1, e, shift, codes.digit(p, shift));
                    float *cr = node_ptr<D>(t.node, c);
                    for (int a = 0; a < D; a++) cr[a] = 0.0f;
                    cr[D] = __uint_as_float(q - p);
                    cr[D + 1] = __uint_as_float(SENTINEL);
                    t.meta[c] = make_uint2((uint32_t)BITS << 8, 0u);
                    t.range[2 * c] = p;
                    t.range[2 * c + 1] = q;
                    p = q;
                }
            }
            nn += total;
        }
        if (tid == 0) round_start[r + 2] = nn;
        rounds = r + 1;
        __syncthreads();
    }
    // centres of mass, deepest round first (arena.rs:261-309)
    for (int r = rounds - 1; r >= 0; r--) {
        const uint32_t f0 = round_start[r], f1 = round_start[r + 1];
        for (uint32_t i = f0 + tid; i < f1; i += ND_BUILD_THREADS) {
            float *nd = node_ptr<D>(t.node, i);
            const uint32_t count = __float_as_uint(nd[D]);
            const uint32_t fc = __float_as_uint(nd[D + 1]);
            float c[D];
#pragma unroll
            for (int a = 0; a < D; a++) c[a] = 0.0f;
            if (fc == SENTINEL) {
                for (uint32_t p = t.range[2 * i]; p < t.range[2 * i + 1]; p++) {
                    const float *pt = y + (size_t)sidx[p] * D;
#pragma unroll
                    for (int a = 0; a < D; a++) c[a] += pt[a];
                }
            } else {
                const uint32_t cc = t.meta[i].x & 0xffu;
                for (uint32_t ch = fc; ch < fc + cc; ch++) {
                    const float *cn = node_ptr<D>(t.node, ch);
                    const float w = (float)__float_as_uint(cn[D]);
#pragma unroll
                    for (int a = 0; a < D; a++) c[a] += cn[a] * w;
                }
            }
            const float inv = 1.0f / (float)count;
#pragma unroll
            for (int a = 0; a < D; a++) nd[a] = c[a] * inv;
        }
        __syncthreads();
    }
    if (tid == 0) t.counts[0] = nn;
}

// compute_non_edge_forces (arena.rs:338-392) for every point, with the child-range stack.
template <int D>
__global__ void __launch_bounds__(ND_TRAV_THREADS) k_nd_traverse(const JobDev *jobs,
                                                                 const uint32_t *map) {
    constexpr int STACK = Nd<D>::STACK, S4 = Nd<D>::S4;
    __shared__ uint2 stack[STACK][ND_TRAV_THREADS];
    const BlockCtx bc = block_ctx(map, jobs, F_TRAV);
    const JobDev &J = jobs[bc.job];
    const float *y = job_y(J);
    const uint32_t *sidx = J.scidx;
    const BhNodes t = J.nodes;
    const float theta_sq = J.theta_sq;
    const int lane = threadIdx.x;
    double zacc = 0.0;
    for (int s = bc.lb * ND_TRAV_THREADS + lane; s < J.n; s += bc.nb * ND_TRAV_THREADS) {
        const uint32_t i = sidx[s];
        float qp[D], f[D];
#pragma unroll
        for (int a = 0; a < D; a++) {
            qp[a] = y[(size_t)i * D + a];
            f[a] = 0.0f;
        }
        float q_sum = 0.0f;
        int top = 1;
        stack[0][lane] = make_uint2(0u, 1u);  // the root, as a range of one
        while (top > 0) {
            uint2 rg = stack[top - 1][lane];
            const uint32_t id = rg.x + rg.y - 1u;
            if (--rg.y == 0u) top--;
            else stack[top - 1][lane] = rg;
            float nd[4 * S4];
#pragma unroll
            for (int k = 0; k < S4; k++) {
                const float4 v = t.node[(size_t)id * S4 + k];
                nd[4 * k] = v.x;
                nd[4 * k + 1] = v.y;
                nd[4 * k + 2] = v.z;
                nd[4 * k + 3] = v.w;
            }
            float disp[D];
            float distance = 0.0f;
#pragma unroll
            for (int a = 0; a < D; a++) {
                const float delta = qp[a] - nd[a];
                disp[a] = delta;
                distance += delta * delta;
            }
            const uint32_t fc = __float_as_uint(nd[D + 1]);
            if (fc == SENTINEL) {
                if (distance == 0.0f) continue;  // the query's own leaf
            } else {
                const uint2 m = t.meta[id];
                if (__uint_as_float(m.y) >= theta_sq * distance) {
                    if (top < STACK) stack[top++][lane] = make_uint2(fc, m.x & 0xffu);
                    continue;
                }
            }
            const float inverse = 1.0f / (1.0f + distance);
            float magnitude = (float)__float_as_uint(nd[D]) * inverse;
            q_sum += magnitude;
            magnitude = magnitude * inverse;
#pragma unroll
            for (int a = 0; a < D; a++) f[a] += magnitude * disp[a];
        }
#pragma unroll
        for (int a = 0; a < D; a++) J.frep[(size_t)i * D + a] = f[a];
        zacc += (double)q_sum;
    }
    zacc = wave_sum_f64(zacc);
    if (lane == 0) J.dpart[bc.lb] = zacc;
}

// ---- attraction (arena.rs:398-427) and the gradient step (mod.rs:470-516) -----------------------

template <int D, bool EPOCH>
__global__ void __launch_bounds__(ND_THREADS) k_nd_attract(const JobDev *jobs, const uint32_t *map,
                                                           float *pos_out) {
    const BlockCtx bc = block_ctx(map, jobs, F_ATT);
    const JobDev &J = jobs[bc.job];
    const int epoch = J.scal->epoch;
    const float *y = J.ybuf[epoch & 1];
    float *out = EPOCH ? J.fattr : pos_out;
    float *pv = J.tval;
    // stop_lying fires after epoch == stop_lying_epoch (lib.rs:1168-1171): the first epoch to
    // read the rescaled values rescales them in place as it reads
    const bool stop_now = EPOCH && J.stop_lying_epoch != 0 && epoch == J.stop_lying_epoch + 1;
    const float inv_exag = J.inv_exag;
    for (int i = bc.lb * ND_THREADS + threadIdx.x; i < J.n; i += bc.nb * ND_THREADS) {
        float yi[D], acc[D];
#pragma unroll
        for (int a = 0; a < D; a++) {
            yi[a] = y[(size_t)i * D + a];
            acc[a] = 0.0f;
        }
        for (uint32_t e = J.rows[i]; e < J.rows[i + 1]; e++) {
            const float *yj = y + (size_t)J.cols[e] * D;
            float p = pv[e];
            if (stop_now) {
                p = p * inv_exag;
                pv[e] = p;
            }
            float disp[D];
            float distance = 0.0f;
#pragma unroll
            for (int a = 0; a < D; a++) {
                const float delta = yi[a] - yj[a];
                disp[a] = delta;
                distance += delta * delta;
            }
            const float factor = p / (distance + 1.0f);
#pragma unroll
            for (int a = 0; a < D; a++) acc[a] += factor * disp[a];
        }
#pragma unroll
        for (int a = 0; a < D; a++) out[(size_t)i * D + a] = acc[a];
    }
}

template <int D>
__global__ void __launch_bounds__(ND_THREADS) k_nd_update(const JobDev *jobs, const uint32_t *map) {
    __shared__ double red[ND_THREADS / 64];
    const BlockCtx bc = block_ctx(map, jobs, F_PT);
    const JobDev &J = jobs[bc.job];
    const int epoch = J.scal->epoch;
    // the momentum switch takes effect after epoch == momentum_switch_epoch (lib.rs:1173-1175)
    const float momentum = epoch > J.momentum_switch_epoch ? J.final_momentum : J.momentum;
    const float inv_z = J.scal->inv_z, lr = J.learning_rate;
    const float *y = J.ybuf[epoch & 1];
    float *y_next = J.ybuf[(epoch + 1) & 1];
    double m[D];
#pragma unroll
    for (int a = 0; a < D; a++) m[a] = 0.0;
    for (int i = bc.lb * ND_THREADS + threadIdx.x; i < J.n; i += bc.nb * ND_THREADS) {
#pragma unroll
        for (int a = 0; a < D; a++) {
            const size_t k = (size_t)i * D + a;
            const float gradient = J.fattr[k] - J.frep[k] * inv_z;
            float uy = J.uy[k];
            float gain = J.gains[k];
            gain = rust_signum(gradient) != rust_signum(uy) ? gain + 0.2f : gain * 0.8f;
            if (gain < 0.01f) gain = 0.01f;
            uy = momentum * uy - lr * gain * gradient;
            const float yn = y[k] + uy;
            J.gains[k] = gain;
            J.uy[k] = uy;
            y_next[k] = yn;
            m[a] += (double)yn;
        }
    }
    double s[D];
#pragma unroll
    for (int a = 0; a < D; a++) s[a] = block_sum_f64(m[a], red);
    if (threadIdx.x == 0) {
        for (int a = 0; a < D; a++) J.mpart[(size_t)D * bc.lb + a] = s[a];
        if (bc.lb == 0) {
            J.scal->epoch_next = epoch + 1;
            J.scal->mean_pending = 1;
        }
    }
}

// ---- launches -----------------------------------------------------------------------------------

template <int D>
void center_d(const LaunchSet &L, hipStream_t s) {
    hipLaunchKernelGGL(k_nd_mean<D>, dim3(L.nb[F_ONE]), dim3(ND_THREADS), 0, s, L.jobs,
                       L.map[F_ONE]);
    hipLaunchKernelGGL(k_nd_center<D>, dim3(L.nb[F_PT]), dim3(ND_THREADS), 0, s, L.jobs,
                       L.map[F_PT]);
    BHT_LAUNCH_CHECK();
}

template <int D>
void bounds_d(const LaunchSet &L, hipStream_t s) {
    hipLaunchKernelGGL(k_nd_bounds<D>, dim3(L.nb[F_ONE]), dim3(ND_THREADS), 0, s, L.jobs,
                       L.map[F_ONE]);
    BHT_LAUNCH_CHECK();
}

template <int D>
const uint32_t *repulsion_d(const LaunchSet &L, hipStream_t s) {
    const dim3 pt(L.nb[F_PT]);
    hipLaunchKernelGGL(k_nd_codes<D>, pt, dim3(ND_THREADS), 0, s, L.jobs, L.map[F_PT]);
    BHT_LAUNCH_CHECK();
    // the job table's sorted-code fields point at the alternate buffers, where both sorts end
    bool alt = sort_group<uint64_t>(L, Nd<D>::WIDE ? 64 : D * Nd<D>::BITS, s);
    if (Nd<D>::WIDE) {
        if (!alt) throw Error(ERR_STATE, "Morton sort: low words did not end in the alternate buffers");
        hipLaunchKernelGGL(k_nd_high_keys, pt, dim3(ND_THREADS), 0, s, L.jobs, L.map[F_PT]);
        BHT_LAUNCH_CHECK();
        alt = sort_group<uint64_t>(L, D * Nd<D>::BITS - 64, s);
        hipLaunchKernelGGL(k_nd_low_words, pt, dim3(ND_THREADS), 0, s, L.jobs, L.map[F_PT]);
        BHT_LAUNCH_CHECK();
    }
    if (!alt) throw Error(ERR_STATE, "Morton sort did not end in the alternate buffers");
    hipLaunchKernelGGL(k_nd_build<D>, dim3(L.nb[F_ONE]), dim3(ND_BUILD_THREADS), 0, s, L.jobs,
                       L.map[F_ONE]);
    hipLaunchKernelGGL(k_nd_traverse<D>, dim3(L.nb[F_TRAV]), dim3(ND_TRAV_THREADS), 0, s, L.jobs,
                       L.map[F_TRAV]);
    BHT_LAUNCH_CHECK();
    // Z is the plain sum of the per-point q sums (repulsion.rs:155)
    finalize_z(L, F_TRAV, s);
    return L.vals_alt;
}

template <int D>
void attract_d(const LaunchSet &L, float *pos, bool epoch, hipStream_t s) {
    if (epoch)
        hipLaunchKernelGGL((k_nd_attract<D, true>), dim3(L.nb[F_ATT]), dim3(ND_THREADS), 0, s,
                           L.jobs, L.map[F_ATT], nullptr);
    else
        hipLaunchKernelGGL((k_nd_attract<D, false>), dim3(L.nb[F_ATT]), dim3(ND_THREADS), 0, s,
                           L.jobs, L.map[F_ATT], pos);
    BHT_LAUNCH_CHECK();
}

template <int D>
void update_d(const LaunchSet &L, hipStream_t s) {
    hipLaunchKernelGGL(k_nd_update<D>, dim3(L.nb[F_PT]), dim3(ND_THREADS), 0, s, L.jobs,
                       L.map[F_PT]);
    BHT_LAUNCH_CHECK();
}

template <typename F>
auto by_dim(int dim, F &&f) {
    switch (dim) {
    case 3: return f(std::integral_constant<int, 3>());
    case 4: return f(std::integral_constant<int, 4>());
    case 5: return f(std::integral_constant<int, 5>());
    case 6: return f(std::integral_constant<int, 6>());
    case 7: return f(std::integral_constant<int, 7>());
    default: throw Error(ERR_INVALID, "the Barnes-Hut path supports d_out in 2..7");
    }
}

}  // namespace

int nd_bits(int dim) {
    return by_dim(dim, [](auto d) { return Nd<decltype(d)::value>::BITS; });
}

int nd_node_floats(int dim) {
    return by_dim(dim, [](auto d) { return 4 * Nd<decltype(d)::value>::S4; });
}

int nd_attract_blocks(int n) {
    return (int)std::min<uint64_t>(ceil_div((uint64_t)(n > 0 ? n : 1), ND_THREADS), 4096);
}

void nd_center_points(const LaunchSet &L, hipStream_t s) {
    by_dim(L.dim, [&](auto d) { center_d<decltype(d)::value>(L, s); });
}

void nd_bounds(const LaunchSet &L, hipStream_t s) {
    by_dim(L.dim, [&](auto d) { bounds_d<decltype(d)::value>(L, s); });
}

const uint32_t *nd_bh_repulsion(const LaunchSet &L, hipStream_t s) {
    return by_dim(L.dim, [&](auto d) { return repulsion_d<decltype(d)::value>(L, s); });
}

void nd_attract(const LaunchSet &L, float *pos, bool epoch, hipStream_t s) {
    by_dim(L.dim, [&](auto d) { attract_d<decltype(d)::value>(L, pos, epoch, s); });
}

void nd_update(const LaunchSet &L, hipStream_t s) {
    by_dim(L.dim, [&](auto d) { update_d<decltype(d)::value>(L, s); });
}

}  // namespace bht
