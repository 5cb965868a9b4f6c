// FIt-SNE repulsion on the device (src/tsne/interpolation.rs:100-332): recentring + joint bounds +
// grid plan, box keys + stable sort, box-segmented charge spreading, FFT convolution (fft.hip),
// potential gather + force / Z reconstruction.
//
// Per-point arithmetic (box location, Lagrange weights, charges, reconstruction) repeats the
// reference's f32 operation order; only the summation order of the spread and of Z differs (Z and
// the recentring mean are summed in FP64).
#include <cfloat>

#include "internal.hpp"

namespace bht {

constexpr int PT_THREADS = 256;
// Point-parallel kernels stride over at most 1024 blocks; a thread's points are taken PT_BATCH
// at a time, their loads issued before the first is computed (one wait for memory per batch).
constexpr int PT_BATCH = 4;
constexpr int GATHER_BATCH = 2;  // the gather's batch (nine grid values per 2-D point)

int point_blocks(int n) {
    int b = (int)ceil_div((uint64_t)(n > 0 ? n : 1), PT_THREADS);
    return b < 1024 ? b : 1024;
}

// The spreading kernels take the box-sorted points in chunks of 64 * spread_rounds(n), one wave
// per chunk in rounds of 64.  Large jobs take 4 rounds, so a box of many points leaves 4x fewer
// chunk partials (1M points: the fixup 29 -> 15 us); small jobs keep one round per wave for the
// parallelism (4 rounds cost C4's plain fits 8%).
__host__ __device__ constexpr int spread_rounds(int n) { return n >= 65536 ? 4 : 1; }

// Two waves per spreading block: the block's LDS slab (20 KiB) then fits a CU beside two of the
// attraction's blocks.
constexpr int SPR_THREADS = 128;

int spread_blocks(int n) {
    const uint64_t nchunks = ceil_div((uint64_t)(n > 0 ? n : 1), 64 * spread_rounds(n));
    return (int)std::min<uint64_t>(ceil_div(nchunks, SPR_THREADS / 64), 4096);
}

// Grid sizes of every kernel family for one job.  A job alone fills the device by itself; in a
// batch the jobs share it, so the grid-bound families (clearing, transforms) take fewer blocks
// per job and stride further.
void job_block_counts(JobDev &J, bool batched) {
    const int n = J.n;
    J.nblk[F_PT] = point_blocks(n);
    J.nblk[F_ONE] = 1;
    J.nblk[F_SPR] = spread_blocks(n);
    J.nblk[F_ATT] = J.dim >= 3 ? nd_attract_blocks(n) : tiled_blocks(n);
    J.nblk[F_TRAV] = traverse_blocks(n, J.dim);
    J.nblk[F_BSUB] = bh_sub_blocks(n);
    J.nblk[F_LINE1] = 2;
    // transform lines per job: 2 * nodes (rows) and 2 * L (columns).  A job alone takes up to
    // 2048 blocks.  In a batch the grid a member reaches grows with its size, and the launch's
    // blocks each ask for the group's whole line buffer in LDS, so a member takes 8 blocks per
    // 1024 points (16 to 512, a multiple of 8 for the column passes' XCD mapping) and its blocks
    // stride over its lines: 256 jobs of 500-50,000 points launched 65,536 mostly idle blocks
    // per pass with 512 each.
    const int per_size = (int)std::min<int64_t>(512, std::max<int64_t>(16, 8 * (((int64_t)n + 1023) / 1024)));
    const int cap = batched ? per_size : 2048;
    const int64_t nodes_total =
        J.dim == 2 ? (int64_t)J.nodes_cap * J.nodes_cap : (int64_t)J.nodes_cap;
    J.nblk[F_ZERO] = (int)std::max<int64_t>(1, std::min<int64_t>(batched ? 64 : 1024,
                                                                 (nodes_total + 255) / 256));
    const int hcap = J.len_cap / 2 + 1;
    J.nblk[F_KSPEC] = std::max(1, std::min(batched ? cap / 2 : 1024, (hcap + 1) / 2));
    J.nblk[F_ROWS] = std::max(1, std::min(cap, 2 * J.nodes_cap));
    J.nblk[F_COLS] = std::max(1, std::min(cap, 2 * J.len_cap));
}

// ---- per-point helpers (interpolation.rs:162-187, 400-456) ---------------------------------

struct Lagrange3 {
    float node[3];
    float inv_denom[3];
    __device__ __forceinline__ Lagrange3() {
        const float half = 0.5f;
        const float p_recip = 1.0f / 3.0f;
#pragma unroll
        for (int k = 0; k < 3; k++) node[k] = ((float)k + half) * p_recip;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float denom = 1.0f;
#pragma unroll
            for (int m = 0; m < 3; m++)
                if (m != k) denom = denom * (node[k] - node[m]);
            inv_denom[k] = 1.0f / denom;
        }
    }
    __device__ __forceinline__ void weights(float local, float *out) const {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float num = 1.0f;
#pragma unroll
            for (int m = 0; m < 3; m++)
                if (m != k) num = num * (local - node[m]);
            out[k] = num * inv_denom[k];
        }
    }
};

__device__ __forceinline__ int locate_box(float v, const GridPlan &p, float &local) {
    float offset = (v - p.lo) * p.inv_box_width;
    float b = floorf(offset);
    int last = p.boxes - 1;
    int bid = 0;
    if (b > 0.0f) bid = b >= (float)last ? last : (int)b;
    local = offset - (float)bid;
    return bid;
}

// ---- 1. recentring + bounds + plan -----------------------------------------------------------
//
// Three small kernels instead of one with a grid-wide "last block" reduction: a device-scope fence
// in every block would write back the XCD's L2 each time.  Kernel boundaries give the ordering.

// Mean of the last update's per-block sums (written by the attraction kernel), applied once;
// advances the job's epoch index (which selects its current embedding buffer).
constexpr int FIN_THREADS = 1024;  // up to 16384 attraction blocks' partials per job

//
// The shift is either applied at once by k_center (the embedding buffer then holds centred
// coordinates and ysub is 0), or left pending (lazy: ysub = mean, and every reader of the epoch's
// embedding takes fl(y - ysub), the value k_center would have stored).
__device__ void mean_body(const JobDev &J, int lazy, double *red) {
    EpochScalars *scal = J.scal;
    const int pending = scal->mean_pending;
    if (!pending) {
        if (threadIdx.x == 0) {
            scal->mean[0] = 0.0f;
            scal->mean[1] = 0.0f;
            scal->ysub[0] = 0.0f;
            scal->ysub[1] = 0.0f;
        }
        return;
    }
    float m0, m1;
    pending_mean2(J, m0, m1);
    if (threadIdx.x == 0) {
        scal->mean[0] = m0;
        scal->mean[1] = m1;
        scal->ysub[0] = lazy ? m0 : 0.0f;
        scal->ysub[1] = lazy ? m1 : 0.0f;
        scal->mean_pending = 0;
        scal->epoch = scal->epoch_next;
    }
}

__global__ void __launch_bounds__(FIN_THREADS) k_mean_finalize(const JobDev *jobs,
                                                               const uint32_t *map) {
    __shared__ double red[FIN_THREADS / 64];
    const BlockCtx bc = block_ctx(map, jobs, F_ONE);
    mean_body(jobs[bc.job], 0, red);
}

// y -= mean; per-block, per-axis min / max.
template <int D>
__global__ void __launch_bounds__(PT_THREADS) k_center(const JobDev *jobs, const uint32_t *map) {
    __shared__ float red[4][PT_THREADS / 64];
    const BlockCtx bc = block_ctx(map, jobs, F_PT);
    const JobDev &J = jobs[bc.job];
    const EpochScalars *scal = J.scal;
    float *y = job_y(J);
    const int n = J.n;
    const float m0 = scal->mean[0], m1 = D == 2 ? scal->mean[1] : 0.0f;
    float mn0 = INFINITY, mn1 = INFINITY, mx0 = -INFINITY, mx1 = -INFINITY;
    const int stride = bc.nb * PT_THREADS;
    for (int i = bc.lb * PT_THREADS + threadIdx.x; i < n; i += stride) {
        if (D == 2) {
            float2 v = reinterpret_cast<float2 *>(y)[i];
            v.x = v.x - m0;
            v.y = v.y - m1;
            reinterpret_cast<float2 *>(y)[i] = v;
            mn0 = fminf(mn0, v.x); mx0 = fmaxf(mx0, v.x);
            mn1 = fminf(mn1, v.y); mx1 = fmaxf(mx1, v.y);
        } else {
            float v = y[i] - m0;
            y[i] = v;
            mn0 = fminf(mn0, v); mx0 = fmaxf(mx0, v);
        }
    }
    mn0 = wave_min_f32(mn0); mx0 = wave_max_f32(mx0);
    mn1 = wave_min_f32(mn1); mx1 = wave_max_f32(mx1);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[0][w] = mn0; red[1][w] = mx0; red[2][w] = mn1; red[3][w] = mx1; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < PT_THREADS / 64; q++) {
            mn0 = fminf(mn0, red[0][q]); mx0 = fmaxf(mx0, red[1][q]);
            mn1 = fminf(mn1, red[2][q]); mx1 = fmaxf(mx1, red[3][q]);
        }
        reinterpret_cast<float4 *>(J.fpart)[bc.lb] = make_float4(mn0, mx0, mn1, mx1);
    }
}

// Combines the bounds (per-block partials of the embedding, less the pending shift); with
// do_plan, decides the FIt-SNE grid of this epoch.  All threads of the block.
template <int D>
__device__ void plan_body(const JobDev &J, const int *len_of_boxes, const int *off_of_boxes,
                          int do_plan, float (*red)[16]) {
    EpochScalars *scal = J.scal;
    const float *fpart = J.fpart;
    const int nparts = J.nblk[F_PT], min_boxes = J.min_boxes, max_boxes = J.max_boxes;
    float mn0 = INFINITY, mx0 = -INFINITY, mn1 = INFINITY, mx1 = -INFINITY;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        float4 pv = reinterpret_cast<const float4 *>(fpart)[b];
        mn0 = fminf(mn0, pv.x); mx0 = fmaxf(mx0, pv.y);
        mn1 = fminf(mn1, pv.z); mx1 = fmaxf(mx1, pv.w);
    }
    mn0 = wave_min_f32(mn0); mx0 = wave_max_f32(mx0);
    mn1 = wave_min_f32(mn1); mx1 = wave_max_f32(mx1);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) { red[0][w] = mn0; red[1][w] = mx0; red[2][w] = mn1; red[3][w] = mx1; }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int q = 1; q < nw; q++) {
        mn0 = fminf(mn0, red[0][q]); mx0 = fmaxf(mx0, red[1][q]);
        mn1 = fminf(mn1, red[2][q]); mx1 = fmaxf(mx1, red[3][q]);
    }
    // fl(. - s) is monotone, so the bounds of the shifted embedding are the shifted bounds
    const float s0 = scal->ysub[0], s1 = scal->ysub[1];
    mn0 = mn0 - s0; mx0 = mx0 - s0;
    mn1 = mn1 - s1; mx1 = mx1 - s1;
    scal->amin[0] = mn0; scal->amax[0] = mx0;
    scal->amin[1] = mn1; scal->amax[1] = mx1;
    if (!do_plan || J.method != M_FITSNE) return;
    // coordinate_bounds + box_count + next_smooth (interpolation.rs:115-128, 337-378)
    float lo = D == 2 ? fminf(mn0, mn1) : mn0;
    float hi = D == 2 ? fmaxf(mx0, mx1) : mx0;
    float extent = fmaxf(hi - lo, FLT_MIN);
    float scaled = ceilf(extent / 1.0f);
    int boxes;
    if (scaled > 0.0f) boxes = scaled >= (float)MAX_BOXES ? MAX_BOXES : (int)scaled;
    else boxes = 0;
    if (boxes < min_boxes) boxes = min_boxes;
    if (boxes > MAX_BOXES) boxes = MAX_BOXES;
    // the job's grid capacity (MAX_BOXES unless the caller lowered it)
    if (boxes > max_boxes) boxes = max_boxes;
    GridPlan p;
    p.lo = lo;
    p.hi = hi;
    p.boxes = boxes;
    p.nodes = boxes * IP;
    p.len = len_of_boxes[boxes];
    p.tw_off = off_of_boxes[boxes];
    p.box_width = extent / (float)boxes;
    p.spacing = p.box_width / (float)IP;
    p.inv_box_width = 1.0f / p.box_width;
    // the charges y and |y|^2 are divided by qscale and qscale^2 inside the FFT so that both
    // halves of each packed complex grid have similar magnitude (powers of two: exact)
    int e2 = 0;
    frexpf(fmaxf(fmaxf(fabsf(lo), fabsf(hi)), 1.0f), &e2);
    p.qscale = ldexpf(1.0f, e2);
    scal->plan = p;
    // 7-bit passes of the box sort: keys are below boxes^D
    const uint32_t nkeys = D == 2 ? (uint32_t)boxes * (uint32_t)boxes : (uint32_t)boxes;
    int kbits = 0;
    while (kbits < 32 && (1u << kbits) < nkeys) kbits++;
    scal->sort_passes = kbits <= 7 ? 1 : (kbits + 6) / 7;
    // the host sizes later FFT launches by the length (a system-scope store to pinned memory)
    if (J.host_len)
        __hip_atomic_store(J.host_len, p.len, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <int D>
__global__ void __launch_bounds__(256) k_plan(const JobDev *jobs, const uint32_t *map,
                                              const int *len_of_boxes, const int *off_of_boxes,
                                              int do_plan) {
    __shared__ float red[4][16];
    const BlockCtx bc = block_ctx(map, jobs, F_ONE);
    plan_body<D>(jobs[bc.job], len_of_boxes, off_of_boxes, do_plan, red);
}

// The head of a lazily centred FIt-SNE epoch: the mean of the last update's sums left pending
// (ysub), then the bounds and the grid plan from the update's per-block bounds.
template <int D>
__global__ void __launch_bounds__(FIN_THREADS) k_epoch_head(const JobDev *jobs,
                                                            const uint32_t *map,
                                                            const int *len_of_boxes,
                                                            const int *off_of_boxes) {
    __shared__ double dred[FIN_THREADS / 64];
    __shared__ float red[4][16];
    const BlockCtx bc = block_ctx(map, jobs, F_ONE);
    const JobDev &J = jobs[bc.job];
    mean_body(J, 1, dred);
    __syncthreads();
    plan_body<D>(J, len_of_boxes, off_of_boxes, 1, red);
}

void epoch_head(const LaunchSet &L, hipStream_t s) {
    const int *lo = L.tw->len_of_boxes, *of = L.tw->off_of_boxes;
    if (L.dim == 2)
        hipLaunchKernelGGL(k_epoch_head<2>, dim3(L.nb[F_ONE]), dim3(FIN_THREADS), 0, s, L.jobs,
                           L.map[F_ONE], lo, of);
    else
        hipLaunchKernelGGL(k_epoch_head<1>, dim3(L.nb[F_ONE]), dim3(FIN_THREADS), 0, s, L.jobs,
                           L.map[F_ONE], lo, of);
    BHT_LAUNCH_CHECK();
}

void center_points(const LaunchSet &L, hipStream_t s) {
    if (L.dim >= 3) {
        nd_center_points(L, s);
        return;
    }
    hipLaunchKernelGGL(k_mean_finalize, dim3(L.nb[F_ONE]), dim3(FIN_THREADS), 0, s, L.jobs,
                       L.map[F_ONE]);
    if (L.dim == 2)
        hipLaunchKernelGGL(k_center<2>, dim3(L.nb[F_PT]), dim3(PT_THREADS), 0, s, L.jobs,
                           L.map[F_PT]);
    else
        hipLaunchKernelGGL(k_center<1>, dim3(L.nb[F_PT]), dim3(PT_THREADS), 0, s, L.jobs,
                           L.map[F_PT]);
    BHT_LAUNCH_CHECK();
}

void plan_grid(const LaunchSet &L, bool do_plan, hipStream_t s) {
    if (L.dim >= 3) {  // Barnes-Hut only: the bounds, no grid
        nd_bounds(L, s);
        return;
    }
    const int *lo = L.tw ? L.tw->len_of_boxes : nullptr, *of = L.tw ? L.tw->off_of_boxes : nullptr;
    const int plan = do_plan && L.tw != nullptr;
    if (L.dim == 2)
        hipLaunchKernelGGL(k_plan<2>, dim3(L.nb[F_ONE]), dim3(256), 0, s, L.jobs, L.map[F_ONE], lo,
                           of, plan);
    else
        hipLaunchKernelGGL(k_plan<1>, dim3(L.nb[F_ONE]), dim3(256), 0, s, L.jobs, L.map[F_ONE], lo,
                           of, plan);
    BHT_LAUNCH_CHECK();
}

void center_plan(const LaunchSet &L, bool do_plan, hipStream_t s) {
    center_points(L, s);
    plan_grid(L, do_plan, s);
}

// Z = sum of the per-block sums (minus the n self-interactions for FIt-SNE, interpolation.rs:331;
// the plain sum of the q sums for Barnes-Hut, repulsion.rs:155).
__global__ void __launch_bounds__(256) k_z_finalize(const JobDev *jobs, const uint32_t *map,
                                                    int part_family) {
    __shared__ double red[4];
    const BlockCtx bc = block_ctx(map, jobs, F_ONE);
    const JobDev &J = jobs[bc.job];
    EpochScalars *scal = J.scal;
    const int nparts = J.nblk[part_family];
    double t = 0.0;
    for (int b = threadIdx.x; b < nparts; b += 256) t += J.dpart[b];
    t = block_sum_f64(t, red);
    if (threadIdx.x == 0) {
        scal->zsum = t;
        const float z = J.method == M_FITSNE ? (float)(t - (double)J.n) : (float)t;
        scal->z = z;
        scal->inv_z = 1.0f / z;
    }
}

void finalize_z(const LaunchSet &L, int part_family, hipStream_t s) {
    hipLaunchKernelGGL(k_z_finalize, dim3(L.nb[F_ONE]), dim3(256), 0, s, L.jobs, L.map[F_ONE],
                       part_family);
    BHT_LAUNCH_CHECK();
}

// ---- 2. box keys ---------------------------------------------------------------------------

template <int D>
__global__ void __launch_bounds__(PT_THREADS) k_box_keys(const JobDev *jobs,
                                                         const uint32_t *map) {
    const BlockCtx bc = block_ctx(map, jobs, F_PT);
    const JobDev &J = jobs[bc.job];
    const GridPlan p = J.scal->plan;
    const float *y = job_y(J);
    const float s0 = J.scal->ysub[0], s1 = J.scal->ysub[1];
    uint32_t *keys = J.keys, *vals = J.kidx;
    const int64_t n = J.n, stride = (int64_t)bc.nb * PT_THREADS;
    // PT_BATCH points per thread with their loads in flight together
    for (int64_t i0 = (int64_t)bc.lb * PT_THREADS + threadIdx.x; i0 < n;
         i0 += PT_BATCH * stride) {
        float2 v[PT_BATCH];
#pragma unroll
        for (int q = 0; q < PT_BATCH; q++) {
            const int64_t i = i0 + q * stride;
            if (i >= n) break;
            if (D == 2) v[q] = reinterpret_cast<const float2 *>(y)[i];
            else v[q].x = y[i];
        }
#pragma unroll
        for (int q = 0; q < PT_BATCH; q++) {
            const int64_t i = i0 + q * stride;
            if (i >= n) break;
            float l;
            uint32_t key;
            if (D == 2) {
                const uint32_t b0 = (uint32_t)locate_box(v[q].x - s0, p, l);
                const uint32_t b1 = (uint32_t)locate_box(v[q].y - s1, p, l);
                key = b0 * (uint32_t)p.boxes + b1;
            } else {
                key = (uint32_t)locate_box(v[q].x - s0, p, l);
            }
            keys[i] = key;
            vals[i] = (uint32_t)i;
        }
    }
}

__global__ void __launch_bounds__(256) k_zero_grid(const JobDev *jobs, const uint32_t *map) {
    const BlockCtx bc = block_ctx(map, jobs, F_ZERO);
    const JobDev &J = jobs[bc.job];
    const GridPlan p = J.scal->plan;
    float4 *grid = J.grid;
    const int total = J.dim == 2 ? p.nodes * p.nodes : p.nodes;
    for (int i = bc.lb * 256 + threadIdx.x; i < total; i += bc.nb * 256)
        grid[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---- 3. box-segmented spreading (interpolation.rs:247-269) ----------------------------------
//
// Points sorted by box form runs; each box's 3^D nodes receive only that run's charges.  One wave
// takes a chunk of 64 * spread_rounds(n) sorted points in rounds of 64: each lane writes its point's
// (3^D nodes x terms) contributions as a row of an LDS slab, then lane t walks the rows summing
// term t, closing a sum at every run end (wave-uniform masks of run ends); a run still open at the
// end of a round carries its sums into the next.  Within a chunk a box's charges are therefore
// added in point order, as the reference's sequential spreading adds them.  A run wholly inside a
// chunk is written directly; a run crossing chunk boundaries leaves partials that k_spread_fixup
// adds in chunk order.

constexpr int NV2 = 36;  // 9 nodes x 4 terms (D = 2)
constexpr int NV1 = 9;   // 3 nodes x 3 terms (D = 1)

template <int D>
__device__ __forceinline__ void point_contrib(float2 pt, float s0, float s1, const GridPlan &p,
                                              const Lagrange3 &lb, float *v) {
    if (D == 2) {
        pt.x = pt.x - s0;
        pt.y = pt.y - s1;
        float l0, l1, w0[3], w1[3];
        locate_box(pt.x, p, l0);
        locate_box(pt.y, p, l1);
        lb.weights(l0, w0);
        lb.weights(l1, w1);
        float c[4];
        c[0] = 1.0f;
        c[1] = pt.x;
        c[2] = pt.y;
        c[3] = pt.x * pt.x + pt.y * pt.y;
#pragma unroll
        for (int k1 = 0; k1 < 3; k1++)
#pragma unroll
            for (int k0 = 0; k0 < 3; k0++) {
                float wt = w0[k0] * w1[k1];
#pragma unroll
                for (int t = 0; t < 4; t++) v[(k0 + 3 * k1) * 4 + t] = wt * c[t];
            }
    } else {
        const float x = pt.x - s0;
        float l0, w0[3];
        locate_box(x, p, l0);
        lb.weights(l0, w0);
        float c[3] = {1.0f, x, x * x};
#pragma unroll
        for (int k0 = 0; k0 < 3; k0++)
#pragma unroll
            for (int t = 0; t < 3; t++) v[k0 * 3 + t] = w0[k0] * c[t];
    }
}

// Term t of a box's node values: node t / 4 of the box's nine, component t % 4 (D = 2), or node
// t / 3 of three, component t % 3 (D = 1).
template <int D>
__device__ __forceinline__ void write_box_term(float4 *grid, const GridPlan &p, uint32_t key,
                                               int t, float value) {
    float *g = reinterpret_cast<float *>(grid);
    if (D == 2) {
        const uint32_t b0 = key / (uint32_t)p.boxes, b1 = key % (uint32_t)p.boxes;
        const int c = t >> 2, k0 = c % 3, k1 = c / 3;
        const size_t node = (size_t)(3 * b0 + k0) * p.nodes + (3 * b1 + k1);
        g[node * 4 + (t & 3)] = value;
    } else {
        g[(size_t)(3 * key + t / 3) * 4 + t % 3] = value;
    }
}

constexpr int SLAB_ROW = 40;             // floats per point row of the spreading slab (>= NV2)
constexpr int SLAB_ROW4 = SLAB_ROW / 4;

constexpr int SPREAD_MAX_ROUNDS = 4;
typedef float spr_f2 __attribute__((ext_vector_type(2)));
static_assert(spread_rounds(1 << 30) <= SPREAD_MAX_ROUNDS, "a chunk's rounds are prefetched");

// acc + rows [r0, r1) of column `lane` of a wave's slab, added in row order, eight reads in
// flight at a time (-0.0 pads the last eight: x + -0.0 == x for every x).
__device__ __forceinline__ float sum_slab_rows(const float *slab, int lane, int r0, int r1,
                                               float acc) {
    for (int r = r0; r < r1; r += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = r + u < r1 ? slab[(r + u) * SLAB_ROW + lane] : -0.0f;
#pragma unroll
        for (int u = 0; u < 8; u++) acc = acc + v[u];
    }
    return acc;
}

template <int D>
__global__ void __launch_bounds__(SPR_THREADS) k_spread(const JobDev *jobs, const uint32_t *map) {
    constexpr int NV = D == 2 ? NV2 : NV1;
    const BlockCtx bc = block_ctx(map, jobs, F_SPR);
    const JobDev &J = jobs[bc.job];
    const GridPlan p = J.scal->plan;
    // global (not flat) loads, so that the slab's LDS waits do not also wait for them
    BHT_GLOBAL const float *y = global_view((const float *)job_y(J));
    const float ys0 = J.scal->ysub[0], ys1 = J.scal->ysub[1];
    BHT_GLOBAL const uint32_t *keys = global_view((const uint32_t *)J.skeys);
    BHT_GLOBAL const uint32_t *vals = global_view((const uint32_t *)J.sidx);
    float4 *grid = J.grid;
    float *part = J.part;
    const int n = J.n;
    const int rounds = spread_rounds(n), chunk = 64 * rounds;
    const int nchunks = (n + chunk - 1) / chunk;
    const Lagrange3 lb;
    const int lane = threadIdx.x & 63;
    const int wave0 = (bc.lb * SPR_THREADS + threadIdx.x) >> 6;
    const int nwaves = (bc.nb * SPR_THREADS) >> 6;
    float *part_first = part;
    float *part_last = part + (size_t)nchunks * NV;
    // a round's contributions, one row per point; lane t then sums term t down the rows
    __shared__ float4 slab_mem[SPR_THREADS / 64][64][SLAB_ROW4];
    float *slab = reinterpret_cast<float *>(slab_mem[threadIdx.x >> 6]);
    for (int c = wave0; c < nchunks; c += nwaves) {
        const int c0 = c * chunk;
        // every round's keys, point indices and positions, and the keys just before and after
        // the chunk, loaded before the first round
        uint32_t rkey[SPREAD_MAX_ROUNDS], ridx[SPREAD_MAX_ROUNDS];
        float2 rpos[SPREAD_MAX_ROUNDS];
#pragma unroll
        for (int q = 0; q < SPREAD_MAX_ROUNDS; q++) {
            const int s = c0 + 64 * q + lane;
            const bool valid = q < rounds && s < n;
            rkey[q] = valid ? keys[s] : 0xffffffffu;
            ridx[q] = valid ? vals[s] : 0u;
        }
        const uint32_t key_before = c0 > 0 ? keys[c0 - 1] : 0xffffffffu;
        const uint32_t key_after = c0 + chunk < n ? keys[c0 + chunk] : 0xffffffffu;
#pragma unroll
        for (int q = 0; q < SPREAD_MAX_ROUNDS; q++) {
            const int s = c0 + 64 * q + lane;
            rpos[q] = make_float2(0.0f, 0.0f);
            if (q < rounds && s < n) {
                if (D == 2) {
                    const spr_f2 v = reinterpret_cast<BHT_GLOBAL const spr_f2 *>(y)[ridx[q]];
                    rpos[q] = make_float2(v.x, v.y);
                } else {
                    rpos[q] = make_float2(y[ridx[q]], 0.0f);
                }
            }
        }
        float acc = 0.0f;      // lane t < NV: term t of the run being summed
        bool open = false;     // a run continues from the previous round
        bool open_rs = false;  // and it started in this chunk
#pragma unroll
        for (int q = 0; q < SPREAD_MAX_ROUNDS; q++) {
            const int s0 = c0 + 64 * q;
            if (q >= rounds || s0 >= n) break;
            const int s = s0 + lane;
            const bool valid = s < n;
            const uint32_t key = rkey[q];
            {
                float v[NV];
                if (valid) {
                    point_contrib<D>(rpos[q], ys0, ys1, p, lb, v);
                } else {
#pragma unroll
                    for (int t = 0; t < NV; t++) v[t] = 0.0f;
                }
                float4 *row = reinterpret_cast<float4 *>(slab + lane * SLAB_ROW);
#pragma unroll
                for (int w = 0; w < (NV + 3) / 4; w++)
                    row[w] = make_float4(v[4 * w], 4 * w + 1 < NV ? v[4 * w + 1] : 0.0f,
                                         4 * w + 2 < NV ? v[4 * w + 2] : 0.0f,
                                         4 * w + 3 < NV ? v[4 * w + 3] : 0.0f);
            }
            // run structure of the round, as wave-uniform masks.  A point's run continues past it
            // when the next sorted key is the same; a run closes at a point where it does not
            // continue, or at the chunk's last point.  Per point, whether its run began in this
            // chunk (rs) and whether it ends here (re).
            const uint32_t next = __shfl_down(key, 1, 64);
            // lane 63's successor: the next round's first key, or the one after the chunk
            const uint32_t succ = q + 1 < rounds ? (uint32_t)__shfl(rkey[q + 1 < SPREAD_MAX_ROUNDS ? q + 1 : q], 0, 64)
                                                 : key_after;
            const uint32_t after = lane == 63 ? (s + 1 < n ? succ : 0xffffffffu) : next;
            const bool cont = valid && s + 1 < n && after == key;
            const bool last_round = q == rounds - 1 || s0 + 64 >= n;
            const bool tail = valid && (!cont || (lane == 63 && last_round));
            const uint32_t prev = __shfl_up(key, 1, 64);
            const bool head = valid && (lane == 0 || prev != key);
            const uint64_t heads = __ballot(head), tails = __ballot(tail);
            bool rs = false;
            if (valid) {
                const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
                const int h = 63 - __clzll(heads & upto);  // the run's first lane in this round
                if (h > 0) rs = true;
                else if (q > 0) rs = open ? open_rs : true;
                else rs = c0 == 0 || key_before != key;
            }
            const bool re = tail && !cont;
            const uint64_t rs_m = __ballot(rs), re_m = __ballot(re);
            __builtin_amdgcn_wave_barrier();
            // lane t < NV: sequential sum of term t over each run (point order within a box)
            uint64_t rest = tails;
            int row = 0;
            while (rest) {
                const int pt = __ffsll((long long)rest) - 1;  // uniform: next run end
                rest &= rest - 1;
                if (lane < NV) acc = sum_slab_rows(slab, lane, row, pt + 1, acc);
                row = pt + 1;
                const uint32_t rkey = __shfl(key, pt, 64);
                const bool r_s = (rs_m >> pt) & 1ull, r_e = (re_m >> pt) & 1ull;
                if (lane < NV) {
                    if (r_s && r_e) {
                        write_box_term<D>(grid, p, rkey, lane, acc);
                    } else {
                        if (!r_s) part_first[(size_t)c * NV + lane] = acc;
                        if (!r_e) part_last[(size_t)c * NV + lane] = acc;
                    }
                }
                acc = 0.0f;
            }
            // the round's last run continues into the next round: carry its sums
            open = s0 + 63 < n && !((tails >> 63) & 1ull);
            if (open) {
                if (lane < NV) acc = sum_slab_rows(slab, lane, row, 64, acc);
                open_rs = (rs_m >> 63) & 1ull;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// One wave per chunk; the chunk where a spanning run starts (its owner) sums the partials of the
// run's later chunks onto its own.  It first probes the first key of the next 64 chunks at once:
// a run that ends within them (every run once the embedding has spread out) is summed by lane t
// adding term t of each chunk in chunk order.  A longer run (the first epochs pack most points
// into a few boxes) is bounded by a 64-way search of the sorted keys and summed by the lanes
// (strided, then a fixed butterfly), so a run of any length costs O(length / 64 / 64) steps.
template <int D>
__global__ void __launch_bounds__(SPR_THREADS) k_spread_fixup(const JobDev *jobs,
                                                             const uint32_t *map) {
    constexpr int NV = D == 2 ? NV2 : NV1;
    const BlockCtx bc = block_ctx(map, jobs, F_SPR);
    const JobDev &J = jobs[bc.job];
    const GridPlan p = J.scal->plan;
    const uint32_t *keys = J.skeys;
    float4 *grid = J.grid;
    const int n = J.n;
    const int chunk = 64 * spread_rounds(n);
    const int nchunks = (n + chunk - 1) / chunk;
    const float *part_first = J.part;
    const float *part_last = J.part + (size_t)nchunks * NV;
    const int lane = threadIdx.x & 63;
    const int wave0 = (bc.lb * SPR_THREADS + threadIdx.x) >> 6;
    const int nwaves = (bc.nb * SPR_THREADS) >> 6;
    for (int c = wave0; c < nchunks; c += nwaves) {
        const int first_s = c * chunk;
        const int last_s = min(first_s + chunk - 1, n - 1);
        // the four keys the ownership test reads, loaded together
        const uint32_t key = keys[last_s];
        const uint32_t after = last_s + 1 < n ? keys[last_s + 1] : ~0u;
        const uint32_t head = keys[first_s];
        const uint32_t before = first_s > 0 ? keys[first_s - 1] : ~0u;
        if (!(last_s + 1 < n && after == key)) continue;  // the run ends in this chunk
        if (head == key && first_s > 0 && before == key) continue;  // not the run's owner
        // chunk j > c holds a point of the run iff its first key is the run's (keys are sorted)
        const int j = c + 1 + lane;
        const int near = __popcll(__ballot(j < nchunks && keys[(size_t)j * chunk] == key));
        if (near < 64) {
            const int c_end = c + near;  // chunk holding the run's last point
            if (lane < NV) {
                float tot = part_last[(size_t)c * NV + lane];
                for (int j0 = c + 1; j0 <= c_end; j0 += 8) {
                    float v[8];
#pragma unroll
                    for (int u = 0; u < 8; u++)
                        v[u] = j0 + u <= c_end ? part_first[(size_t)(j0 + u) * NV + lane] : -0.0f;
#pragma unroll
                    for (int u = 0; u < 8; u++) tot = tot + v[u];
                }
                write_box_term<D>(grid, p, key, lane, tot);
            }
            continue;
        }
        // first sorted position past the run, in [lo, hi]: a 64-way search, each lane probing one
        // point of the range per step (keys[lo - 1] == key; keys[hi] > key or hi == n)
        int lo = (c + 64) * chunk + 1, hi = n;
        while (lo < hi) {
            const int step = (hi - lo + 63) / 64;
            const int probe = lo + lane * step;
            const bool le = probe < hi && keys[probe] <= key;
            const int c_le = __popcll(__ballot(le));  // the probes at or below the key lead
            if (c_le == 0) break;
            hi = min(hi, lo + c_le * step);
            lo = lo + (c_le - 1) * step + 1;
        }
        const int c_end = (lo - 1) / chunk;
        float tot[NV];
#pragma unroll
        for (int t = 0; t < NV; t++) tot[t] = 0.0f;
        for (int jj = c + 1 + lane; jj <= c_end; jj += 64)
#pragma unroll
            for (int t = 0; t < NV; t++) tot[t] += part_first[(size_t)jj * NV + t];
#pragma unroll
        for (int t = 0; t < NV; t++) tot[t] = wave_sum_f32(tot[t]);
        // lane t keeps term t
        float mine = 0.0f;
#pragma unroll
        for (int t = 0; t < NV; t++)
            if (lane == t) mine = tot[t];
        if (lane < NV) write_box_term<D>(grid, p, key, lane, part_last[(size_t)c * NV + lane] + mine);
    }
}

// ---- 5. gather + reconstruct (interpolation.rs:282-332) -------------------------------------

template <int D>
__global__ void __launch_bounds__(PT_THREADS) k_gather(const JobDev *jobs, const uint32_t *map) {
    __shared__ double red[PT_THREADS / 64];
    const BlockCtx bc = block_ctx(map, jobs, F_PT);
    const JobDev &J = jobs[bc.job];
    const GridPlan p = J.scal->plan;
    const float *y = job_y(J);
    const float s0 = J.scal->ysub[0], s1 = J.scal->ysub[1];
    const float4 *grid = J.grid;
    float *frep = J.frep;
    const int64_t n = J.n, stride = (int64_t)bc.nb * PT_THREADS;
    const Lagrange3 lb;
    double zacc = 0.0;
    constexpr int NB = GATHER_BATCH;
    constexpr int NG = D == 2 ? 9 : 3;  // grid nodes per point
    // NB points per thread: their positions, then the grid values of all of them, are loaded
    // before the first is reconstructed; Z is summed in the points' order.
    for (int64_t i0 = (int64_t)bc.lb * PT_THREADS + threadIdx.x; i0 < n; i0 += NB * stride) {
        float2 pt[NB];
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int64_t i = i0 + q * stride;
            if (i >= n) break;
            if (D == 2) pt[q] = reinterpret_cast<const float2 *>(y)[i];
            else pt[q] = make_float2(y[i], 0.0f);
        }
        float w0[NB][3], w1[NB][3];
        float4 g[NB][NG];
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int64_t i = i0 + q * stride;
            if (i >= n) break;
            pt[q].x = pt[q].x - s0;
            float l0, l1;
            const int b0 = locate_box(pt[q].x, p, l0);
            lb.weights(l0, w0[q]);
            if (D == 2) {
                pt[q].y = pt[q].y - s1;
                const int b1 = locate_box(pt[q].y, p, l1);
                lb.weights(l1, w1[q]);
#pragma unroll
                for (int k1 = 0; k1 < 3; k1++)
#pragma unroll
                    for (int k0 = 0; k0 < 3; k0++)
                        g[q][k0 + 3 * k1] = grid[(size_t)(3 * b0 + k0) * p.nodes + (3 * b1 + k1)];
            } else {
#pragma unroll
                for (int k0 = 0; k0 < 3; k0++) g[q][k0] = grid[3 * b0 + k0];
            }
        }
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int64_t i = i0 + q * stride;
            if (i >= n) break;
            if (D == 2) {
                float phi0 = 0.f, phi1 = 0.f, phi2 = 0.f, phi3 = 0.f;
#pragma unroll
                for (int k1 = 0; k1 < 3; k1++)
#pragma unroll
                    for (int k0 = 0; k0 < 3; k0++) {
                        const float wt = w0[q][k0] * w1[q][k1];
                        const float4 gv = g[q][k0 + 3 * k1];
                        phi0 = phi0 + wt * gv.x;
                        phi1 = phi1 + wt * gv.y;
                        phi2 = phi2 + wt * gv.z;
                        phi3 = phi3 + wt * gv.w;
                    }
                const float px = pt[q].x, py = pt[q].y;
                float2 f;
                f.x = px * phi0 - phi1;
                f.y = py * phi0 - phi2;
                float norm_sq = 0.0f, cross = 0.0f;
                norm_sq = norm_sq + px * px;
                cross = cross + px * phi1;
                norm_sq = norm_sq + py * py;
                cross = cross + py * phi2;
                reinterpret_cast<float2 *>(frep)[i] = f;
                const float zt = (1.0f + norm_sq) * phi0 - 2.0f * cross + phi3;
                zacc += (double)zt;
            } else {
                const float px = pt[q].x;
                float phi0 = 0.f, phi1 = 0.f, phi2 = 0.f;
#pragma unroll
                for (int k0 = 0; k0 < 3; k0++) {
                    const float4 gv = g[q][k0];
                    phi0 = phi0 + w0[q][k0] * gv.x;
                    phi1 = phi1 + w0[q][k0] * gv.y;
                    phi2 = phi2 + w0[q][k0] * gv.w;
                }
                frep[i] = px * phi0 - phi1;
                const float norm_sq = 0.0f + px * px;
                const float cross = 0.0f + px * phi1;
                const float zt = (1.0f + norm_sq) * phi0 - 2.0f * cross + phi2;
                zacc += (double)zt;
            }
        }
    }
    double bs = block_sum_f64(zacc, red);
    if (threadIdx.x == 0) J.dpart[bc.lb] = bs;
}

// ---- driver ----------------------------------------------------------------------------------

const uint32_t *fitsne_repulsion(const LaunchSet &L, hipStream_t s) {
    const dim3 pt(L.nb[F_PT]), spr(L.nb[F_SPR]), zero(L.nb[F_ZERO]);
    if (L.dim == 2)
        hipLaunchKernelGGL(k_box_keys<2>, pt, dim3(PT_THREADS), 0, s, L.jobs, L.map[F_PT]);
    else
        hipLaunchKernelGGL(k_box_keys<1>, pt, dim3(PT_THREADS), 0, s, L.jobs, L.map[F_PT]);
    BHT_LAUNCH_CHECK();
    const bool alt = sort_group<uint32_t>(L, box_key_bits(L.dim), s);
    hipLaunchKernelGGL(k_zero_grid, zero, dim3(256), 0, s, L.jobs, L.map[F_ZERO]);
    if (L.dim == 2) {
        hipLaunchKernelGGL(k_spread<2>, spr, dim3(SPR_THREADS), 0, s, L.jobs, L.map[F_SPR]);
        hipLaunchKernelGGL(k_spread_fixup<2>, spr, dim3(SPR_THREADS), 0, s, L.jobs, L.map[F_SPR]);
    } else {
        hipLaunchKernelGGL(k_spread<1>, spr, dim3(SPR_THREADS), 0, s, L.jobs, L.map[F_SPR]);
        hipLaunchKernelGGL(k_spread_fixup<1>, spr, dim3(SPR_THREADS), 0, s, L.jobs, L.map[F_SPR]);
    }
    BHT_LAUNCH_CHECK();
    fft_kernel_spectrum(L, s);
    fft_convolve(L, s);
    if (L.dim == 2)
        hipLaunchKernelGGL(k_gather<2>, pt, dim3(PT_THREADS), 0, s, L.jobs, L.map[F_PT]);
    else
        hipLaunchKernelGGL(k_gather<1>, pt, dim3(PT_THREADS), 0, s, L.jobs, L.map[F_PT]);
    BHT_LAUNCH_CHECK();
    finalize_z(L, F_PT, s);
    // point indices of job 0 in box order: a spatial order of the current embedding
    if (L.sort_npass) return L.vals_fin;
    return alt ? L.vals_alt : L.vals;
}

}  // namespace bht
