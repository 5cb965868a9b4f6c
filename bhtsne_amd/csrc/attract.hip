// Attractive forces over the CSR graph fused with the gains / momentum update and the partial
// sums of the next recentring (src/tsne/arena.rs:398-427, src/tsne/mod.rs:470-516, 534-568).
//
// Work unit: a row group, the rows whose first edge lies in one GROUP_EDGES-edge window of the
// CSR (built once per CSR layout by k_row_groups; at most a few rows, since a symmetrised row has
// at least k entries).  One wave streams the group's whole edge span, lane l taking edges
// l, l + 64, .. of each pass of 64 * GROUP_UNROLL edges: every column, value and gathered y load of
// a pass is independent of the others and of the row boundaries, so a wave keeps eight gathers per
// lane in flight with no idle lanes, where a wave per row would wait on each row's offsets and
// leave a third of its lanes empty.  Each edge adds into its row's accumulator (a uniform
// comparison against the group's row offsets picks it); the rows are reduced by fixed butterflies
// and lane q applies the update of the group's q-th row.  A group of more than MAX_GROUP_ROWS rows
// (short rows, from caller-supplied affinities) is processed row by row.  The new embedding goes
// to the other buffer of the pair because other rows still read y this epoch.
//
// The kernel writes each row's attraction sums; k_update then applies the gradient step per point
// (with the repulsion, which runs concurrently on another stream), so that the attraction does
// not wait for the repulsion.
//
// Each block owns a contiguous run of groups and first copies the positions of an index window
// centred on its rows into LDS (WIN_FLOATS floats).  The points are numbered so that P's
// well-connected groups are contiguous, so most gathered columns fall inside the window and are
// read from LDS; the rest are read from global memory.  Gathers are single 8-byte reads from
// scattered lines, so it is their number of L2 requests, not their bytes, that the window saves.
#include "internal.hpp"

namespace bht {

constexpr int UP_THREADS = 512;
constexpr int UPD_THREADS = 256;  // k_update runs over the point-parallel (F_PT) block partition
constexpr int GROUP_UNROLL = 8;
constexpr int GROUP_EDGES = 64 * GROUP_UNROLL;  // edge window of a row group
constexpr int MAX_ATT_BLOCKS = MAX_ATTRACT_BLOCKS;
// 80,000 bytes of LDS per block: two blocks (16 waves) per CU.  The LDS of a CU is 160 KiB; a
// block asking for the whole of it ran at half the speed, as if only one block fitted.
constexpr int WIN_FLOATS = 20000;
// A block takes ngroups / 1024 groups, at least 8 (one per wave) and at most 64: large jobs get
// long runs (few window copies), small ones enough blocks to keep their chains of dependent loads
// short.
constexpr int MIN_GROUPS_PER_BLOCK = UP_THREADS / 64, MAX_GROUPS_PER_BLOCK = 64;

uint32_t row_groups_count(uint64_t nnz) {
    return (uint32_t)((nnz + GROUP_EDGES - 1) / GROUP_EDGES);
}

int attract_blocks(int ngroups) {
    const uint64_t ng = (uint64_t)(ngroups > 0 ? ngroups : 1);
    const uint64_t per = std::min<uint64_t>(
        std::max<uint64_t>(ng / 1024, MIN_GROUPS_PER_BLOCK), MAX_GROUPS_PER_BLOCK);
    const uint64_t b = ceil_div(ng, per);
    return (int)(b < (uint64_t)MAX_ATT_BLOCKS ? b : (uint64_t)MAX_ATT_BLOCKS);
}

size_t row_groups_bytes(uint64_t nnz) {
    return ((size_t)row_groups_count(nnz) + 1) * GROUP_DESC_WORDS * sizeof(uint32_t);
}

// first row whose first edge is at or after `target` (rows are nondecreasing), n if none
__device__ uint32_t first_row_at(const uint32_t *rows, int n, uint64_t target) {
    uint32_t lo = 0, hi = (uint32_t)n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if ((uint64_t)rows[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Group m is rows [first_row_at(m * GROUP_EDGES), first_row_at((m + 1) * GROUP_EDGES)).
__global__ void k_row_groups(const uint32_t *rows, int n, uint32_t ngroups, uint32_t *groups) {
    const uint32_t m = blockIdx.x * 256 + threadIdx.x;
    if (m > ngroups) return;
    uint32_t *d = groups + (size_t)GROUP_DESC_WORDS * m;
    if (m == ngroups) {
        d[0] = (uint32_t)n;
        for (int q = 1; q < GROUP_DESC_WORDS; q++) d[q] = q == 1 ? 0u : 0xffffffffu;
        return;
    }
    const uint32_t r0 = first_row_at(rows, n, (uint64_t)m * GROUP_EDGES);
    const uint32_t r1 = m + 1 == ngroups ? (uint32_t)n
                                         : first_row_at(rows, n, (uint64_t)(m + 1) * GROUP_EDGES);
    const uint32_t nr = r1 - r0;
    d[0] = r0;
    d[1] = nr;
    for (int q = 0; q <= MAX_GROUP_ROWS; q++)
        d[2 + q] = (uint32_t)q <= nr && nr <= MAX_GROUP_ROWS ? rows[r0 + q] : 0xffffffffu;
    d[GROUP_DESC_WORDS - 1] = 0u;
}

void build_row_groups(const uint32_t *rows, int n, uint64_t nnz, uint32_t *groups,
                      hipStream_t s) {
    const uint32_t ng = row_groups_count(nnz);
    hipLaunchKernelGGL(k_row_groups, dim3(ceil_div((uint64_t)ng + 1, 256)), dim3(256), 0, s, rows,
                       n, ng, groups);
    BHT_LAUNCH_CHECK();
}

struct StepCtx {
    float inv_z, momentum, learning_rate;
};

__device__ __forceinline__ StepCtx step_ctx(const JobDev &J, int epoch) {
    // the momentum switch takes effect after epoch == momentum_switch_epoch (lib.rs:1173-1175)
    return StepCtx{J.scal->inv_z, epoch > J.momentum_switch_epoch ? J.final_momentum : J.momentum,
                   J.learning_rate};
}

// The job's arrays as global-memory pointers (the kernel mixes them with LDS reads).
struct AttractMem {
    BHT_GLOBAL const uint32_t *rows, *cols, *groups;
    BHT_GLOBAL float *vals, *qv, *uy, *gains, *y_next;
    BHT_GLOBAL const float *frep;
};

// Per-row operands of the gradient step: the repulsive force, the velocity and the gains.
template <int D>
struct RowState {
    float frep[D], uy[D], gain[D];
};

template <int D>
__device__ __forceinline__ RowState<D> load_row_state(const AttractMem &M, int i) {
    RowState<D> r;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const size_t k = (size_t)i * D + d;
        r.frep[d] = M.frep[k];
        r.uy[d] = M.uy[k];
        r.gain[d] = M.gains[k];
    }
    return r;
}

// Gradient step of row i from its attraction sums (mod.rs:470-516): stores the new coordinates,
// velocity and gains, and adds the new coordinates to the recentring sums.
template <int D>
__device__ __forceinline__ void apply_update(const AttractMem &M, const StepCtx &c, int i,
                                             const float *acc, const float *yi,
                                             const RowState<D> &st, double &m0, double &m1) {
#pragma unroll
    for (int d = 0; d < D; d++) {
        const size_t k = (size_t)i * D + d;
        const float gradient = acc[d] - st.frep[d] * c.inv_z;
        float uy = st.uy[d];
        float gain = st.gain[d];
        gain = rust_signum(gradient) != rust_signum(uy) ? gain + 0.2f : gain * 0.8f;
        if (gain < 0.01f) gain = 0.01f;
        uy = c.momentum * uy - c.learning_rate * gain * gradient;
        const float yn = yi[d] + uy;
        M.gains[k] = gain;
        M.uy[k] = uy;
        M.y_next[k] = yn;
        if (d == 0) m0 += (double)yn;
        else m1 += (double)yn;
    }
}

// clang vector types: loads through address-space-qualified pointers (HIP's float2 / float4
// cannot be copy-constructed from such references)
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));

// Positions of points [lo, lo + count) copied to LDS; lookups outside it go to global memory.
template <int D>
struct YWindow {
    BHT_GLOBAL const float *y;
    BHT_LDS const float *win;
    uint32_t lo, count;
    __device__ __forceinline__ bool holds(uint32_t i) const { return i - lo < count; }
    __device__ __forceinline__ float2 lds(uint32_t i) const {
        if (D == 2) {
            const f2v v = ((BHT_LDS const f2v *)win)[i - lo];
            return make_float2(v.x, v.y);
        }
        return make_float2(win[i - lo], 0.0f);
    }
    __device__ __forceinline__ float2 global(uint32_t i) const {
        if (D == 2) {
            const f2v v = ((BHT_GLOBAL const f2v *)y)[i];
            return make_float2(v.x, v.y);
        }
        return make_float2(y[i], 0.0f);
    }
    __device__ __forceinline__ float2 operator()(uint32_t i) const {
        if (holds(i)) return lds(i);
        return global(i);
    }
};

// Window of at most WIN_FLOATS / D points centred on rows [r_begin, r_end), its start a multiple
// of 4 / D points so that the copy runs in 16-byte loads; the whole block copies it.
template <int D>
__device__ YWindow<D> stage_window(BHT_GLOBAL const float *y, int n, int r_begin, int r_end,
                                   BHT_LDS float *win) {
    constexpr int WPTS = WIN_FLOATS / D;
    constexpr int ALIGN = 4 / D;
    int count = WPTS < n ? WPTS : n;
    int lo = ((r_begin + r_end) >> 1) - count / 2;
    if (lo > n - count) lo = n - count;
    if (lo < 0) lo = 0;
    lo -= lo % ALIGN;
    if (count > n - lo) count = n - lo;
    const int nf = count * D;
    BHT_GLOBAL const float *src = y + (size_t)lo * D;
    for (int k = threadIdx.x; 4 * k < nf; k += blockDim.x) {
        if (4 * k + 3 < nf) {
            ((BHT_LDS f4v *)win)[k] = ((BHT_GLOBAL const f4v *)src)[k];
        } else {
            for (int t = 4 * k; t < nf; t++) win[t] = src[t];
        }
    }
    __syncthreads();
    return YWindow<D>{y, win, (uint32_t)lo, (uint32_t)count};
}

// EPOCH: the sums go to the job's fattr buffer, and the first epoch after stop-lying rescales P as
// it reads it; otherwise (the forces export) they go to pos_out and nothing else is written.
template <int D, bool EPOCH>
__global__ void __launch_bounds__(UP_THREADS) k_attract(const JobDev *jobs, const uint32_t *map,
                                                        float *pos_out) {
    __shared__ double red[UP_THREADS / 64];
    __shared__ __align__(16) float win[WIN_FLOATS];
    const BlockCtx bc = block_ctx(map, jobs, F_ATT);
    const JobDev &J = jobs[bc.job];
    const int lane = threadIdx.x & 63;
    constexpr int NW = UP_THREADS / 64;
    const int epoch = J.scal->epoch;
    AttractMem M;
    M.rows = global_view(J.rows);
    M.cols = global_view(J.cols);
    M.groups = global_view(J.groups);
    M.vals = global_view(J.vals);
    M.qv = global_view(J.qv);
    M.uy = global_view(J.uy);
    M.gains = global_view(J.gains);
    BHT_GLOBAL float *out = global_view(EPOCH ? J.fattr : pos_out);
    BHT_GLOBAL const uint32_t *rows = M.rows, *cols = M.cols, *groups = M.groups;
    BHT_GLOBAL float *vals = M.vals;
    // this block's groups [g_begin, g_end) and the window around their rows
    const int g_begin = (int)((int64_t)J.ngroups * bc.lb / bc.nb);
    const int g_end = (int)((int64_t)J.ngroups * (bc.lb + 1) / bc.nb);
    const int r_begin = g_begin < g_end ? (int)groups[(size_t)GROUP_DESC_WORDS * g_begin] : 0;
    const int r_end = g_begin < g_end ? (int)groups[(size_t)GROUP_DESC_WORDS * g_end] : 0;
    const YWindow<D> y = stage_window<D>(global_view((const float *)J.ybuf[epoch & 1]), J.n,
                                         r_begin, r_end, lds_view(win));
    const float inv_exag = J.inv_exag;
    // stop_lying fires after epoch == stop_lying_epoch (lib.rs:1168-1171): the next epoch is the
    // first to read the rescaled values, and rescales them in place as it reads (with the copies
    // of the reference-position values that KL weights by).
    const bool stop_now = EPOCH && J.stop_lying_epoch != 0 && epoch == J.stop_lying_epoch + 1;
    // the wave index as a uniform value, so that the group loop and its descriptor loads are scalar
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (int g = g_begin + wid; g < g_end; g += NW) {
        BHT_GLOBAL const uint32_t *desc = groups + (size_t)GROUP_DESC_WORDS * g;
        uint32_t dw[2 + MAX_GROUP_ROWS + 1];
#pragma unroll
        for (int q = 0; q < 2 + MAX_GROUP_ROWS + 1; q++) dw[q] = desc[q];
        const int r0 = (int)dw[0], nr = (int)dw[1], r1 = r0 + nr;
        if (nr <= 0) continue;
        if (stop_now)
            for (int i = r0 + lane; i < r1 && i < J.qv_n; i += 64) M.qv[i] = M.qv[i] * inv_exag;
        if (nr > MAX_GROUP_ROWS) {
            // many short rows: one row at a time, lanes over its edges
            for (int i = r0; i < r1; i++) {
                const float2 yi = y(i);
                float a0 = 0.0f, a1 = 0.0f;
                for (uint32_t e = rows[i] + lane; e < rows[i + 1]; e += 64) {
                    float pv = vals[e];
                    if (stop_now) {
                        pv = pv * inv_exag;
                        vals[e] = pv;
                    }
                    const float2 yj = y(cols[e]);
                    const float d0 = yi.x - yj.x;
                    const float d1 = D == 2 ? yi.y - yj.y : 0.0f;
                    float dist = 0.0f;
                    dist += d0 * d0;
                    if (D == 2) dist += d1 * d1;
                    const float f = pv / (dist + 1.0f);
                    a0 += f * d0;
                    if (D == 2) a1 += f * d1;
                }
                a0 = wave_sum_f32(a0);
                if (D == 2) a1 = wave_sum_f32(a1);
                if (lane != 0) continue;
                out[(size_t)i * D] = a0;
                if (D == 2) out[(size_t)i * D + 1] = a1;
            }
            continue;
        }
        // the group's row offsets and positions (uniform across the wave)
        const uint32_t *bnd = dw + 2;
        float2 yr[MAX_GROUP_ROWS];
#pragma unroll
        for (int q = 0; q < MAX_GROUP_ROWS; q++)
            yr[q] = q < nr ? y(r0 + q) : make_float2(0.0f, 0.0f);
        float acc[MAX_GROUP_ROWS][2];
#pragma unroll
        for (int q = 0; q < MAX_GROUP_ROWS; q++) acc[q][0] = acc[q][1] = 0.0f;
        const uint32_t e_end = bnd[nr];
        for (uint32_t base = bnd[0]; base < e_end; base += GROUP_EDGES) {
            uint32_t col[GROUP_UNROLL];
            float pv[GROUP_UNROLL];
#pragma unroll
            for (int u = 0; u < GROUP_UNROLL; u++) {
                const uint32_t e = base + lane + 64 * u;
                const bool ok = e < e_end;
                col[u] = ok ? __builtin_nontemporal_load(cols + e) : (uint32_t)r0;
                pv[u] = ok ? __builtin_nontemporal_load(vals + e) : 0.0f;
                if (stop_now && ok) {
                    pv[u] = pv[u] * inv_exag;
                    vals[e] = pv[u];
                }
            }
            // window reads for all lanes (clamped), then global reads for the columns outside
            float2 yj[GROUP_UNROLL];
            bool inside[GROUP_UNROLL];
#pragma unroll
            for (int u = 0; u < GROUP_UNROLL; u++) {
                inside[u] = y.holds(col[u]);
                yj[u] = y.lds(inside[u] ? col[u] : y.lo);
            }
#pragma unroll
            for (int u = 0; u < GROUP_UNROLL; u++)
                if (!inside[u]) yj[u] = y.global(col[u]);
#pragma unroll
            for (int u = 0; u < GROUP_UNROLL; u++) {
                const uint32_t e = base + lane + 64 * u;
                int q = 0;
#pragma unroll
                for (int t = 1; t < MAX_GROUP_ROWS; t++) q += e >= bnd[t] ? 1 : 0;
                float2 yi = yr[0];
#pragma unroll
                for (int t = 1; t < MAX_GROUP_ROWS; t++) yi = q == t ? yr[t] : yi;
                const float d0 = yi.x - yj[u].x;
                const float d1 = D == 2 ? yi.y - yj[u].y : 0.0f;
                float dist = 0.0f;
                dist += d0 * d0;
                if (D == 2) dist += d1 * d1;
                const float f = pv[u] / (dist + 1.0f);
                const float f0 = f * d0, f1 = f * d1;
#pragma unroll
                for (int t = 0; t < MAX_GROUP_ROWS; t++) {
                    acc[t][0] += q == t ? f0 : 0.0f;
                    if (D == 2) acc[t][1] += q == t ? f1 : 0.0f;
                }
            }
        }
        // row sums: lane q keeps the sums of the group's q-th row
        float mine0 = 0.0f, mine1 = 0.0f;
#pragma unroll
        for (int t = 0; t < MAX_GROUP_ROWS; t++) {
            if (t >= nr) break;
            const float s0 = wave_sum_f32(acc[t][0]);
            const float s1 = D == 2 ? wave_sum_f32(acc[t][1]) : 0.0f;
            if (lane == t) {
                mine0 = s0;
                mine1 = s1;
            }
        }
        if (lane >= nr) continue;
        const int i = r0 + lane;
        out[(size_t)i * D] = mine0;
        if (D == 2) out[(size_t)i * D + 1] = mine1;
    }
}

// Gradient step of every point from its attraction sums (fattr) and the repulsion (frep, Z), and
// the per-block sums of the new embedding for the next recentring.
template <int D>
__global__ void __launch_bounds__(UPD_THREADS) k_update(const JobDev *jobs, const uint32_t *map) {
    __shared__ double red[UPD_THREADS / 64];
    const BlockCtx bc = block_ctx(map, jobs, F_PT);
    const JobDev &J = jobs[bc.job];
    const int epoch = J.scal->epoch;
    const StepCtx sc = step_ctx(J, epoch);
    AttractMem M;
    M.uy = global_view(J.uy);
    M.gains = global_view(J.gains);
    M.y_next = global_view(J.ybuf[(epoch + 1) & 1]);
    M.frep = global_view((const float *)J.frep);
    BHT_GLOBAL const float *y = global_view((const float *)J.ybuf[epoch & 1]);
    BHT_GLOBAL const float *fattr = global_view((const float *)J.fattr);
    double m0 = 0.0, m1 = 0.0;
    const int stride = bc.nb * UPD_THREADS;
    for (int i = bc.lb * UPD_THREADS + threadIdx.x; i < J.n; i += stride) {
        float acc[2] = {0.0f, 0.0f}, yi[2] = {0.0f, 0.0f};
#pragma unroll
        for (int d = 0; d < D; d++) {
            acc[d] = fattr[(size_t)i * D + d];
            yi[d] = y[(size_t)i * D + d];
        }
        apply_update<D>(M, sc, i, acc, yi, load_row_state<D>(M, i), m0, m1);
    }
    const double s0 = block_sum_f64(m0, red);
    const double s1 = D == 2 ? block_sum_f64(m1, red) : 0.0;
    if (threadIdx.x == 0) {
        J.mpart[2 * bc.lb] = s0;
        J.mpart[2 * bc.lb + 1] = s1;
        if (bc.lb == 0) {
            J.scal->epoch_next = epoch + 1;
            J.scal->mean_pending = 1;
        }
    }
}

void attract_forces(const LaunchSet &L, hipStream_t s) {
    const dim3 g(L.nb[F_ATT]);
    if (L.dim == 2)
        hipLaunchKernelGGL((k_attract<2, true>), g, dim3(UP_THREADS), 0, s, L.jobs, L.map[F_ATT],
                           nullptr);
    else
        hipLaunchKernelGGL((k_attract<1, true>), g, dim3(UP_THREADS), 0, s, L.jobs, L.map[F_ATT],
                           nullptr);
    BHT_LAUNCH_CHECK();
}

void update_points(const LaunchSet &L, hipStream_t s) {
    const dim3 g(L.nb[F_PT]);
    if (L.dim == 2)
        hipLaunchKernelGGL(k_update<2>, g, dim3(UPD_THREADS), 0, s, L.jobs, L.map[F_PT]);
    else
        hipLaunchKernelGGL(k_update<1>, g, dim3(UPD_THREADS), 0, s, L.jobs, L.map[F_PT]);
    BHT_LAUNCH_CHECK();
}

void attract_only(const LaunchSet &L, float *pos, hipStream_t s) {
    const dim3 g(L.nb[F_ATT]);
    if (L.dim == 2)
        hipLaunchKernelGGL((k_attract<2, false>), g, dim3(UP_THREADS), 0, s, L.jobs, L.map[F_ATT],
                           pos);
    else
        hipLaunchKernelGGL((k_attract<1, false>), g, dim3(UP_THREADS), 0, s, L.jobs, L.map[F_ATT],
                           pos);
    BHT_LAUNCH_CHECK();
}

// ---- P scaling and sum (src/tsne/mod.rs:284-293, 525-531) ----------------------------------

__global__ void k_scale(float *v, uint64_t n, float scale) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        v[i] *= scale;
}

void scale_values(float *vals, uint64_t nnz, float scale, hipStream_t s) {
    if (nnz == 0) return;
    int nb = (int)std::min<uint64_t>(ceil_div(nnz, 256), 4096);
    hipLaunchKernelGGL(k_scale, dim3(nb), dim3(256), 0, s, vals, nnz, scale);
    BHT_LAUNCH_CHECK();
}

__global__ void __launch_bounds__(256) k_sum(const float *v, uint64_t n, EpochScalars *scal,
                                             double *dpart) {
    __shared__ double red[4];
    double acc = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        acc += (double)v[i];
    double b = block_sum_f64(acc, red);
    if (threadIdx.x == 0) dpart[blockIdx.x] = b;
    if (!arrive_last(&scal->ctr[CTR_PSUM], gridDim.x)) return;
    double t = 0.0;
    for (int q = threadIdx.x; q < (int)gridDim.x; q += 256) t += dpart[q];
    t = block_sum_f64(t, red);
    if (threadIdx.x == 0) {
        scal->psum = t;
        scal->ctr[CTR_PSUM] = 0;
    }
}

void sum_values_f64(const float *vals, uint64_t nnz, EpochScalars *scal, double *dpart,
                    hipStream_t s) {
    int nb = (int)std::min<uint64_t>(ceil_div(nnz > 0 ? nnz : 1, 256), 1024);
    hipLaunchKernelGGL(k_sum, dim3(nb), dim3(256), 0, s, vals, nnz, scal, dpart);
    BHT_LAUNCH_CHECK();
}

}  // namespace bht
