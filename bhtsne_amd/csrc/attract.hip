// The gradient step of every point from its attraction sums (tiled.hip) and the repulsion, and
// the partial sums of the next recentring (src/tsne/mod.rs:470-516, 534-568); P scaling and sums.
//
// The attraction kernel writes each row's sums; k_update applies the step per point once both the
// attraction and the repulsion (which runs concurrently on another stream) are done.  The new
// embedding goes to the other buffer of the pair.
#include "internal.hpp"

namespace bht {

typedef float upd_f2 __attribute__((ext_vector_type(2)));

constexpr int UPD_THREADS = 256;  // k_update runs over the point-parallel (F_PT) block partition
constexpr int UPD_BATCH = 4;      // points per thread whose loads are in flight together

struct StepCtx {
    float inv_z, momentum, learning_rate;
};

__device__ __forceinline__ StepCtx step_ctx(const JobDev &J, int epoch) {
    // the momentum switch takes effect after epoch == momentum_switch_epoch (lib.rs:1173-1175)
    return StepCtx{J.scal->inv_z, epoch > J.momentum_switch_epoch ? J.final_momentum : J.momentum,
                   J.learning_rate};
}

// The arrays of the gradient step, as global-memory pointers.
struct StepMem {
    BHT_GLOBAL float *uy, *gains, *y_next;
    BHT_GLOBAL const float *frep;
};

// Per-row operands of the gradient step: the repulsive force, the velocity and the gains.
template <int D>
struct RowState {
    float frep[D], uy[D], gain[D];
};

template <int D>
__device__ __forceinline__ RowState<D> load_row_state(const StepMem &M, int i) {
    RowState<D> r;
    if (D == 2) {  // one 8-byte load per array
        const upd_f2 f = reinterpret_cast<BHT_GLOBAL const upd_f2 *>(M.frep)[i];
        const upd_f2 u = reinterpret_cast<BHT_GLOBAL const upd_f2 *>(M.uy)[i];
        const upd_f2 g = reinterpret_cast<BHT_GLOBAL const upd_f2 *>(M.gains)[i];
        r.frep[0] = f.x; r.frep[D - 1] = f.y;
        r.uy[0] = u.x; r.uy[D - 1] = u.y;
        r.gain[0] = g.x; r.gain[D - 1] = g.y;
        return r;
    }
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
__device__ __forceinline__ void apply_update(const StepMem &M, const StepCtx &c, int i,
                                             const float *acc, const float *yi,
                                             const RowState<D> &st, double &m0, double &m1,
                                             float *ynew) {
    float ngain[D], nuy[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        const float gradient = acc[d] - st.frep[d] * c.inv_z;
        float uy = st.uy[d];
        float gain = st.gain[d];
        gain = rust_signum(gradient) != rust_signum(uy) ? gain + 0.2f : gain * 0.8f;
        if (gain < 0.01f) gain = 0.01f;
        uy = c.momentum * uy - c.learning_rate * gain * gradient;
        const float yn = yi[d] + uy;
        ngain[d] = gain;
        nuy[d] = uy;
        ynew[d] = yn;
        if (d == 0) m0 += (double)yn;
        else m1 += (double)yn;
    }
    if (D == 2) {  // one 8-byte store per array
        reinterpret_cast<BHT_GLOBAL upd_f2 *>(M.gains)[i] = upd_f2{ngain[0], ngain[D - 1]};
        reinterpret_cast<BHT_GLOBAL upd_f2 *>(M.uy)[i] = upd_f2{nuy[0], nuy[D - 1]};
        reinterpret_cast<BHT_GLOBAL upd_f2 *>(M.y_next)[i] = upd_f2{ynew[0], ynew[D - 1]};
    } else {
#pragma unroll
        for (int d = 0; d < D; d++) {
            const size_t k = (size_t)i * D + d;
            M.gains[k] = ngain[d];
            M.uy[k] = nuy[d];
            M.y_next[k] = ynew[d];
        }
    }
}

// Gradient step of every point from its attraction sums (fattr) and the repulsion (frep, Z), and
// the per-block sums of the new embedding for the next recentring.
template <int D>
__global__ void __launch_bounds__(UPD_THREADS) k_update(const JobDev *jobs, const uint32_t *map) {
    __shared__ double red[UPD_THREADS / 64];
    __shared__ float fred[4][UPD_THREADS / 64];
    const BlockCtx bc = block_ctx(map, jobs, F_PT);
    const JobDev &J = jobs[bc.job];
    const int epoch = J.scal->epoch;
    const StepCtx sc = step_ctx(J, epoch);
    StepMem M;
    M.uy = global_view(J.uy);
    M.gains = global_view(J.gains);
    M.y_next = global_view(J.ybuf[(epoch + 1) & 1]);
    M.frep = global_view((const float *)J.frep);
    BHT_GLOBAL const float *y = global_view((const float *)J.ybuf[epoch & 1]);
    BHT_GLOBAL const float *fattr = global_view((const float *)J.fattr);
    const float sh[2] = {J.scal->ysub[0], J.scal->ysub[1]};
    double m0 = 0.0, m1 = 0.0;
    float mn[2] = {INFINITY, INFINITY}, mx[2] = {-INFINITY, -INFINITY};
    const int64_t n = J.n, stride = (int64_t)bc.nb * UPD_THREADS;
    // A thread's points i0, i0 + stride, ... are taken UPD_BATCH at a time: every load of the
    // batch is issued before the first point is computed, so a thread waits for memory once per
    // batch rather than once per point.  The points are still summed in the same order.
    for (int64_t i0 = (int64_t)bc.lb * UPD_THREADS + threadIdx.x; i0 < n;
         i0 += UPD_BATCH * stride) {
        float acc[UPD_BATCH][2], yi[UPD_BATCH][2];
        RowState<D> st[UPD_BATCH];
#pragma unroll
        for (int q = 0; q < UPD_BATCH; q++) {
            const int64_t i = i0 + q * stride;
            if (i >= n) break;
            if (D == 2) {
                const upd_f2 a = reinterpret_cast<BHT_GLOBAL const upd_f2 *>(fattr)[i];
                const upd_f2 v = reinterpret_cast<BHT_GLOBAL const upd_f2 *>(y)[i];
                acc[q][0] = a.x;
                acc[q][1] = a.y;
                yi[q][0] = v.x - sh[0];
                yi[q][1] = v.y - sh[1];
            } else {
#pragma unroll
                for (int d = 0; d < D; d++) {
                    acc[q][d] = fattr[(size_t)i * D + d];
                    yi[q][d] = y[(size_t)i * D + d] - sh[d];
                }
            }
            st[q] = load_row_state<D>(M, (int)i);
        }
#pragma unroll
        for (int q = 0; q < UPD_BATCH; q++) {
            const int64_t i = i0 + q * stride;
            if (i >= n) break;
            float yn[2];
            apply_update<D>(M, sc, (int)i, acc[q], yi[q], st[q], m0, m1, yn);
#pragma unroll
            for (int d = 0; d < D; d++) {
                mn[d] = fminf(mn[d], yn[d]);
                mx[d] = fmaxf(mx[d], yn[d]);
            }
        }
    }
    const double s0 = block_sum_f64(m0, red);
    const double s1 = D == 2 ? block_sum_f64(m1, red) : 0.0;
    // per-block bounds of the new embedding (the next epoch's grid plan when it is centred lazily)
    float b[4] = {wave_min_f32(mn[0]), wave_max_f32(mx[0]), wave_min_f32(mn[1]),
                  wave_max_f32(mx[1])};
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < 4; q++) fred[q][w] = b[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < UPD_THREADS / 64; q++) {
            b[0] = fminf(b[0], fred[0][q]); b[1] = fmaxf(b[1], fred[1][q]);
            b[2] = fminf(b[2], fred[2][q]); b[3] = fmaxf(b[3], fred[3][q]);
        }
        reinterpret_cast<float4 *>(J.fpart)[bc.lb] = make_float4(b[0], b[1], b[2], b[3]);
        J.mpart[2 * bc.lb] = s0;
        J.mpart[2 * bc.lb + 1] = s1;
        if (bc.lb == 0) {
            J.scal->att_queue = 0;  // the attraction's work queue, for the next epoch
            J.scal->epoch_next = epoch + 1;
            J.scal->mean_pending = 1;
        }
    }
}

void update_points(const LaunchSet &L, hipStream_t s) {
    if (L.dim >= 3) {
        nd_update(L, s);
        return;
    }
    const dim3 g(L.nb[F_PT]);
    if (L.dim == 2)
        hipLaunchKernelGGL(k_update<2>, g, dim3(UPD_THREADS), 0, s, L.jobs, L.map[F_PT]);
    else
        hipLaunchKernelGGL(k_update<1>, g, dim3(UPD_THREADS), 0, s, L.jobs, L.map[F_PT]);
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
