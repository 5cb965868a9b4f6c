// Attractive forces over the CSR graph fused with the gains / momentum update and the partial
// sums of the next recentring (src/tsne/arena.rs:398-427, src/tsne/mod.rs:470-516, 534-568).
//
// One wave owns one CSR row at a time.  Each lane takes up to four edges of the row per step
// (issuing their column, value and gathered y loads before using any), so a row of ~150 edges
// needs one or two dependent memory round trips instead of ten.  The wave reduces with a fixed
// butterfly and lane 0 applies the update.  The new embedding goes to a second buffer (y_next)
// because other rows still read y this epoch.
#include "internal.hpp"

namespace bht {

constexpr int UP_THREADS = 256;
constexpr int UNROLL = 4;

int attract_blocks(int n) {
    uint64_t b = ceil_div((uint64_t)(n > 0 ? n : 1), UP_THREADS / 64);
    return (int)(b < 4096 ? b : 4096);
}

template <int D, bool UPDATE>
__global__ void __launch_bounds__(UP_THREADS) k_attract(const JobDev *jobs, const uint32_t *map,
                                                        float *pos_out) {
    __shared__ double red[UP_THREADS / 64];
    const BlockCtx bc = block_ctx(map, jobs, F_ATT);
    const JobDev &J = jobs[bc.job];
    const int lane = threadIdx.x & 63;
    const int wave = (bc.lb * UP_THREADS + threadIdx.x) >> 6;
    const int nwaves = (bc.nb * UP_THREADS) >> 6;
    const int epoch = J.scal->epoch;
    const int n = J.n;
    const float *y = J.ybuf[epoch & 1];
    float *y_next = J.ybuf[(epoch + 1) & 1];
    const uint32_t *rows = J.rows, *cols = J.cols;
    float *vals = J.vals;
    const float inv_exag = J.inv_exag;
    float inv_z = 0.0f, momentum = J.momentum;
    bool stop_now = false;
    if (UPDATE) {
        inv_z = J.scal->inv_z;
        // stop_lying fires after epoch == stop_lying_epoch (lib.rs:1168-1171): the next epoch is
        // the first to read the rescaled values, and rescales them in place as it reads (with
        // the copies of the reference-position values that KL weights by).
        stop_now = J.stop_lying_epoch != 0 && epoch == J.stop_lying_epoch + 1;
        // the momentum switch likewise takes effect after epoch == momentum_switch_epoch
        momentum = epoch > J.momentum_switch_epoch ? J.final_momentum : J.momentum;
    }
    double m0 = 0.0, m1 = 0.0;
    for (int i = wave; i < n; i += nwaves) {
        float yi0, yi1 = 0.0f;
        if (D == 2) {
            const float2 v = reinterpret_cast<const float2 *>(y)[i];
            yi0 = v.x;
            yi1 = v.y;
        } else {
            yi0 = y[i];
        }
        if (UPDATE && stop_now && lane == 0 && i < J.qv_n) J.qv[i] = J.qv[i] * inv_exag;
        float acc0 = 0.0f, acc1 = 0.0f;
        const uint32_t e0 = rows[i], e1 = rows[i + 1];
        for (uint32_t base = e0; base < e1; base += 64 * UNROLL) {
            uint32_t col[UNROLL];
            float pv[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                const uint32_t e = base + lane + 64 * u;
                const bool ok = e < e1;
                col[u] = ok ? __builtin_nontemporal_load(cols + e) : (uint32_t)i;
                pv[u] = ok ? __builtin_nontemporal_load(vals + e) : 0.0f;
                if (UPDATE && stop_now && ok) {
                    pv[u] = pv[u] * inv_exag;
                    vals[e] = pv[u];
                }
            }
            float yj0[UNROLL], yj1[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                if (D == 2) {
                    const float2 v = reinterpret_cast<const float2 *>(y)[col[u]];
                    yj0[u] = v.x;
                    yj1[u] = v.y;
                } else {
                    yj0[u] = y[col[u]];
                    yj1[u] = 0.0f;
                }
            }
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                const float d0 = yi0 - yj0[u];
                float dist = 0.0f;
                dist += d0 * d0;
                float d1 = 0.0f;
                if (D == 2) {
                    d1 = yi1 - yj1[u];
                    dist += d1 * d1;
                }
                const float factor = pv[u] / (dist + 1.0f);
                acc0 += factor * d0;
                if (D == 2) acc1 += factor * d1;
            }
        }
        acc0 = wave_sum_f32(acc0);
        if (D == 2) acc1 = wave_sum_f32(acc1);
        if (lane != 0) continue;
        if (!UPDATE) {
            pos_out[(size_t)i * D] = acc0;
            if (D == 2) pos_out[(size_t)i * D + 1] = acc1;
            continue;
        }
        const float acc[2] = {acc0, acc1};
        const float yi[2] = {yi0, yi1};
#pragma unroll
        for (int d = 0; d < D; d++) {
            const size_t k = (size_t)i * D + d;
            const float gradient = acc[d] - J.frep[k] * inv_z;
            float uy = J.uy[k];
            float gain = J.gains[k];
            gain = rust_signum(gradient) != rust_signum(uy) ? gain + 0.2f : gain * 0.8f;
            if (gain < 0.01f) gain = 0.01f;
            uy = momentum * uy - J.learning_rate * gain * gradient;
            const float yn = yi[d] + uy;
            J.gains[k] = gain;
            J.uy[k] = uy;
            y_next[k] = yn;
            if (d == 0) m0 += (double)yn;
            else m1 += (double)yn;
        }
    }
    if (!UPDATE) return;
    // per-block sums of the new embedding for the next recentring
    double s0 = block_sum_f64(m0, red);
    double s1 = D == 2 ? block_sum_f64(m1, red) : 0.0;
    if (threadIdx.x == 0) {
        J.mpart[2 * bc.lb] = s0;
        J.mpart[2 * bc.lb + 1] = s1;
        if (bc.lb == 0) {
            J.scal->epoch_next = epoch + 1;
            J.scal->mean_pending = 1;
        }
    }
}

void attract_update(const LaunchSet &L, hipStream_t s) {
    const dim3 g(L.nb[F_ATT]);
    if (L.dim == 2)
        hipLaunchKernelGGL((k_attract<2, true>), g, dim3(UP_THREADS), 0, s, L.jobs, L.map[F_ATT],
                           nullptr);
    else
        hipLaunchKernelGGL((k_attract<1, true>), g, dim3(UP_THREADS), 0, s, L.jobs, L.map[F_ATT],
                           nullptr);
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
