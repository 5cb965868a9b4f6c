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
__global__ void __launch_bounds__(UP_THREADS) k_attract(UpdateArgs a, float *pos_out) {
    __shared__ double red[UP_THREADS / 64];
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * UP_THREADS + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * UP_THREADS) >> 6;
    int epoch = 0;
    float inv_z = 0.0f, momentum = a.momentum;
    bool stop_now = false;
    if (UPDATE) {
        epoch = a.scal->epoch;
        inv_z = a.scal->inv_z;
        // stop_lying fires after epoch == stop_lying_epoch (lib.rs:1168-1171): the next epoch is
        // the first to read the rescaled values, and rescales them in place as it reads.
        stop_now = a.stop_lying_epoch != 0 && epoch == a.stop_lying_epoch + 1;
        // the momentum switch likewise takes effect after epoch == momentum_switch_epoch
        momentum = epoch > a.momentum_switch_epoch ? a.final_momentum : a.momentum;
    }
    double m0 = 0.0, m1 = 0.0;
    for (int i = wave; i < a.n; i += nwaves) {
        float yi0, yi1 = 0.0f;
        if (D == 2) {
            const float2 v = reinterpret_cast<const float2 *>(a.y)[i];
            yi0 = v.x;
            yi1 = v.y;
        } else {
            yi0 = a.y[i];
        }
        float acc0 = 0.0f, acc1 = 0.0f;
        const uint32_t e0 = a.rows[i], e1 = a.rows[i + 1];
        for (uint32_t base = e0; base < e1; base += 64 * UNROLL) {
            uint32_t col[UNROLL];
            float pv[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                const uint32_t e = base + lane + 64 * u;
                const bool ok = e < e1;
                col[u] = ok ? __builtin_nontemporal_load(a.cols + e) : (uint32_t)i;
                pv[u] = ok ? __builtin_nontemporal_load(a.vals + e) : 0.0f;
                if (UPDATE && stop_now && ok) {
                    pv[u] = pv[u] * a.inv_exag;
                    a.vals[e] = pv[u];
                }
            }
            float yj0[UNROLL], yj1[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                if (D == 2) {
                    const float2 v = reinterpret_cast<const float2 *>(a.y)[col[u]];
                    yj0[u] = v.x;
                    yj1[u] = v.y;
                } else {
                    yj0[u] = a.y[col[u]];
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
            const float gradient = acc[d] - a.frep[k] * inv_z;
            float uy = a.uy[k];
            float gain = a.gains[k];
            gain = rust_signum(gradient) != rust_signum(uy) ? gain + 0.2f : gain * 0.8f;
            if (gain < 0.01f) gain = 0.01f;
            uy = momentum * uy - a.learning_rate * gain * gradient;
            const float yn = yi[d] + uy;
            a.gains[k] = gain;
            a.uy[k] = uy;
            a.y_next[k] = yn;
            if (d == 0) m0 += (double)yn;
            else m1 += (double)yn;
        }
    }
    if (!UPDATE) return;
    // per-block sums of the new embedding for the next recentring
    double s0 = block_sum_f64(m0, red);
    double s1 = D == 2 ? block_sum_f64(m1, red) : 0.0;
    if (threadIdx.x == 0) {
        a.mpart[2 * blockIdx.x] = s0;
        a.mpart[2 * blockIdx.x + 1] = s1;
        if (blockIdx.x == 0) {
            a.scal->epoch_next = epoch + 1;
            a.scal->mean_pending = 1;
        }
    }
}

void attract_update(const UpdateArgs &a, hipStream_t s) {
    const int nb = attract_blocks(a.n);
    if (a.dim == 2)
        hipLaunchKernelGGL((k_attract<2, true>), dim3(nb), dim3(UP_THREADS), 0, s, a, nullptr);
    else
        hipLaunchKernelGGL((k_attract<1, true>), dim3(nb), dim3(UP_THREADS), 0, s, a, nullptr);
    BHT_LAUNCH_CHECK();
}

void attract_only(int n, int dim, const uint32_t *rows, const uint32_t *cols, const float *vals,
                  const float *y, float *pos, hipStream_t s) {
    UpdateArgs a{};
    a.n = n;
    a.dim = dim;
    a.rows = rows;
    a.cols = cols;
    a.vals = const_cast<float *>(vals);
    a.y = y;
    const int nb = attract_blocks(n);
    if (dim == 2)
        hipLaunchKernelGGL((k_attract<2, false>), dim3(nb), dim3(UP_THREADS), 0, s, a, pos);
    else
        hipLaunchKernelGGL((k_attract<1, false>), dim3(nb), dim3(UP_THREADS), 0, s, a, pos);
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
