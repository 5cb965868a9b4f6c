// Attractive forces over the CSR graph fused with the gains / momentum update and the partial
// sums of the next recentring (src/tsne/arena.rs:398-427, src/tsne/mod.rs:470-516, 534-568).
//
// A group of 16 lanes owns one CSR row: lanes stride over the row's edges (coalesced column and
// value loads, gathered y[col]), reduce with shuffles, and lane 0 applies the update.  The new
// embedding goes to a second buffer (y_next) because other rows still read y this epoch.
#include "internal.hpp"

namespace bht {

constexpr int UP_THREADS = 256;
constexpr int GROUP = 16;

template <int D, bool UPDATE>
__global__ void __launch_bounds__(UP_THREADS) k_attract(UpdateArgs a, float *pos_out) {
    __shared__ double red[UP_THREADS / 64];
    const int sub = threadIdx.x & (GROUP - 1);
    const int grp = (blockIdx.x * UP_THREADS + threadIdx.x) / GROUP;
    const int ngrp = gridDim.x * UP_THREADS / GROUP;
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
    for (int i = grp; i < a.n; i += ngrp) {
        float yi[2];
        yi[0] = a.y[(size_t)i * D];
        yi[1] = D == 2 ? a.y[(size_t)i * D + 1] : 0.0f;
        float acc0 = 0.0f, acc1 = 0.0f;
        const uint32_t e1 = a.rows[i + 1];
        for (uint32_t e = a.rows[i] + sub; e < e1; e += GROUP) {
            const uint32_t j = a.cols[e];
            float pv = a.vals[e];
            if (UPDATE && stop_now) {
                pv = pv * a.inv_exag;
                a.vals[e] = pv;
            }
            float d0 = yi[0] - a.y[(size_t)j * D];
            float dist = 0.0f;
            dist += d0 * d0;
            float d1 = 0.0f;
            if (D == 2) {
                d1 = yi[1] - a.y[(size_t)j * D + 1];
                dist += d1 * d1;
            }
            const float factor = pv / (dist + 1.0f);
            acc0 += factor * d0;
            if (D == 2) acc1 += factor * d1;
        }
#pragma unroll
        for (int off = GROUP / 2; off > 0; off >>= 1) {
            acc0 += __shfl_xor(acc0, off, GROUP);
            if (D == 2) acc1 += __shfl_xor(acc1, off, GROUP);
        }
        if (sub != 0) continue;
        if (!UPDATE) {
            pos_out[(size_t)i * D] = acc0;
            if (D == 2) pos_out[(size_t)i * D + 1] = acc1;
            continue;
        }
        float acc[2] = {acc0, acc1};
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
    // partial sums of the new embedding for the next recentring
    double s0 = block_sum_f64(m0, red);
    double s1 = D == 2 ? block_sum_f64(m1, red) : 0.0;
    if (threadIdx.x == 0) {
        a.dpart[2 * blockIdx.x] = s0;
        a.dpart[2 * blockIdx.x + 1] = s1;
    }
    if (!arrive_last(&a.scal->ctr[CTR_UPDATE], gridDim.x)) return;
    double t0 = 0.0, t1 = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += UP_THREADS) {
        t0 += a.dpart[2 * b];
        t1 += a.dpart[2 * b + 1];
    }
    t0 = block_sum_f64(t0, red);
    t1 = block_sum_f64(t1, red);
    if (threadIdx.x == 0) {
        a.scal->mean[0] = (float)(t0 / (double)a.n);
        a.scal->mean[1] = D == 2 ? (float)(t1 / (double)a.n) : 0.0f;
        a.scal->epoch = epoch + 1;
        a.scal->ctr[CTR_UPDATE] = 0;
    }
}

static int update_blocks(int n) {
    uint64_t b = ceil_div((uint64_t)(n > 0 ? n : 1) * GROUP, UP_THREADS);
    return (int)(b < 2048 ? b : 2048);
}

void attract_update(const UpdateArgs &a, hipStream_t s) {
    const int nb = update_blocks(a.n);
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
    const int nb = update_blocks(n);
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
