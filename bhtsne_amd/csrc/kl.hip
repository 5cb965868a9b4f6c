// Sparse KL divergence over the CSR graph (src/tsne/mod.rs:738-775).
//
// One thread per sample accumulates its row in f32 in CSR order exactly as the reference,
// including its weighting of every edge of row i by p_values[i] (the value stored at CSR position
// i, not the edge's own value); the true KL, sum over edges of p_e log(p_e / q_e), is accumulated
// beside it in FP64.  Row partials are summed in FP64.
#include <cfloat>

#include "internal.hpp"

namespace bht {

// qv[s] is the value at CSR position s of the reference's layout (s < n); perm maps the engine's
// point order to sample indices (null when the points are in input order).
template <int D>
__global__ void __launch_bounds__(256) k_sparse_kl(const float *y, int n, const uint32_t *rows,
                                                   const uint32_t *cols, const float *vals,
                                                   const float *qv, const uint32_t *perm,
                                                   EpochScalars *scal, double *dpart) {
    __shared__ double red[4];
    const float inv_z = scal->inv_z;
    double ref = 0.0, tru = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float pi = qv[perm ? perm[i] : (uint32_t)i];
        float a[D];
#pragma unroll
        for (int t = 0; t < D; t++) a[t] = y[(size_t)i * D + t];
        float cost = 0.0f;
        for (uint32_t e = rows[i]; e < rows[i + 1]; e++) {
            const uint32_t j = cols[e];
            float q = 0.0f;
#pragma unroll
            for (int t = 0; t < D; t++) {
                const float dt = a[t] - y[(size_t)j * D + t];
                q = q + dt * dt;
            }
            q = (1.0f / (1.0f + q)) * inv_z;
            cost += pi * logf((pi + FLT_MIN) / (q + FLT_MIN));
            const float pe = vals[e];
            if (pe > 0.0f) tru += (double)pe * log((double)pe / (double)q);
        }
        ref += (double)cost;
    }
    double r = block_sum_f64(ref, red);
    double t = block_sum_f64(tru, red);
    if (threadIdx.x == 0) {
        dpart[2 * blockIdx.x] = r;
        dpart[2 * blockIdx.x + 1] = t;
    }
    if (!arrive_last(&scal->ctr[CTR_KL], gridDim.x)) return;
    r = 0.0;
    t = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += 256) {
        r += dpart[2 * b];
        t += dpart[2 * b + 1];
    }
    r = block_sum_f64(r, red);
    t = block_sum_f64(t, red);
    if (threadIdx.x == 0) {
        scal->kl_ref = r;
        scal->kl_true = t;
        scal->ctr[CTR_KL] = 0;
    }
}

void sparse_kl(const float *y, int n, int dim, const uint32_t *rows, const uint32_t *cols,
               const float *vals, const float *qv, const uint32_t *perm, EpochScalars *scal,
               double *dpart, hipStream_t s) {
    int nb = (int)std::min<uint64_t>(ceil_div((uint64_t)(n > 0 ? n : 1), 256), 2048);
    auto launch = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(nb), dim3(256), 0, s, y, n, rows, cols, vals, qv, perm,
                           scal, dpart);
    };
    switch (dim) {
    case 1: launch(k_sparse_kl<1>); break;
    case 2: launch(k_sparse_kl<2>); break;
    case 3: launch(k_sparse_kl<3>); break;
    case 4: launch(k_sparse_kl<4>); break;
    case 5: launch(k_sparse_kl<5>); break;
    case 6: launch(k_sparse_kl<6>); break;
    case 7: launch(k_sparse_kl<7>); break;
    default: throw Error(ERR_INVALID, "KL divergence: d_out must be in 1..7");
    }
    BHT_LAUNCH_CHECK();
}

}  // namespace bht
