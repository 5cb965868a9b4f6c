// FP64 brute-force repulsion: the non-edge terms of the reference's exact gradient
// (src/lib.rs:570-622: grad_i = sum_j (p_ij - q_ij / Z) q_ij (y_i - y_j), q_ij = 1 / (1 + d_ij^2),
// Z = sum over ordered pairs of q_ij), evaluated for every point against all others in double
// precision.  It is the O(N^2) yardstick for the error of the approximate gradients (FIt-SNE
// interpolation, Barnes-Hut) at full problem sizes, where a CPU evaluation would take minutes.
//
// One thread per point; the other points stream through LDS in tiles of 256.  Every point's
// terms are summed in tile order and Z is combined from per-block partials on the host in block
// order, so the result is deterministic.
#include "internal.hpp"

namespace bht {

constexpr int EX_THREADS = 256;

template <int D>
__global__ void __launch_bounds__(EX_THREADS) k_exact_rep(const float *y, int n, double *neg,
                                                          double *zpart) {
    __shared__ float tile[EX_THREADS * D];
    __shared__ double red[EX_THREADS / 64];
    const int i = blockIdx.x * EX_THREADS + threadIdx.x;
    const bool valid = i < n;
    double yi[2] = {0.0, 0.0};
    if (valid)
        for (int a = 0; a < D; a++) yi[a] = (double)y[(size_t)i * D + a];
    double f[2] = {0.0, 0.0}, z = 0.0;
    for (int j0 = 0; j0 < n; j0 += EX_THREADS) {
        const int m = min(EX_THREADS, n - j0);
        __syncthreads();
        if ((int)threadIdx.x < m)
            for (int a = 0; a < D; a++)
                tile[threadIdx.x * D + a] = y[(size_t)(j0 + threadIdx.x) * D + a];
        __syncthreads();
        for (int t = 0; t < m; t++) {
            double dl[2] = {0.0, 0.0}, d2 = 0.0;
            for (int a = 0; a < D; a++) {
                dl[a] = yi[a] - (double)tile[t * D + a];
                d2 += dl[a] * dl[a];
            }
            const double q = 1.0 / (1.0 + d2);
            const bool self = j0 + t == i;
            z += self ? 0.0 : q;
            const double w = self ? 0.0 : q * q;
            for (int a = 0; a < D; a++) f[a] += w * dl[a];
        }
    }
    if (valid)
        for (int a = 0; a < D; a++) neg[(size_t)i * D + a] = f[a];
    const double b = block_sum_f64(valid ? z : 0.0, red);
    if (threadIdx.x == 0) zpart[blockIdx.x] = b;
}

int exact_repulsion_blocks(int64_t n) { return (int)ceil_div((uint64_t)(n > 0 ? n : 1), EX_THREADS); }

void exact_repulsion_f64(const float *y, int64_t n, int dim, double *neg, double *zpart,
                         hipStream_t s) {
    const dim3 g(exact_repulsion_blocks(n));
    if (dim == 2)
        hipLaunchKernelGGL(k_exact_rep<2>, g, dim3(EX_THREADS), 0, s, y, (int)n, neg, zpart);
    else
        hipLaunchKernelGGL(k_exact_rep<1>, g, dim3(EX_THREADS), 0, s, y, (int)n, neg, zpart);
    BHT_LAUNCH_CHECK();
}

}  // namespace bht
