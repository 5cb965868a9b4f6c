// The exact O(N^2) fit (src/lib.rs:489-646): dense input distances, a dense bandwidth search per
// row (affinity.hip), the dense symmetrisation, and per epoch the exact gradient
//   dy_i = sum_j (y_i - y_j) (p_ij - q_ij / Z) q_ij,   q_ij = 1 / (1 + |y_i - y_j|^2),
// with Z the sum of q over all n^2 ordered pairs, the n diagonal terms (q = 1) included, as the
// reference sums its dense Q (src/lib.rs:583-592).  Dense rows include each point's own entry,
// as the reference's do: d_ii = 0, so p_ii is that row's largest conditional value.  The optimiser
// step and the recentring are the approximate paths' kernels.
//
// Layout: P is n x n f32, row-major, in HBM (the fit's only O(N^2) state); the input distances
// use a second n x n buffer that is released once P is built.  Each epoch reads P once (the
// gradient pass) and y twice (the Z pass and the gradient pass), one block per row; the other
// points stream through the block's threads in index order.  Per-row sums are FP64 and each
// block reduces its threads' sums in a fixed order, so a fit repeats bit for bit.
#include <cfloat>

#include "internal.hpp"

namespace bht {

constexpr int XF_THREADS = 256;

// Distances between the n samples (d_in f32 features each): sqrt of the left-to-right sum of
// squared differences (metric 0, Euclidean) or that sum itself (metric 1, squared Euclidean, the
// distance the reference's exact tests pass).  Block b covers row b / cb, columns (b % cb) * 256
// onwards (cb column blocks per row).
__global__ void __launch_bounds__(XF_THREADS) k_dense_dist(const float *x, int64_t n, int d_in,
                                                           int metric, float *dist) {
    const int64_t cb = (n + XF_THREADS - 1) / XF_THREADS;
    const int64_t i = blockIdx.x / cb;
    const int64_t j = (blockIdx.x % cb) * XF_THREADS + threadIdx.x;
    if (j >= n) return;
    const float *xi = x + i * d_in, *xj = x + j * d_in;
    float s = 0.0f;
    for (int t = 0; t < d_in; t++) {
        const float diff = xi[t] - xj[t];
        s = s + diff * diff;
    }
    dist[i * n + j] = metric == 0 ? sqrtf(s) : s;
}

// P_ij = P_ji = p_ij + p_ji for i < j (src/lib.rs:546-552); the diagonal is kept.
__global__ void __launch_bounds__(XF_THREADS) k_sym_dense(float *p, int64_t n) {
    const int64_t cb = (n + XF_THREADS - 1) / XF_THREADS;
    const int64_t i = blockIdx.x / cb;
    const int64_t j = (blockIdx.x % cb) * XF_THREADS + threadIdx.x;
    if (j <= i || j >= n) return;
    const float s = p[i * n + j] + p[j * n + i];
    p[i * n + j] = s;
    p[j * n + i] = s;
}

// |y_i - y_j|^2 with the reference's operation order (sum of powi(2) terms, left to right).
template <int D>
__device__ __forceinline__ float sq_dist(const float *yi, const float *y, int64_t j) {
    float s = 0.0f;
#pragma unroll
    for (int a = 0; a < D; a++) {
        const float diff = yi[a] - y[j * D + a];
        s = s + diff * diff;
    }
    return s;
}

// Per row i: the sum over j of q_ij (FP64).
template <int D>
__global__ void __launch_bounds__(XF_THREADS) k_exact_zrows(const JobDev *jobs, double *zrow) {
    __shared__ double red[XF_THREADS / 64];
    const JobDev &J = jobs[0];
    const float *y = job_y(J);
    const int64_t n = J.n, i = blockIdx.x;
    float yi[D];
#pragma unroll
    for (int a = 0; a < D; a++) yi[a] = y[i * D + a];
    double z = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += XF_THREADS)
        z += (double)(1.0f / (1.0f + sq_dist<D>(yi, y, j)));
    const double b = block_sum_f64(z, red);
    if (threadIdx.x == 0) zrow[i] = b;
}

// Z = the row sums in row order; inv_z = 1 / Z in f32 (the reference's inverse_q_sum).
__global__ void __launch_bounds__(1024) k_exact_zfinal(const JobDev *jobs, const double *zrow) {
    __shared__ double red[16];
    const JobDev &J = jobs[0];
    double t = 0.0;
    for (int64_t i = threadIdx.x; i < J.n; i += 1024) t += zrow[i];
    t = block_sum_f64(t, red);
    if (threadIdx.x == 0) {
        J.scal->zsum = t;
        J.scal->z = (float)t;
        J.scal->inv_z = 1.0f / (float)t;
    }
}

// Per row i: dy_i = sum_j (y_i - y_j) m_ij, m_ij = (p_ij - q_ij inv_z) q_ij in f32 as the
// reference forms it, summed in FP64; P is read divided by the exaggeration once stop-lying has
// fired (the epochs after stop_lying_epoch).  The gradient goes to the attraction sums, which
// the update kernel takes as the whole gradient (the repulsion buffer stays zero).
template <int D>
__global__ void __launch_bounds__(XF_THREADS) k_exact_grad(const JobDev *jobs, const float *P) {
    __shared__ double red[XF_THREADS / 64];
    const JobDev &J = jobs[0];
    const float *y = job_y(J);
    const int64_t n = J.n, i = blockIdx.x;
    const int epoch = J.scal->epoch;
    const bool lied_off = J.stop_lying_epoch != 0 && epoch > J.stop_lying_epoch;
    const float inv_exag = J.inv_exag, inv_z = J.scal->inv_z;
    float yi[D];
#pragma unroll
    for (int a = 0; a < D; a++) yi[a] = y[i * D + a];
    double g[D];
#pragma unroll
    for (int a = 0; a < D; a++) g[a] = 0.0;
    const float *prow = P + i * n;
    for (int64_t j = threadIdx.x; j < n; j += XF_THREADS) {
        const float q = 1.0f / (1.0f + sq_dist<D>(yi, y, j));
        float p = prow[j];
        if (lied_off) p = p * inv_exag;
        const float m = (p - q * inv_z) * q;
#pragma unroll
        for (int a = 0; a < D; a++) g[a] += (double)((yi[a] - y[j * D + a]) * m);
    }
#pragma unroll
    for (int a = 0; a < D; a++) {
        const double b = block_sum_f64(g[a], red);
        if (threadIdx.x == 0) J.fattr[i * D + a] = (float)b;
    }
}

// KL of the current embedding (src/tsne/mod.rs:579-618): per row, sum_j p ln((p + min) /
// (q inv_z + min)) over all n^2 pairs (FP64 sums; q and q inv_z in f32).
template <int D>
__global__ void __launch_bounds__(XF_THREADS) k_exact_kl(const JobDev *jobs, const float *P,
                                                         double *klrow) {
    __shared__ double red[XF_THREADS / 64];
    const JobDev &J = jobs[0];
    const float *y = job_y(J);
    const int64_t n = J.n, i = blockIdx.x;
    const float inv_z = J.scal->inv_z;
    float yi[D];
#pragma unroll
    for (int a = 0; a < D; a++) yi[a] = y[i * D + a];
    double kl = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += XF_THREADS) {
        const float q = (1.0f / (1.0f + sq_dist<D>(yi, y, j))) * inv_z;
        const float p = P[i * n + j];
        kl += (double)p * log(((double)p + (double)FLT_MIN) / ((double)q + (double)FLT_MIN));
    }
    const double b = block_sum_f64(kl, red);
    if (threadIdx.x == 0) klrow[i] = b;
}

__global__ void __launch_bounds__(1024) k_exact_klfinal(const JobDev *jobs, const double *klrow) {
    __shared__ double red[16];
    const JobDev &J = jobs[0];
    double t = 0.0;
    for (int64_t i = threadIdx.x; i < J.n; i += 1024) t += klrow[i];
    t = block_sum_f64(t, red);
    if (threadIdx.x == 0) {
        J.scal->kl_true = t;
        J.scal->kl_ref = (double)(float)t;
    }
}

void exact_distances(const float *x, int64_t n, int d_in, int metric, float *dist,
                     hipStream_t s) {
    if (n == 0) return;
    const dim3 g((unsigned)(ceil_div((uint64_t)n, XF_THREADS) * (uint64_t)n));
    hipLaunchKernelGGL(k_dense_dist, g, dim3(XF_THREADS), 0, s, x, n, d_in, metric, dist);
    BHT_LAUNCH_CHECK();
}

void exact_symmetrize(float *p, int64_t n, hipStream_t s) {
    if (n == 0) return;
    const dim3 g((unsigned)(ceil_div((uint64_t)n, XF_THREADS) * (uint64_t)n));
    hipLaunchKernelGGL(k_sym_dense, g, dim3(XF_THREADS), 0, s, p, n);
    BHT_LAUNCH_CHECK();
}

// Instantiates f for the output dimension (1..7).
template <typename F>
void with_dim(int dim, F &&f) {
    switch (dim) {
    case 1: f(std::integral_constant<int, 1>()); break;
    case 2: f(std::integral_constant<int, 2>()); break;
    case 3: f(std::integral_constant<int, 3>()); break;
    case 4: f(std::integral_constant<int, 4>()); break;
    case 5: f(std::integral_constant<int, 5>()); break;
    case 6: f(std::integral_constant<int, 6>()); break;
    case 7: f(std::integral_constant<int, 7>()); break;
    default: throw Error(ERR_INVALID, "the exact path supports d_out in 1..7");
    }
}

void exact_z(const JobDev *jobs, int64_t n, int dim, double *zrow, hipStream_t s) {
    if (n == 0) return;
    with_dim(dim, [&](auto d) {
        hipLaunchKernelGGL(k_exact_zrows<decltype(d)::value>, dim3((unsigned)n),
                           dim3(XF_THREADS), 0, s, jobs, zrow);
    });
    hipLaunchKernelGGL(k_exact_zfinal, dim3(1), dim3(1024), 0, s, jobs, (const double *)zrow);
    BHT_LAUNCH_CHECK();
}

void exact_gradient(const JobDev *jobs, const float *P, int64_t n, int dim, double *zrow,
                    hipStream_t s) {
    if (n == 0) return;
    exact_z(jobs, n, dim, zrow, s);
    with_dim(dim, [&](auto d) {
        hipLaunchKernelGGL(k_exact_grad<decltype(d)::value>, dim3((unsigned)n), dim3(XF_THREADS),
                           0, s, jobs, P);
    });
    BHT_LAUNCH_CHECK();
}

void exact_kl(const JobDev *jobs, const float *P, int64_t n, int dim, double *zrow,
              hipStream_t s) {
    if (n == 0) return;
    exact_z(jobs, n, dim, zrow, s);
    double *klrow = zrow + n;
    with_dim(dim, [&](auto d) {
        hipLaunchKernelGGL(k_exact_kl<decltype(d)::value>, dim3((unsigned)n), dim3(XF_THREADS),
                           0, s, jobs, P, klrow);
    });
    hipLaunchKernelGGL(k_exact_klfinal, dim3(1), dim3(1024), 0, s, jobs, (const double *)klrow);
    BHT_LAUNCH_CHECK();
}

}  // namespace bht
