// FP64 brute-force repulsion: the non-edge terms of the reference's exact gradient
// (src/lib.rs:570-622: grad_i = sum_j (p_ij - q_ij / Z) q_ij (y_i - y_j), q_ij = 1 / (1 + d_ij^2),
// Z = sum over ordered pairs of q_ij), evaluated for every point against all others in double
// precision.  It is the O(N^2) yardstick for the error of the approximate gradients (FIt-SNE
// interpolation, Barnes-Hut) at full problem sizes, where a CPU evaluation would take minutes.
//
// One thread per point; the other points stream through LDS in tiles of 256.  Every point's
// terms are summed in tile order and Z is combined from per-block partials on the host in block
// order, so the result is deterministic.
#include <map>
#include <mutex>

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

// ---- the variant that allocates nothing ---------------------------------------------------------
//
// Its per-block partials of Z live in `scratch`, a file-scope device array (each GPU holds its
// own copy for the life of the process), instead of a buffer the call allocates.  In host code
// the array's name only stands for the symbol: its device address is looked up once per device
// (exact_scratch_init, with that device current) and only that address is handed to kernels.
// Every call on one GPU shares the array, so calls that overlap on one GPU race; every partial is
// written before it is read, so nothing needs clearing between calls.
constexpr int NOALLOC_MAX_BLOCKS = 65536;  // 512 KiB: n <= 65536 * 256 = 16,777,216 points
static __device__ __align__(256) double scratch[NOALLOC_MAX_BLOCKS];

namespace {
std::mutex g_scratch_mu;
std::map<int, double *> g_scratch_of_device;

// Z = the partials summed in block order by one thread (the order the allocating entry point
// sums them in on the host, so both give the same bits).
__global__ void k_sum_partials(const double *part, int nparts, double *z_out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double z = 0.0;
    for (int b = 0; b < nparts; b++) z += part[b];
    *z_out = z;
}
}  // namespace

uint64_t exact_noalloc_max_points() { return (uint64_t)NOALLOC_MAX_BLOCKS * EX_THREADS; }

void exact_scratch_init(int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_of_device.count(device)) return;
    BHT_HIP(hipSetDevice(device));
    void *ptr = nullptr;
    BHT_HIP(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    g_scratch_of_device[device] = (double *)ptr;
}

void exact_repulsion_f64_noalloc(int device, const float *y, int64_t n, int dim, double *neg,
                                 double *z_out, hipStream_t s) {
    double *part = nullptr;
    {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        auto it = g_scratch_of_device.find(device);
        if (it == g_scratch_of_device.end())
            throw Error(ERR_STATE, "the device's scratch was not set up (bht_ctx_create)");
        part = it->second;
    }
    const int nb = exact_repulsion_blocks(n);
    if (nb > NOALLOC_MAX_BLOCKS)
        throw Error(ERR_INVALID, "at most 16777216 points without allocation");
    exact_repulsion_f64(y, n, dim, neg, part, s);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, part, nb, z_out);
    BHT_LAUNCH_CHECK();
}

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
