// Exact k nearest neighbours on the device (replaces the VP-tree search, src/tsne/vptree.rs:193-319,
// for the Euclidean metric).
//
// Distances use the reference metric's own f32 arithmetic: sqrt of the left-to-right sum of
// (a_t - b_t)^2 with no FMA contraction, so every distance is bit-identical to the CPU metric and
// neighbour index sets are exact, with ties broken by ascending index (src/test.rs:866-881).
//
// A block owns 32 query rows and streams every candidate in tiles of 64 through LDS; each thread
// computes a 2 x 4 register tile of distances.  Candidates beating the row's current k-th key
// (distance bits << 32 | index: ordering keys orders (distance, index)) are appended to an LDS
// buffer that a wave compacts by a bitonic sort whenever it fills.
#include "affinity.hpp"

namespace bht {

constexpr int KQ = 32;     // query rows per block
constexpr int KC = 64;     // candidates per tile
constexpr int KD = 64;     // dimensions per LDS chunk
constexpr int KBUF = 256;  // candidate slots per row (k + KC must fit)
constexpr int KNN_THREADS = 256;

__device__ __forceinline__ void wave_bitonic_u64(uint64_t *a, int lane) {
    for (int size = 2; size <= KBUF; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
            for (int t = lane; t < KBUF / 2; t += 64) {
                int i = 2 * t - (t & (stride - 1));
                int j = i + stride;
                bool up = (i & size) == 0;
                uint64_t x = a[i], y = a[j];
                if ((x > y) == up) { a[i] = y; a[j] = x; }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

__global__ void __launch_bounds__(KNN_THREADS) k_knn(const float *x, int n, int D, int k,
                                                     uint32_t *idx_out, float *dist_out) {
    __shared__ float qs[KQ][KD + 1];
    __shared__ float cs[KC][KD + 1];
    __shared__ uint64_t buf[KQ][KBUF];
    __shared__ uint32_t cnt[KQ];
    __shared__ uint64_t tau[KQ];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int q0 = blockIdx.x * KQ;
    const int tq = t >> 4, tc = t & 15;
    if (t < KQ) {
        cnt[t] = 0;
        tau[t] = ~0ull;
    }
    const bool one_chunk = D <= KD;
    if (one_chunk) {
        for (int e = t; e < KQ * KD; e += KNN_THREADS) {
            int r = e / KD, c = e % KD;
            qs[r][c] = (q0 + r < n && c < D) ? x[(size_t)(q0 + r) * D + c] : 0.0f;
        }
    }
    __syncthreads();
    for (int c0 = 0; c0 < n; c0 += KC) {
        float acc[2][4];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) acc[a][b] = 0.0f;
        for (int d0 = 0; d0 < D; d0 += KD) {
            const int dl = min(KD, D - d0);
            if (!one_chunk) {
                for (int e = t; e < KQ * KD; e += KNN_THREADS) {
                    int r = e / KD, c = e % KD;
                    qs[r][c] = (q0 + r < n && c < dl) ? x[(size_t)(q0 + r) * D + d0 + c] : 0.0f;
                }
            }
            for (int e = t; e < KC * KD; e += KNN_THREADS) {
                int r = e / KD, c = e % KD;
                cs[r][c] = (c0 + r < n && c < dl) ? x[(size_t)(c0 + r) * D + d0 + c] : 0.0f;
            }
            __syncthreads();
            for (int dd = 0; dd < dl; dd++) {
                const float a0 = qs[tq][dd], a1 = qs[tq + 16][dd];
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const float cv = cs[tc + 16 * b][dd];
                    const float e0 = cv - a0, e1 = cv - a1;
                    acc[0][b] = acc[0][b] + e0 * e0;
                    acc[1][b] = acc[1][b] + e1 * e1;
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int a = 0; a < 2; a++) {
            const int qi = tq + 16 * a, gq = q0 + qi;
            if (gq >= n) continue;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int j = c0 + tc + 16 * b;
                if (j >= n || j == gq) continue;
                const float dv = sqrtf(acc[a][b]);
                const uint64_t key = ((uint64_t)__float_as_uint(dv) << 32) | (uint32_t)j;
                if (key < tau[qi]) {
                    uint32_t pos = atomicAdd(&cnt[qi], 1u);
                    buf[qi][pos] = key;
                }
            }
        }
        __syncthreads();
        for (int qi = w; qi < KQ; qi += KNN_THREADS / 64) {
            const uint32_t c = cnt[qi];
            if (c <= (uint32_t)(KBUF - KC)) continue;
            for (int s = lane; s < KBUF; s += 64)
                if (s >= (int)c) buf[qi][s] = ~0ull;
            __builtin_amdgcn_wave_barrier();
            wave_bitonic_u64(buf[qi], lane);
            if (lane == 0) {
                uint32_t keep = c < (uint32_t)k ? c : (uint32_t)k;
                cnt[qi] = keep;
                if (keep == (uint32_t)k) tau[qi] = buf[qi][k - 1];
            }
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
    }
    for (int qi = w; qi < KQ; qi += KNN_THREADS / 64) {
        const int gq = q0 + qi;
        if (gq >= n) continue;
        const uint32_t c = cnt[qi];
        for (int s = lane; s < KBUF; s += 64)
            if (s >= (int)c) buf[qi][s] = ~0ull;
        __builtin_amdgcn_wave_barrier();
        wave_bitonic_u64(buf[qi], lane);
        for (int s = lane; s < k; s += 64) {
            uint64_t key = buf[qi][s];
            bool ok = s < (int)c;
            idx_out[(size_t)gq * k + s] = ok ? (uint32_t)(key & 0xffffffffu) : 0u;
            dist_out[(size_t)gq * k + s] = ok ? __uint_as_float((uint32_t)(key >> 32)) : 0.0f;
        }
    }
}

void knn_exact(const float *x, int64_t n, int d, int k, uint32_t *idx, float *dist,
               hipStream_t s) {
    if (k <= 0 || n <= 0) return;
    if (k > KBUF - KC) throw Error(ERR_INVALID, "at most 192 neighbours per row are supported");
    hipLaunchKernelGGL(k_knn, dim3(ceil_div((uint64_t)n, KQ)), dim3(KNN_THREADS), 0, s, x,
                       (int)n, d, k, idx, dist);
    BHT_LAUNCH_CHECK();
}

}  // namespace bht
