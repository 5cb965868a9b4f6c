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

// Query row r of block b: the rows listed in `rows` (the rows an approximate pass could not
// certify), or all rows in order.
__device__ __forceinline__ int query_row(const uint32_t *rows, int nrows, int q) {
    if (rows == nullptr) return q;
    return q < nrows ? (int)rows[q] : INT32_MAX;
}

__global__ void __launch_bounds__(KNN_THREADS) k_knn(const float *x, int n, int D, int k,
                                                     uint32_t *idx_out, float *dist_out,
                                                     const uint32_t *rows, int nrows) {
    __shared__ float qs[KQ][KD + 1];
    __shared__ float cs[KC][KD + 1];
    __shared__ uint64_t buf[KQ][KBUF];
    __shared__ uint32_t cnt[KQ];
    __shared__ uint64_t tau[KQ];
    __shared__ int qrow[KQ];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int q0 = blockIdx.x * KQ;
    const int tq = t >> 4, tc = t & 15;
    if (t < KQ) {
        const int r = query_row(rows, nrows, q0 + t);
        qrow[t] = r < n ? r : INT32_MAX;
    }
    if (t < KQ) {
        cnt[t] = 0;
        tau[t] = ~0ull;
    }
    __syncthreads();
    const bool one_chunk = D <= KD;
    if (one_chunk) {
        for (int e = t; e < KQ * KD; e += KNN_THREADS) {
            int r = e / KD, c = e % KD;
            qs[r][c] = (qrow[r] < n && c < D) ? x[(size_t)qrow[r] * D + c] : 0.0f;
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
                    qs[r][c] = (qrow[r] < n && c < dl) ? x[(size_t)qrow[r] * D + d0 + c] : 0.0f;
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
            const int qi = tq + 16 * a, gq = qrow[qi];
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
        const int gq = qrow[qi];
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
                       (int)n, d, k, idx, dist, nullptr, 0);
    BHT_LAUNCH_CHECK();
}

// ---- matrix-core candidate pass, exact re-rank, certificate ---------------------------------
//
// 1. Candidates: the squared distance |q|^2 + |c|^2 - 2 q.c with the dot products from FP32 MFMA
//    (v_mfma_f32_16x16x4_f32; a wave computes two 16 x 16 tiles: 32 queries x 16 candidates);
//    each row keeps its KCAND smallest (approximate distance, index) keys in the LDS buffer
//    scheme of the exact kernel.
// 2. Re-rank: the exact reference distance sqrt(sum (c_t - q_t)^2) (left to right, no FMA) of
//    every candidate, ordered by (distance, index); the first k are the row's neighbours.
// 3. Certificate: every point outside the candidates has an approximate squared distance of at
//    least tau (the KCAND-th kept one), hence an exact one of at least tau - delta, where delta
//    bounds the rounding of both computations.  When tau - delta exceeds the k-th exact squared
//    distance (widened by 2^-20 so that no f32 square root can tie), no outside point can be a
//    neighbour and the row is exact.  Rows without the certificate are recomputed by the exact
//    kernel, so the result is the exact kNN for every row.
constexpr int MQ = 32;      // query rows per block: two 16-row tiles
constexpr int MC = 64;      // candidates per step: four 16-column tiles, one per wave
constexpr int MKS = 16;     // k-steps of 4 dimensions (d <= 64)
constexpr int KCAND = 128;  // candidates kept per row
static_assert(KCAND + MC <= KBUF, "candidate buffer");

typedef float v4f __attribute__((ext_vector_type(4)));

// Element (row, d) of the samples, 0 outside [0, n) x [0, D).  The load address is clamped into
// the array so that the load is unconditional (no branch per operand).
__device__ __forceinline__ float operand(const float *x, int n, int D, int row, int d) {
    const bool ok = row < n && d < D;
    const float v = x[(size_t)(row < n ? row : n - 1) * D + (d < D ? d : D - 1)];
    return ok ? v : 0.0f;
}

// Block barrier ordering LDS accesses only: global loads issued before it stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__global__ void __launch_bounds__(256) k_sqnorm(const float *x, int n, int D, float *nrm,
                                                uint32_t *max_bits) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    float s = 0.0f;
    if (i < n) {
        for (int t = 0; t < D; t++) {
            const float v = x[(size_t)i * D + t];
            s = s + v * v;
        }
        nrm[i] = s;
    }
    // non-negative floats order like their bit patterns
    const float m = wave_max_f32(s);
    if ((threadIdx.x & 63) == 0) atomicMax(max_bits, __float_as_uint(m));
}

__global__ void __launch_bounds__(KNN_THREADS) k_knn_cand(const float *x, const float *nrm, int n,
                                                          int D, uint32_t *cand, float *tau_out) {
    __shared__ uint64_t buf[MQ][KBUF];
    __shared__ uint32_t cnt[MQ];
    __shared__ uint64_t tau[MQ];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int q0 = blockIdx.x * MQ;
    const int ksteps = (D + 3) / 4;
    if (t < MQ) {
        cnt[t] = 0;
        tau[t] = ~0ull;
    }
    // A operands (16 x 4 per step): lane holds query row (lane % 16) of the tile, dimension
    // 4 * step + lane / 16
    float a[2][MKS];
#pragma unroll
    for (int rt = 0; rt < 2; rt++) {
        const int row = q0 + rt * 16 + (lane & 15);
#pragma unroll
        for (int kb = 0; kb < MKS; kb++) a[rt][kb] = operand(x, n, D, row, kb * 4 + (lane >> 4));
    }
    // output element r of a lane is row 4 * (lane / 16) + r of the tile, column lane % 16
    float qn[2][4];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = q0 + rt * 16 + 4 * (lane >> 4) + r;
            qn[rt][r] = row < n ? nrm[row] : 0.0f;
        }
    __syncthreads();
    // B operands (4 x 16 per step): lane holds dimension 4 * step + lane / 16 of candidate column
    // lane % 16.  The next tile's operands and norms are loaded while this tile is reduced; the
    // block barriers order LDS only, so those loads stay in flight across them.
    float b[MKS];
    float cn = 0.0f;
    {
        const int col = w * 16 + (lane & 15);
#pragma unroll
        for (int kb = 0; kb < MKS; kb++) b[kb] = operand(x, n, D, col, kb * 4 + (lane >> 4));
        cn = col < n ? nrm[col] : 0.0f;
    }
    for (int c0 = 0; c0 < n; c0 += MC) {
        const int col = c0 + w * 16 + (lane & 15);
        const int ncol = col + MC;
        float bn[MKS];
#pragma unroll
        for (int kb = 0; kb < MKS; kb++) bn[kb] = operand(x, n, D, ncol, kb * 4 + (lane >> 4));
        const float cnn = nrm[ncol < n ? ncol : n - 1];
        v4f acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < MKS; kb++) {
            if (kb < ksteps) {
                acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][kb], b[kb], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1][kb], b[kb], acc1, 0, 0, 0);
            }
        }
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int qi = rt * 16 + 4 * (lane >> 4) + r, gq = q0 + qi;
                if (gq >= n || col >= n || col == gq) continue;
                const float dot = rt == 0 ? acc0[r] : acc1[r];
                const float approx = fmaxf((qn[rt][r] + cn) - 2.0f * dot, 0.0f);
                const uint64_t key = ((uint64_t)__float_as_uint(approx) << 32) | (uint32_t)col;
                if (key < tau[qi]) {
                    const uint32_t pos = atomicAdd(&cnt[qi], 1u);
                    buf[qi][pos] = key;
                }
            }
        lds_barrier();
        for (int qi = w; qi < MQ; qi += KNN_THREADS / 64) {
            const uint32_t c = cnt[qi];
            if (c <= (uint32_t)(KBUF - MC)) continue;
            for (int s = lane; s < KBUF; s += 64)
                if (s >= (int)c) buf[qi][s] = ~0ull;
            __builtin_amdgcn_wave_barrier();
            wave_bitonic_u64(buf[qi], lane);
            if (lane == 0) {
                const uint32_t keep = c < (uint32_t)KCAND ? c : (uint32_t)KCAND;
                cnt[qi] = keep;
                if (keep == (uint32_t)KCAND) tau[qi] = buf[qi][KCAND - 1];
            }
            __builtin_amdgcn_wave_barrier();
        }
        lds_barrier();
#pragma unroll
        for (int kb = 0; kb < MKS; kb++) b[kb] = bn[kb];
        cn = cnn;
    }
    for (int qi = w; qi < MQ; qi += KNN_THREADS / 64) {
        const int gq = q0 + qi;
        if (gq >= n) continue;
        const uint32_t c = cnt[qi];
        for (int s = lane; s < KBUF; s += 64)
            if (s >= (int)c) buf[qi][s] = ~0ull;
        __builtin_amdgcn_wave_barrier();
        wave_bitonic_u64(buf[qi], lane);
        for (int s = lane; s < KCAND; s += 64)
            cand[(size_t)gq * KCAND + s] =
                s < (int)c ? (uint32_t)(buf[qi][s] & 0xffffffffu) : 0xffffffffu;
        // fewer than KCAND kept: every other point is a candidate
        if (lane == 0)
            tau_out[gq] = c >= (uint32_t)KCAND ? __uint_as_float((uint32_t)(buf[qi][KCAND - 1] >> 32))
                                               : INFINITY;
    }
}

__global__ void __launch_bounds__(256) k_knn_rerank(const float *x, const float *nrm,
                                                    const uint32_t *max_bits, int n, int D, int k,
                                                    const uint32_t *cand, const float *tau_in,
                                                    uint32_t *idx_out, float *dist_out,
                                                    uint32_t *uncertain) {
    __shared__ uint64_t sb[4][KBUF];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + w;
    if (i >= n) return;  // whole wave: no block barrier follows
    uint64_t *s = sb[w];
    for (int q = lane; q < KBUF; q += 64) {
        uint64_t key = ~0ull;
        const uint32_t j = q < KCAND ? cand[(size_t)i * KCAND + q] : 0xffffffffu;
        if (j < (uint32_t)n) {
            float acc = 0.0f;
            for (int d = 0; d < D; d++) {
                const float e = x[(size_t)j * D + d] - x[(size_t)i * D + d];
                acc = acc + e * e;
            }
            key = ((uint64_t)__float_as_uint(sqrtf(acc)) << 32) | j;
        }
        s[q] = key;
    }
    __builtin_amdgcn_wave_barrier();
    wave_bitonic_u64(s, lane);
    for (int q = lane; q < k; q += 64) {
        const uint64_t key = s[q];
        idx_out[(size_t)i * k + q] = (uint32_t)(key & 0xffffffffu);
        dist_out[(size_t)i * k + q] = __uint_as_float((uint32_t)(key >> 32));
    }
    if (lane == 0) {
        const float tau = tau_in[i];
        bool ok = true;
        if (tau != INFINITY) {
            const double dk = (double)__uint_as_float((uint32_t)(s[k - 1] >> 32));
            const double dk2 = dk * dk * (1.0 + 1.0 / 1048576.0);
            const double delta = (3.0 * D + 8.0) * (1.0 / 16777216.0) *
                                 ((double)nrm[i] + (double)__uint_as_float(*max_bits) + (double)tau);
            ok = (double)tau - delta > dk2;
        }
        uncertain[i] = ok ? 0u : 1u;
    }
}

__global__ void k_list_rows(const uint32_t *flag, const uint32_t *pos, int n, uint32_t *list) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n && flag[i]) list[pos[i]] = (uint32_t)i;
}

// Workspace pieces, each starting on a 256-byte boundary: norms, candidates, tau, flags,
// positions, row list, misc words, scan temp.
static size_t knn_layout(size_t nn, size_t *off) {
    const size_t piece[8] = {nn * 4, nn * KCAND * 4, nn * 4, nn * 4, nn * 4, nn * 4, 256,
                             scan_temp_bytes(nn)};
    size_t at = 0;
    for (int i = 0; i < 8; i++) {
        if (off) off[i] = at;
        at += (piece[i] + 255) & ~(size_t)255;
    }
    return at;
}

size_t knn_temp_bytes(int64_t n, int d, int k) {
    if (d > 4 * MKS || k > KCAND || n <= KCAND) return 0;
    return knn_layout((size_t)n, nullptr);
}

bool knn_fast(const float *x, int64_t n, int d, int k, uint32_t *idx, float *dist, void *temp,
              hipStream_t s, uint64_t *exact_rows) {
    if (k <= 0 || n <= 0) return true;
    if (knn_temp_bytes(n, d, k) == 0 || temp == nullptr) {
        knn_exact(x, n, d, k, idx, dist, s);
        if (exact_rows) *exact_rows = (uint64_t)n;
        return false;
    }
    const size_t nn = (size_t)n;
    char *p = (char *)temp;
    size_t off[8];
    knn_layout(nn, off);
    float *nrm = (float *)(p + off[0]);
    uint32_t *cand = (uint32_t *)(p + off[1]);
    float *tau = (float *)(p + off[2]);
    uint32_t *flag = (uint32_t *)(p + off[3]);
    uint32_t *pos = (uint32_t *)(p + off[4]);
    uint32_t *list = (uint32_t *)(p + off[5]);
    uint32_t *misc = (uint32_t *)(p + off[6]);  // [0] max norm bits, [1] uncertain rows
    void *stemp = p + off[7];
    BHT_HIP(hipMemsetAsync(misc, 0, 256, s));
    const unsigned nb = ceil_div(nn, 256);
    hipLaunchKernelGGL(k_sqnorm, dim3(nb), dim3(256), 0, s, x, (int)n, d, nrm, misc);
    hipLaunchKernelGGL(k_knn_cand, dim3(ceil_div(nn, MQ)), dim3(KNN_THREADS), 0, s, x, nrm,
                       (int)n, d, cand, tau);
    hipLaunchKernelGGL(k_knn_rerank, dim3(ceil_div(nn, 4)), dim3(256), 0, s, x, nrm, misc,
                       (int)n, d, k, cand, tau, idx, dist, flag);
    scan_exclusive<uint32_t>(flag, pos, nn, stemp, misc + 1, s);
    hipLaunchKernelGGL(k_list_rows, dim3(nb), dim3(256), 0, s, flag, pos, (int)n, list);
    BHT_LAUNCH_CHECK();
    uint32_t nbad = 0;
    BHT_HIP(hipMemcpyAsync(&nbad, misc + 1, 4, hipMemcpyDeviceToHost, s));
    BHT_HIP(hipStreamSynchronize(s));
    if (nbad > 0) {
        hipLaunchKernelGGL(k_knn, dim3(ceil_div(nbad, KQ)), dim3(KNN_THREADS), 0, s, x, (int)n, d,
                           k, idx, dist, list, (int)nbad);
        BHT_LAUNCH_CHECK();
    }
    if (exact_rows) *exact_rows = nbad;
    return true;
}

}  // namespace bht
