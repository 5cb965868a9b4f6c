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

// Packed operands.  For the MFMA's 16 x 4 operand of k-step kb, lane l holds row (l & 15) of a
// 16-point tile and dimension 4 kb + (l >> 4) (A: query rows; B: candidate columns, the same
// layout).  xp stores, per tile t and lane l, that lane's KSP values (KSP = the k-steps rounded up
// to a multiple of 4) contiguously: xp[(t * 64 + l) * KSP + kb], 0 outside the data, so a lane
// loads its operands of a tile as KSP / 4 aligned float4s.  The norms of points past n are +inf:
// such a column's approximate distance is +inf and never passes a row's threshold.
__global__ void __launch_bounds__(256) k_knn_pack(const float *x, int n, int D, int ksp,
                                                  uint32_t ntiles, float *xp, float *nrm) {
    const uint64_t total = (uint64_t)ntiles * 64 * (uint64_t)ksp;
    for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256) {
        const int kb = (int)(e % (uint64_t)ksp);
        const uint64_t tl = e / (uint64_t)ksp;
        const int l = (int)(tl & 63);
        const uint64_t row = (tl >> 6) * 16 + (uint64_t)(l & 15);
        const int d = 4 * kb + (l >> 4);
        xp[e] = (row < (uint64_t)n && d < D) ? x[row * D + d] : 0.0f;
    }
    const uint64_t npad = (uint64_t)ntiles * 16;
    for (uint64_t r = (uint64_t)n + blockIdx.x * 256ull + threadIdx.x; r < npad;
         r += (uint64_t)gridDim.x * 256)
        nrm[r] = INFINITY;
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

// The KS k-steps of a lane's operands of tile `tile` (KSP values per lane, float4 aligned).
template <int KS>
__device__ __forceinline__ void load_operands(const float *xp, uint32_t tile, int lane, float *v) {
    constexpr int KSP = (KS + 3) & ~3;
    const float4 *p = reinterpret_cast<const float4 *>(xp + ((size_t)tile * 64 + lane) * KSP);
#pragma unroll
    for (int g = 0; g < KSP / 4; g++) {
        const float4 q = p[g];
        if (4 * g + 0 < KS) v[4 * g + 0] = q.x;
        if (4 * g + 1 < KS) v[4 * g + 1] = q.y;
        if (4 * g + 2 < KS) v[4 * g + 2] = q.z;
        if (4 * g + 3 < KS) v[4 * g + 3] = q.w;
    }
}

// A block owns MQ = 32 query rows (two 16-row tiles); per step each of its four waves takes one
// 16-candidate tile, so a step covers 64 candidates.  Per row, the candidates whose approximate
// squared distance is below the row's threshold (the KCAND-th kept one, +inf until KCAND are
// kept; a register copy per lane, reloaded after each compaction phase) are appended to the row's
// LDS buffer; a row whose buffer may not take another step's appends is sorted by (distance,
// index) and cut to KCAND.  A pair is only compared in f32 against the threshold: the append
// path runs for the waves that have a pair below it.
template <int KS>
__global__ void __launch_bounds__(KNN_THREADS) k_knn_cand(const float *xp, const float *nrm,
                                                          int n, uint32_t ntiles, uint32_t *cand,
                                                          float *tau_out) {
    __shared__ uint64_t buf[MQ][KBUF];
    __shared__ uint32_t cnt[MQ];
    __shared__ __align__(16) float tauf[MQ];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int q0 = blockIdx.x * MQ;
    const uint32_t qt0 = (uint32_t)q0 / 16;
    if (t < MQ) {
        cnt[t] = 0;
        tauf[t] = INFINITY;
    }
    float a[2][KS];
    load_operands<KS>(xp, qt0, lane, a[0]);
    load_operands<KS>(xp, qt0 + 1, lane, a[1]);
    // output element r of a lane is row 4 * (lane / 16) + r of the tile, column lane % 16
    float qn[2][4];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int r = 0; r < 4; r++) qn[rt][r] = nrm[q0 + rt * 16 + 4 * (lane >> 4) + r];
    float b[KS];
    load_operands<KS>(xp, (uint32_t)w, lane, b);
    float cn = nrm[w * 16 + (lane & 15)];
    __syncthreads();
    const uint32_t nsteps = ntiles / 4;
    for (uint32_t st = 0; st < nsteps; st++) {
        const uint32_t ct = st * 4 + (uint32_t)w;
        const int col = (int)(ct * 16) + (lane & 15);
        // the next step's operands are loaded while this step is reduced
        float bn[KS];
        const uint32_t nt = st + 1 < nsteps ? ct + 4 : ct;
        load_operands<KS>(xp, nt, lane, bn);
        const float cnn = nrm[nt * 16 + (lane & 15)];
        // the rows' thresholds (settled by the last compaction phase) load beside the MFMAs
        float4 tq[2];
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
            tq[rt] = *reinterpret_cast<const float4 *>(&tauf[rt * 16 + 4 * (lane >> 4)]);
        v4f acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < KS; kb++) {
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][kb], b[kb], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1][kb], b[kb], acc1, 0, 0, 0);
        }
        // the step whose columns hold the block's own rows excludes each row's own point
        const bool diag = (int)(st * 64) <= q0 && q0 < (int)(st * 64 + 64);
        float approx[2][4];
        bool any = false;
#pragma unroll
        for (int rt = 0; rt < 2; rt++) {
            const float tqv[4] = {tq[rt].x, tq[rt].y, tq[rt].z, tq[rt].w};
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float dot = rt == 0 ? acc0[r] : acc1[r];
                float ap = fmaxf((qn[rt][r] + cn) - 2.0f * dot, 0.0f);
                if (diag && col == q0 + rt * 16 + 4 * (lane >> 4) + r) ap = INFINITY;
                approx[rt][r] = ap;
                any |= ap < tqv[r];
                if (!(ap < tqv[r])) approx[rt][r] = INFINITY;
            }
        }
        if (__ballot(any) != 0ull) {
#pragma unroll
            for (int rt = 0; rt < 2; rt++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    if (approx[rt][r] == INFINITY) continue;
                    const int qi = rt * 16 + 4 * (lane >> 4) + r;
                    const uint64_t key =
                        ((uint64_t)__float_as_uint(approx[rt][r]) << 32) | (uint32_t)col;
                    const uint32_t pos = atomicAdd(&cnt[qi], 1u);
                    buf[qi][pos] = key;
                }
        }
        lds_barrier();
        // this wave compacts rows w, w + 4, ..., w + 28: their counts in one read (lane j holds
        // row w + 4 j), then only the rows that may not take another step's appends
        const uint32_t cl = lane < MQ / 4 ? cnt[w + 4 * lane] : 0u;
        uint64_t over = __ballot(cl > (uint32_t)(KBUF - MC));
        while (over != 0ull) {
            const int j = __builtin_ctzll(over);
            over &= over - 1ull;
            const int qi = w + 4 * j;
            const uint32_t c = __builtin_amdgcn_readlane(cl, j);
            for (int s = lane; s < KBUF; s += 64)
                if (s >= (int)c) buf[qi][s] = ~0ull;
            __builtin_amdgcn_wave_barrier();
            wave_bitonic_u64(buf[qi], lane);
            if (lane == 0) {
                const uint32_t keep = c < (uint32_t)KCAND ? c : (uint32_t)KCAND;
                cnt[qi] = keep;
                if (keep == (uint32_t)KCAND)
                    tauf[qi] = __uint_as_float((uint32_t)(buf[qi][KCAND - 1] >> 32));
            }
            __builtin_amdgcn_wave_barrier();
        }
        lds_barrier();
#pragma unroll
        for (int kb = 0; kb < KS; kb++) b[kb] = bn[kb];
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
    // a lane's KCAND / 64 candidates at once, eight dimensions of each per batch of loads (the
    // sums still run left to right)
    constexpr int PER = KCAND / 64;
    static_assert(KCAND % 64 == 0, "candidates per lane");
    uint32_t j[PER];
    float acc[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        j[u] = cand[(size_t)i * KCAND + lane + 64 * u];
        acc[u] = 0.0f;
    }
    const float *qr = x + (size_t)i * D;
    int d = 0;
    for (; d + 8 <= D; d += 8) {
        float qv[8], cv[PER][8];
#pragma unroll
        for (int t = 0; t < 8; t++) qv[t] = qr[d + t];
#pragma unroll
        for (int u = 0; u < PER; u++)
#pragma unroll
            for (int t = 0; t < 8; t++)
                cv[u][t] = j[u] < (uint32_t)n ? x[(size_t)j[u] * D + d + t] : 0.0f;
#pragma unroll
        for (int u = 0; u < PER; u++)
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const float e = cv[u][t] - qv[t];
                acc[u] = acc[u] + e * e;
            }
    }
    for (; d < D; d++) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const float c = j[u] < (uint32_t)n ? x[(size_t)j[u] * D + d] : 0.0f;
            const float e = c - qr[d];
            acc[u] = acc[u] + e * e;
        }
    }
#pragma unroll
    for (int u = 0; u < PER; u++)
        s[lane + 64 * u] = j[u] < (uint32_t)n
                               ? ((uint64_t)__float_as_uint(sqrtf(acc[u])) << 32) | j[u]
                               : ~0ull;
    for (int q = KCAND + lane; q < KBUF; q += 64) s[q] = ~0ull;
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

// Candidate tiles of 16 points: a multiple of 4 (a step's four tiles), at least two per block of
// 32 query rows.
static uint32_t knn_tiles(size_t nn) { return 4 * (uint32_t)((nn + 63) / 64); }
static int knn_ksteps(int d) { return (d + 3) / 4; }

// Workspace pieces, each starting on a 256-byte boundary: norms (padded to the tiles), candidates,
// tau, flags, positions, row list, misc words, scan temp, packed operands.
static size_t knn_layout(size_t nn, int d, size_t *off) {
    const size_t nt = knn_tiles(nn);
    const size_t ksp = (size_t)((knn_ksteps(d) + 3) & ~3);
    const size_t piece[9] = {nt * 16 * 4, nn * KCAND * 4, nn * 4, nn * 4, nn * 4, nn * 4, 256,
                             scan_temp_bytes(nn), nt * 64 * ksp * 4};
    size_t at = 0;
    for (int i = 0; i < 9; i++) {
        if (off) off[i] = at;
        at += (piece[i] + 255) & ~(size_t)255;
    }
    return at;
}

size_t knn_temp_bytes(int64_t n, int d, int k) {
    if (d > 4 * MKS || k > KCAND || n <= KCAND) return 0;
    return knn_layout((size_t)n, d, nullptr);
}

template <int KS>
static void launch_cand(const float *xp, const float *nrm, int n, uint32_t ntiles, uint32_t *cand,
                        float *tau, hipStream_t s) {
    hipLaunchKernelGGL(k_knn_cand<KS>, dim3(ceil_div((uint64_t)n, MQ)), dim3(KNN_THREADS), 0, s,
                       xp, nrm, n, ntiles, cand, tau);
}

template <int... KS>
static void cand_by_ksteps(int ks, std::integer_sequence<int, KS...>, const float *xp,
                           const float *nrm, int n, uint32_t ntiles, uint32_t *cand, float *tau,
                           hipStream_t s) {
    ((ks == KS + 1 ? launch_cand<KS + 1>(xp, nrm, n, ntiles, cand, tau, s) : void()), ...);
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
    size_t off[9];
    knn_layout(nn, d, off);
    float *nrm = (float *)(p + off[0]);
    uint32_t *cand = (uint32_t *)(p + off[1]);
    float *tau = (float *)(p + off[2]);
    uint32_t *flag = (uint32_t *)(p + off[3]);
    uint32_t *pos = (uint32_t *)(p + off[4]);
    uint32_t *list = (uint32_t *)(p + off[5]);
    uint32_t *misc = (uint32_t *)(p + off[6]);  // [0] max norm bits, [1] uncertain rows
    void *stemp = p + off[7];
    float *xp = (float *)(p + off[8]);
    const uint32_t ntiles = knn_tiles(nn);
    const int ks = knn_ksteps(d);
    BHT_HIP(hipMemsetAsync(misc, 0, 256, s));
    const unsigned nb = ceil_div(nn, 256);
    hipLaunchKernelGGL(k_sqnorm, dim3(nb), dim3(256), 0, s, x, (int)n, d, nrm, misc);
    hipLaunchKernelGGL(k_knn_pack, dim3(4096), dim3(256), 0, s, x, (int)n, d, (ks + 3) & ~3,
                       ntiles, xp, nrm);
    cand_by_ksteps(ks, std::make_integer_sequence<int, MKS>(), xp, nrm, (int)n, ntiles, cand,
                   tau, s);
    BHT_LAUNCH_CHECK();
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
