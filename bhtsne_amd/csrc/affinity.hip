// Affinity construction on the device: per-row bandwidth search (src/tsne/mod.rs:201-275) and
// symmetrisation into CSR with the reference's row order (src/tsne/mod.rs:306-406).
#include <cfloat>

#include "internal.hpp"
#include "affinity.hpp"

namespace bht {

constexpr int AF_THREADS = 256;
constexpr int AF_ROWS = AF_THREADS / 64;  // one wave per row
constexpr int KMAX = 256;                 // neighbours per row handled by one wave
constexpr int KPL = KMAX / 64;            // values per lane

// ---- bandwidth search ------------------------------------------------------------------------
//
// One wave per row.  Every lane evaluates p_j = exp(-beta d_j^2) with the reference's f32
// argument; exp and log are evaluated in FP64 and rounded to f32, the correctly rounded f32
// value.  The host libm's expf / logf (which the reference's f32 exp / ln call) return that value
// for all but about 6 in 10,000 arguments, where they are one ulp off, and a one-ulp change of a
// single p_j rarely moves the row's entropy, so the bisection takes the reference's steps.
// mode FP64: the row sum and the entropy are reduced in FP64 across the wave.
// mode F32_SEQ: lane 0 sums in f32 in the reference's left-to-right order from LDS.
// beta, its brackets and the bisection steps are f32 as in the reference in both modes.
__device__ __forceinline__ float exp_f32_rn(float x) { return (float)exp((double)x); }
__device__ __forceinline__ float log_f32_rn(float x) { return (float)log((double)x); }

template <int MODE>
__global__ void __launch_bounds__(AF_THREADS) k_search_beta(const float *dist, float *p_out,
                                                            int64_t n, int k, float perplexity) {
    __shared__ float sp[AF_ROWS][KMAX];
    __shared__ float sd[AF_ROWS][KMAX];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t row = (int64_t)blockIdx.x * AF_ROWS + w;
    if (row >= n) return;
    float d[KPL], p[KPL];
#pragma unroll
    for (int t = 0; t < KPL; t++) {
        int j = lane + 64 * t;
        d[t] = j < k ? dist[row * k + j] : 0.0f;
        if (MODE == 1 && j < k) sd[w][j] = d[t];
    }
    bool found = false;
    float beta = 1.0f, min_beta = -FLT_MAX, max_beta = FLT_MAX;
    const float tolerance = 1e-5f;
    const float log_perp = log_f32_rn(perplexity);
    float sum_f = 0.0f;
    double sum_d = 0.0;
    int iteration = 0;
    while (!found && iteration < 200) {
        float diff;
#pragma unroll
        for (int t = 0; t < KPL; t++) {
            int j = lane + 64 * t;
            p[t] = j < k ? exp_f32_rn(-beta * (d[t] * d[t])) : 0.0f;
        }
        if (MODE == 0) {
            double s = 0.0, e = 0.0;
#pragma unroll
            for (int t = 0; t < KPL; t++) {
                s += (double)p[t];
                e += (double)(beta * p[t]) * (double)(d[t] * d[t]);
            }
            s = wave_sum_f64(s) + (double)FLT_MIN;
            e = wave_sum_f64(e);
            sum_d = s;
            double entropy = e / s + log(s);
            diff = (float)(entropy - (double)log_perp);
        } else {
#pragma unroll
            for (int t = 0; t < KPL; t++) {
                int j = lane + 64 * t;
                if (j < k) sp[w][j] = p[t];
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            float df = 0.0f, sf = 0.0f;
            if (lane == 0) {
                float s = 0.0f;
                for (int j = 0; j < k; j++) s = s + sp[w][j];
                s = s + FLT_MIN;
                float entropy = 0.0f;
                for (int j = 0; j < k; j++)
                    entropy = entropy + beta * sp[w][j] * (sd[w][j] * sd[w][j]);
                entropy = entropy / s + log_f32_rn(s);
                df = entropy - log_perp;
                sf = s;
            }
            diff = __shfl(df, 0, 64);
            sum_f = __shfl(sf, 0, 64);
            __builtin_amdgcn_wave_barrier();
        }
        if (diff < tolerance && -diff < tolerance) {
            found = true;
        } else {
            if (diff > 0.0f) {
                min_beta = beta;
                if (max_beta == FLT_MAX || max_beta == -FLT_MAX) beta *= 2.0f;
                else beta = (beta + max_beta) / 2.0f;
            } else {
                max_beta = beta;
                if (min_beta == -FLT_MAX || min_beta == FLT_MAX) beta /= 2.0f;
                else beta = (beta + min_beta) / 2.0f;
            }
            if (isinf(beta) && beta > 0.0f) beta = FLT_MAX;
            if (isinf(beta) && beta < 0.0f) beta = -FLT_MAX;
        }
        iteration += 1;
    }
#pragma unroll
    for (int t = 0; t < KPL; t++) {
        int j = lane + 64 * t;
        if (j >= k) continue;
        float v;
        if (MODE == 0) v = (float)((double)p[t] / (sum_d + (double)FLT_EPSILON));
        else v = p[t] / (sum_f + FLT_EPSILON);
        p_out[row * k + j] = v;
    }
}

void search_beta(const float *dist, float *p, int64_t n, int k, float perplexity, int mode,
                 hipStream_t s) {
    if (k > KMAX) throw Error(ERR_INVALID, "at most 256 neighbours per row are supported");
    unsigned nb = ceil_div((uint64_t)n, AF_ROWS);
    if (nb == 0) return;
    if (mode == 0)
        hipLaunchKernelGGL(k_search_beta<0>, dim3(nb), dim3(AF_THREADS), 0, s, dist, p, n, k,
                           perplexity);
    else
        hipLaunchKernelGGL(k_search_beta<1>, dim3(nb), dim3(AF_THREADS), 0, s, dist, p, n, k,
                           perplexity);
    BHT_LAUNCH_CHECK();
}

// The same search over dense rows (the exact path, src/lib.rs:530-542): row i holds the distances
// to all n points, its own (0) included, as the reference's dense rows do.  One wave per row; the
// row is read from global memory in every iteration, and in mode F32_SEQ the row's p values are
// staged in p_out for lane 0's left-to-right sums.
template <int MODE>
__global__ void __launch_bounds__(AF_THREADS) k_search_beta_dense(const float *dist, float *p_out,
                                                                  int64_t n, float perplexity) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t row = (int64_t)blockIdx.x * AF_ROWS + w;
    if (row >= n) return;
    const float *d = dist + row * n;
    float *p = p_out + row * n;
    bool found = false;
    float beta = 1.0f, min_beta = -FLT_MAX, max_beta = FLT_MAX;
    const float tolerance = 1e-5f;
    const float log_perp = log_f32_rn(perplexity);
    float sum_f = 0.0f;
    double sum_d = 0.0;
    int iteration = 0;
    float beta_used = beta;  // the beta of the last p values computed
    while (!found && iteration < 200) {
        float diff;
        beta_used = beta;
        if (MODE == 0) {
            double s = 0.0, e = 0.0;
            for (int64_t j = lane; j < n; j += 64) {
                const float dj = d[j];
                const float pj = exp_f32_rn(-beta * (dj * dj));
                s += (double)pj;
                e += (double)(beta * pj) * (double)(dj * dj);
            }
            s = wave_sum_f64(s) + (double)FLT_MIN;
            e = wave_sum_f64(e);
            sum_d = s;
            diff = (float)(e / s + log(s) - (double)log_perp);
        } else {
            for (int64_t j = lane; j < n; j += 64) {
                const float dj = d[j];
                p[j] = exp_f32_rn(-beta * (dj * dj));
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            float df = 0.0f, sf = 0.0f;
            if (lane == 0) {
                float s = 0.0f;
                for (int64_t j = 0; j < n; j++) s = s + p[j];
                s = s + FLT_MIN;
                float entropy = 0.0f;
                for (int64_t j = 0; j < n; j++) entropy = entropy + beta * p[j] * (d[j] * d[j]);
                entropy = entropy / s + log_f32_rn(s);
                df = entropy - log_perp;
                sf = s;
            }
            diff = __shfl(df, 0, 64);
            sum_f = __shfl(sf, 0, 64);
            __builtin_amdgcn_wave_barrier();
        }
        if (diff < tolerance && -diff < tolerance) {
            found = true;
        } else {
            if (diff > 0.0f) {
                min_beta = beta;
                if (max_beta == FLT_MAX || max_beta == -FLT_MAX) beta *= 2.0f;
                else beta = (beta + max_beta) / 2.0f;
            } else {
                max_beta = beta;
                if (min_beta == -FLT_MAX || min_beta == FLT_MAX) beta /= 2.0f;
                else beta = (beta + min_beta) / 2.0f;
            }
            if (isinf(beta) && beta > 0.0f) beta = FLT_MAX;
            if (isinf(beta) && beta < 0.0f) beta = -FLT_MAX;
        }
        iteration += 1;
    }
    // the last iteration's p (recomputed for the FP64 mode, which does not keep them)
    for (int64_t j = lane; j < n; j += 64) {
        const float dj = d[j];
        float v;
        if (MODE == 0) {
            const float pj = exp_f32_rn(-beta_used * (dj * dj));
            v = (float)((double)pj / (sum_d + (double)FLT_EPSILON));
        } else {
            v = p[j] / (sum_f + FLT_EPSILON);
        }
        p[j] = v;
    }
}

void search_beta_dense(const float *dist, float *p, int64_t n, float perplexity, int mode,
                       hipStream_t s) {
    unsigned nb = ceil_div((uint64_t)n, AF_ROWS);
    if (nb == 0) return;
    if (mode == 0)
        hipLaunchKernelGGL(k_search_beta_dense<0>, dim3(nb), dim3(AF_THREADS), 0, s, dist, p, n,
                           perplexity);
    else
        hipLaunchKernelGGL(k_search_beta_dense<1>, dim3(nb), dim3(AF_THREADS), 0, s, dist, p, n,
                           perplexity);
    BHT_LAUNCH_CHECK();
}

// ---- symmetrisation --------------------------------------------------------------------------

// 1. sort every row's (column, p) pairs by column: bitonic sort of 256 slots by one wave in LDS.
__global__ void __launch_bounds__(AF_THREADS) k_row_sort(uint32_t *cols, float *vals, int64_t n,
                                                         int k) {
    __shared__ uint32_t sc[AF_ROWS][KMAX];
    __shared__ float sv[AF_ROWS][KMAX];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t row = (int64_t)blockIdx.x * AF_ROWS + w;
    if (row >= n) return;
    int m = 1;
    while (m < k) m <<= 1;
    for (int j = lane; j < m; j += 64) {
        sc[w][j] = j < k ? cols[row * k + j] : 0xffffffffu;
        sv[w][j] = j < k ? vals[row * k + j] : 0.0f;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int size = 2; size <= m; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < m / 2; t += 64) {
                int i = 2 * t - (t & (stride - 1));
                int j = i + stride;
                bool up = (i & size) == 0;
                uint32_t a = sc[w][i], b = sc[w][j];
                if ((a > b) == up) {
                    sc[w][i] = b; sc[w][j] = a;
                    float va = sv[w][i]; sv[w][i] = sv[w][j]; sv[w][j] = va;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
    for (int j = lane; j < k; j += 64) {
        cols[row * k + j] = sc[w][j];
        vals[row * k + j] = sv[w][j];
    }
}

// 2. mutual test per edge (binary search of r in the sorted row of c, mod.rs:346-349), the summed
// value of the pair, and the count of non-mutual edges landing in each row.
__global__ void k_mutual(const uint32_t *cols, const float *vals, int64_t n, int k, float *pair,
                         uint32_t *nonmutual, uint32_t *in_count) {
    const int64_t ne = n * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne;
         e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)(e / k);
        const uint32_t c = cols[e];
        const uint32_t *row = cols + (int64_t)c * k;
        int lo = 0, hi = k, pos = -1;
        while (lo < hi) {
            int mid = lo + (hi - lo) / 2;
            uint32_t v = row[mid];
            if (v < r) lo = mid + 1;
            else if (v > r) hi = mid;
            else { pos = mid; break; }
        }
        if (pos >= 0) {
            pair[e] = vals[e] + vals[(int64_t)c * k + pos];
            nonmutual[e] = 0u;
        } else {
            pair[e] = vals[e];
            nonmutual[e] = 1u;
            atomicAdd(&in_count[c], 1u);
        }
    }
}

__global__ void k_row_counts(const uint32_t *in_count, int64_t n, int k, uint32_t *row_count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        row_count[i] = (uint32_t)k + in_count[i];
}

// 4. compact the non-mutual edges (in row-major order) as (destination, edge id).
__global__ void k_compact(const uint32_t *cols, const uint32_t *nonmutual, const uint32_t *pos,
                          int64_t ne, uint32_t *dest, uint32_t *edge) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (nonmutual[e]) {
            dest[pos[e]] = cols[e];
            edge[pos[e]] = (uint32_t)e;
        }
    }
}

// Row r of the symmetrised matrix, in the order of mod.rs:365-397, values halved (mod.rs:400-401):
//   (a) every m < r whose list holds r, ascending: the mutual own entries c < r merged with the
//       incoming non-mutual sources m < r (both ascending, and disjoint: a mutual pair is never an
//       incoming edge);
//   (b) the own entries except the mutual c < r, ascending;
//   (c) the incoming non-mutual sources m > r, ascending.
// One wave per row.  Every entry's position follows from ranks: its rank in its own list plus, in
// (a), the number of the other list's entries below it (a search of the row's (a) own entries,
// compacted in LDS, or of the sorted incoming sources), so the wave writes its row's segment
// together (a thread writing a whole row took 20.8 ms at 1M points, the wave 1.8 ms).
__global__ void __launch_bounds__(AF_THREADS) k_fill_wave(
    const uint32_t *cols, const float *vals, const float *pair, const uint32_t *nonmutual,
    int64_t n, int k, const uint32_t *in_off, const uint32_t *in_count, const uint32_t *in_edge,
    const uint32_t *rows, uint32_t *out_cols, float *out_vals) {
    __shared__ uint32_t acomp[AF_ROWS][KMAX];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * AF_ROWS + w;
    if (r >= n) return;  // whole wave: no block barrier follows
    const uint32_t ur = (uint32_t)r, uk = (uint32_t)k;
    const uint32_t *own = cols + r * k;
    const float *own_pair = pair + r * k;
    const uint32_t *own_nm = nonmutual + r * k;
    const uint32_t ib = in_off[r], ic = in_count[r];
    const uint32_t o0 = rows[r];
    const uint64_t below = (1ull << lane) - 1ull;
    // own entries j = lane + 64 q: in A (mutual, c < r) or not; ranks within A and within the rest
    uint32_t c[KPL], rank[KPL];
    bool in_a[KPL];
    uint32_t n_a = 0, n_rest = 0;
#pragma unroll
    for (int q = 0; q < KPL; q++) {
        const int j = lane + 64 * q;
        const bool valid = j < k;
        c[q] = valid ? own[j] : 0xffffffffu;
        in_a[q] = valid && c[q] < ur && !own_nm[j];
        const uint64_t ma = __ballot(in_a[q]), mr = __ballot(valid && !in_a[q]);
        rank[q] = in_a[q] ? n_a + (uint32_t)__popcll(ma & below)
                          : n_rest + (uint32_t)__popcll(mr & below);
        n_a += (uint32_t)__popcll(ma);
        n_rest += (uint32_t)__popcll(mr);
        if (in_a[q]) acomp[w][rank[q]] = c[q];
    }
    // incoming sources below r: a prefix of the (ascending) sources
    uint32_t n_b = 0;
    for (uint32_t b0 = 0; b0 < ic; b0 += 64) {
        const uint32_t b = b0 + (uint32_t)lane;
        const bool lowb = b < ic && in_edge[ib + b] / uk < ur;
        n_b += (uint32_t)__popcll(__ballot(lowb));
    }
    __builtin_amdgcn_wave_barrier();
    // (a) own entries: rank in A plus the incoming sources below them
#pragma unroll
    for (int q = 0; q < KPL; q++) {
        const int j = lane + 64 * q;
        if (j >= k) continue;
        uint32_t pos;
        if (in_a[q]) {
            uint32_t lo = 0, hi = n_b;  // first source >= c (none equals c)
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (in_edge[ib + mid] / uk < c[q]) lo = mid + 1;
                else hi = mid;
            }
            pos = rank[q] + lo;
        } else {
            pos = n_a + n_b + rank[q];  // (b)
        }
        out_cols[o0 + pos] = c[q];
        out_vals[o0 + pos] = own_pair[j] * 0.5f;
    }
    // incoming: (a) rank plus the A entries below; (c) after the row's own entries
    for (uint32_t b = (uint32_t)lane; b < ic; b += 64) {
        const uint32_t e = in_edge[ib + b];
        const uint32_t m = e / uk;
        uint32_t pos;
        if (b < n_b) {
            uint32_t lo = 0, hi = n_a;  // A entries below m
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (acomp[w][mid] < m) lo = mid + 1;
                else hi = mid;
            }
            pos = b + lo;
        } else {
            pos = uk + b;
        }
        out_cols[o0 + pos] = m;
        out_vals[o0 + pos] = vals[e] * 0.5f;
    }
}

static int grid_for(int64_t work, int threads = 256, int cap = 8192) {
    int64_t b = (work + threads - 1) / threads;
    if (b < 1) b = 1;
    return (int)(b < cap ? b : cap);
}

void sym_sort_rows(uint32_t *cols, float *vals, int64_t n, int k, hipStream_t s) {
    if (k > KMAX) throw Error(ERR_INVALID, "at most 256 neighbours per row are supported");
    hipLaunchKernelGGL(k_row_sort, dim3(ceil_div((uint64_t)n, AF_ROWS)), dim3(AF_THREADS), 0, s,
                       cols, vals, n, k);
    BHT_LAUNCH_CHECK();
}

// Full symmetrisation.  cols / vals (n * k) are sorted in place.  Returns nnz; with out == false
// only the count is produced (rows_u32 filled).  Device outputs: rows (n + 1, u32), out_cols,
// out_vals (allocated by the caller through `alloc`).
uint64_t symmetrize_device(uint32_t *cols, float *vals, int64_t n, int k, SymWork &wk,
                           hipStream_t s) {
    const int64_t ne = n * k;
    sym_sort_rows(cols, vals, n, k, s);
    BHT_HIP(hipMemsetAsync(wk.in_count, 0, sizeof(uint32_t) * n, s));
    hipLaunchKernelGGL(k_mutual, dim3(grid_for(ne)), dim3(256), 0, s, cols, vals, n, k, wk.pair,
                       wk.nonmutual, wk.in_count);
    hipLaunchKernelGGL(k_row_counts, dim3(grid_for(n)), dim3(256), 0, s, wk.in_count, n, k,
                       wk.row_count);
    BHT_LAUNCH_CHECK();
    // row offsets and their total
    scan_exclusive<uint32_t>(wk.row_count, wk.rows, (uint64_t)n, wk.scan_temp, wk.rows + n, s);
    // compaction positions of the non-mutual edges, then bucket them by destination (stable)
    scan_exclusive<uint32_t>(wk.nonmutual, wk.nm_pos, (uint64_t)ne, wk.scan_temp, wk.total, s);
    uint32_t n_nm = 0;
    BHT_HIP(hipMemcpyAsync(&n_nm, wk.total, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    BHT_HIP(hipStreamSynchronize(s));
    hipLaunchKernelGGL(k_compact, dim3(grid_for(ne)), dim3(256), 0, s, cols, wk.nonmutual,
                       wk.nm_pos, ne, wk.dest, wk.edge);
    BHT_LAUNCH_CHECK();
    int bits = 0;
    while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)n) bits++;
    bits = ((bits + 7) / 8) * 8;
    if (bits == 0) bits = 8;
    bool alt = radix_sort_pairs<uint32_t>(wk.dest, wk.edge, wk.dest_alt, wk.edge_alt,
                                          (uint64_t)n_nm, 0, bits, wk.sort_temp, s);
    const uint32_t *in_edge = alt ? wk.edge_alt : wk.edge;
    scan_exclusive<uint32_t>(wk.in_count, wk.in_off, (uint64_t)n, wk.scan_temp, wk.total, s);
    uint32_t nnz = 0;
    BHT_HIP(hipMemcpyAsync(&nnz, wk.rows + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    BHT_HIP(hipStreamSynchronize(s));
    wk.nnz = nnz;
    wk.alloc_out(nnz);
    hipLaunchKernelGGL(k_fill_wave, dim3(ceil_div((uint64_t)n, AF_ROWS)), dim3(AF_THREADS), 0, s,
                       cols, vals, wk.pair, wk.nonmutual, n, k, wk.in_off, wk.in_count, in_edge,
                       wk.rows, wk.out_cols, wk.out_vals);
    BHT_LAUNCH_CHECK();
    return nnz;
}

}  // namespace bht
