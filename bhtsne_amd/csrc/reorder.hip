// Spatial renumbering of the points of a job.
//
// The attraction gathers y[column] for every edge; with the points in input order those gathers
// land on random cache lines of an array larger than an XCD's L2.  Before the first epoch a large
// job renumbers its points in the diffusion order of P (below), so that a row's neighbours sit at
// nearby indices.  Per-point state is permuted, CSR rows are moved and their columns relabelled
// with the edge order inside every row unchanged, so each row's attraction is the same sum as
// before.  The library keeps the permutation and undoes it at every output.
#include "internal.hpp"
#include "reorder.hpp"

namespace bht {

__global__ void k_inverse_perm(const uint32_t *order, int n, uint32_t *new_of_old) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        new_of_old[order[i]] = (uint32_t)i;
}

// dst[i] = src[order[i]] for rows of `width` floats
__global__ void k_gather_rows(const float *src, const uint32_t *order, int n, int width,
                              float *dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t o = order[i];
        for (int d = 0; d < width; d++) dst[(size_t)i * width + d] = src[(size_t)o * width + d];
    }
}

// dst[order[i]] = src[i]
__global__ void k_scatter_rows(const float *src, const uint32_t *order, int n, int width,
                               float *dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t o = order[i];
        for (int d = 0; d < width; d++) dst[(size_t)o * width + d] = src[(size_t)i * width + d];
    }
}

__global__ void k_gather_u32(const uint32_t *src, const uint32_t *order, int n, uint32_t *dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[i] = src[order[i]];
}

__global__ void k_row_lengths(const uint32_t *rows, const uint32_t *order, int n, uint32_t *len) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t o = order[i];
        len[i] = rows[o + 1] - rows[o];
    }
}

// 16 lanes per new row copy the old row's entries in order, relabelling columns.
__global__ void __launch_bounds__(256) k_csr_move(const uint32_t *rows, const uint32_t *cols,
                                                  const float *vals, const uint32_t *order,
                                                  const uint32_t *relabel, int n,
                                                  const uint32_t *rows_new, uint32_t *cols_new,
                                                  float *vals_new) {
    const int sub = threadIdx.x & 15;
    const int grp = (blockIdx.x * 256 + threadIdx.x) >> 4;
    const int ngrp = (gridDim.x * 256) >> 4;
    for (int i = grp; i < n; i += ngrp) {
        const uint32_t o = order[i];
        const uint32_t s = rows[o], e = rows[o + 1], d = rows_new[i];
        for (uint32_t t = s + sub; t < e; t += 16) {
            cols_new[d + (t - s)] = relabel[cols[t]];
            vals_new[d + (t - s)] = vals[t];
        }
    }
}

static int nb_for(int n) {
    int b = (int)ceil_div((uint64_t)(n > 0 ? n : 1), 256);
    return b < 4096 ? b : 4096;
}

void inverse_perm(const uint32_t *order, int n, uint32_t *new_of_old, hipStream_t s) {
    hipLaunchKernelGGL(k_inverse_perm, dim3(nb_for(n)), dim3(256), 0, s, order, n, new_of_old);
    BHT_LAUNCH_CHECK();
}

void gather_rows(const float *src, const uint32_t *order, int n, int width, float *dst,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_gather_rows, dim3(nb_for(n)), dim3(256), 0, s, src, order, n, width,
                       dst);
    BHT_LAUNCH_CHECK();
}

void scatter_rows(const float *src, const uint32_t *order, int n, int width, float *dst,
                  hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_rows, dim3(nb_for(n)), dim3(256), 0, s, src, order, n, width,
                       dst);
    BHT_LAUNCH_CHECK();
}

void gather_u32(const uint32_t *src, const uint32_t *order, int n, uint32_t *dst,
                hipStream_t s) {
    hipLaunchKernelGGL(k_gather_u32, dim3(nb_for(n)), dim3(256), 0, s, src, order, n, dst);
    BHT_LAUNCH_CHECK();
}

void csr_permute(const uint32_t *rows, const uint32_t *cols, const float *vals,
                 const uint32_t *order, const uint32_t *relabel, int n, uint32_t *len_tmp,
                 void *scan_temp, uint32_t *rows_new, uint32_t *cols_new, float *vals_new,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_row_lengths, dim3(nb_for(n)), dim3(256), 0, s, rows, order, n, len_tmp);
    scan_exclusive<uint32_t>(len_tmp, rows_new, (uint64_t)n, scan_temp, rows_new + n, s);
    int nb = (int)std::min<uint64_t>(ceil_div((uint64_t)(n > 0 ? n : 1) * 16, 256), 8192);
    hipLaunchKernelGGL(k_csr_move, dim3(nb), dim3(256), 0, s, rows, cols, vals, order, relabel,
                       n, rows_new, cols_new, vals_new);
    BHT_LAUNCH_CHECK();
}

// ---- diffusion order ---------------------------------------------------------------------------

__global__ void k_diffuse_init(int n, float *v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        // splitmix-style hash of the index to [0, 1)
        uint64_t z = (uint64_t)i * 0x9e3779b97f4a7c15ull + 0x632be59bd9b4e019ull;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        v[i] = (float)(z >> 40) * (1.0f / 16777216.0f);
    }
}

// v_out[i] = sum_j p_ij v_in[j] / sum_j p_ij, one wave per row
__global__ void __launch_bounds__(256) k_diffuse_step(const uint32_t *rows, const uint32_t *cols,
                                                      const float *vals, int n, const float *v_in,
                                                      float *v_out) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * 256) >> 6;
    for (int i = wave; i < n; i += nwaves) {
        float num = 0.0f, den = 0.0f;
        for (uint32_t e = rows[i] + lane; e < rows[i + 1]; e += 64) {
            const float p = vals[e];
            num += p * v_in[cols[e]];
            den += p;
        }
        num = wave_sum_f32(num);
        den = wave_sum_f32(den);
        if (lane == 0) v_out[i] = den > 0.0f ? num / den : v_in[i];
    }
}

// sort keys: the bits of a non-negative float order like the float
__global__ void k_float_keys(const float *v, int n, uint32_t *keys, uint32_t *idx) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        keys[i] = __float_as_uint(fmaxf(v[i], 0.0f));
        idx[i] = (uint32_t)i;
    }
}

static void diffusion_layout(int n, size_t *off, size_t *total) {
    const size_t nn = (size_t)(n > 0 ? n : 1);
    const size_t piece[7] = {nn * 4, nn * 4, nn * 4, nn * 4, nn * 4, nn * 4,
                             radix_sort_temp_bytes(nn)};
    size_t at = 0;
    for (int i = 0; i < 7; i++) {
        off[i] = at;
        at += (piece[i] + 255) & ~(size_t)255;
    }
    *total = at;
}

size_t diffusion_order_temp_bytes(int n) {
    size_t off[7], total;
    diffusion_layout(n, off, &total);
    return total;
}

const uint32_t *diffusion_order(const uint32_t *rows, const uint32_t *cols, const float *vals,
                                int n, int iters, void *temp, hipStream_t s) {
    size_t off[7], total;
    diffusion_layout(n, off, &total);
    char *p = (char *)temp;
    float *v0 = (float *)(p + off[0]), *v1 = (float *)(p + off[1]);
    uint32_t *keys = (uint32_t *)(p + off[2]), *idx = (uint32_t *)(p + off[3]);
    uint32_t *keys_alt = (uint32_t *)(p + off[4]), *idx_alt = (uint32_t *)(p + off[5]);
    void *stemp = p + off[6];
    hipLaunchKernelGGL(k_diffuse_init, dim3(nb_for(n)), dim3(256), 0, s, n, v0);
    const int nb = (int)std::min<uint64_t>(ceil_div((uint64_t)(n > 0 ? n : 1), 4), 8192);
    for (int t = 0; t < iters; t++) {
        hipLaunchKernelGGL(k_diffuse_step, dim3(nb), dim3(256), 0, s, rows, cols, vals, n, v0, v1);
        float *tmp = v0;
        v0 = v1;
        v1 = tmp;
    }
    hipLaunchKernelGGL(k_float_keys, dim3(nb_for(n)), dim3(256), 0, s, v0, n, keys, idx);
    BHT_LAUNCH_CHECK();
    const bool alt = radix_sort_pairs<uint32_t>(keys, idx, keys_alt, idx_alt, (uint64_t)n, 0, 32,
                                                stemp, s);
    return alt ? idx_alt : idx;
}

}  // namespace bht
