// Spatial renumbering of the points of a job.
//
// The attraction gathers y[column] for every edge; with the points in input order those gathers
// land on random cache lines of an array larger than an XCD's L2.  Every so often the job
// renumbers its points in the order of the current box (or Morton) sort, i.e. spatially, so that a
// row's neighbours sit at nearby indices.  Per-point state is permuted, CSR rows are moved and
// their columns relabelled with the edge order inside every row unchanged, so each row's
// attraction is the same sum as before.  The library keeps the permutation and undoes it at every
// output.
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

}  // namespace bht
