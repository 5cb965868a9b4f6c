// Point renumbering helpers (reorder.hip).
#pragma once

#include "internal.hpp"

namespace bht {

void inverse_perm(const uint32_t *order, int n, uint32_t *new_of_old, hipStream_t s);
void gather_rows(const float *src, const uint32_t *order, int n, int width, float *dst,
                 hipStream_t s);
void scatter_rows(const float *src, const uint32_t *order, int n, int width, float *dst,
                  hipStream_t s);
void gather_u32(const uint32_t *src, const uint32_t *order, int n, uint32_t *dst, hipStream_t s);
// CSR with rows taken in `order` (new row i = old row order[i]) and columns mapped by `relabel`.
void csr_permute(const uint32_t *rows, const uint32_t *cols, const float *vals,
                 const uint32_t *order, const uint32_t *relabel, int n, uint32_t *len_tmp,
                 void *scan_temp, uint32_t *rows_new, uint32_t *cols_new, float *vals_new,
                 hipStream_t s);

// A point order from the affinity graph alone: `iters` steps of v <- D^-1 P v from a hashed start
// vector make v nearly constant within each well-connected group of points (diffusion mixes
// fast inside a group and slowly across groups), so sorting by v puts each group's points next
// to each other.  Workspace: diffusion_order_temp_bytes(n).  Returns the order (device, n
// indices: new point i = current point order[i]) inside the workspace.
size_t diffusion_order_temp_bytes(int n);
const uint32_t *diffusion_order(const uint32_t *rows, const uint32_t *cols, const float *vals,
                                int n, int iters, void *temp, hipStream_t s);

}  // namespace bht
