// Affinity-construction workspaces and launchers (affinity.hip, knn.hip).
#pragma once

#include <functional>

#include "internal.hpp"

namespace bht {

struct SymWork {
    float *pair = nullptr;          // n*k: p_rc (+ p_cr when mutual)
    uint32_t *nonmutual = nullptr;  // n*k flags
    uint32_t *nm_pos = nullptr;     // n*k compaction positions
    uint32_t *in_count = nullptr;   // n: non-mutual edges landing in each row
    uint32_t *row_count = nullptr;  // n
    uint32_t *in_off = nullptr;     // n
    uint32_t *rows = nullptr;       // n + 1 (u32 offsets), output
    uint32_t *dest = nullptr, *edge = nullptr, *dest_alt = nullptr, *edge_alt = nullptr;  // n*k
    uint32_t *total = nullptr;      // 1
    void *scan_temp = nullptr;
    void *sort_temp = nullptr;
    uint32_t *out_cols = nullptr;   // nnz, set by alloc_out
    float *out_vals = nullptr;      // nnz, set by alloc_out
    uint64_t nnz = 0;
    std::function<void(uint64_t)> alloc_out;
};

void search_beta(const float *dist, float *p, int64_t n, int k, float perplexity, int mode,
                 hipStream_t s);
// The same search over n dense rows of n distances (each row's own 0 included).
void search_beta_dense(const float *dist, float *p, int64_t n, float perplexity, int mode,
                       hipStream_t s);
void sym_sort_rows(uint32_t *cols, float *vals, int64_t n, int k, hipStream_t s);
uint64_t symmetrize_device(uint32_t *cols, float *vals, int64_t n, int k, SymWork &wk,
                           hipStream_t s);

// Exact k nearest neighbours (Euclidean, the reference metric's f32 arithmetic), self excluded,
// ascending (distance, index).
void knn_exact(const float *x, int64_t n, int d, int k, uint32_t *idx, float *dist,
               hipStream_t s);
// The same result through a matrix-core candidate pass, an exact re-rank and a certificate per
// row (uncertified rows go through knn_exact).  temp: knn_temp_bytes(n, d, k) bytes; 0 means the
// shape is outside the fast path (d > 64, k > 128 or n <= 128) and knn_exact runs instead.
// Synchronises the stream.  exact_rows (optional) receives the rows recomputed exactly.
size_t knn_temp_bytes(int64_t n, int d, int k);
bool knn_fast(const float *x, int64_t n, int d, int k, uint32_t *idx, float *dist, void *temp,
              hipStream_t s, uint64_t *exact_rows);

}  // namespace bht
