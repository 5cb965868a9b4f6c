// Internal interfaces between the C ABI (api.cpp) and the HIP kernels (*.hip).
#pragma once

#include "common.hpp"

namespace bht {

// ---- device-resident per-job scalars ---------------------------------------------------------

// FIt-SNE grid of one epoch (src/tsne/interpolation.rs:112-128), decided on the device from the
// joint bounding box so that an epoch needs no host round trip.
struct GridPlan {
    float lo, hi;          // joint min / max over all coordinates
    float box_width;       // extent / boxes
    float inv_box_width;   // 1 / box_width
    float spacing;         // box_width / 3
    int boxes;             // boxes per axis
    int nodes;             // 3 * boxes
    int len;               // FFT length per axis (7-smooth >= 2 * nodes - 1)
    int tw_off;            // offset of this length's twiddle table
    float qscale;          // power of two >= max |coordinate|: balances the packed FFT pairs
};

struct EpochScalars {
    float mean[2];         // pending recentring shift (applied by the next centring kernel)
    float amin[2], amax[2];// per-axis bounding box after centring (Barnes-Hut)
    GridPlan plan;
    double zsum;           // sum of the per-point Z terms (FP64)
    float z, inv_z;        // Z and 1/Z of this epoch
    int epoch;             // index of the epoch being executed
    int pad0;
    double psum;           // FP64 sum of P (normalisation)
    double kl_ref, kl_true;// KL divergence: reference formula, true formula
    int mean_pending;      // the last update left recentring partials to be applied
    int epoch_next;        // epoch index after the last update
    unsigned ctr[8];       // arrival counters of the single-pass reductions
};

enum CtrSlot { CTR_CENTER = 0, CTR_GATHER = 1, CTR_UPDATE = 2, CTR_PSUM = 3, CTR_KL = 4, CTR_BH = 5 };

// Per-length twiddle tables exp(-2 pi i j / L), one per distinct length reachable from
// boxes = 1..MAX_BOXES (context-wide, built once).
constexpr int MAX_BOXES = 1000;  // interpolation.rs:42
constexpr int IP = 3;            // interpolation.rs:24
struct TwiddleSet {
    float2 *table = nullptr;     // concatenated tables
    int *len_of_boxes = nullptr; // [MAX_BOXES + 1]
    int *off_of_boxes = nullptr; // [MAX_BOXES + 1]
    size_t total = 0;
};

// Device view of one FIt-SNE job's epoch workspace.
struct FitsneDev {
    int n;
    int dim;           // 1 or 2
    int min_boxes;     // 10 = reference grid; larger = finer
    int nodes_cap;     // capacity of the node grids (3 * max boxes)
    int len_cap;       // capacity of the FFT length
    float *y;          // current embedding (n * dim)
    float *frep;       // repulsive forces (n * dim)
    uint32_t *keys, *vals, *keys_alt, *vals_alt;  // box sort
    float4 *grid;      // spread charges / potentials, dense nodes x nodes (D = 2)
    float2 *spec;      // 2 x nodes_cap x len_cap complex
    float *rker;       // kernel row transform, (len_cap/2+1)^2
    float *kspec;      // kernel spectrum quarter, (len_cap/2+1)^2
    float *part;       // spread partials, 2 x nchunks x 36
    double *dpart;     // per-block FP64 partials (Z)
    double *mpart;     // per-block FP64 partials of the update (recentring sums)
    float *fpart;      // per-block f32 partials (min/max)
    void *sort_temp;
    EpochScalars *scal;
    const TwiddleSet *tw;
};

// ---- scan / sort (sort.hip) ----------------------------------------------------------------
size_t scan_temp_bytes(uint64_t n);
template <typename T>
void scan_exclusive(const uint32_t *in, T *out, uint64_t n, void *temp, T *total_out,
                    hipStream_t s);
size_t radix_sort_temp_bytes(uint64_t n);
template <typename K>
bool radix_sort_pairs(K *keys, uint32_t *vals, K *keys_alt, uint32_t *vals_alt, uint64_t n,
                      int begin_bit, int end_bit, void *temp, hipStream_t s);

// ---- FIt-SNE epoch (fitsne.hip, fft.hip) ---------------------------------------------------
int point_blocks(int n);    // grid size of the point-parallel kernels
int attract_blocks(int n);  // grid size of the attraction kernel (= its partial count)
// Applies the pending recentring of the last update, then bounds and the grid plan.
void fitsne_center_plan(const FitsneDev &d, hipStream_t s);
// Applies the pending recentring and writes the per-axis bounds (no grid plan).
void center_only(float *y, int n, int dim, EpochScalars *scal, float *fpart, const double *mpart,
                 hipStream_t s);
// Z and 1/Z from per-block partial sums (minus n for FIt-SNE).
void finalize_z(EpochScalars *scal, const double *dpart, int nparts, int n, bool subtract_n,
                hipStream_t s);
// after center_plan: frep, Z; returns the point indices in box order (device)
const uint32_t *fitsne_repulsion(FitsneDev &d, hipStream_t s);
void fft_kernel_spectrum(const FitsneDev &d, hipStream_t s);
void fft_convolve(const FitsneDev &d, hipStream_t s);

// Attraction + fused update (attract.hip).  rows are u32 CSR offsets (nnz < 2^32).
struct UpdateArgs {
    int n, dim;
    const uint32_t *rows;
    const uint32_t *cols;
    float *vals;
    const float *y;
    float *y_next;
    float *uy, *gains;
    const float *frep;
    float learning_rate, momentum, final_momentum, inv_exag;
    int momentum_switch_epoch, stop_lying_epoch;
    EpochScalars *scal;
    double *mpart;  // per-block sums of the new embedding (2 per block)
};
void attract_update(const UpdateArgs &a, hipStream_t s);
void attract_only(int n, int dim, const uint32_t *rows, const uint32_t *cols, const float *vals,
                  const float *y, float *pos, hipStream_t s);
void scale_values(float *vals, uint64_t nnz, float scale, hipStream_t s);
// Sparse KL (kl.hip): writes scal->kl_ref / kl_true using scal->inv_z.
void sparse_kl(const float *y, int n, int dim, const uint32_t *rows, const uint32_t *cols,
               const float *vals, const float *qv, const uint32_t *perm, EpochScalars *scal,
               double *dpart, hipStream_t s);
void sum_values_f64(const float *vals, uint64_t nnz, EpochScalars *scal, double *dpart,
                    hipStream_t s);

}  // namespace bht
