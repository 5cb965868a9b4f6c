// Internal interfaces between the C ABI (api.cpp, batch.cpp) and the HIP kernels (*.hip).
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

// Two cache-line groups: what the repulsion writes (on the second stream of a forked epoch) and
// what the rest of the epoch writes, so that the two streams never write into one 128-byte line.
struct alignas(128) EpochScalars {
    // -- written by the repulsion's kernels (grid plan, bounds, Z)
    float amin[8], amax[8];// per-axis bounding box after centring (Barnes-Hut, D <= 7)
    GridPlan plan;
    double zsum;           // sum of the per-point Z terms (FP64)
    float z, inv_z;        // Z and 1/Z of this epoch
    int sort_passes;       // 7-bit passes the box sort needs for this epoch's grid
    // -- written by the centring, attraction and update kernels and at set-up
    alignas(128) float mean[8]; // pending recentring shift (applied by the next centring kernel)
    float ysub[2];         // shift every reader subtracts from the epoch's embedding buffer
    int epoch;             // index of the epoch being executed
    int pad0;
    double psum;           // FP64 sum of P (normalisation)
    double kl_ref, kl_true;// KL divergence: reference formula, true formula
    int mean_pending;      // the last update left recentring partials to be applied
    int epoch_next;        // epoch index after the last update
    unsigned ctr[8];       // arrival counters of the single-pass reductions
    // -- the attraction's work queue (row blocks taken so far), written by the attraction alone
    // while the repulsion's head writes the group above on the other stream; reset by the update
    alignas(128) unsigned att_queue;
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

// Barnes-Hut tree in breadth-first order, as the reference's arena (src/tsne/arena.rs:36-62).
struct BhNodes {
    float4 *node;     // [2n] com.x, com.y, count (u32 bits), first_child (u32 bits, ~0 = leaf);
                      // D >= 3: nd_node_floats(D) floats per node, com[D], count, first_child
    uint2 *meta;      // [2n] child_count | level << 8; the squared half-width of an internal
                      // node's level (f32 bits), read beside the node by the walk
    uint32_t *range;  // start, end in the sorted order, per node
    float *hw2;       // [33] squared half-width per level
    uint32_t *counts; // build bookkeeping: top nodes, top rounds, deferred roots, round starts
    uint32_t *defer;  // [n] D = 2: the roots of the subtrees built by several workgroups
};

// ---- job table -------------------------------------------------------------------------------
//
// Every epoch kernel reads its job from a device-resident table.  A launch covers one job or
// many: block b of a launch belongs to job map[b] >> 16 as its block map[b] & 0xffff, out of the
// job's nblk[family] blocks of that kernel family.  Without a map the launch is job 0 alone and
// its grid is the job's block count.

enum Family {
    F_PT = 0,   // point-parallel passes (centring, box keys, gather, Morton codes)
    F_ONE,      // one block per job (finalisers, grid plan, tree build)
    F_SPR,      // box-segmented spreading and its fix-up
    F_ZERO,     // grid clearing
    F_KSPEC,    // kernel-spectrum row / column transforms
    F_ROWS,     // forward / inverse row transforms of the charge grids
    F_COLS,     // column transforms with the spectral product
    F_LINE1,    // D = 1 line convolution (two blocks)
    F_ATT,      // attraction + update
    F_TRAV,     // Barnes-Hut traversal
    F_BSUB,     // Barnes-Hut subtree builds (several blocks per job)
    NFAM
};

enum Method { M_BH = 0, M_FITSNE = 1, M_EXACT = 2 };

struct JobDev {
    int n, dim, method;
    int min_boxes, max_boxes, nodes_cap, len_cap;
    int nblk[NFAM];
    float theta_sq;
    // optimiser state; the embedding of epoch e is ybuf[e & 1]
    float *ybuf[2];
    float *uy, *gains, *frep;
    float *fattr;       // attraction sums of the epoch (k_attract -> k_update)
    const uint32_t *rows, *cols;
    float *vals;
    // column-tiled copy of P for the attraction (tiled.hip): rows in blocks of tile_rows, columns
    // in tiles of tile_points; tval is the copy the epochs read (and rescale at stop-lying)
    uint32_t tile_rows, tile_points, tile_nnz;
    const uint16_t *tkey;
    float *tval;
    const uint32_t *blk_seg, *seg_tile, *seg_e0, *seg_g0, *grp_e0, *grp_row;
    float *qv;          // values at the reference's CSR positions [0, qv_n) (KL weights)
    int qv_n;
    EpochScalars *scal;
    double *dpart;      // per-block FP64 partials (Z)
    double *mpart;      // per-block FP64 partials of the update (recentring sums)
    float *fpart;       // per-block f32 partials (min / max)
    // gradient step (src/tsne/mod.rs:470-516, src/lib.rs:1162-1179)
    float learning_rate, momentum, final_momentum, inv_exag;
    int momentum_switch_epoch, stop_lying_epoch;
    // FIt-SNE: box sort input and output, grids
    uint32_t *keys, *kidx;
    const uint32_t *skeys, *sidx;
    float4 *grid;       // spread charges / potentials, dense nodes x nodes (D = 2)
    float2 *spec;       // 2 x nodes_cap x len_cap complex
    float *rker;        // kernel row transform, (len_cap/2+1)^2
    float *kspec;       // kernel spectrum quarter, (len_cap/2+1)^2
    float *part;        // spread partials, 2 x nchunks x 36
    // FFT line buffers for launches whose LDS is smaller than the grid has grown (two lines of
    // len_cap points per fallback block; null when the launches always hold len_cap), and the
    // host-visible
    // word the grid plan reports its FFT length to (null: none)
    float2 *fft_fb;
    int *host_len;
    // Barnes-Hut: Morton sort input and output, tree.  Two-word codes (D >= 5) keep their low and
    // high words per point (clo, chi) and the low words in sorted order (slo); scodes then holds
    // the sorted high words.
    uint64_t *codes;
    uint32_t *cidx;
    const uint64_t *scodes;
    const uint32_t *scidx;
    uint64_t *clo, *chi, *slo;
    BhNodes nodes;
};

struct BlockCtx {
    int job, lb, nb;
};

__device__ __forceinline__ BlockCtx block_ctx(const uint32_t *map, const JobDev *jobs, int fam) {
    if (map == nullptr) return {0, (int)blockIdx.x, (int)gridDim.x};
    const uint32_t m = map[blockIdx.x];
    const int j = (int)(m >> 16);
    return {j, (int)(m & 0xffffu), jobs[j].nblk[fam]};
}

__device__ __forceinline__ int job_epoch(const JobDev &J) { return J.scal->epoch; }

// The recentring shift the last update left pending (zero_mean, src/tsne/mod.rs:534-568): the
// mean of its per-block FP64 sums (mpart, two per block).  Evaluated by a whole wave: lane-strided
// sums, then a butterfly whose every lane holds the same bits (FP64 addition commutes), so every
// kernel that evaluates it, whatever its block size, gets the same shift.
__device__ __forceinline__ void pending_mean2(const JobDev &J, float &m0, float &m1) {
    const int lane = threadIdx.x & 63;
    const int nparts = J.nblk[F_PT];
    double t0 = 0.0, t1 = 0.0;
    for (int b = lane; b < nparts; b += 64) {
        t0 += J.mpart[2 * b];
        t1 += J.mpart[2 * b + 1];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        t0 += __shfl_xor(t0, off, 64);
        t1 += __shfl_xor(t1, off, 64);
    }
    m0 = (float)(t0 / (double)J.n);
    m1 = (float)(t1 / (double)J.n);
}
__device__ __forceinline__ float *job_y(const JobDev &J) { return J.ybuf[J.scal->epoch & 1]; }

constexpr int FFT_FB_BLOCKS = 64;  // blocks of an FFT launch that transform lines too long for LDS

// Block counts of one job (the same formulas size the single-job grids and the batch maps).
int point_blocks(int n);
constexpr int MAX_UPDATE_PARTS = 1024;  // point_blocks() cap: partial sums of the update
// Column-tiled layout of P (tiled.hip).
struct TiledSizes {
    uint64_t max_segs, max_groups;  // bounds of the segment and group tables
    uint64_t padded_cap;            // bound of the padded edge count (tkey, tval entries)
    uint64_t persistent;            // bytes of the layout itself
    uint64_t temp;                  // bytes of workspace of tiled_build
};
struct TiledOut {
    uint16_t *tkey;
    float *tval;
    uint32_t *blk_seg, *seg_tile, *seg_e0, *seg_g0, *grp_e0, *grp_row;
};
uint32_t tiled_rows_per_block(int n);
uint32_t tiled_tile_points(int dim);
int tiled_blocks(int n);
TiledSizes tiled_sizes(int n, int dim, uint64_t nnz);
// Builds the layout; returns the padded edge count (tkey, tval entries).  Synchronises s.
uint64_t tiled_build(const uint32_t *rows, const uint32_t *cols, const float *vals, int n,
                     int dim, uint64_t nnz, const TiledOut &o, void *temp, hipStream_t s);
int spread_blocks(int n);
int traverse_blocks(int n, int dim);
int bh_sub_blocks(int n);
void job_block_counts(JobDev &J, bool batched);

// Segment of a segmented sort: the keys of one job at [base, base + n) of the arena, sorted in
// place of the same range; its tiles are [tile0, tile0 + ntiles) of the launch.
struct SortSeg {
    uint64_t base;     // offset of the segment in the arena (a multiple of the tile size)
    uint64_t dense0;   // keys of all earlier segments
    uint32_t n;
    uint32_t tile0;
    uint32_t ntiles;
    uint32_t pad;
};

// One launch group: jobs of one method and one output dimension, advanced together.
struct LaunchSet {
    const JobDev *jobs = nullptr;       // device table
    const uint32_t *map[NFAM] = {};     // device block maps (null: job 0 alone)
    int nb[NFAM] = {};                  // grid size of each family
    int dim = 2;
    int njobs = 1;
    int fft_cap = 0;                    // FFT line points per block the launches' LDS holds
    const TwiddleSet *tw = nullptr;
    // sort arena of the group (u32 box keys for FIt-SNE, u64 Morton codes for Barnes-Hut)
    void *keys = nullptr, *keys_alt = nullptr;
    uint32_t *vals = nullptr, *vals_alt = nullptr;
    uint64_t sort_n = 0;                // arena length (padded segments)
    void *sort_temp = nullptr;
    const SortSeg *segs = nullptr;      // null: one unpadded segment of sort_n keys
    const uint32_t *tile_seg = nullptr; // segment of each tile
    uint32_t sort_tiles = 0;            // tiles of the segmented sort
    uint32_t max_seg_tiles = ~0u;       // tiles of its largest segment
    // a single job's box sort with its pass count decided on the device (the grid plan's
    // sort_passes); its result is then in (keys_fin, vals_fin)
    const int *sort_npass = nullptr;
    void *keys_fin = nullptr;
    uint32_t *vals_fin = nullptr;
};

// ---- cross-stream cache maintenance (sync.hip) -----------------------------------------------
// Writes back and invalidates the L2 of every XCD, in stream order on s: the first launch on a
// stream that consumes data another stream (or a host copy) produced.
void l2_sync(hipStream_t s);

// ---- scan / sort (sort.hip) ----------------------------------------------------------------
size_t scan_temp_bytes(uint64_t n);
template <typename T>
void scan_exclusive(const uint32_t *in, T *out, uint64_t n, void *temp, T *total_out,
                    hipStream_t s);
constexpr int RS_TILE_KEYS = 2048;      // keys per radix-sort tile (segment alignment)
// Stable LSD radix sort of (key, u32 value) pairs, 8-bit digits (affinity construction).
size_t radix_sort_temp_bytes(uint64_t n);
template <typename K>
bool radix_sort_pairs(K *keys, uint32_t *vals, K *keys_alt, uint32_t *vals_alt, uint64_t n,
                      int begin_bit, int end_bit, void *temp, hipStream_t s);
// The per-epoch sorts of a launch group (box keys, Morton codes) with wider digits: the plain
// sort for one job, the segmented one for a batch.  Temp: group_sort_temp_bytes of the arena.
constexpr int GROUP_SORT_BITS = 10;
size_t group_sort_temp_bytes(uint64_t n);
template <typename K>
bool sort_group(const LaunchSet &L, int end_bit, hipStream_t s);
// Key bits of the per-epoch sorts and whether their result lands in the alternate buffers.
inline int box_key_bits(int dim) { return dim == 2 ? 20 : 10; }  // boxes^D <= 1000^D
constexpr int MORTON_KEY_BITS = 64;
// Digit width of a per-epoch sort.  A single job of SORT7_MIN_KEYS points or more sorts its 2-D
// box keys (20 bits) in three passes of 7-bit digits, few enough bins that a tile's keys land in
// runs (1M points: 2065 against 2040 epochs/s); smaller jobs and batches, bound by launches, keep
// two 10-bit passes (7-bit digits cost C4's plain fits 12%).  1-D box keys take one 10-bit pass,
// Morton codes 10-bit passes.
constexpr uint64_t SORT7_MIN_KEYS = 262144;
inline int group_sort_bits(int end_bit, bool large) {
    return end_bit == 20 && large ? 7 : GROUP_SORT_BITS;
}
inline bool sort_result_alt(int end_bit, bool large) {
    const int b = group_sort_bits(end_bit, large);
    return (((end_bit + b - 1) / b) & 1) != 0;
}

// ---- FIt-SNE epoch (fitsne.hip, fft.hip) ---------------------------------------------------
// Applies the pending recentring of the last update, then bounds and (do_plan) the grid plan.
void center_plan(const LaunchSet &L, bool do_plan, hipStream_t s);
// Its two halves: the recentring (mean + shift + per-block bounds), and the bounds' combination
// with the grid plan (which only the repulsion reads).
void center_points(const LaunchSet &L, hipStream_t s);
void plan_grid(const LaunchSet &L, bool do_plan, hipStream_t s);
// The exact fit (exactfit.hip): dense input distances (metric 0: Euclidean, 1: squared
// Euclidean), dense symmetrisation, and per epoch Z (zrow: n doubles) and the gradient of job 0
// of `jobs` into its attraction sums; the KL of the current embedding (zrow: 2n doubles).
void exact_distances(const float *x, int64_t n, int d_in, int metric, float *dist, hipStream_t s);
void exact_symmetrize(float *p, int64_t n, hipStream_t s);
void exact_gradient(const JobDev *jobs, const float *P, int64_t n, int dim, double *zrow,
                    hipStream_t s);
void exact_kl(const JobDev *jobs, const float *P, int64_t n, int dim, double *zrow,
              hipStream_t s);

// Head of a lazily centred FIt-SNE epoch (the recentring shift left to the readers) and its grid
// plan, from the last update's sums and bounds.
void epoch_head(const LaunchSet &L, hipStream_t s);
// Z and 1/Z from the per-block partial sums (minus n for FIt-SNE).
void finalize_z(const LaunchSet &L, int part_family, hipStream_t s);
// after center_plan: frep, Z; returns the sorted point indices of job 0 (a spatial order)
const uint32_t *fitsne_repulsion(const LaunchSet &L, hipStream_t s);
void fft_kernel_spectrum(const LaunchSet &L, hipStream_t s);
void fft_convolve(const LaunchSet &L, hipStream_t s);

// Barnes-Hut (bh.hip): after center_plan (no plan), frep and Z; returns job 0's sorted indices.
const uint32_t *bh_repulsion(const LaunchSet &L, hipStream_t s);

// Barnes-Hut epochs with D = 3..7 (bhnd.hip): the recentring, the bounds, the repulsion, the
// attraction over the CSR (tval: the values in CSR order) and the update, for a launch group of
// that dimension.  center_points, plan_grid, bh_repulsion, attract_tiled and update_points hand
// such groups to these.
int nd_bits(int dim);             // bits per axis of the Morton code
int nd_node_floats(int dim);      // floats per tree node
int nd_attract_blocks(int n);     // F_ATT blocks of a job
void nd_center_points(const LaunchSet &L, hipStream_t s);
void nd_bounds(const LaunchSet &L, hipStream_t s);
const uint32_t *nd_bh_repulsion(const LaunchSet &L, hipStream_t s);
void nd_attract(const LaunchSet &L, float *pos, bool epoch, hipStream_t s);
void nd_update(const LaunchSet &L, hipStream_t s);

// Attraction sums over the tiled layout (tiled.hip): epoch = true writes fattr (and applies the
// stop-lying rescale to tval), epoch = false writes job 0's sums into pos.
// queue_blocks > 0 (single job with its repulsion running concurrently on another stream): that
// many blocks take the row blocks from a work queue, leaving the rest of the device to the
// repulsion; 0: one block per row block.  The sums are the same either way.
// self_head (a lazily centred epoch whose head runs on the repulsion's stream beside it): the
// kernel takes the epoch index from epoch_next and the shift from pending_mean2 itself, the values
// the head derives from the same data.
void attract_tiled(const LaunchSet &L, float *pos, bool epoch, hipStream_t s,
                   int queue_blocks = 0, bool self_head = false);
int attract_default_queue_blocks();
// The gradient step from fattr, frep and Z (attract.hip).
void update_points(const LaunchSet &L, hipStream_t s);
void scale_values(float *vals, uint64_t nnz, float scale, hipStream_t s);
// Sparse KL (kl.hip): writes scal->kl_ref / kl_true using scal->inv_z.
void sparse_kl(const float *y, int n, int dim, const uint32_t *rows, const uint32_t *cols,
               const float *vals, const float *qv, const uint32_t *perm, EpochScalars *scal,
               double *dpart, hipStream_t s);
void sum_values_f64(const float *vals, uint64_t nnz, EpochScalars *scal, double *dpart,
                    hipStream_t s);

// FP64 brute-force repulsion (exact.hip): neg (n * dim doubles) and one partial Z per block.
int exact_repulsion_blocks(int64_t n);
void exact_repulsion_f64(const float *y, int64_t n, int dim, double *neg, double *zpart,
                         hipStream_t s);
// The same with its partials in the file's static device scratch (no allocation): y, neg and
// z_out are device pointers; Z (one double) is written to z_out on the device.
// exact_scratch_init(device) looks the scratch up once per device (at context creation).
void exact_scratch_init(int device);
uint64_t exact_noalloc_max_points();
void exact_repulsion_f64_noalloc(int device, const float *y, int64_t n, int dim, double *neg,
                                 double *z_out, hipStream_t s);

}  // namespace bht
