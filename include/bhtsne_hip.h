/*
 * bhtsne_hip.h -- C ABI of the MI355X (gfx950) t-SNE engine: sparse affinity construction and the
 * per-epoch gradient of the bhtsne crate, for the Euclidean case, as a drop-in behind the crate's
 * builder.
 *
 * Reference interface each entry point replaces (paths relative to the bhtsne 0.7.10 crate root):
 *
 *   tSNE::new(&data)                         src/lib.rs:241-266   -> bht_job_create
 *   learning_rate .. perplexity setters      src/lib.rs:277-371   -> bht_params fields
 *   initial_embedding(..)                    src/lib.rs:431-435   -> bht_job_set_initial_embedding
 *   barnes_hut(theta, euclidean)             src/lib.rs:701-742   -> bht_job_barnes_hut
 *   barnes_hut_with_neighbors(theta, nbrs)   src/lib.rs:757-772   -> bht_job_barnes_hut_with_neighbors
 *   fit_sne(euclidean)                       src/lib.rs:835-871   -> bht_job_fit_sne
 *   fit_sne_with_neighbors(nbrs)             src/lib.rs:884-894   -> bht_job_fit_sne_with_neighbors
 *   epoch_callback(..)                       src/lib.rs:415-422   -> bht_epoch_cb argument
 *   embedding()                              src/lib.rs:438-440   -> bht_job_get_embedding
 *   kl_divergence()                          src/lib.rs:448-477   -> bht_job_kl_divergence
 *   affinities() / with_affinities(..)       src/lib.rs:1019-1051 -> bht_job_get_affinities /
 *                                                                    bht_job_set_affinities
 *   VPTree::search (k nearest neighbours)    src/tsne/vptree.rs:274-319 -> bht_knn_f32
 *   search_beta (bandwidth bisection)        src/tsne/mod.rs:201-275    -> bht_search_beta_f32
 *   symmetrize_sparse_matrix                 src/tsne/mod.rs:306-406    -> bht_symmetrize
 *
 * Every function returns a bht_status.  On failure bht_last_error() (thread-local) holds a message;
 * for the inputs the reference rejects with a panic the message text is the reference's own
 * ("error: the provided perplexity is too large ...", "... must have the same length ...",
 * "... out of range ...", "error: initial embedding has ...", "error: theta value must be ...").
 *
 * All pointers are host pointers owned by the caller; the library copies in and out during the
 * call.  Device state (data, P, embedding, optimiser buffers) lives in the job between calls.
 * A context and its jobs may be used from one host thread at a time.
 */
#ifndef BHTSNE_HIP_H
#define BHTSNE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t bht_status;
#define BHT_OK 0
#define BHT_ERR_INVALID 1 /* input rejected (reference panic message in bht_last_error) */
#define BHT_ERR_HIP 2     /* HIP runtime error */
#define BHT_ERR_OOM 3     /* device allocation failed or HBM budget exceeded */
#define BHT_ERR_STATE 4   /* call out of order */

typedef struct bht_ctx bht_ctx;
typedef struct bht_job bht_job;

/* Called after every epoch with the epoch index and the (recentred) embedding, n * d_out values
 * (src/lib.rs:1175-1178).  Runs on the calling thread. */
typedef void (*bht_epoch_cb)(uint64_t epoch, const float *embedding, uint64_t len, void *user);

#define BHT_BANDWIDTH_FP64 0u     /* bandwidth-search sums in FP64 (default) */
#define BHT_BANDWIDTH_F32_SEQ 1u  /* sums in f32, sequential, as the reference's loop */

/* Builder state of tSNE (src/lib.rs:149-176) with the defaults of tSNE::new (src/lib.rs:241-266). */
typedef struct bht_params {
    uint32_t d_out;                 /* embedding dimension: 2..7 (Barnes-Hut), 1 or 2 (FIt-SNE),
                                       1..7 (exact) */
    uint32_t bandwidth_mode;        /* BHT_BANDWIDTH_* */
    uint64_t epochs;                /* 1000 */
    uint64_t momentum_switch_epoch; /* 250 */
    uint64_t stop_lying_epoch;      /* 250 */
    float learning_rate;            /* <= 0: auto, max(n / early_exaggeration / 4, 50) */
    float momentum;                 /* 0.5 */
    float final_momentum;           /* 0.8 */
    float early_exaggeration;       /* 12 */
    float perplexity;               /* 20; k = (int)(3 * perplexity) neighbours per row, at most
                                       192 from sample vectors (perplexity < 64.34) and 256
                                       from caller neighbours: larger k is rejected with
                                       BHT_ERR_INVALID ("... neighbours per row ...") */
    uint32_t min_boxes;             /* FIt-SNE boxes per axis lower bound; 10 is the reference's */
    uint32_t max_boxes;             /* FIt-SNE grid capacity per axis, <= 1000 (the reference's) */
    uint32_t flags;                 /* BHT_FLAG_* */
} bht_params;

/* Jobs of 32768 points or more are renumbered once, before the first epoch, in an order of
 * their affinity graph that keeps well-connected points together, so that the attraction's
 * gathers stay in cache (results are unchanged up to summation order; outputs are always in
 * sample order).  This flag keeps the input order. */
#define BHT_FLAG_NO_REORDER 1u
/* Jobs of 32768 points or more run each epoch's repulsion on a second stream beside the
 * attraction (with the cache maintenance the handover between streams needs, the results are
 * the same bits as in order).  This flag keeps every epoch in order on the job's stream, the
 * order a batch runs its members in. */
#define BHT_FLAG_SERIAL_EPOCH 2u
/* A job of 32768 points or more takes the form of its epochs (the two streams or one, and how
 * many workgroups the attraction keeps beside the repulsion) from a per-context autotuner, which
 * times the candidates on the first run of each job size class (a few epochs from a snapshot of
 * the job's state, which is then restored; every candidate gives the same result bits).  This
 * flag uses the built-in form instead. */
#define BHT_FLAG_NO_AUTOTUNE 4u
/* A 2-D FIt-SNE job of 32768 points or more sizes the LDS of its FFT launches for the lengths its
 * grid has reached (so that they fit a CU beside the attraction), and transforms a line that has
 * outgrown them in global memory with the same arithmetic.  This flag sizes them for 64 points, so
 * that every line takes the global-memory path (a check of that path; results are the same bits). */
#define BHT_FLAG_FFT_GLOBAL_LINES 8u

void bht_params_default(bht_params *p);

/* ---- context --------------------------------------------------------------------------------- */
bht_status bht_ctx_create(int device, uint64_t hbm_budget_bytes, bht_ctx **out);
void bht_ctx_destroy(bht_ctx *ctx);
const char *bht_last_error(void);
/* Bytes of device memory currently held by the context's jobs. */
uint64_t bht_ctx_bytes_in_use(const bht_ctx *ctx);

/* ---- job (one tSNE instance) ----------------------------------------------------------------- */
bht_status bht_job_create(bht_ctx *ctx, uint64_t n_samples, const bht_params *params,
                          bht_job **out);
void bht_job_destroy(bht_job *job);
bht_status bht_job_set_params(bht_job *job, const bht_params *params);
bht_status bht_job_get_params(const bht_job *job, bht_params *out);

/* Sample vectors for the Euclidean metric: n_samples rows of d_in f32, row-major. */
bht_status bht_job_set_data_f32(bht_job *job, const float *x, uint64_t n, uint32_t d_in);
/* Seed embedding for the next fit, consumed by it (src/lib.rs:672-684). */
bht_status bht_job_set_initial_embedding(bht_job *job, const float *y, uint64_t len);

/* Fits.  theta > 0 (src/lib.rs:1182-1188).  Neighbours are given as concatenated rows:
 * row i holds row_lens[i] entries of (index, distance), like &[Vec<Neighbor<f32>>]. */
bht_status bht_job_barnes_hut(bht_job *job, float theta, bht_epoch_cb cb, void *user);
bht_status bht_job_barnes_hut_with_neighbors(bht_job *job, float theta, const uint64_t *nbr_index,
                                             const float *nbr_distance, const uint64_t *row_lens,
                                             uint64_t n_rows, bht_epoch_cb cb, void *user);
bht_status bht_job_fit_sne(bht_job *job, bht_epoch_cb cb, void *user);
/* The exact O(n^2) fit (tSNE::exact, src/lib.rs:489-646) with the distance `metric` between the
 * sample vectors: dense P over all pairs (each row's own entry included, as the reference's dense
 * rows do), the exact gradient with Z summed over all n^2 ordered pairs, n <= 65536.  The
 * reference's exact tests pass the squared Euclidean distance.  affinities() is None after it. */
#define BHT_METRIC_EUCLIDEAN 0u
#define BHT_METRIC_SQEUCLIDEAN 1u
bht_status bht_job_exact(bht_job *job, uint32_t metric, bht_epoch_cb cb, void *user);
bht_status bht_job_fit_sne_with_neighbors(bht_job *job, const uint64_t *nbr_index,
                                          const float *nbr_distance, const uint64_t *row_lens,
                                          uint64_t n_rows, bht_epoch_cb cb, void *user);

/* Results. */
bht_status bht_job_get_embedding(const bht_job *job, float *out, uint64_t len);
/* KL divergence of the last fit with the fit's own Z estimate: the reference's value
 * (src/tsne/mod.rs:738-775, which weights every edge of row i by p_values[i]) and the true
 * sum over edges of p log(p / q).  BHT_ERR_STATE before any fit (the reference returns None). */
bht_status bht_job_kl_divergence(bht_job *job, double *reference_kl, double *true_kl);

/* Affinity cache (src/lib.rs:1019-1051): two-phase.  With rows == NULL, *nnz receives the number
 * of entries; otherwise rows (n+1, usize offsets), cols and vals (nnz each) are filled.  Values are
 * divided by early_exaggeration unless stop-lying fired, as affinities() returns them (so with
 * stop_lying_epoch == 0, which never sets the flag, they are divided once more).
 * set_affinities checks the CSR on the host before any upload (rows[0] == 0, non-decreasing,
 * rows[n] == nnz, every column < n) and rejects a malformed one with BHT_ERR_INVALID. */
bht_status bht_job_get_affinities(const bht_job *job, uint64_t *rows, uint32_t *cols, float *vals,
                                  uint64_t *nnz, float *perplexity);
bht_status bht_job_set_affinities(bht_job *job, const uint64_t *rows, const uint32_t *cols,
                                  const float *vals, uint64_t n_rows_plus_one, uint64_t nnz,
                                  float perplexity);
/* The structure check set_affinities makes, on its own (host only; no device or context needed).
 * No counterpart in the reference, whose with_affinities (src/lib.rs:1043-1051) takes the CSR as
 * it is; for callers that want to vet an affinity graph before handing it over. */
bht_status bht_check_affinities(const uint64_t *rows, const uint32_t *cols,
                                uint64_t n_rows_plus_one, uint64_t nnz);

/* ---- stepping interface (the epoch loop of src/lib.rs:961-1014 in pieces) ---------------------
 * prepare: P from data / cached affinities normalised and exaggerated, optimiser reset, embedding
 * seeded (src/lib.rs:663-685, 1227-1236).  method: BHT_METHOD_*.  step: run n epochs.
 * finish: recentre, apply a pending stop-lying, record the fit for kl_divergence(). */
#define BHT_METHOD_BARNES_HUT 0u
#define BHT_METHOD_FITSNE 1u
bht_status bht_job_prepare(bht_job *job, uint32_t method, float theta);
bht_status bht_job_step(bht_job *job, uint64_t n_epochs, bht_epoch_cb cb, void *user);
bht_status bht_job_finish(bht_job *job);
/* Blocks until the job's queued device work is done. */
bht_status bht_job_synchronize(bht_job *job);

/* State access for parity checks: forces of the current state (attractive, repulsive, 1/Z;
 * src/repulsion.rs:95-158, 206-234) without updating, and the optimiser state (n * d_out each). */
bht_status bht_job_forces(bht_job *job, float *attractive, float *repulsive, float *inverse_z);
bht_status bht_job_get_state(const bht_job *job, float *y, float *uy, float *gains);
bht_status bht_job_set_state(bht_job *job, const float *y, const float *uy, const float *gains);
/* Position of the prepared run: the next epoch index, and whether the next epoch's P is the
 * exported affinities times early_exaggeration (while P is still exaggerated, and throughout a
 * run with stop_lying_epoch == 0, whose export divides P by the exaggeration once more as
 * affinities() does, src/lib.rs:1024-1029). */
bht_status bht_job_run_state(const bht_job *job, uint64_t *next_epoch, int *exaggerated);
/* The form the last run's epochs took (no counterpart in the reference): -1 in order on one
 * stream, 0 the repulsion beside the attraction with a workgroup per row block, k > 0 beside an
 * attraction on k workgroups taking row blocks from a queue. */
bht_status bht_job_epoch_form(const bht_job *job, int32_t *attraction_blocks);
/* FIt-SNE grid of the last repulsion: boxes per axis, FFT length, joint min and max. */
bht_status bht_job_grid_info(const bht_job *job, uint32_t *boxes, uint32_t *fft_len, float *lo,
                             float *hi);

/* Per-phase device time of the epochs run while timing is enabled, measured with HIP events
 * (the events add a little time to each epoch; time throughput with timing off).  After the
 * recentring an epoch runs the repulsion and the attraction, concurrently on two streams for a
 * job of 32768 points or more, else in order, and then the update.  out[0] = repulsion ms (box
 * sort .. Z), out[1] = attraction ms (the attraction kernel alone), summed over the timed
 * epochs; out[2] = epochs timed.  Enabling resets the counters. */
bht_status bht_job_set_timing(bht_job *job, int enable);
bht_status bht_job_get_timing(bht_job *job, double *out3);

/* ---- operators on host buffers (the reference's internal steps, for parity checks) ----------- */
/* k nearest neighbours of every row (self excluded), ascending (distance, index), with the
 * reference metric's f32 distances.  The fits use the same routine. */
bht_status bht_knn_f32(bht_ctx *ctx, const float *x, uint64_t n, uint32_t d_in, uint32_t k,
                       uint32_t *idx_out, float *dist_out);
/* The same with the path chosen: BHT_KNN_AUTO = matrix-core candidates + exact re-rank +
 * per-row certificate (exact kernel for uncertified rows; the result is identical), or
 * BHT_KNN_EXACT = the exact kernel for every row.  exact_rows (optional) receives the number of
 * rows the exact kernel computed. */
#define BHT_KNN_AUTO 0u
#define BHT_KNN_EXACT 1u
bht_status bht_knn_f32_ex(bht_ctx *ctx, const float *x, uint64_t n, uint32_t d_in, uint32_t k,
                          uint32_t mode, uint32_t *idx_out, float *dist_out,
                          uint64_t *exact_rows);
/* The non-edge terms of the exact gradient in FP64 (src/lib.rs:570-622): for every point i,
 * neg_i = sum_j q_ij^2 (y_i - y_j) over all other points and Z = sum of q_ij over ordered pairs,
 * q_ij = 1 / (1 + |y_i - y_j|^2).  O(n^2); the yardstick for the approximate gradients' error.
 * y: n * d f32 (d = 1 or 2); neg_out: n * d doubles. */
bht_status bht_exact_repulsion_f64(bht_ctx *ctx, const float *y, uint64_t n, uint32_t d,
                                   double *neg_out, double *z_out);
/* The same computation without any allocation, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  y_dev (n * d f32), neg_dev (n * d doubles) and z_dev (one double) are DEVICE
 * pointers on the context's device; the call only enqueues work on `stream` (a hipStream_t; NULL:
 * the context's stream) and returns.  The per-block partials of Z live in a static device array
 * of 512 KiB that each GPU holds for the life of the process, so n <= 16,777,216 (larger n is
 * rejected with BHT_ERR_INVALID), and calls that overlap on one GPU race.  The results equal
 * bht_exact_repulsion_f64's bit for bit. */
bht_status bht_exact_repulsion_f64_noalloc(bht_ctx *ctx, const float *y_dev, uint64_t n,
                                           uint32_t d, double *neg_dev, double *z_dev,
                                           void *stream);
/* Per-row bandwidth search: conditional P (n * k) from neighbour distances (n * k). */
bht_status bht_search_beta_f32(bht_ctx *ctx, const float *dist, uint64_t n, uint32_t k,
                               float perplexity, uint32_t mode, float *p_out);
/* Symmetrisation into CSR with the reference's row order.  Two-phase like get_affinities. */
bht_status bht_symmetrize(bht_ctx *ctx, const uint32_t *cols, const float *vals, uint64_t n,
                          uint32_t k, uint64_t *rows_out, uint32_t *cols_out, float *vals_out,
                          uint64_t *nnz);

/* ---- batched and served jobs ------------------------------------------------------------------
 *
 * No counterpart in the reference, which runs one tSNE per call on the CPU.  A batch advances
 * many independent jobs together: every epoch kernel is one launch for all members (of the same
 * method and d_out), and the epoch is replayed from a captured graph.  Each member computes
 * exactly what bht_job_fit_sne / bht_job_barnes_hut computes for it alone (same kernels, same
 * block partition, same summation order, same point renumbering).  A request loop admits jobs with bht_batch_add as others finish, so a batch serves a
 * stream of requests. */
typedef struct bht_batch bht_batch;

bht_status bht_batch_create(bht_ctx *ctx, bht_batch **out);
/* Members still running are dropped (their runs are left unfinished). */
void bht_batch_destroy(bht_batch *batch);
/* The fit preamble of the job (affinities from its data unless cached, P normalisation, the
 * initial embedding; the same validation and messages as the fit calls), then membership.
 * method: BHT_METHOD_FITSNE or BHT_METHOD_BARNES_HUT (theta > 0).  Each job runs params.epochs. */
bht_status bht_batch_add(bht_batch *batch, bht_job *job, uint32_t method, float theta);
/* Advances every member by up to max_epochs, never past the member with the fewest epochs left;
 * members whose runs complete are finished (as the fit calls finish them) and queued for
 * bht_batch_take_finished.  Jobs must not be destroyed while they are members. */
bht_status bht_batch_step(bht_batch *batch, uint64_t max_epochs, uint64_t *epochs_advanced);
bht_status bht_batch_take_finished(bht_batch *batch, bht_job **out, uint64_t capacity,
                                   uint64_t *count);
uint64_t bht_batch_size(const bht_batch *batch);
/* Graph replay of the epoch (default on) or plain launches. */
bht_status bht_batch_set_graph(bht_batch *batch, int enable);
/* The groups' repulsions on a second stream beside their attractions, or every kernel in order
 * on one stream (the default: a batch's grouped launches already fill the device); the results
 * are the same bits either way. */
bht_status bht_batch_set_concurrent(bht_batch *batch, int enable);
bht_status bht_batch_synchronize(bht_batch *batch);
/* Upper bound of the device bytes one job of n samples (d_in features) needs: what a request loop
 * admits against the context's HBM budget. */
bht_status bht_job_footprint(uint64_t n, uint32_t d_in, const bht_params *params, uint32_t method,
                             uint64_t *bytes);

#ifdef __cplusplus
}
#endif

#endif /* BHTSNE_HIP_H */
