//! Raw declarations of the C ABI in `include/bhtsne_hip.h`, one per entry point, in the header's
//! order.  Every function that replaces a step of the `bhtsne` crate cites it in the header; the
//! safe builder over these is the `bhtsne-hip` crate.
//!
//! Pointers to device memory appear only in `bht_exact_repulsion_f64_noalloc`; every other
//! pointer argument is host memory of the length the header states.
#![allow(non_camel_case_types)]
#![no_std]

use core::ffi::{c_char, c_int, c_void};

pub type bht_status = i32;
pub const BHT_OK: bht_status = 0;
/// Input rejected; `bht_last_error()` holds the reference's panic message.
pub const BHT_ERR_INVALID: bht_status = 1;
pub const BHT_ERR_HIP: bht_status = 2;
pub const BHT_ERR_OOM: bht_status = 3;
pub const BHT_ERR_STATE: bht_status = 4;

pub const BHT_BANDWIDTH_FP64: u32 = 0;
pub const BHT_BANDWIDTH_F32_SEQ: u32 = 1;
pub const BHT_FLAG_NO_REORDER: u32 = 1;
pub const BHT_FLAG_SERIAL_EPOCH: u32 = 2;
pub const BHT_FLAG_NO_AUTOTUNE: u32 = 4;
pub const BHT_FLAG_FFT_GLOBAL_LINES: u32 = 8;
pub const BHT_METHOD_BARNES_HUT: u32 = 0;
pub const BHT_METHOD_FITSNE: u32 = 1;
pub const BHT_KNN_AUTO: u32 = 0;
pub const BHT_KNN_EXACT: u32 = 1;
pub const BHT_METRIC_EUCLIDEAN: u32 = 0;
pub const BHT_METRIC_SQEUCLIDEAN: u32 = 1;

#[repr(C)]
pub struct bht_ctx {
    _private: [u8; 0],
}
#[repr(C)]
pub struct bht_job {
    _private: [u8; 0],
}
#[repr(C)]
pub struct bht_batch {
    _private: [u8; 0],
}

/// Builder state of `tSNE` (src/lib.rs:149-176); `bht_params_default` fills the defaults of
/// `tSNE::new` (src/lib.rs:241-266).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct bht_params {
    pub d_out: u32,
    pub bandwidth_mode: u32,
    pub epochs: u64,
    pub momentum_switch_epoch: u64,
    pub stop_lying_epoch: u64,
    /// <= 0: the automatic rate, max(n / early_exaggeration / 4, 50)
    pub learning_rate: f32,
    pub momentum: f32,
    pub final_momentum: f32,
    pub early_exaggeration: f32,
    pub perplexity: f32,
    pub min_boxes: u32,
    pub max_boxes: u32,
    pub flags: u32,
}

/// Called on the calling thread after every epoch with a host copy of the recentred embedding
/// (src/lib.rs:1175-1178).
pub type bht_epoch_cb =
    Option<unsafe extern "C" fn(epoch: u64, embedding: *const f32, len: u64, user: *mut c_void)>;

extern "C" {
    pub fn bht_params_default(p: *mut bht_params);

    // ---- context
    pub fn bht_ctx_create(device: c_int, hbm_budget_bytes: u64, out: *mut *mut bht_ctx)
        -> bht_status;
    pub fn bht_ctx_destroy(ctx: *mut bht_ctx);
    pub fn bht_last_error() -> *const c_char;
    pub fn bht_ctx_bytes_in_use(ctx: *const bht_ctx) -> u64;

    // ---- job (one tSNE instance)
    pub fn bht_job_create(ctx: *mut bht_ctx, n_samples: u64, params: *const bht_params,
                          out: *mut *mut bht_job) -> bht_status;
    pub fn bht_job_destroy(job: *mut bht_job);
    pub fn bht_job_set_params(job: *mut bht_job, params: *const bht_params) -> bht_status;
    pub fn bht_job_get_params(job: *const bht_job, out: *mut bht_params) -> bht_status;
    pub fn bht_job_set_data_f32(job: *mut bht_job, x: *const f32, n: u64, d_in: u32)
        -> bht_status;
    pub fn bht_job_set_initial_embedding(job: *mut bht_job, y: *const f32, len: u64)
        -> bht_status;
    pub fn bht_job_barnes_hut(job: *mut bht_job, theta: f32, cb: bht_epoch_cb, user: *mut c_void)
        -> bht_status;
    pub fn bht_job_barnes_hut_with_neighbors(job: *mut bht_job, theta: f32,
                                             nbr_index: *const u64, nbr_distance: *const f32,
                                             row_lens: *const u64, n_rows: u64, cb: bht_epoch_cb,
                                             user: *mut c_void) -> bht_status;
    pub fn bht_job_fit_sne(job: *mut bht_job, cb: bht_epoch_cb, user: *mut c_void) -> bht_status;
    pub fn bht_job_exact(job: *mut bht_job, metric: u32, cb: bht_epoch_cb, user: *mut c_void)
        -> bht_status;
    pub fn bht_job_fit_sne_with_neighbors(job: *mut bht_job, nbr_index: *const u64,
                                          nbr_distance: *const f32, row_lens: *const u64,
                                          n_rows: u64, cb: bht_epoch_cb, user: *mut c_void)
        -> bht_status;
    pub fn bht_job_get_embedding(job: *const bht_job, out: *mut f32, len: u64) -> bht_status;
    pub fn bht_job_kl_divergence(job: *mut bht_job, reference_kl: *mut f64, true_kl: *mut f64)
        -> bht_status;
    pub fn bht_job_get_affinities(job: *const bht_job, rows: *mut u64, cols: *mut u32,
                                  vals: *mut f32, nnz: *mut u64, perplexity: *mut f32)
        -> bht_status;
    pub fn bht_job_set_affinities(job: *mut bht_job, rows: *const u64, cols: *const u32,
                                  vals: *const f32, n_rows_plus_one: u64, nnz: u64,
                                  perplexity: f32) -> bht_status;
    pub fn bht_check_affinities(rows: *const u64, cols: *const u32, n_rows_plus_one: u64,
                                nnz: u64) -> bht_status;

    // ---- stepping interface
    pub fn bht_job_prepare(job: *mut bht_job, method: u32, theta: f32) -> bht_status;
    pub fn bht_job_step(job: *mut bht_job, n_epochs: u64, cb: bht_epoch_cb, user: *mut c_void)
        -> bht_status;
    pub fn bht_job_finish(job: *mut bht_job) -> bht_status;
    pub fn bht_job_synchronize(job: *mut bht_job) -> bht_status;
    pub fn bht_job_forces(job: *mut bht_job, attractive: *mut f32, repulsive: *mut f32,
                          inverse_z: *mut f32) -> bht_status;
    pub fn bht_job_get_state(job: *const bht_job, y: *mut f32, uy: *mut f32, gains: *mut f32)
        -> bht_status;
    pub fn bht_job_set_state(job: *mut bht_job, y: *const f32, uy: *const f32,
                             gains: *const f32) -> bht_status;
    pub fn bht_job_run_state(job: *const bht_job, next_epoch: *mut u64, exaggerated: *mut c_int)
        -> bht_status;
    pub fn bht_job_epoch_form(job: *const bht_job, attraction_blocks: *mut i32) -> bht_status;
    pub fn bht_job_grid_info(job: *const bht_job, boxes: *mut u32, fft_len: *mut u32,
                             lo: *mut f32, hi: *mut f32) -> bht_status;
    pub fn bht_job_set_timing(job: *mut bht_job, enable: c_int) -> bht_status;
    pub fn bht_job_get_timing(job: *mut bht_job, out3: *mut f64) -> bht_status;

    // ---- operators on host buffers
    pub fn bht_knn_f32(ctx: *mut bht_ctx, x: *const f32, n: u64, d_in: u32, k: u32,
                       idx_out: *mut u32, dist_out: *mut f32) -> bht_status;
    pub fn bht_knn_f32_ex(ctx: *mut bht_ctx, x: *const f32, n: u64, d_in: u32, k: u32,
                          mode: u32, idx_out: *mut u32, dist_out: *mut f32,
                          exact_rows: *mut u64) -> bht_status;
    pub fn bht_exact_repulsion_f64(ctx: *mut bht_ctx, y: *const f32, n: u64, d: u32,
                                   neg_out: *mut f64, z_out: *mut f64) -> bht_status;
    /// Device pointers and a hipStream_t (null: the context's stream); allocates nothing.
    pub fn bht_exact_repulsion_f64_noalloc(ctx: *mut bht_ctx, y_dev: *const f32, n: u64, d: u32,
                                           neg_dev: *mut f64, z_dev: *mut f64,
                                           stream: *mut c_void) -> bht_status;
    pub fn bht_search_beta_f32(ctx: *mut bht_ctx, dist: *const f32, n: u64, k: u32,
                               perplexity: f32, mode: u32, p_out: *mut f32) -> bht_status;
    pub fn bht_symmetrize(ctx: *mut bht_ctx, cols: *const u32, vals: *const f32, n: u64, k: u32,
                          rows_out: *mut u64, cols_out: *mut u32, vals_out: *mut f32,
                          nnz: *mut u64) -> bht_status;

    // ---- batched and served jobs
    pub fn bht_batch_create(ctx: *mut bht_ctx, out: *mut *mut bht_batch) -> bht_status;
    pub fn bht_batch_destroy(batch: *mut bht_batch);
    pub fn bht_batch_add(batch: *mut bht_batch, job: *mut bht_job, method: u32, theta: f32)
        -> bht_status;
    pub fn bht_batch_step(batch: *mut bht_batch, max_epochs: u64, epochs_advanced: *mut u64)
        -> bht_status;
    pub fn bht_batch_take_finished(batch: *mut bht_batch, out: *mut *mut bht_job, capacity: u64,
                                   count: *mut u64) -> bht_status;
    pub fn bht_batch_size(batch: *const bht_batch) -> u64;
    pub fn bht_batch_set_graph(batch: *mut bht_batch, enable: c_int) -> bht_status;
    pub fn bht_batch_set_concurrent(batch: *mut bht_batch, enable: c_int) -> bht_status;
    pub fn bht_batch_synchronize(batch: *mut bht_batch) -> bht_status;
    pub fn bht_job_footprint(n: u64, d_in: u32, params: *const bht_params, method: u32,
                             bytes: *mut u64) -> bht_status;
}
