//! The `bhtsne` builder for `f32` samples and the Euclidean metric, run by libbhtsne_hip on an
//! MI355X.  The surface follows `tSNE<'data, f32, &[f32], D>` (src/lib.rs:149-1268 of the
//! `bhtsne` crate): the same setters and defaults, `barnes_hut` / `fit_sne` and their
//! `*_with_neighbors` variants, `embedding`, `kl_divergence`, `affinities` / `with_affinities`,
//! and the same panics with the same messages (the library reports the reference's texts).
//!
//! The metric closure of the reference cannot be inspected, so the Euclidean fits are named
//! `barnes_hut_euclidean` / `fit_sne_euclidean`; the `*_with_neighbors` fits take the caller's
//! neighbours as the reference does.  `Batch` advances many jobs together (no counterpart in the
//! reference).

use std::collections::HashMap;
use std::ffi::{c_void, CStr};
use std::ptr;
use std::sync::{Arc, OnceLock};

pub use bhtsne_hip_sys as sys;

/// A sample's nearest neighbour (src/lib.rs:107-111).
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct Neighbor<T> {
    pub index: usize,
    pub distance: T,
}

/// Symmetrised affinities in CSR form (src/lib.rs:116-144): usize row offsets, u32 columns.
#[derive(Clone, Debug, PartialEq)]
pub struct SparseAffinities<T> {
    rows: Vec<usize>,
    columns: Vec<u32>,
    values: Vec<T>,
    perplexity: T,
}

impl<T> SparseAffinities<T> {
    pub fn new(rows: Vec<usize>, columns: Vec<u32>, values: Vec<T>, perplexity: T) -> Self {
        SparseAffinities { rows, columns, values, perplexity }
    }
    pub fn rows(&self) -> &[usize] {
        &self.rows
    }
    pub fn columns(&self) -> &[u32] {
        &self.columns
    }
    pub fn values(&self) -> &[T] {
        &self.values
    }
    pub fn perplexity(&self) -> &T {
        &self.perplexity
    }
}

/// Panics with the library's message unless `st` is `BHT_OK` (the reference panics with the
/// same texts, which its `should_panic` tests match).
fn check(st: sys::bht_status) {
    if st != sys::BHT_OK {
        let msg = unsafe { CStr::from_ptr(sys::bht_last_error()) };
        panic!("{}", msg.to_string_lossy());
    }
}

/// One device context: twiddle tables, the caching allocator and the HBM budget.
pub struct Context {
    raw: *mut sys::bht_ctx,
}

// The library locks the context's allocator and autotuner; a context may be shared by threads.
unsafe impl Send for Context {}
unsafe impl Sync for Context {}

impl Context {
    /// A context on `device`; `hbm_budget` bytes (0: no limit) bound what its jobs may hold.
    pub fn new(device: i32, hbm_budget: u64) -> Arc<Context> {
        let mut raw = ptr::null_mut();
        check(unsafe { sys::bht_ctx_create(device, hbm_budget, &mut raw) });
        Arc::new(Context { raw })
    }

    /// Bytes of device memory the context's jobs hold.
    pub fn bytes_in_use(&self) -> u64 {
        unsafe { sys::bht_ctx_bytes_in_use(self.raw) }
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        // jobs and batches hold the context: the library frees it after the last of them
        unsafe { sys::bht_ctx_destroy(self.raw) }
    }
}

fn default_context() -> Arc<Context> {
    static DEFAULT: OnceLock<Arc<Context>> = OnceLock::new();
    DEFAULT.get_or_init(|| Context::new(0, 0)).clone()
}

type Callback<'cb> = Box<dyn FnMut(usize, &[f32]) + 'cb>;

unsafe extern "C" fn epoch_trampoline(epoch: u64, y: *const f32, len: u64, user: *mut c_void) {
    let cb = &mut *(user as *mut Callback<'_>);
    cb(epoch as usize, std::slice::from_raw_parts(y, len as usize));
}

/// t-SNE of `f32` samples with the Euclidean metric; `D` is the embedding dimension (2 for
/// Barnes-Hut, 1 or 2 for FIt-SNE).
pub struct TSne<'data, const D: usize = 2> {
    ctx: Arc<Context>,
    raw: *mut sys::bht_job,
    data: &'data [f32],
    d_in: usize,
    n: usize,
    params: sys::bht_params,
    callback: Option<Callback<'data>>,
    fitted: bool,
}

unsafe impl<'data, const D: usize> Send for TSne<'data, D> {}

impl<'data, const D: usize> TSne<'data, D> {
    /// `tSNE::new(data)` (src/lib.rs:241-266): `data` holds n samples of `d_in` features,
    /// row-major; the defaults are the reference's.
    pub fn new(data: &'data [f32], d_in: usize) -> Self {
        Self::with_context(default_context(), data, d_in)
    }

    pub fn with_context(ctx: Arc<Context>, data: &'data [f32], d_in: usize) -> Self {
        let n = if d_in == 0 { 0 } else { data.len() / d_in };
        assert_eq!(n * d_in, data.len(), "data length must be a multiple of d_in");
        let mut params = sys::bht_params::default();
        unsafe { sys::bht_params_default(&mut params) };
        params.d_out = D as u32;
        let mut raw = ptr::null_mut();
        check(unsafe { sys::bht_job_create(ctx.raw, n as u64, &params, &mut raw) });
        if n > 0 && d_in > 0 {
            check(unsafe { sys::bht_job_set_data_f32(raw, data.as_ptr(), n as u64, d_in as u32) });
        }
        TSne { ctx, raw, data, d_in, n, params, callback: None, fitted: false }
    }

    // ---- setters (src/lib.rs:277-431)
    pub fn learning_rate(&mut self, learning_rate: f32) -> &mut Self {
        self.params.learning_rate = learning_rate;
        self
    }
    pub fn epochs(&mut self, epochs: usize) -> &mut Self {
        self.params.epochs = epochs as u64;
        self
    }
    pub fn momentum(&mut self, momentum: f32) -> &mut Self {
        self.params.momentum = momentum;
        self
    }
    pub fn final_momentum(&mut self, final_momentum: f32) -> &mut Self {
        self.params.final_momentum = final_momentum;
        self
    }
    pub fn momentum_switch_epoch(&mut self, epoch: usize) -> &mut Self {
        self.params.momentum_switch_epoch = epoch as u64;
        self
    }
    pub fn stop_lying_epoch(&mut self, epoch: usize) -> &mut Self {
        self.params.stop_lying_epoch = epoch as u64;
        self
    }
    pub fn early_exaggeration(&mut self, early_exaggeration: f32) -> &mut Self {
        self.params.early_exaggeration = early_exaggeration;
        self
    }
    pub fn perplexity(&mut self, perplexity: f32) -> &mut Self {
        self.params.perplexity = perplexity;
        self
    }
    /// FIt-SNE grid: at least `min` and at most `max` boxes per axis (10 and 1000 by default).
    pub fn grid_boxes(&mut self, min: u32, max: u32) -> &mut Self {
        self.params.min_boxes = min;
        self.params.max_boxes = max;
        self
    }
    /// Called after every epoch with the recentred embedding (src/lib.rs:415-429).
    pub fn epoch_callback<C: FnMut(usize, &[f32]) + 'data>(&mut self, callback: C) -> &mut Self {
        self.callback = Some(Box::new(callback));
        self
    }
    /// The seed of the next fit, consumed by it (src/lib.rs:431-436).
    pub fn initial_embedding(&mut self, embedding: impl Into<Vec<f32>>) -> &mut Self {
        let y: Vec<f32> = embedding.into();
        check(unsafe { sys::bht_job_set_initial_embedding(self.raw, y.as_ptr(), y.len() as u64) });
        self
    }

    fn push_params(&mut self) {
        check(unsafe { sys::bht_job_set_params(self.raw, &self.params) });
    }

    fn cb_parts(&mut self) -> (sys::bht_epoch_cb, *mut c_void) {
        match self.callback.as_mut() {
            Some(cb) => (Some(epoch_trampoline), cb as *mut Callback<'data> as *mut c_void),
            None => (None, ptr::null_mut()),
        }
    }

    // ---- fits (src/lib.rs:701-894)
    /// `barnes_hut(theta, euclidean)`.
    pub fn barnes_hut_euclidean(&mut self, theta: f32) -> &mut Self {
        self.push_params();
        let (cb, user) = self.cb_parts();
        check(unsafe { sys::bht_job_barnes_hut(self.raw, theta, cb, user) });
        self.fitted = true;
        self
    }

    /// `barnes_hut_with_neighbors(theta, neighbors)`.
    pub fn barnes_hut_with_neighbors(&mut self, theta: f32, neighbors: &[Vec<Neighbor<f32>>])
        -> &mut Self {
        self.push_params();
        let (idx, dist, lens) = flatten(neighbors);
        let (cb, user) = self.cb_parts();
        check(unsafe {
            sys::bht_job_barnes_hut_with_neighbors(self.raw, theta, idx.as_ptr(), dist.as_ptr(),
                                                   lens.as_ptr(), lens.len() as u64, cb, user)
        });
        self.fitted = true;
        self
    }

    /// `exact(euclidean)`: the O(n^2) fit with the Euclidean distance.
    pub fn exact_euclidean(&mut self) -> &mut Self {
        self.exact_with(sys::BHT_METRIC_EUCLIDEAN)
    }

    /// `exact(squared euclidean)`, the distance the reference's exact tests pass.
    pub fn exact_squared_euclidean(&mut self) -> &mut Self {
        self.exact_with(sys::BHT_METRIC_SQEUCLIDEAN)
    }

    fn exact_with(&mut self, metric: u32) -> &mut Self {
        self.push_params();
        let (cb, user) = self.cb_parts();
        check(unsafe { sys::bht_job_exact(self.raw, metric, cb, user) });
        self.fitted = true;
        self
    }

    /// `fit_sne(euclidean)`.
    pub fn fit_sne_euclidean(&mut self) -> &mut Self {
        self.push_params();
        let (cb, user) = self.cb_parts();
        check(unsafe { sys::bht_job_fit_sne(self.raw, cb, user) });
        self.fitted = true;
        self
    }

    /// `fit_sne_with_neighbors(neighbors)`.
    pub fn fit_sne_with_neighbors(&mut self, neighbors: &[Vec<Neighbor<f32>>]) -> &mut Self {
        self.push_params();
        let (idx, dist, lens) = flatten(neighbors);
        let (cb, user) = self.cb_parts();
        check(unsafe {
            sys::bht_job_fit_sne_with_neighbors(self.raw, idx.as_ptr(), dist.as_ptr(),
                                                lens.as_ptr(), lens.len() as u64, cb, user)
        });
        self.fitted = true;
        self
    }

    // ---- results (src/lib.rs:438-477, 1019-1051)
    pub fn embedding(&self) -> Vec<f32> {
        let mut out = vec![0.0f32; self.n * D];
        check(unsafe { sys::bht_job_get_embedding(self.raw, out.as_mut_ptr(), out.len() as u64) });
        out
    }

    /// The reference's KL formula of the last fit (None before any fit).
    pub fn kl_divergence(&self) -> Option<f32> {
        self.kl_divergence_pair().map(|(r, _)| r as f32)
    }

    /// (reference formula, true KL) of the last fit.
    pub fn kl_divergence_pair(&self) -> Option<(f64, f64)> {
        if !self.fitted {
            return None;
        }
        let (mut r, mut t) = (0.0f64, 0.0f64);
        let st = unsafe { sys::bht_job_kl_divergence(self.raw, &mut r, &mut t) };
        if st == sys::BHT_ERR_STATE {
            return None;
        }
        check(st);
        Some((r, t))
    }

    pub fn affinities(&self) -> Option<SparseAffinities<f32>> {
        let (mut nnz, mut perp) = (0u64, 0.0f32);
        let st = unsafe {
            sys::bht_job_get_affinities(self.raw, ptr::null_mut(), ptr::null_mut(),
                                        ptr::null_mut(), &mut nnz, &mut perp)
        };
        if st == sys::BHT_ERR_STATE {
            return None;
        }
        check(st);
        let mut rows = vec![0u64; self.n + 1];
        let mut cols = vec![0u32; nnz as usize];
        let mut vals = vec![0.0f32; nnz as usize];
        check(unsafe {
            sys::bht_job_get_affinities(self.raw, rows.as_mut_ptr(), cols.as_mut_ptr(),
                                        vals.as_mut_ptr(), &mut nnz, &mut perp)
        });
        let rows = rows.into_iter().map(|r| r as usize).collect();
        Some(SparseAffinities::new(rows, cols, vals, perp))
    }

    pub fn with_affinities(&mut self, affinities: SparseAffinities<f32>) -> &mut Self {
        let rows: Vec<u64> = affinities.rows.iter().map(|&r| r as u64).collect();
        check(unsafe {
            sys::bht_job_set_affinities(self.raw, rows.as_ptr(), affinities.columns.as_ptr(),
                                        affinities.values.as_ptr(), rows.len() as u64,
                                        affinities.columns.len() as u64, affinities.perplexity)
        });
        self
    }

    /// The samples this builder was created over.
    pub fn data(&self) -> (&'data [f32], usize) {
        (self.data, self.d_in)
    }

    pub fn context(&self) -> &Arc<Context> {
        &self.ctx
    }
}

impl<'data, const D: usize> Drop for TSne<'data, D> {
    fn drop(&mut self) {
        unsafe { sys::bht_job_destroy(self.raw) }
    }
}

/// Neighbours as concatenated rows (index u64, distance f32, row lengths).
fn flatten(neighbors: &[Vec<Neighbor<f32>>]) -> (Vec<u64>, Vec<f32>, Vec<u64>) {
    let total: usize = neighbors.iter().map(|r| r.len()).sum();
    let mut idx = Vec::with_capacity(total);
    let mut dist = Vec::with_capacity(total);
    let mut lens = Vec::with_capacity(neighbors.len());
    for row in neighbors {
        lens.push(row.len() as u64);
        for nb in row {
            idx.push(nb.index as u64);
            dist.push(nb.distance);
        }
    }
    (idx, dist, lens)
}

/// Which fit a batch member runs.
#[derive(Clone, Copy, Debug)]
pub enum Method {
    BarnesHut { theta: f32 },
    FitSne,
}

/// Many independent jobs advanced together: every epoch kernel is one launch for all members of
/// a method and dimension, replayed from a captured graph.  Each member's result equals its fit
/// alone bit for bit.  No counterpart in the reference.
pub struct Batch<'data, const D: usize = 2> {
    raw: *mut sys::bht_batch,
    _ctx: Arc<Context>,
    members: HashMap<usize, TSne<'data, D>>,
}

impl<'data, const D: usize> Batch<'data, D> {
    pub fn new(ctx: Arc<Context>) -> Self {
        let mut raw = ptr::null_mut();
        check(unsafe { sys::bht_batch_create(ctx.raw, &mut raw) });
        Batch { raw, _ctx: ctx, members: HashMap::new() }
    }

    /// Builds the job's affinities (unless cached) and joins it; it runs its `epochs`.
    pub fn add(&mut self, mut job: TSne<'data, D>, method: Method) {
        job.push_params();
        let (m, theta) = match method {
            Method::BarnesHut { theta } => (sys::BHT_METHOD_BARNES_HUT, theta),
            Method::FitSne => (sys::BHT_METHOD_FITSNE, 0.5),
        };
        check(unsafe { sys::bht_batch_add(self.raw, job.raw, m, theta) });
        self.members.insert(job.raw as usize, job);
    }

    pub fn len(&self) -> usize {
        unsafe { sys::bht_batch_size(self.raw) as usize }
    }

    pub fn is_empty(&self) -> bool {
        self.len() == 0
    }

    /// Advances every member by up to `max_epochs`; returns the epochs advanced.
    pub fn step(&mut self, max_epochs: u64) -> u64 {
        let mut advanced = 0u64;
        check(unsafe { sys::bht_batch_step(self.raw, max_epochs, &mut advanced) });
        advanced
    }

    /// The members whose runs completed since the last call.
    pub fn take_finished(&mut self) -> Vec<TSne<'data, D>> {
        let mut out = Vec::new();
        let mut buf = [ptr::null_mut::<sys::bht_job>(); 64];
        loop {
            let mut count = 0u64;
            check(unsafe {
                sys::bht_batch_take_finished(self.raw, buf.as_mut_ptr(), buf.len() as u64,
                                             &mut count)
            });
            for &raw in &buf[..count as usize] {
                if let Some(mut job) = self.members.remove(&(raw as usize)) {
                    job.fitted = true;
                    out.push(job);
                }
            }
            if (count as usize) < buf.len() {
                return out;
            }
        }
    }

    /// Steps until every member has finished; the jobs in completion order.
    pub fn run(&mut self) -> Vec<TSne<'data, D>> {
        let mut done = Vec::new();
        while !self.is_empty() {
            self.step(u64::MAX);
            done.extend(self.take_finished());
        }
        done.extend(self.take_finished());
        done
    }
}

impl<'data, const D: usize> Drop for Batch<'data, D> {
    fn drop(&mut self) {
        // members still running are dropped by the library before their jobs are destroyed
        unsafe { sys::bht_batch_destroy(self.raw) }
    }
}
