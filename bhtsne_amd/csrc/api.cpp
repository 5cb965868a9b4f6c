// C ABI of the engine (include/bhtsne_hip.h): context, jobs and the fit driver.
//
// The driver mirrors the reference builder's control flow (src/lib.rs:652-1268): parameter
// validation with the same messages, the affinity cache rules, P normalisation / exaggeration /
// stop-lying, the momentum switch and the epoch loop; the numeric work runs in the HIP kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/bhtsne_hip.h"
#include "affinity.hpp"
#include "internal.hpp"
#include "reorder.hpp"

using namespace bht;

namespace {

thread_local std::string g_last_error;

bht_status fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

// Runs an entry point's body; a failure becomes a status and a message, and the runtime's
// last-error slot is cleared so that it is not reported again by a later, unrelated call.
template <typename F>
bht_status guarded(F &&f) {
    try {
        f();
        return BHT_OK;
    } catch (const Error &e) {
        (void)hipGetLastError();
        return fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        (void)hipGetLastError();
        return fail(BHT_ERR_OOM, "host allocation failed");
    } catch (const std::exception &e) {
        (void)hipGetLastError();
        return fail(BHT_ERR_HIP, e.what());
    }
}

int64_t next_smooth(int64_t n) {  // interpolation.rs:364-378
    int64_t c = n < 1 ? 1 : n;
    for (;;) {
        int64_t v = c;
        for (int p : {2, 3, 5, 7})
            while (v % p == 0) v /= p;
        if (v == 1) return c;
        c++;
    }
}

}  // namespace

// ---- context ---------------------------------------------------------------------------------

// Allocation accounting is shared by the jobs and batches of a context, which a request loop may
// drive from two host threads (one preparing arriving jobs, one stepping the batch).
struct TuneCache;

struct bht_ctx {
    int device = 0;
    uint64_t budget = 0;
    uint64_t in_use = 0;
    TwiddleSet tw;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::unique_ptr<TuneCache> tune;  // the autotuner's choices per job size class
    // Jobs and batches hold their context: destroying a context that still has them only marks
    // it, and the last of them frees it, so that destruction in any order is safe.
    int users = 0;
    bool closed = false;
    ~bht_ctx();

    // Freed blocks are kept for reuse instead of being returned with hipFree, which waits for
    // the whole device: a request loop frees a preparing job's temporaries while a batch runs,
    // and each free would stall on the batch's queued epochs.  Live bytes (what the budget and
    // the admission accounting see) are the requested sizes; cached blocks are trimmed when
    // they would exceed CACHE_LIMIT or, with a budget, live + cached bytes would exceed it.
    static constexpr uint64_t CACHE_LIMIT = 16ull << 30;
    std::multimap<size_t, void *> cached;         // capacity -> free block
    std::map<void *, size_t> capacity;            // live block -> capacity
    uint64_t cached_bytes = 0;

    void *alloc(size_t bytes) {
        if (bytes == 0) bytes = 16;
        {
            std::lock_guard<std::mutex> g(mu);
            if (budget && in_use + bytes > budget)
                throw Error(ERR_OOM, "HBM budget exceeded: " + std::to_string(in_use + bytes) +
                                         " bytes requested of " + std::to_string(budget));
            in_use += bytes;
            // the smallest cached block that fits without wasting more than its request
            auto it = cached.lower_bound(bytes);
            if (it != cached.end() && it->first <= 2 * bytes + (1u << 20)) {
                void *p = it->second;
                capacity[p] = it->first;
                cached_bytes -= it->first;
                cached.erase(it);
                return p;
            }
        }
        void *p = nullptr;
        hipError_t e = hipSetDevice(device);
        if (e == hipSuccess) e = hipMalloc(&p, bytes);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            trim(0);
            e = hipMalloc(&p, bytes);
        }
        std::lock_guard<std::mutex> g(mu);
        if (e != hipSuccess) {
            in_use -= bytes;
            BHT_HIP(e);
        }
        capacity[p] = bytes;
        return p;
    }
    void release(void *p, size_t bytes) {
        if (!p) return;
        bool over;
        {
            std::lock_guard<std::mutex> g(mu);
            in_use -= bytes ? bytes : 16;
            auto it = capacity.find(p);
            const size_t cap = it != capacity.end() ? it->second : (bytes ? bytes : 16);
            if (it != capacity.end()) capacity.erase(it);
            cached.emplace(cap, p);
            cached_bytes += cap;
            over = cached_bytes > CACHE_LIMIT || (budget && in_use + cached_bytes > budget);
        }
        if (over) trim(budget ? budget - std::min(budget, used()) : CACHE_LIMIT / 2);
    }
    // Returns cached blocks to the device (largest first) until at most `keep` bytes stay cached.
    void trim(uint64_t keep) {
        std::vector<void *> out;
        {
            std::lock_guard<std::mutex> g(mu);
            while (cached_bytes > keep && !cached.empty()) {
                auto it = std::prev(cached.end());
                cached_bytes -= it->first;
                out.push_back(it->second);
                cached.erase(it);
            }
        }
        for (void *p : out) hipFree(p);
    }
    uint64_t used() {
        std::lock_guard<std::mutex> g(mu);
        return in_use;
    }
};

namespace {

// Device buffer owned by a context.
struct DevBuf {
    bht_ctx *ctx = nullptr;
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(bht_ctx *c) : ctx(c) {}
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { reset(); }
    void reset() {
        if (p && ctx) ctx->release(p, bytes);
        p = nullptr;
        bytes = 0;
    }
    void ensure(bht_ctx *c, size_t b) {
        ctx = c;
        if (p && bytes >= b) return;
        reset();
        p = ctx->alloc(b);
        bytes = b ? b : 16;
    }
    template <typename T>
    T *as() const { return reinterpret_cast<T *>(p); }
};

void build_twiddles(bht_ctx *ctx) {
    std::map<int, int> off_of_len;
    std::vector<float2> table;
    std::vector<int> len_of(MAX_BOXES + 1, 0), off_of(MAX_BOXES + 1, 0);
    for (int b = 1; b <= MAX_BOXES; b++) {
        int L = (int)next_smooth(6 * (int64_t)b - 1);
        auto it = off_of_len.find(L);
        if (it == off_of_len.end()) {
            int off = (int)table.size();
            off_of_len[L] = off;
            for (int j = 0; j < L; j++) {
                // exp(-2 pi i j / L) from an exactly reduced angle in double
                long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)j /
                                (long double)L;
                table.push_back(make_float2((float)cosl(a), (float)sinl(a)));
            }
            it = off_of_len.find(L);
        }
        len_of[b] = L;
        off_of[b] = it->second;
    }
    ctx->tw.total = table.size();
    ctx->tw.table = (float2 *)ctx->alloc(table.size() * sizeof(float2));
    ctx->tw.len_of_boxes = (int *)ctx->alloc(len_of.size() * sizeof(int));
    ctx->tw.off_of_boxes = (int *)ctx->alloc(off_of.size() * sizeof(int));
    BHT_HIP(hipMemcpy(ctx->tw.table, table.data(), table.size() * sizeof(float2),
                      hipMemcpyHostToDevice));
    BHT_HIP(hipMemcpy(ctx->tw.len_of_boxes, len_of.data(), len_of.size() * sizeof(int),
                      hipMemcpyHostToDevice));
    BHT_HIP(hipMemcpy(ctx->tw.off_of_boxes, off_of.data(), off_of.size() * sizeof(int),
                      hipMemcpyHostToDevice));
}

const char *MSG_PERPLEXITY =
    "error: the provided perplexity is too large for the number of data points.\n";
const char *MSG_THETA =
    "error: theta value must be greater than 0.0.\n"
    "            A value of 0.0 corresponds to using the exact version of the algorithm.";

}  // namespace

// ---- job -------------------------------------------------------------------------------------

// Events that order work between streams (or let the host wait for it) and carry no data: the
// data a stream hands to another goes through the cross-XCD cache maintenance of l2_sync, and what
// the host reads is written with system-scope stores.  They skip the system-scope fence (an L2
// write-back and invalidate) that an event otherwise performs when it completes.
constexpr unsigned HANDOVER_EVENT = hipEventDisableTiming | hipEventDisableSystemFence;

enum FitKind { FIT_NONE = -1, FIT_BH = 0, FIT_FITSNE = 1, FIT_EXACT = 2 };

struct bht_job {
    bht_ctx *ctx;
    uint64_t n;
    bht_params params;
    hipStream_t stream = nullptr;

    // sample vectors (Euclidean)
    DevBuf x;
    uint32_t d_in = 0;
    bool have_data = false;

    // affinities: CSR with u32 offsets on the device, and its column-tiled copy (the attraction's
    // layout: edges, values, block / segment / group tables)
    DevBuf rows, cols, vals;
    DevBuf tkey, tval, tblk, tseg_tile, tseg_e0, tseg_g0, tgrp, tgrp_row;
    uint64_t nnz = 0;
    uint64_t tile_nnz = 0;  // edges of the tiled copy, each group padded to a multiple of 8
    bool have_p = false;
    bool have_cached_perplexity = false;
    float cached_perplexity = 0.0f;
    bool stop_lying_fired = false;

    // optimiser state; y[cur] is the embedding of the next epoch (cur = epoch parity, which the
    // kernels read from the device epoch counter)
    DevBuf y[2], uy, gains, frep, fattr, pos;
    int cur = 0;
    bool have_init = false;
    std::vector<float> init;
    uint64_t seed_counter = 0;  // random initialisations drawn by this job
    uint64_t serial = 0;        // creation index of the job in the process

    // epoch workspace
    DevBuf keys, svals, keys_alt, svals_alt, grid, spec, rker, kspec, part, dpart, mpart, fpart,
        sort_temp, scal;
    int nodes_cap = 0, len_cap = 0;
    // A large 2-D FIt-SNE job sizes its FFT launches' LDS for the lengths its grid has reached
    // (reported by the grid plan into host_len, pinned host memory), so that they fit a CU beside
    // the attraction; lines that outgrow a launch use fft_fb.  The host keeps at most
    // 16 epochs queued so that the reported length is recent.
    int *host_len = nullptr;
    DevBuf fft_fb;
    // the box sort's final buffers (2-D jobs sorted in 7-bit passes, whose pass count the grid
    // plan decides on the device)
    DevBuf keys_fin, svals_fin;
    // the device holds per-block bounds of the current embedding buffer (from the last update or
    // centring), so that the next FIt-SNE epoch can leave the recentring to its readers
    bool lazy_ready = false;
    // the exact fit: dense P (n x n) and per-row sums (Z, KL)
    DevBuf pdense, zrow;
    hipEvent_t ahead[4] = {};
    bool ahead_rec[4] = {};
    // Barnes-Hut workspace
    DevBuf codes, codes_alt, bidx, bidx_alt, bsort_temp, bnode, bmeta, brange, bhw2, bcount;
    DevBuf bclo, bchi, bslo;  // two-word Morton codes (D >= 5): low / high words, sorted lows
    DevBuf bdefer;            // D = 2: roots of the subtrees built by several workgroups
    // the job's entry of the device job table (host copy + device copy)
    JobDev hd{};
    DevBuf jd;
    bool in_batch = false;

    // point renumbering: perm[i] = sample index of the engine's point i (identity until the
    // first reorder); qv[s] = P value at position s of the reference's CSR layout (s < n)
    DevBuf perm, perm_tmp, iperm, rows2, cols2, vals2, qv, ptmp, lentmp, pscan;
    bool permuted = false;
    uint64_t reorders = 0;

    uint64_t knn_exact_rows = 0;  // rows of the last kNN recomputed by the exact kernel

    // run state
    int method = -1;
    float theta = 0.5f;
    bool prepared = false;
    uint64_t epoch = 0;          // next epoch index of the prepared run
    float run_momentum = 0.5f;   // momentum at the start of the prepared run
    int fit_kind = FIT_NONE;
    float fit_theta = 0.5f;

    // An epoch forks after the recentring: the repulsion runs on `side` while the attraction runs
    // on `stream`, and the update on `stream` waits for both (fork / join events).
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    // per-phase timing with events (4 per timed epoch: fork, repulsion done, attraction done,
    // update done)
    int last_choice = -1;     // the last run's epoch form: attraction queue blocks, -1 in order
    bool timing = false;
    bool timed_fork = false;  // the timed epochs forked (else every phase ran in order)
    std::vector<hipEvent_t> events;
    size_t events_used = 0;

    explicit bht_job(bht_ctx *c, uint64_t n_, const bht_params &p) : ctx(c), n(n_), params(p) {
        BHT_HIP(hipSetDevice(ctx->device));
        BHT_HIP(hipStreamCreate(&stream));
        BHT_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        BHT_HIP(hipEventCreateWithFlags(&ev_fork, HANDOVER_EVENT));
        BHT_HIP(hipEventCreateWithFlags(&ev_join, HANDOVER_EVENT));
        for (auto &e : ahead) BHT_HIP(hipEventCreateWithFlags(&e, HANDOVER_EVENT));
    }
    ~bht_job() {
        for (auto e : events) hipEventDestroy(e);
        for (auto e : ahead)
            if (e) hipEventDestroy(e);
        if (stream) hipStreamSynchronize(stream);
        if (host_len) hipHostFree(host_len);
        if (ev_fork) hipEventDestroy(ev_fork);
        if (ev_join) hipEventDestroy(ev_join);
        if (side) hipStreamDestroy(side);
        if (stream) hipStreamDestroy(stream);
    }
    // Records timing event `slot` (0: epoch start after the recentring, 1: repulsion done,
    // 2: attraction done, 3: update done) of the epoch being enqueued.
    void mark(hipStream_t s, int slot) {
        if (!timing) return;
        while (events.size() < events_used + 4) {
            hipEvent_t e;
            BHT_HIP(hipEventCreate(&e));
            events.push_back(e);
        }
        BHT_HIP(hipEventRecord(events[events_used + slot], s));
    }
    uint32_t dim() const { return params.d_out; }
    float *ycur() const { return y[cur].as<float>(); }
};

namespace {

void check_perplexity(const bht_job *j) {  // src/tsne/mod.rs:55-59
    float perp = j->params.perplexity;
    uint64_t p = perp > 0.0f ? (uint64_t)perp : 0;
    if (j->n == 0 || j->n - 1 < 3 * p) throw Error(ERR_INVALID, MSG_PERPLEXITY);
}

uint32_t n_neighbors_of(const bht_job *j) {  // lib.rs:721, 852
    float k = 3.0f * j->params.perplexity;
    return k > 0.0f ? (uint32_t)k : 0u;
}

void ensure_state_buffers(bht_job *j) {
    size_t b = (size_t)j->n * j->dim() * sizeof(float);
    j->y[0].ensure(j->ctx, b);
    j->y[1].ensure(j->ctx, b);
    j->uy.ensure(j->ctx, b);
    j->gains.ensure(j->ctx, b);
    j->frep.ensure(j->ctx, b);
    j->fattr.ensure(j->ctx, b);
    j->pos.ensure(j->ctx, b);
    j->scal.ensure(j->ctx, sizeof(EpochScalars));
    j->dpart.ensure(j->ctx, 4096 * sizeof(double));
    // update partials (the recentring sums, one per axis and block) and per-block bounds (min and
    // max per axis, at least 4 floats per block)
    const size_t axes = std::max<size_t>(2, j->dim());
    j->mpart.ensure(j->ctx, axes * (size_t)MAX_UPDATE_PARTS * sizeof(double));
    j->fpart.ensure(j->ctx, 2 * axes * (size_t)MAX_UPDATE_PARTS * sizeof(float));
}

// Jobs from this size on run their epochs forked and size their FFT launches dynamically.
constexpr uint64_t DYN_FFT_MIN_POINTS = 32768;

// The global-memory FFT lines and the grid-length report of a 2-D FIt-SNE job: with them its FFT
// launches may hold less LDS than its grid capacity needs (lines that outgrow a launch take the
// global lines, with the same arithmetic).
void ensure_fft_lines(bht_job *j) {
    if (j->dim() != 2) return;
    j->fft_fb.ensure(j->ctx, (size_t)FFT_FB_BLOCKS * 2 * j->len_cap * sizeof(float2));
    if (!j->host_len) {
        BHT_HIP(hipHostMalloc((void **)&j->host_len, 64, hipHostMallocDefault));
        *j->host_len = 0;
    }
    j->hd.fft_fb = j->fft_fb.as<float2>();
    j->hd.host_len = j->host_len;
}

void ensure_fitsne_workspace(bht_job *j) {
    ensure_state_buffers(j);
    bht_ctx *c = j->ctx;
    int maxb = (int)std::min<uint32_t>(std::max<uint32_t>(j->params.max_boxes, 1), MAX_BOXES);
    int minb = (int)std::max<uint32_t>(j->params.min_boxes, 1);
    if (minb > maxb) maxb = minb > MAX_BOXES ? MAX_BOXES : minb;
    j->nodes_cap = 3 * maxb;
    j->len_cap = (int)next_smooth(6 * (int64_t)maxb - 1);
    const uint64_t n = j->n;
    j->keys.ensure(c, n * 4);
    j->svals.ensure(c, n * 4);
    j->keys_alt.ensure(c, n * 4);
    j->svals_alt.ensure(c, n * 4);
    j->sort_temp.ensure(c, group_sort_temp_bytes(n));
    const size_t nodes_total = j->dim() == 2 ? (size_t)j->nodes_cap * j->nodes_cap : j->nodes_cap;
    j->grid.ensure(c, nodes_total * sizeof(float4));
    if (j->dim() == 2) j->spec.ensure(c, 2 * (size_t)j->nodes_cap * j->len_cap * sizeof(float2));
    const size_t h = j->len_cap / 2 + 1;
    j->rker.ensure(c, h * h * sizeof(float));
    j->kspec.ensure(c, h * h * sizeof(float));
    const size_t nchunks = (n + 63) / 64;
    j->part.ensure(c, 2 * nchunks * 36 * sizeof(float));
    JobDev &d = j->hd;
    d.min_boxes = minb > MAX_BOXES ? MAX_BOXES : minb;
    d.max_boxes = maxb;
    d.nodes_cap = j->nodes_cap;
    d.len_cap = j->len_cap;
    d.grid = j->grid.as<float4>();
    d.spec = j->spec.as<float2>();
    d.rker = j->rker.as<float>();
    d.kspec = j->kspec.as<float>();
    d.part = j->part.as<float>();
    if (j->dim() == 2 && n >= DYN_FFT_MIN_POINTS) {
        ensure_fft_lines(j);
        if (n >= SORT7_MIN_KEYS) {
            j->keys_fin.ensure(c, n * 4);
            j->svals_fin.ensure(c, n * 4);
        }
    } else {
        d.fft_fb = nullptr;
        d.host_len = nullptr;
    }
}

void ensure_bh_workspace(bht_job *j) {
    ensure_state_buffers(j);
    bht_ctx *c = j->ctx;
    const uint64_t n = std::max<uint64_t>(j->n, 1);
    j->codes.ensure(c, n * 8);
    j->codes_alt.ensure(c, n * 8);
    j->bidx.ensure(c, n * 4);
    j->bidx_alt.ensure(c, n * 4);
    j->bsort_temp.ensure(c, group_sort_temp_bytes(n));
    // D = 2: the top nodes in [0, 2n), each deferred subtree in a region of its own in [2n, 4n)
    const size_t node_bytes = j->dim() >= 3 ? (size_t)nd_node_floats((int)j->dim()) * 4 : 16;
    const size_t nodes = (j->dim() == 2 ? 4 : 2) * n;
    j->bnode.ensure(c, nodes * node_bytes);
    j->bmeta.ensure(c, nodes * 8);
    if (j->dim() == 2) j->bdefer.ensure(c, n * 4);
    if (j->dim() >= 5) {
        j->bclo.ensure(c, n * 8);
        j->bchi.ensure(c, n * 8);
        j->bslo.ensure(c, n * 8);
    }
    j->brange.ensure(c, 2 * nodes * 4);
    j->bhw2.ensure(c, 64 * 4);
    j->bcount.ensure(c, 128 * 4);
    JobDev &d = j->hd;
    d.nodes.node = j->bnode.as<float4>();
    d.nodes.meta = j->bmeta.as<uint2>();
    d.nodes.range = j->brange.as<uint32_t>();
    d.nodes.hw2 = j->bhw2.as<float>();
    d.nodes.counts = j->bcount.as<uint32_t>();
    d.nodes.defer = j->bdefer.as<uint32_t>();
}

float resolve_learning_rate(const bht_job *j) {  // lib.rs:652-659
    if (j->params.learning_rate > 0.0f) return j->params.learning_rate;
    float a = (float)j->n / j->params.early_exaggeration / 4.0f;
    return std::max(a, 50.0f);
}

// Fills the job's table entry from its buffers and parameters.  A job alone sorts in its own
// buffers; a batch points the sort fields at its arena afterwards.
// Whether the job's box sort takes its pass count from the device's grid plan (a 2-D FIt-SNE job
// large enough for 7-bit passes and with the final buffers; batches sort with fixed passes).
bool device_sort_passes(const bht_job *j) {
    return j->method == FIT_FITSNE && j->dim() == 2 && j->n >= SORT7_MIN_KEYS &&
           j->keys_fin.p != nullptr;
}

void fill_jobdev(bht_job *j) {
    JobDev &d = j->hd;
    d.n = (int)j->n;
    d.dim = (int)j->dim();
    d.method = j->method == FIT_BH ? M_BH : j->method == FIT_EXACT ? M_EXACT : M_FITSNE;
    d.theta_sq = j->theta * j->theta;
    d.ybuf[0] = j->y[0].as<float>();
    d.ybuf[1] = j->y[1].as<float>();
    d.uy = j->uy.as<float>();
    d.gains = j->gains.as<float>();
    d.frep = j->frep.as<float>();
    d.fattr = j->fattr.as<float>();
    d.rows = j->rows.as<uint32_t>();
    d.cols = j->cols.as<uint32_t>();
    d.vals = j->vals.as<float>();
    d.tile_rows = tiled_rows_per_block((int)j->n);
    d.tile_points = tiled_tile_points((int)j->dim());
    d.tile_nnz = (uint32_t)j->tile_nnz;
    d.tkey = j->tkey.as<uint16_t>();
    d.tval = j->tval.as<float>();
    d.blk_seg = j->tblk.as<uint32_t>();
    d.seg_tile = j->tseg_tile.as<uint32_t>();
    d.seg_e0 = j->tseg_e0.as<uint32_t>();
    d.seg_g0 = j->tseg_g0.as<uint32_t>();
    d.grp_e0 = j->tgrp.as<uint32_t>();
    d.grp_row = j->tgrp_row.as<uint32_t>();
    d.qv = j->qv.as<float>();
    d.qv_n = (int)std::min<uint64_t>(j->n, j->nnz);
    d.scal = j->scal.as<EpochScalars>();
    d.dpart = j->dpart.as<double>();
    d.mpart = j->mpart.as<double>();
    d.fpart = j->fpart.as<float>();
    d.learning_rate = resolve_learning_rate(j);
    d.momentum = j->run_momentum;
    d.final_momentum = j->params.final_momentum;
    d.inv_exag = 1.0f / j->params.early_exaggeration;
    d.momentum_switch_epoch =
        (int)std::min<uint64_t>(j->params.momentum_switch_epoch, (uint64_t)INT32_MAX - 2);
    d.stop_lying_epoch =
        (int)std::min<uint64_t>(j->params.stop_lying_epoch, (uint64_t)INT32_MAX - 2);
    if (d.method == M_FITSNE) {
        const bool alt = sort_result_alt(box_key_bits(d.dim), j->n >= SORT7_MIN_KEYS);
        d.keys = j->keys.as<uint32_t>();
        d.kidx = j->svals.as<uint32_t>();
        d.skeys = alt ? j->keys_alt.as<uint32_t>() : j->keys.as<uint32_t>();
        d.sidx = alt ? j->svals_alt.as<uint32_t>() : j->svals.as<uint32_t>();
        if (device_sort_passes(j)) {
            d.skeys = j->keys_fin.as<uint32_t>();
            d.sidx = j->svals_fin.as<uint32_t>();
        }
    } else {
        const bool alt = sort_result_alt(MORTON_KEY_BITS, false);
        d.codes = j->codes.as<uint64_t>();
        d.cidx = j->bidx.as<uint32_t>();
        d.scodes = alt ? j->codes_alt.as<uint64_t>() : j->codes.as<uint64_t>();
        d.scidx = alt ? j->bidx_alt.as<uint32_t>() : j->bidx.as<uint32_t>();
        d.clo = j->bclo.as<uint64_t>();
        d.chi = j->bchi.as<uint64_t>();
        d.slo = j->bslo.as<uint64_t>();
    }
    job_block_counts(d, false);
}

// Column-tiled copy of the current P (after any change of the CSR or of its values before a run:
// build, injection, normalisation, renumbering).
void refresh_tiled(bht_job *j) {
    bht_ctx *c = j->ctx;
    const int n = (int)j->n, D = (int)j->dim();
    if (D >= 3) {
        // the attraction of D = 3..7 reads the CSR rows; tval is its copy of the values (rescaled
        // in place at stop-lying)
        j->tval.ensure(c, std::max<uint64_t>(j->nnz, 1) * 4);
        if (j->nnz)
            BHT_HIP(hipMemcpyAsync(j->tval.p, j->vals.p, j->nnz * 4, hipMemcpyDeviceToDevice,
                                   j->stream));
        j->tile_nnz = j->nnz;
        return;
    }
    const TiledSizes z = tiled_sizes(n, D, j->nnz);
    j->tkey.ensure(c, z.padded_cap * 2);
    j->tval.ensure(c, z.padded_cap * 4);
    j->tblk.ensure(c, ((size_t)tiled_blocks(n) + 1) * 4);
    j->tseg_tile.ensure(c, z.max_segs * 4);
    j->tseg_e0.ensure(c, z.max_segs * 4);
    j->tseg_g0.ensure(c, z.max_segs * 4);
    j->tgrp.ensure(c, (z.max_groups + 1) * 4);
    j->tgrp_row.ensure(c, (z.max_groups + 1) * 4);
    DevBuf temp(c);
    temp.ensure(c, z.temp);
    TiledOut o{j->tkey.as<uint16_t>(), j->tval.as<float>(), j->tblk.as<uint32_t>(),
               j->tseg_tile.as<uint32_t>(), j->tseg_e0.as<uint32_t>(), j->tseg_g0.as<uint32_t>(),
               j->tgrp.as<uint32_t>(),     j->tgrp_row.as<uint32_t>()};
    j->tile_nnz = tiled_build(j->rows.as<uint32_t>(), j->cols.as<uint32_t>(),
                              j->vals.as<float>(), n, D, j->nnz, o, temp.p, j->stream);
}

// Refreshes the job's own device table entry (single-job launches); with `layout`, also the
// tiled copy of P.
void sync_jobdev(bht_job *j, bool layout = true) {
    if (layout) refresh_tiled(j);
    fill_jobdev(j);
    j->jd.ensure(j->ctx, sizeof(JobDev));
    BHT_HIP(hipMemcpyAsync(j->jd.p, &j->hd, sizeof(JobDev), hipMemcpyHostToDevice, j->stream));
    BHT_HIP(hipStreamSynchronize(j->stream));
}

// Launch description of the job alone.
LaunchSet single_set(bht_job *j) {
    LaunchSet L;
    L.jobs = j->jd.as<JobDev>();
    for (int f = 0; f < NFAM; f++) L.nb[f] = j->hd.nblk[f];
    L.dim = (int)j->dim();
    L.njobs = 1;
    L.tw = &j->ctx->tw;
    L.fft_cap = j->len_cap;
    if (j->method == FIT_BH) {
        L.keys = j->codes.p;
        L.keys_alt = j->codes_alt.p;
        L.vals = j->bidx.as<uint32_t>();
        L.vals_alt = j->bidx_alt.as<uint32_t>();
        L.sort_temp = j->bsort_temp.p;
    } else {
        L.keys = j->keys.p;
        L.keys_alt = j->keys_alt.p;
        L.vals = j->svals.as<uint32_t>();
        L.vals_alt = j->svals_alt.as<uint32_t>();
        L.sort_temp = j->sort_temp.p;
        if (device_sort_passes(j)) {
            L.sort_npass = &j->scal.as<EpochScalars>()->sort_passes;
            L.keys_fin = j->keys_fin.p;
            L.vals_fin = j->svals_fin.as<uint32_t>();
        }
    }
    L.sort_n = j->n;
    return L;
}

// Applies the recentring left pending by the last update (zero_mean, mod.rs:534-568) and writes
// the per-axis bounds.
void recenter(bht_job *j) {
    center_plan(single_set(j), false, j->stream);
    j->lazy_ready = false;
}

// ---- point order -----------------------------------------------------------------------------

// A freshly built or injected CSR is in the reference's layout: identity order, and the values
// at positions [0, n) are what kl_divergence weights rows by (mod.rs:768-770).
void reset_order(bht_job *j) {
    j->permuted = false;
    const uint64_t n = std::max<uint64_t>(j->n, 1);
    j->qv.ensure(j->ctx, n * 4);
    BHT_HIP(hipMemsetAsync(j->qv.p, 0, n * 4, j->stream));
    const uint64_t m = std::min<uint64_t>(j->n, j->nnz);
    if (m) BHT_HIP(hipMemcpyAsync(j->qv.p, j->vals.p, m * 4, hipMemcpyDeviceToDevice, j->stream));
}

void ensure_order_buffers(bht_job *j) {
    bht_ctx *c = j->ctx;
    const uint64_t n = std::max<uint64_t>(j->n, 1);
    j->perm.ensure(c, n * 4);
    j->perm_tmp.ensure(c, n * 4);
    j->iperm.ensure(c, n * 4);
    j->lentmp.ensure(c, n * 4);
    j->pscan.ensure(c, scan_temp_bytes(n));
    j->ptmp.ensure(c, n * j->dim() * 4);
}

// host (sample order) -> device (engine order)
void upload_points(bht_job *j, const float *host, float *dev) {
    const uint64_t len = j->n * j->dim();
    if (!j->permuted) {
        BHT_HIP(hipMemcpyAsync(dev, host, len * 4, hipMemcpyHostToDevice, j->stream));
        return;
    }
    ensure_order_buffers(j);
    BHT_HIP(hipMemcpyAsync(j->ptmp.p, host, len * 4, hipMemcpyHostToDevice, j->stream));
    gather_rows(j->ptmp.as<float>(), j->perm.as<uint32_t>(), (int)j->n, (int)j->dim(), dev,
                j->stream);
}

// device (engine order) -> host (sample order); synchronises the job's stream
void download_points(bht_job *j, const float *dev, float *host) {
    const uint64_t len = j->n * j->dim();
    const float *src = dev;
    if (j->permuted) {
        ensure_order_buffers(j);
        scatter_rows(dev, j->perm.as<uint32_t>(), (int)j->n, (int)j->dim(), j->ptmp.as<float>(),
                     j->stream);
        src = j->ptmp.as<float>();
    }
    BHT_HIP(hipMemcpyAsync(host, src, len * 4, hipMemcpyDeviceToHost, j->stream));
    BHT_HIP(hipStreamSynchronize(j->stream));
}

// CSR in the reference's layout (sample order, original row order) on the host
void download_csr(bht_job *j, std::vector<uint32_t> &rows, std::vector<uint32_t> &cols,
                  std::vector<float> *vals) {
    rows.resize(j->n + 1);
    cols.resize(j->nnz);
    if (vals) vals->resize(j->nnz);
    const uint32_t *r = j->rows.as<uint32_t>();
    const uint32_t *c = j->cols.as<uint32_t>();
    const float *v = j->vals.as<float>();
    if (j->permuted) {
        ensure_order_buffers(j);
        bht_ctx *cx = j->ctx;
        j->rows2.ensure(cx, (j->n + 1) * 4);
        j->cols2.ensure(cx, std::max<uint64_t>(j->nnz, 1) * 4);
        j->vals2.ensure(cx, std::max<uint64_t>(j->nnz, 1) * 4);
        // sample s is engine point iperm[s]; engine label c is sample perm[c]
        inverse_perm(j->perm.as<uint32_t>(), (int)j->n, j->iperm.as<uint32_t>(), j->stream);
        csr_permute(r, c, v, j->iperm.as<uint32_t>(), j->perm.as<uint32_t>(), (int)j->n,
                    j->lentmp.as<uint32_t>(), j->pscan.p, j->rows2.as<uint32_t>(),
                    j->cols2.as<uint32_t>(), j->vals2.as<float>(), j->stream);
        r = j->rows2.as<uint32_t>();
        c = j->cols2.as<uint32_t>();
        v = j->vals2.as<float>();
    }
    BHT_HIP(hipMemcpyAsync(rows.data(), r, (j->n + 1) * 4, hipMemcpyDeviceToHost, j->stream));
    if (j->nnz) {
        BHT_HIP(hipMemcpyAsync(cols.data(), c, j->nnz * 4, hipMemcpyDeviceToHost, j->stream));
        if (vals)
            BHT_HIP(hipMemcpyAsync(vals->data(), v, j->nnz * 4, hipMemcpyDeviceToHost, j->stream));
    }
    BHT_HIP(hipStreamSynchronize(j->stream));
}

// Renumber the points in `order` (new point i = current point order[i]).
void reorder_points(bht_job *j, const uint32_t *order) {
    ensure_order_buffers(j);
    bht_ctx *c = j->ctx;
    hipStream_t s = j->stream;
    const int n = (int)j->n, D = (int)j->dim();
    if (!j->permuted) {
        std::vector<uint32_t> id(j->n);
        for (uint64_t i = 0; i < j->n; i++) id[i] = (uint32_t)i;
        BHT_HIP(hipMemcpyAsync(j->perm.p, id.data(), j->n * 4, hipMemcpyHostToDevice, s));
        BHT_HIP(hipStreamSynchronize(s));
    }
    j->rows2.ensure(c, (j->n + 1) * 4);
    j->cols2.ensure(c, std::max<uint64_t>(j->nnz, 1) * 4);
    j->vals2.ensure(c, std::max<uint64_t>(j->nnz, 1) * 4);
    inverse_perm(order, n, j->iperm.as<uint32_t>(), s);
    csr_permute(j->rows.as<uint32_t>(), j->cols.as<uint32_t>(), j->vals.as<float>(), order,
                j->iperm.as<uint32_t>(), n, j->lentmp.as<uint32_t>(), j->pscan.p,
                j->rows2.as<uint32_t>(), j->cols2.as<uint32_t>(), j->vals2.as<float>(), s);
    std::swap(j->rows.p, j->rows2.p);
    std::swap(j->rows.bytes, j->rows2.bytes);
    std::swap(j->cols.p, j->cols2.p);
    std::swap(j->cols.bytes, j->cols2.bytes);
    std::swap(j->vals.p, j->vals2.p);
    std::swap(j->vals.bytes, j->vals2.bytes);
    // per-point state: embedding through the idle ping-pong buffer (the buffers then trade
    // places, so the epoch parity still selects the current one), uy and gains through ptmp
    gather_rows(j->ycur(), order, n, D, j->y[1 - j->cur].as<float>(), s);
    std::swap(j->y[0].p, j->y[1].p);
    std::swap(j->y[0].bytes, j->y[1].bytes);
    gather_rows(j->uy.as<float>(), order, n, D, j->ptmp.as<float>(), s);
    std::swap(j->uy.p, j->ptmp.p);
    std::swap(j->uy.bytes, j->ptmp.bytes);
    j->ptmp.ensure(c, j->n * D * 4);
    gather_rows(j->gains.as<float>(), order, n, D, j->ptmp.as<float>(), s);
    std::swap(j->gains.p, j->ptmp.p);
    std::swap(j->gains.bytes, j->ptmp.bytes);
    gather_u32(j->perm.as<uint32_t>(), order, n, j->perm_tmp.as<uint32_t>(), s);
    std::swap(j->perm.p, j->perm_tmp.p);
    std::swap(j->perm.bytes, j->perm_tmp.bytes);
    j->permuted = true;
    j->reorders++;
    sync_jobdev(j);
    // the previous CSR is not read again (an affinity export re-creates these buffers)
    j->rows2.reset();
    j->cols2.reset();
    j->vals2.reset();
}

// Structure of an injected CSR (the reference takes it unchecked and would panic on the first
// out-of-bounds index; here a bad offset or column would be a device memory fault).
void check_csr(const uint64_t *rows, const uint32_t *cols, uint64_t n_rows_plus_one,
               uint64_t nnz) {
    if (n_rows_plus_one == 0) throw Error(ERR_INVALID, "rows must hold n + 1 offsets");
    if (nnz >= (1ull << 32)) throw Error(ERR_INVALID, "nnz must be below 2^32");
    if (rows[0] != 0) throw Error(ERR_INVALID, "rows[0] must be 0");
    if (rows[n_rows_plus_one - 1] != nnz) throw Error(ERR_INVALID, "rows[n] must equal nnz");
    for (uint64_t i = 1; i < n_rows_plus_one; i++)
        if (rows[i] < rows[i - 1])
            throw Error(ERR_INVALID,
                        "rows must be non-decreasing (row " + std::to_string(i - 1) + ")");
    const uint64_t n = n_rows_plus_one - 1;
    for (uint64_t e = 0; e < nnz; e++)
        if (cols[e] >= n)
            throw Error(ERR_INVALID, "column " + std::to_string(cols[e]) + " at entry " +
                                         std::to_string(e) + " is out of range (n = " +
                                         std::to_string(n) + ")");
}

// Back to sample order: the current embedding, uy and gains (engine point i is sample perm[i]).
// The CSR is left to the caller, who replaces it.
void unpermute_state(bht_job *j) {
    if (!j->permuted) return;
    ensure_order_buffers(j);
    hipStream_t s = j->stream;
    const int n = (int)j->n, D = (int)j->dim();
    if (j->y[j->cur].p) {
        scatter_rows(j->ycur(), j->perm.as<uint32_t>(), n, D, j->y[1 - j->cur].as<float>(), s);
        std::swap(j->y[0].p, j->y[1].p);
        std::swap(j->y[0].bytes, j->y[1].bytes);
    }
    for (DevBuf *b : {&j->uy, &j->gains}) {
        if (!b->p) continue;
        j->ptmp.ensure(j->ctx, j->n * D * 4);
        scatter_rows(b->as<float>(), j->perm.as<uint32_t>(), n, D, j->ptmp.as<float>(), s);
        std::swap(b->p, j->ptmp.p);
        std::swap(b->bytes, j->ptmp.bytes);
    }
    BHT_HIP(hipStreamSynchronize(s));
    j->permuted = false;
}

// Large jobs (y outgrows an XCD's L2) are renumbered once, before the first epoch.
constexpr uint64_t REORDER_MIN_POINTS = 32768;
constexpr uint64_t FORK_MIN_POINTS = 32768;

bool reorder_enabled(const bht_job *j) {
    // the renumbering serves the column-tiled attraction of D <= 2
    return !(j->params.flags & BHT_FLAG_NO_REORDER) && j->n >= REORDER_MIN_POINTS &&
           j->dim() <= 2;
}

// Renumbers the points in the diffusion order of the affinity graph: the attraction gathers
// y[col] over the edges of P, so an order that keeps P's well-connected groups together keeps
// each row's gathers within a narrow index range.  Measured at N = 1M it also beats the
// embedding's spatial (box) order at every stage of the run, so the order is kept for the run.
void reorder_by_diffusion(bht_job *j) {
    DevBuf temp(j->ctx);
    temp.ensure(j->ctx, diffusion_order_temp_bytes((int)j->n));
    const uint32_t *order =
        diffusion_order(j->rows.as<uint32_t>(), j->cols.as<uint32_t>(), j->vals.as<float>(),
                        (int)j->n, 16, temp.p, j->stream);
    reorder_points(j, order);  // synchronises the stream before temp is released
}

// Builds P (CSR) from per-row neighbour indices / distances already on the device.
void affinities_from_knn(bht_job *j, uint32_t *d_idx, float *d_dist, uint32_t k) {
    bht_ctx *c = j->ctx;
    const int64_t n = (int64_t)j->n;
    const uint64_t ne = (uint64_t)n * k;
    if (ne >= (1ull << 32)) throw Error(ERR_INVALID, "n * k must be below 2^32");
    DevBuf p(c);
    p.ensure(c, ne * 4);
    search_beta(d_dist, p.as<float>(), n, (int)k, j->params.perplexity,
                (int)j->params.bandwidth_mode, j->stream);
    DevBuf pair(c), nm(c), nmpos(c), inc(c), rowc(c), inoff(c), rws(c), dest(c), edge(c),
        dest2(c), edge2(c), tot(c), stemp(c), rtemp(c);
    pair.ensure(c, ne * 4);
    nm.ensure(c, ne * 4);
    nmpos.ensure(c, ne * 4);
    inc.ensure(c, n * 4);
    rowc.ensure(c, n * 4);
    inoff.ensure(c, n * 4);
    dest.ensure(c, ne * 4);
    edge.ensure(c, ne * 4);
    dest2.ensure(c, ne * 4);
    edge2.ensure(c, ne * 4);
    tot.ensure(c, 16);
    stemp.ensure(c, scan_temp_bytes(ne));
    rtemp.ensure(c, radix_sort_temp_bytes(ne));
    j->rows.ensure(c, (n + 1) * 4);
    SymWork wk;
    wk.pair = pair.as<float>();
    wk.nonmutual = nm.as<uint32_t>();
    wk.nm_pos = nmpos.as<uint32_t>();
    wk.in_count = inc.as<uint32_t>();
    wk.row_count = rowc.as<uint32_t>();
    wk.in_off = inoff.as<uint32_t>();
    wk.rows = j->rows.as<uint32_t>();
    wk.dest = dest.as<uint32_t>();
    wk.edge = edge.as<uint32_t>();
    wk.dest_alt = dest2.as<uint32_t>();
    wk.edge_alt = edge2.as<uint32_t>();
    wk.total = tot.as<uint32_t>();
    wk.scan_temp = stemp.p;
    wk.sort_temp = rtemp.p;
    wk.alloc_out = [&](uint64_t nnz) {
        j->cols.ensure(c, nnz * 4);
        j->vals.ensure(c, nnz * 4);
        wk.out_cols = j->cols.as<uint32_t>();
        wk.out_vals = j->vals.as<float>();
    };
    uint64_t nnz = symmetrize_device(d_idx, p.as<float>(), n, (int)k, wk, j->stream);
    BHT_HIP(hipStreamSynchronize(j->stream));
    j->nnz = nnz;
    j->have_p = true;
    j->cached_perplexity = j->params.perplexity;  // lib.rs:1123-1124
    j->have_cached_perplexity = true;
    reset_order(j);
}

// Neighbours per row the device kernels handle: the kNN from sample vectors keeps a row's
// candidates in one wave's LDS buffer, the bandwidth search a row's values in one wave.
constexpr uint32_t MAX_K_DATA = 192;
constexpr uint32_t MAX_K_NEIGHBORS = 256;

void check_k(uint32_t k, uint32_t cap, const char *what) {
    if (k > cap)
        throw Error(ERR_INVALID, "k = " + std::to_string(k) + " neighbours per row (" + what +
                                     "): at most " + std::to_string(cap) +
                                     " neighbours per row are supported");
}

void build_affinities_from_data(bht_job *j) {
    if (!j->have_data) throw Error(ERR_STATE, "no sample vectors: call bht_job_set_data_f32");
    uint32_t k = n_neighbors_of(j);
    check_k(k, MAX_K_DATA, "3 * perplexity, from sample vectors");
    bht_ctx *c = j->ctx;
    DevBuf idx(c), dist(c), temp(c);
    idx.ensure(c, (size_t)j->n * k * 4);
    dist.ensure(c, (size_t)j->n * k * 4);
    const size_t tb = knn_temp_bytes((int64_t)j->n, (int)j->d_in, (int)k);
    if (tb) temp.ensure(c, tb);
    knn_fast(j->x.as<float>(), (int64_t)j->n, (int)j->d_in, (int)k, idx.as<uint32_t>(),
             dist.as<float>(), tb ? temp.p : nullptr, j->stream, &j->knn_exact_rows);
    temp.reset();
    affinities_from_knn(j, idx.as<uint32_t>(), dist.as<float>(), k);
}

// finalize_p_and_seed / prepare_cached (lib.rs:663-685, 1227-1236) + buffer prep.
// Random or given seed, zero velocity, unit gains, cleared scalars (src/lib.rs:671-684,
// tsne/mod.rs:96-105).
void seed_state(bht_job *j) {
    const uint64_t len = j->n * j->dim();
    hipStream_t s = j->stream;
    std::vector<float> host(len);
    if (j->have_init) {
        host = j->init;
        j->init.clear();
        j->have_init = false;
    } else {
        // random_init (mod.rs:142-147) draws from an unseeded N(0, 1e-4), so separate runs start
        // apart; here each draw is seeded by the job's creation index in the process and its
        // draw count, so that separate jobs start apart too and a program repeats its runs.
        std::mt19937_64 rng(0x6268747366ull + 0x9e3779b97f4a7c15ull * j->serial +
                            7919ull * (j->seed_counter++));
        std::normal_distribution<double> nd(0.0, 1e-4);
        for (auto &v : host) v = (float)nd(rng);
    }
    j->cur = 0;
    upload_points(j, host.data(), j->ycur());
    BHT_HIP(hipMemsetAsync(j->uy.p, 0, len * 4, s));
    std::vector<float> ones(len, 1.0f);
    BHT_HIP(hipMemcpyAsync(j->gains.p, ones.data(), len * 4, hipMemcpyHostToDevice, s));
    BHT_HIP(hipMemsetAsync(j->scal.p, 0, sizeof(EpochScalars), s));
    l2_sync(s);  // after the seed's host copies
    BHT_HIP(hipStreamSynchronize(s));
}

void prepare_run(bht_job *j, int method, float theta, bool cached) {
    if (cached) j->stop_lying_fired = false;
    if (method == FIT_FITSNE) ensure_fitsne_workspace(j);
    else ensure_bh_workspace(j);
    hipStream_t s = j->stream;
    // P and the data may have been written by host copies or on another stream
    l2_sync(s);
    // normalise P with the exaggeration folded in (mod.rs:284-293)
    EpochScalars *sc = j->scal.as<EpochScalars>();
    BHT_HIP(hipMemsetAsync(sc, 0, sizeof(EpochScalars), s));
    sum_values_f64(j->vals.as<float>(), j->nnz, sc, j->dpart.as<double>(), s);
    double psum = 0.0;
    BHT_HIP(hipMemcpyAsync(&psum, &sc->psum, sizeof(double), hipMemcpyDeviceToHost, s));
    BHT_HIP(hipStreamSynchronize(s));
    const float exag = j->params.early_exaggeration;
    const float scale = exag / ((float)psum + FLT_EPSILON);
    scale_values(j->vals.as<float>(), j->nnz, scale, s);
    scale_values(j->qv.as<float>(), std::min<uint64_t>(j->n, j->nnz), scale, s);
    if (j->params.stop_lying_epoch == 0) {
        scale_values(j->vals.as<float>(), j->nnz, 1.0f / exag, s);
        scale_values(j->qv.as<float>(), std::min<uint64_t>(j->n, j->nnz), 1.0f / exag, s);
    }
    seed_state(j);
    j->method = method;
    j->theta = theta;
    j->epoch = 0;
    j->lazy_ready = false;
    j->run_momentum = j->params.momentum;
    j->prepared = true;
    // the tiled copy of P is built once the points have their final numbering
    const bool renumber = reorder_enabled(j);
    sync_jobdev(j, !renumber);
    if (renumber) reorder_by_diffusion(j);
}

// Repulsive forces and Z of the current embedding.
void repulsion(bht_job *j) {
    const LaunchSet L = single_set(j);
    if (j->method == FIT_FITSNE) {
        center_plan(L, true, j->stream);
        fitsne_repulsion(L, j->stream);
        return;
    }
    // recentre + per-axis bounds (arena.rs:65-94), then the arena and its traversal
    center_plan(L, false, j->stream);
    if (j->n > 0) bh_repulsion(L, j->stream);
}

// Host bookkeeping after epoch `epoch` of a job (lib.rs:1162-1179); the device side of the
// schedule (stop-lying rescale, momentum switch) follows the epoch counter by itself.
void epoch_tail(bht_job *j, uint64_t epoch) {
    if (epoch == j->params.stop_lying_epoch && epoch != 0) j->stop_lying_fired = true;
    if (epoch == j->params.momentum_switch_epoch) j->params.momentum = j->params.final_momentum;
}

// FFT line points the epoch's FFT launches hold in LDS: four times the latest length the job's grid
// plan reported (rounded to 64), at most the grid capacity; the whole capacity until the plan has
// reported, and for jobs without the fallback buffer.  The report is up to 16 epochs old (the host
// enqueues ahead), and the length can grow that much meanwhile: early in a run a point thrown out
// of the cloud took it from 300 to 1080 within four epochs.  Lines longer than the launch's LDS
// take the global-memory fallback, which is several times slower (with a quarter more than the
// report, a 300-epoch run at 1M points took 7% longer).
int fft_cap_for(const bht_job *j) {
    if (!j->hd.host_len || !j->hd.fft_fb) return j->len_cap;
    if (j->params.flags & BHT_FLAG_FFT_GLOBAL_LINES) return 64;
    const int seen = __atomic_load_n(j->host_len, __ATOMIC_ACQUIRE);
    if (seen <= 0) return j->len_cap;
    const int want = (4 * seen + 63) / 64 * 64;
    return std::min(want, j->len_cap);
}

// LDS points per block of a batch's 2-D FFT launches for member j: four times the length its grid
// last reported (before any report, the length of its first grid, min_boxes per axis), at most
// its capacity.  The batch grows a group's launches when a member's report outgrows them.
int batch_fft_cap(const bht_job *j) {
    if (j->params.flags & BHT_FLAG_FFT_GLOBAL_LINES) return 64;
    int seen = j->hd.host_len ? __atomic_load_n(j->hd.host_len, __ATOMIC_ACQUIRE) : 0;
    if (seen <= 0) seen = (int)next_smooth(6 * (int64_t)std::max<int>(j->hd.min_boxes, 1) - 1);
    return std::min((4 * seen + 63) / 64 * 64, j->len_cap);
}

// One epoch's launches (src/lib.rs:961-1014: repulsion + attraction, gradient step, zero-mean):
// the recentring, then the repulsion and the attraction (side by side on two streams when
// `fork`, the attraction on `queue_blocks` blocks), then the update.
void enqueue_epoch(bht_job *j, const LaunchSet &L0, bool fit, bool fork, int queue_blocks) {
    LaunchSet L = L0;
    if (fit) L.fft_cap = fft_cap_for(j);
    hipStream_t rs = fork ? j->side : j->stream;
    // A FIt-SNE epoch after an update leaves the recentring to the readers of the embedding (one
    // head kernel: the shift and the grid plan); otherwise the embedding is centred in place.
    // A forked one runs its head on the repulsion's stream, the repulsion being the only reader
    // of the grid plan: the attraction starts at once and derives the shift and the epoch index
    // from the last update's sums itself (the same values, computed the same way).
    const bool lazy = fit && j->lazy_ready;
    const bool head_aside = fork && lazy;
    if (lazy && !head_aside) epoch_head(L, j->stream);
    if (fork) {
        if (!lazy) center_points(L, j->stream);
        BHT_HIP(hipEventRecord(j->ev_fork, j->stream));
        BHT_HIP(hipStreamWaitEvent(j->side, j->ev_fork, 0));
        // the repulsion reads y, the scalars and the bounds' partials written on `stream`
        l2_sync(j->side);
        if (head_aside) epoch_head(L, j->side);
        else if (!lazy) plan_grid(L, fit, j->side);
    } else if (!lazy) {
        center_plan(L, fit, j->stream);
    }
    j->mark(j->stream, 0);
    if (fork) {
        // the attraction is enqueued first: the host takes a while to enqueue the repulsion's
        // chain of kernels, and the attraction must not wait for that
        attract_tiled(L, nullptr, true, j->stream, queue_blocks, head_aside);
        j->mark(j->stream, 2);
        if (fit) fitsne_repulsion(L, rs);
        else if (j->n > 0) bh_repulsion(L, rs);
        // The update reads F- and Z written on `side`: the side writes its lines back and
        // drops stale copies on every XCD before the join, while the attraction still runs
        // (which neither reads nor writes them).
        l2_sync(rs);
        j->mark(rs, 1);
        BHT_HIP(hipEventRecord(j->ev_join, j->side));
        BHT_HIP(hipStreamWaitEvent(j->stream, j->ev_join, 0));
    } else {
        if (fit) fitsne_repulsion(L, rs);
        else if (j->n > 0) bh_repulsion(L, rs);
        j->mark(rs, 1);
        attract_tiled(L, nullptr, true, j->stream, 0);
        j->mark(j->stream, 2);
    }
    update_points(L, j->stream);
    j->lazy_ready = fit;
    j->mark(j->stream, 3);
    if (j->timing) j->events_used += 4;
}

// ---- autotuner ---------------------------------------------------------------------------------
//
// For a job large enough to run its repulsion beside its attraction, how many blocks the
// attraction keeps (the rest of the device goes to the repulsion's chain of small kernels), or
// whether the epoch is better run in order on one stream, depends on the job's size and on the
// device.  None of these choices changes any result bit (the attraction's sums do not depend on
// which block takes which row block, and the two streams compute what one stream computes), so
// the choice is made by timing: the first run of a job in a size class (per context; method,
// output dimension and half-octave of n) times a few epochs of each candidate from a snapshot of
// the job's state, restores the snapshot, and keeps the fastest.

struct TuneChoice {
    int fork = 1;
    int queue_blocks = 0;
};

}  // namespace

struct TuneCache {
    std::mutex mu;
    std::map<uint64_t, TuneChoice> choices;
};

bht_ctx::~bht_ctx() = default;

namespace {

uint64_t tune_key(const bht_job *j) {
    int lg = 0;  // floor(log2 n)
    while (lg < 62 && (2ull << lg) <= j->n) lg++;
    const uint64_t half = j->n >= (3ull << lg) / 2 ? 1 : 0;  // upper half of the octave
    return ((uint64_t)j->method << 40) | ((uint64_t)j->dim() << 32) | (2ull * lg + half);
}

// Per candidate and round: TUNE_WARM untimed epochs (then a synchronisation), one epoch that lets
// the host enqueue ahead of the device, and TUNE_TIMED timed epochs; TUNE_ROUNDS rounds over the
// candidates, each candidate's best round counts.
constexpr int TUNE_WARM = 1, TUNE_LEAD = 1, TUNE_TIMED = 3, TUNE_ROUNDS = 2;

// Epoch-carried device state: both embedding buffers, uy, gains, the scalars (epoch counter,
// pending recentring) and the update's partial sums (the next recentring reads them).
struct StateSnapshot {
    DevBuf y0, y1, uy, gains, scal, mpart, fpart;
    int cur = 0;
    uint64_t epoch = 0;
    bool lazy_ready = false;
};

void snapshot_take(bht_job *j, StateSnapshot &sn) {
    bht_ctx *c = j->ctx;
    const size_t b = (size_t)j->n * j->dim() * 4;
    auto copy = [&](DevBuf &dst, const DevBuf &src, size_t bytes) {
        dst.ensure(c, bytes);
        BHT_HIP(hipMemcpyAsync(dst.p, src.p, bytes, hipMemcpyDeviceToDevice, j->stream));
    };
    copy(sn.y0, j->y[0], b);
    copy(sn.y1, j->y[1], b);
    copy(sn.uy, j->uy, b);
    copy(sn.gains, j->gains, b);
    copy(sn.scal, j->scal, sizeof(EpochScalars));
    copy(sn.mpart, j->mpart, 2 * (size_t)MAX_UPDATE_PARTS * sizeof(double));
    copy(sn.fpart, j->fpart, (size_t)MAX_UPDATE_PARTS * sizeof(float4));
    sn.cur = j->cur;
    sn.epoch = j->epoch;
    sn.lazy_ready = j->lazy_ready;
}

void snapshot_restore(bht_job *j, StateSnapshot &sn) {
    const size_t b = (size_t)j->n * j->dim() * 4;
    auto copy = [&](DevBuf &dst, const DevBuf &src, size_t bytes) {
        BHT_HIP(hipMemcpyAsync(dst.p, src.p, bytes, hipMemcpyDeviceToDevice, j->stream));
    };
    copy(j->y[0], sn.y0, b);
    copy(j->y[1], sn.y1, b);
    copy(j->uy, sn.uy, b);
    copy(j->gains, sn.gains, b);
    copy(j->scal, sn.scal, sizeof(EpochScalars));
    copy(j->mpart, sn.mpart, 2 * (size_t)MAX_UPDATE_PARTS * sizeof(double));
    copy(j->fpart, sn.fpart, (size_t)MAX_UPDATE_PARTS * sizeof(float4));
    l2_sync(j->stream);
    BHT_HIP(hipStreamSynchronize(j->stream));
    j->cur = sn.cur;
    j->epoch = sn.epoch;
    j->lazy_ready = sn.lazy_ready;
}

TuneChoice autotune(bht_job *j, const LaunchSet &L, bool fit) {
    TuneChoice best;
    best.queue_blocks = attract_default_queue_blocks();
    // every candidate runs epochs [epoch, epoch + TUNE_WARM + TUNE_LEAD + TUNE_TIMED) from the
    // snapshot; none may be the epoch after stop-lying (it rescales the tiled P in place)
    const uint64_t sl = j->params.stop_lying_epoch;
    const uint64_t trials_end = j->epoch + TUNE_WARM + TUNE_LEAD + TUNE_TIMED;
    if (sl != 0 && j->epoch <= sl + 1 && sl + 1 < trials_end) return best;
    std::vector<TuneChoice> cands;
    const int nblk = L.nb[F_ATT];
    for (int q : {192, 224, 256, 288, 320, 384, 448, 512})
        if (q < nblk) cands.push_back(TuneChoice{1, q});
    cands.push_back(TuneChoice{1, 0});  // forked, one block per row block
    cands.push_back(TuneChoice{0, 0});  // in order on one stream
    StateSnapshot sn;
    snapshot_take(j, sn);
    BHT_HIP(hipStreamSynchronize(j->stream));
    hipEvent_t t0, t1;
    BHT_HIP(hipEventCreate(&t0));
    BHT_HIP(hipEventCreate(&t1));
    std::vector<float> cand_ms(cands.size(), 0.0f);
    const bool was_timing = j->timing;
    j->timing = false;
    try {
        for (int r = 0; r < TUNE_ROUNDS; r++) {
            for (size_t c = 0; c < cands.size(); c++) {
                auto epoch_of = [&]() {
                    enqueue_epoch(j, L, fit, cands[c].fork != 0, cands[c].queue_blocks);
                    j->cur = 1 - j->cur;
                };
                for (int e = 0; e < TUNE_WARM; e++) epoch_of();
                // the later epochs size their FFT launches by the grid the warm epoch reported
                BHT_HIP(hipStreamSynchronize(j->stream));
                for (int e = 0; e < TUNE_LEAD; e++) epoch_of();
                BHT_HIP(hipEventRecord(t0, j->stream));
                for (int e = 0; e < TUNE_TIMED; e++) epoch_of();
                BHT_HIP(hipEventRecord(t1, j->stream));
                BHT_HIP(hipEventSynchronize(t1));
                float ms = 0.0f;
                BHT_HIP(hipEventElapsedTime(&ms, t0, t1));
                if (r == 0 || ms < cand_ms[c]) cand_ms[c] = ms;
                snapshot_restore(j, sn);
            }
        }
        size_t pick = 0;
        for (size_t c = 1; c < cands.size(); c++)
            if (cand_ms[c] < cand_ms[pick]) pick = c;
        best = cands[pick];
    } catch (...) {
        j->timing = was_timing;
        hipEventDestroy(t0);
        hipEventDestroy(t1);
        throw;
    }
    j->timing = was_timing;
    BHT_HIP(hipEventDestroy(t0));
    BHT_HIP(hipEventDestroy(t1));
    return best;
}

TuneChoice epoch_choice(bht_job *j, const LaunchSet &L, bool fit) {
    TuneChoice ch;
    ch.fork = j->n >= FORK_MIN_POINTS && j->dim() <= 2 &&
                      !(j->params.flags & BHT_FLAG_SERIAL_EPOCH)
                  ? 1
                  : 0;
    ch.queue_blocks = ch.fork ? attract_default_queue_blocks() : 0;
    if (!ch.fork || (j->params.flags & BHT_FLAG_NO_AUTOTUNE)) return ch;
    TuneCache &tc = *j->ctx->tune;
    const uint64_t key = tune_key(j);
    {
        std::lock_guard<std::mutex> g(tc.mu);
        auto it = tc.choices.find(key);
        if (it != tc.choices.end()) return it->second;
    }
    ch = autotune(j, L, fit);
    std::lock_guard<std::mutex> g(tc.mu);
    tc.choices[key] = ch;
    return ch;
}

void run_epochs(bht_job *j, uint64_t count, bht_epoch_cb cb, void *user) {
    if (!j->prepared) throw Error(ERR_STATE, "bht_job_prepare must precede bht_job_step");
    if (j->in_batch) throw Error(ERR_STATE, "the job is advanced by its batch");
    const uint64_t len = j->n * j->dim();
    std::vector<float> snap(cb ? len : 0);
    const LaunchSet L = single_set(j);
    const bool fit = j->method == FIT_FITSNE;
    // Small jobs are bound by launch latency, which the fork's events would only add to; they
    // run the same kernels in order on one stream.  Large ones use the autotuner's choice.
    const TuneChoice ch = count > 0 ? epoch_choice(j, L, fit) : TuneChoice{0, 0};
    j->last_choice = ch.fork ? ch.queue_blocks : -1;
    if (j->timing) j->timed_fork = ch.fork != 0;
    const bool bounded = fit && j->hd.host_len != nullptr;
    for (auto &r : j->ahead_rec) r = false;
    for (uint64_t e = 0; e < count; e++) {
        // at most 16 epochs in flight (every 4th epoch records an event)
        const int slot = (int)((e / 4) % 4);
        if (bounded && e % 4 == 0 && j->ahead_rec[slot])
            BHT_HIP(hipEventSynchronize(j->ahead[slot]));
        enqueue_epoch(j, L, fit, ch.fork != 0, ch.queue_blocks);
        if (bounded && e % 4 == 0) {
            BHT_HIP(hipEventRecord(j->ahead[slot], j->stream));
            j->ahead_rec[slot] = true;
        }
        j->cur = 1 - j->cur;
        const uint64_t epoch = j->epoch++;
        epoch_tail(j, epoch);
        if (cb) {
            recenter(j);
            download_points(j, j->ycur(), snap.data());
            cb(epoch, snap.data(), len, user);
        }
    }
}

void finish_run(bht_job *j) {
    if (!j->prepared) throw Error(ERR_STATE, "no prepared run");
    // a batch member's epochs ran on the batch's stream
    l2_sync(j->stream);
    recenter(j);
    // The epochs read the tiled copy of P and rescale it (tval) at stop-lying; the CSR values
    // (exports, KL) are rescaled here once the schedule has stopped lying.  If the run ended
    // right after stop_lying_epoch, no epoch has rescaled the tiled copy yet either.
    const uint64_t sl = j->params.stop_lying_epoch;
    if (j->method == FIT_EXACT) {
        // the exact epochs read P divided by the exaggeration once stop-lying fired; the stored P
        // (KL) is divided here
        if (sl != 0 && j->epoch >= sl + 1)
            scale_values(j->pdense.as<float>(), j->n * j->n,
                         1.0f / j->params.early_exaggeration, j->stream);
    } else if (sl != 0 && j->epoch >= sl + 1) {
        const float inv = 1.0f / j->params.early_exaggeration;
        scale_values(j->vals.as<float>(), j->nnz, inv, j->stream);
        scale_values(j->qv.as<float>(), std::min<uint64_t>(j->n, j->nnz), inv, j->stream);
        if (j->epoch == sl + 1) scale_values(j->tval.as<float>(), j->tile_nnz, inv, j->stream);
    }
    BHT_HIP(hipStreamSynchronize(j->stream));
    j->fit_kind = j->method;
    j->fit_theta = j->theta;
    j->prepared = false;
}

// Whether the epoch's P (what the next epoch's forces use) is the exported affinities times the
// early exaggeration.  During a run the epoch after stop_lying_epoch applies the rescale as it
// reads P, so until that epoch has run the device values are exaggerated although the reference's
// schedule has already rescaled them; with stop_lying_epoch == 0 the export divides P by the
// exaggeration once more (lib.rs:1024-1029).
bool device_p_exaggerated(const bht_job *j) {
    const uint64_t sl = j->params.stop_lying_epoch;
    if (j->prepared) return sl == 0 || j->epoch <= sl + 1;
    return !j->stop_lying_fired;
}

bool has_cached_affinities(const bht_job *j) {  // lib.rs:1192-1196
    return j->have_p && j->have_cached_perplexity &&
           j->cached_perplexity == j->params.perplexity;
}

void validate_init(const bht_job *j) {
    if (j->have_init && j->init.size() != j->n * j->dim())
        throw Error(ERR_INVALID, "error: initial embedding has " +
                                     std::to_string(j->init.size()) +
                                     " values, expected n_samples * D = " +
                                     std::to_string(j->n * j->dim()));
}

// ---- the exact fit (src/lib.rs:489-646) -----------------------------------------------------

constexpr uint64_t EXACT_MAX_POINTS = 65536;

// Dense P from the samples: distances, a bandwidth search per dense row, symmetrisation, and the
// normalisation with the exaggeration folded in; then the seed.
void prepare_exact(bht_job *j, int metric) {
    const uint64_t n = j->n;
    if (n > EXACT_MAX_POINTS)
        throw Error(ERR_INVALID, "the exact path supports at most " +
                                     std::to_string(EXACT_MAX_POINTS) + " samples");
    if (!j->have_data) throw Error(ERR_STATE, "no sample vectors: call bht_job_set_data_f32");
    ensure_state_buffers(j);
    bht_ctx *c = j->ctx;
    hipStream_t s = j->stream;
    const uint64_t nn = std::max<uint64_t>(n * n, 1);
    j->pdense.ensure(c, nn * sizeof(float));
    j->zrow.ensure(c, 2 * std::max<uint64_t>(n, 1) * sizeof(double));
    {
        DevBuf dist(c);
        dist.ensure(c, nn * sizeof(float));
        l2_sync(s);  // the samples may come from host copies
        exact_distances(j->x.as<float>(), (int64_t)n, (int)j->d_in, metric, dist.as<float>(), s);
        search_beta_dense(dist.as<float>(), j->pdense.as<float>(), (int64_t)n,
                          j->params.perplexity, (int)j->params.bandwidth_mode, s);
        BHT_HIP(hipStreamSynchronize(s));  // before the distances are released
    }
    exact_symmetrize(j->pdense.as<float>(), (int64_t)n, s);
    // normalise with the exaggeration folded in (tsne/mod.rs:284-293)
    EpochScalars *sc = j->scal.as<EpochScalars>();
    BHT_HIP(hipMemsetAsync(sc, 0, sizeof(EpochScalars), s));
    sum_values_f64(j->pdense.as<float>(), n * n, sc, j->dpart.as<double>(), s);
    double psum = 0.0;
    BHT_HIP(hipMemcpyAsync(&psum, &sc->psum, sizeof(double), hipMemcpyDeviceToHost, s));
    BHT_HIP(hipStreamSynchronize(s));
    const float exag = j->params.early_exaggeration;
    scale_values(j->pdense.as<float>(), n * n, exag / ((float)psum + FLT_EPSILON), s);
    if (j->params.stop_lying_epoch == 0)
        scale_values(j->pdense.as<float>(), n * n, 1.0f / exag, s);
    // the update takes the attraction sums as the whole gradient: the repulsion stays zero
    BHT_HIP(hipMemsetAsync(j->frep.p, 0, n * j->dim() * sizeof(float), s));
    seed_state(j);
    j->method = FIT_EXACT;
    j->theta = 0.0f;
    j->epoch = 0;
    j->lazy_ready = false;
    j->run_momentum = j->params.momentum;
    j->prepared = true;
    // an exact fit keeps no sparse affinities (the reference's affinities() is None after it)
    j->have_p = false;
    sync_jobdev(j, false);
}

void run_exact_epochs(bht_job *j, uint64_t count, bht_epoch_cb cb, void *user) {
    const uint64_t len = j->n * j->dim();
    std::vector<float> snap(cb ? len : 0);
    const LaunchSet L = single_set(j);
    for (uint64_t e = 0; e < count; e++) {
        // zero_mean of the previous epoch's update (src/lib.rs:631), then the exact gradient
        center_points(L, j->stream);
        exact_gradient(L.jobs, j->pdense.as<float>(), (int64_t)j->n, (int)j->dim(),
                       j->zrow.as<double>(), j->stream);
        update_points(L, j->stream);
        j->cur = 1 - j->cur;
        const uint64_t epoch = j->epoch++;
        epoch_tail(j, epoch);
        if (cb) {
            recenter(j);
            download_points(j, j->ycur(), snap.data());
            cb(epoch, snap.data(), len, user);
        }
    }
    j->last_choice = -1;
}

void fit(bht_job *j, int method, float theta, bht_epoch_cb cb, void *user, bool cached) {
    validate_init(j);
    prepare_run(j, method, theta, cached);
    run_epochs(j, j->params.epochs, cb, user);
    finish_run(j);
}

void check_dims(const bht_job *j, int method) {
    if (method == FIT_BH && (j->dim() < 2 || j->dim() > 7))
        throw Error(ERR_INVALID, "the Barnes-Hut path supports d_out in 2..7");
    // the 2-D walk packs a child index (below 4n: the subtrees' regions follow the top nodes')
    // in 29 bits
    if (method == FIT_BH && j->dim() == 2 && j->n >= (1ull << 27))
        throw Error(ERR_INVALID, "the Barnes-Hut path supports fewer than 2^27 points");
    if (method == FIT_FITSNE && j->dim() != 1 && j->dim() != 2)
        throw Error(ERR_INVALID, "the FIt-SNE path supports d_out in {1, 2}");
}

// check_neighbors + cache match + range check (lib.rs:781-815, 922-951)
void fit_with_neighbors(bht_job *j, int method, float theta, const uint64_t *idx,
                        const float *dist, const uint64_t *row_lens, uint64_t n_rows,
                        bht_epoch_cb cb, void *user) {
    if (n_rows != j->n)
        throw Error(ERR_INVALID, "error: neighbors has " + std::to_string(n_rows) +
                                     " rows, expected one per sample = " + std::to_string(j->n));
    uint64_t k = n_rows ? row_lens[0] : 0;
    bool ok = k > 0;
    for (uint64_t i = 0; ok && i < n_rows; i++) ok = row_lens[i] == k;
    if (!ok)
        throw Error(ERR_INVALID,
                    "error: every neighbors row must have the same length, greater than zero.");
    check_k((uint32_t)std::min<uint64_t>(k, UINT32_MAX), MAX_K_NEIGHBORS, "caller neighbours");
    // cached_affinities_match_neighbors (lib.rs:1201-1220)
    if (j->have_p && j->have_cached_perplexity &&
        j->cached_perplexity == j->params.perplexity) {
        std::vector<uint32_t> hrows, hcols;
        download_csr(j, hrows, hcols, nullptr);
        bool match = true;
        for (uint64_t i = 0; match && i < j->n; i++) {
            if (hrows[i + 1] - hrows[i] != k) { match = false; break; }
            for (uint64_t t = 0; t < k; t++)
                if (hcols[hrows[i] + t] != (uint32_t)idx[i * k + t]) { match = false; break; }
        }
        if (match) {
            fit(j, method, theta, cb, user, true);
            return;
        }
    }
    for (uint64_t e = 0; e < n_rows * k; e++)
        if (idx[e] >= j->n)
            throw Error(ERR_INVALID,
                        "error: a neighbor index is out of range, every index must be < "
                        "n_samples = " + std::to_string(j->n) + ".");
    bht_ctx *c = j->ctx;
    DevBuf didx(c), ddist(c);
    didx.ensure(c, n_rows * k * 4);
    ddist.ensure(c, n_rows * k * 4);
    std::vector<uint32_t> hidx(n_rows * k);
    for (uint64_t e = 0; e < n_rows * k; e++) hidx[e] = (uint32_t)idx[e];
    BHT_HIP(hipMemcpyAsync(didx.p, hidx.data(), hidx.size() * 4, hipMemcpyHostToDevice,
                           j->stream));
    BHT_HIP(hipMemcpyAsync(ddist.p, dist, n_rows * k * 4, hipMemcpyHostToDevice, j->stream));
    affinities_from_knn(j, didx.as<uint32_t>(), ddist.as<float>(), (uint32_t)k);
    fit(j, method, theta, cb, user, false);
}

}  // namespace

// ---- batch: many jobs advanced by the same launches -------------------------------------------
//
// The members' table entries live in one device array.  Jobs of one method and output dimension
// form a group; every epoch kernel of a group is one launch whose blocks are dealt out to the
// group's jobs by a block map (job index << 16 | block of the job).  The per-epoch sorts of a
// group run as one segmented sort over an arena holding every member's keys at a tile-aligned
// offset.  The epoch's launch sequence depends only on the membership, so it is captured once
// as a graph and replayed.

namespace {

struct BatchGroup {
    int method = M_FITSNE, dim = 2;
    std::vector<bht_job *> jobs;
    std::vector<int> slots;  // table index of each member
    LaunchSet L;
    DevBuf keys, keys_alt, vals, vals_alt, sort_temp, segs, tile_seg, maps;
};

}  // namespace

struct bht_batch {
    bht_ctx *ctx;
    hipStream_t stream = nullptr;
    std::vector<bht_job *> members;
    std::vector<bht_job *> finished;
    std::vector<std::unique_ptr<BatchGroup>> groups;
    DevBuf table;
    bool dirty = true;
    bool use_graph = true;
    // The groups' repulsions on `side` beside their attractions.  Off by default: a batch's
    // grouped launches already fill the device (C4: 5.18 against 5.25 s of epochs; C5 serving
    // 17.7 against 21.3 jobs/s).
    bool fork = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t epochs_launched = 0;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // a batch with 2-D FIt-SNE groups keeps at most 16 epochs queued (an event every 4th epoch),
    // so that the members' grid-length reports are recent when their FFT launches are resized
    hipEvent_t ahead[4] = {};
    bool ahead_rec[4] = {};

    // non-blocking: synchronous copies on the null stream by another host thread (a request
    // loop preparing jobs) neither wait for the batch nor disturb a capture in progress
    explicit bht_batch(bht_ctx *c) : ctx(c) {
        BHT_HIP(hipSetDevice(ctx->device));
        BHT_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        BHT_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        BHT_HIP(hipEventCreateWithFlags(&ev_fork, HANDOVER_EVENT));
        BHT_HIP(hipEventCreateWithFlags(&ev_join, HANDOVER_EVENT));
        for (auto &e : ahead) BHT_HIP(hipEventCreateWithFlags(&e, HANDOVER_EVENT));
    }
    ~bht_batch() {
        drop_graph();
        for (auto &e : ahead)
            if (e) hipEventDestroy(e);
        if (ev_fork) hipEventDestroy(ev_fork);
        if (ev_join) hipEventDestroy(ev_join);
        if (side) hipStreamDestroy(side);
        if (stream) hipStreamDestroy(stream);
    }
    void drop_graph() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

namespace {

uint64_t pad_tile(uint64_t n) {
    return (n + RS_TILE_KEYS - 1) / RS_TILE_KEYS * RS_TILE_KEYS;
}

// Rebuilds the table, the groups' arenas and block maps after a membership change.
void batch_rebuild(bht_batch *b) {
    bht_ctx *c = b->ctx;
    // epochs launched earlier may still read the old table, maps and arenas
    BHT_HIP(hipStreamSynchronize(b->stream));
    b->drop_graph();
    b->groups.clear();
    std::vector<JobDev> table(b->members.size());
    for (size_t m = 0; m < b->members.size(); m++) {
        bht_job *j = b->members[m];
        const int method = j->method == FIT_BH ? M_BH : M_FITSNE;
        BatchGroup *g = nullptr;
        for (auto &q : b->groups)
            if (q->method == method && q->dim == (int)j->dim()) g = q.get();
        if (!g) {
            b->groups.emplace_back(new BatchGroup());
            g = b->groups.back().get();
            g->method = method;
            g->dim = (int)j->dim();
        }
        g->jobs.push_back(j);
        g->slots.push_back((int)m);
    }
    for (auto &gp : b->groups) {
        BatchGroup &g = *gp;
        const bool bh = g.method == M_BH;
        const size_t ksize = bh ? 8 : 4;
        const int key_bits = bh ? MORTON_KEY_BITS : box_key_bits(g.dim);
        const bool alt = sort_result_alt(key_bits, false);  // batches keep 10-bit digits
        // arena layout: one tile-aligned segment per member
        std::vector<SortSeg> segs(g.jobs.size());
        std::vector<uint32_t> tile_seg;
        uint64_t base = 0, dense = 0;
        for (size_t q = 0; q < g.jobs.size(); q++) {
            const uint64_t n = g.jobs[q]->n;
            const uint32_t nt = (uint32_t)(pad_tile(n) / RS_TILE_KEYS);
            segs[q] = SortSeg{base, dense, (uint32_t)n, (uint32_t)tile_seg.size(), nt, 0};
            for (uint32_t t = 0; t < nt; t++) tile_seg.push_back((uint32_t)q);
            base += pad_tile(n);
            dense += n;
        }
        const uint64_t arena = std::max<uint64_t>(base, RS_TILE_KEYS);
        g.keys.ensure(c, arena * ksize);
        g.keys_alt.ensure(c, arena * ksize);
        g.vals.ensure(c, arena * 4);
        g.vals_alt.ensure(c, arena * 4);
        g.sort_temp.ensure(c, group_sort_temp_bytes(arena));
        g.segs.ensure(c, segs.size() * sizeof(SortSeg));
        g.tile_seg.ensure(c, std::max<size_t>(tile_seg.size(), 1) * 4);
        BHT_HIP(hipMemcpyAsync(g.segs.p, segs.data(), segs.size() * sizeof(SortSeg),
                               hipMemcpyHostToDevice, b->stream));
        if (!tile_seg.empty())
            BHT_HIP(hipMemcpyAsync(g.tile_seg.p, tile_seg.data(), tile_seg.size() * 4,
                                   hipMemcpyHostToDevice, b->stream));
        // table entries: batch grid sizes, sort fields into the arena
        int len_cap = 1, dyn_cap = 1;
        for (size_t q = 0; q < g.jobs.size(); q++) {
            bht_job *j = g.jobs[q];
            fill_jobdev(j);
            if (!bh && g.dim == 2) dyn_cap = std::max(dyn_cap, batch_fft_cap(j));
            JobDev d = j->hd;
            job_block_counts(d, true);
            const uint64_t o = segs[q].base;
            if (bh) {
                d.codes = g.keys.as<uint64_t>() + o;
                d.cidx = g.vals.as<uint32_t>() + o;
                d.scodes = (alt ? g.keys_alt.as<uint64_t>() : g.keys.as<uint64_t>()) + o;
                d.scidx = (alt ? g.vals_alt.as<uint32_t>() : g.vals.as<uint32_t>()) + o;
            } else {
                d.keys = g.keys.as<uint32_t>() + o;
                d.kidx = g.vals.as<uint32_t>() + o;
                d.skeys = (alt ? g.keys_alt.as<uint32_t>() : g.keys.as<uint32_t>()) + o;
                d.sidx = (alt ? g.vals_alt.as<uint32_t>() : g.vals.as<uint32_t>()) + o;
            }
            table[g.slots[q]] = d;
            len_cap = std::max(len_cap, j->len_cap);
        }
        // block maps, family by family
        std::vector<uint32_t> maps;
        size_t off[NFAM];
        int count[NFAM];
        for (int f = 0; f < NFAM; f++) {
            off[f] = maps.size();
            for (size_t q = 0; q < g.jobs.size(); q++) {
                const int nb = table[g.slots[q]].nblk[f];
                for (int lb = 0; lb < nb; lb++)
                    maps.push_back(((uint32_t)g.slots[q] << 16) | (uint32_t)lb);
            }
            count[f] = (int)(maps.size() - off[f]);
        }
        g.maps.ensure(c, std::max<size_t>(maps.size(), 1) * 4);
        BHT_HIP(hipMemcpyAsync(g.maps.p, maps.data(), maps.size() * 4, hipMemcpyHostToDevice,
                               b->stream));
        // the host vectors are released at the end of this iteration
        BHT_HIP(hipStreamSynchronize(b->stream));
        LaunchSet &L = g.L;
        for (int f = 0; f < NFAM; f++) {
            L.map[f] = g.maps.as<uint32_t>() + off[f];
            L.nb[f] = count[f];
        }
        L.dim = g.dim;
        L.njobs = (int)g.jobs.size();
        L.tw = &c->tw;
        // 2-D groups: LDS for the lengths reached (every member has the global lines); the 1-D
        // line convolution holds its capacity
        L.fft_cap = !bh && g.dim == 2 ? dyn_cap : len_cap;
        L.keys = g.keys.p;
        L.keys_alt = g.keys_alt.p;
        L.vals = g.vals.as<uint32_t>();
        L.vals_alt = g.vals_alt.as<uint32_t>();
        L.sort_n = arena;
        L.sort_temp = g.sort_temp.p;
        L.segs = g.segs.as<SortSeg>();
        L.tile_seg = g.tile_seg.as<uint32_t>();
        L.sort_tiles = (uint32_t)tile_seg.size();
        L.max_seg_tiles = 0;
        for (const SortSeg &sg : segs) L.max_seg_tiles = std::max(L.max_seg_tiles, sg.ntiles);
    }
    b->table.ensure(c, std::max<size_t>(table.size(), 1) * sizeof(JobDev));
    if (!table.empty())
        BHT_HIP(hipMemcpyAsync(b->table.p, table.data(), table.size() * sizeof(JobDev),
                               hipMemcpyHostToDevice, b->stream));
    BHT_HIP(hipStreamSynchronize(b->stream));
    for (auto &g : b->groups) g->L.jobs = b->table.as<JobDev>();
    b->dirty = false;
}

// One epoch of every member: per group the recentring, the repulsion, the attraction and the
// update, in order on the batch stream.  (The many jobs of a batch fill the device without the
// two-stream fork of a single large job; batches of many jobs in two groups that ran the
// repulsion beside the attraction came out different from their members run alone, in graph
// replay as well as in plain launches, so the batch does not fork.)
void batch_epoch_launches(bht_batch *b) {
    if (!b->fork) {
        for (auto &gp : b->groups) {
            const LaunchSet &L = gp->L;
            if (L.njobs == 0) continue;
            const bool fit = gp->method == M_FITSNE;
            center_plan(L, fit, b->stream);
            if (fit) fitsne_repulsion(L, b->stream);
            else bh_repulsion(L, b->stream);
            attract_tiled(L, nullptr, true, b->stream);
            update_points(L, b->stream);
        }
        return;
    }
    // every group's recentring and grid plan, then the repulsions (on `side`) beside the
    // attractions, then the updates: the same kernels on the same data as in order, so the
    // same bits (the handover between the streams goes through an L2 write-back/invalidate)
    for (auto &gp : b->groups)
        if (gp->L.njobs) center_plan(gp->L, gp->method == M_FITSNE, b->stream);
    BHT_HIP(hipEventRecord(b->ev_fork, b->stream));
    BHT_HIP(hipStreamWaitEvent(b->side, b->ev_fork, 0));
    l2_sync(b->side);
    for (auto &gp : b->groups)
        if (gp->L.njobs) attract_tiled(gp->L, nullptr, true, b->stream);
    for (auto &gp : b->groups) {
        if (!gp->L.njobs) continue;
        if (gp->method == M_FITSNE) fitsne_repulsion(gp->L, b->side);
        else bh_repulsion(gp->L, b->side);
    }
    BHT_HIP(hipEventRecord(b->ev_join, b->side));
    BHT_HIP(hipStreamWaitEvent(b->stream, b->ev_join, 0));
    l2_sync(b->stream);
    for (auto &gp : b->groups)
        if (gp->L.njobs) update_points(gp->L, b->stream);
}

// One epoch as a graph replay once the membership has held for a whole step (a request loop
// that admits jobs every step runs plain launches instead of re-capturing each time).  A capture
// that fails (another host thread made a call the runtime will not record beside a capture) is
// discarded and the epoch launched plainly.
void batch_launch_epoch(bht_batch *b, bool may_capture) {
    if (!b->use_graph || (!b->exec && !may_capture)) {
        batch_epoch_launches(b);
        return;
    }
    if (!b->exec) {
        hipGraph_t g = nullptr;
        BHT_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
        bool recorded = true;
        try {
            batch_epoch_launches(b);
        } catch (const Error &) {
            recorded = false;  // launches refused by an invalidated capture
        }
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (!recorded || e != hipSuccess || g == nullptr) {
            if (g) hipGraphDestroy(g);
            (void)hipGetLastError();
            batch_epoch_launches(b);
            return;
        }
        b->graph = g;
        BHT_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    BHT_HIP(hipGraphLaunch(b->exec, b->stream));
}

// Validates and prepares a job for batched stepping (the fit's preamble, lib.rs:652-685).  A job
// already prepared for this method at epoch 0 (bht_job_prepare, e.g. on another host thread
// while the batch runs) joins as it is.
void batch_admit(bht_batch *b, bht_job *j, int method, float theta) {
    if (j->in_batch) throw Error(ERR_STATE, "the job already belongs to a batch");
    if (j->ctx != b->ctx) throw Error(ERR_INVALID, "the job belongs to another context");
    if (method == FIT_BH && !(theta > 0.0f)) throw Error(ERR_INVALID, MSG_THETA);
    check_dims(j, method);
    check_perplexity(j);
    if (j->n >= (1u << 31) / 2) throw Error(ERR_INVALID, "batched jobs must have n < 2^30");
    const bool ready = j->prepared && j->method == method && j->epoch == 0 &&
                       (method != FIT_BH || j->theta == theta);
    if (!ready) {
        const bool cached = has_cached_affinities(j);
        if (!cached) build_affinities_from_data(j);
        validate_init(j);
        prepare_run(j, method, theta, cached);
    }
    // a 2-D FIt-SNE member's FFT launches are sized by the lengths the grids reach
    if (method == FIT_FITSNE) ensure_fft_lines(j);
    j->in_batch = true;
    j->lazy_ready = false;
    b->members.push_back(j);
    b->dirty = true;
}

// Grows a 2-D FIt-SNE group's FFT launches when a member's reported grid length has outgrown
// them (its lines meanwhile took the global lines); the next epoch captures a new graph.
void batch_grow_fft(bht_batch *b) {
    for (auto &gp : b->groups) {
        BatchGroup &g = *gp;
        if (g.method != M_FITSNE || g.dim != 2) continue;
        int need = 1;
        for (bht_job *j : g.jobs) need = std::max(need, batch_fft_cap(j));
        if (need > g.L.fft_cap) {
            g.L.fft_cap = need;
            b->drop_graph();
        }
    }
}

// Advances every member by up to max_epochs (never past the member with the fewest epochs
// left), then retires the members whose runs are complete.  Returns the epochs advanced.
uint64_t batch_step(bht_batch *b, uint64_t max_epochs) {
    if (b->members.empty()) return 0;
    uint64_t k = max_epochs;
    for (bht_job *j : b->members) {
        const uint64_t left = j->params.epochs > j->epoch ? j->params.epochs - j->epoch : 0;
        k = std::min(k, left);
    }
    if (k > 0) {
        const bool rebuilt = b->dirty;
        if (b->dirty) {
            batch_rebuild(b);
            // new members were prepared on their own streams
            l2_sync(b->stream);
        } else {
            batch_grow_fft(b);
        }
        bool dyn = false;  // a group sizes its FFT launches by its members' reports
        for (auto &gp : b->groups) dyn |= gp->method == M_FITSNE && gp->dim == 2;
        for (auto &r : b->ahead_rec) r = false;
        for (uint64_t e = 0; e < k; e++) {
            const int slot = (int)((e / 4) % 4);
            if (dyn && e % 4 == 0 && b->ahead_rec[slot]) {
                BHT_HIP(hipEventSynchronize(b->ahead[slot]));
                batch_grow_fft(b);
            }
            batch_launch_epoch(b, !rebuilt || k >= 64);
            if (dyn && e % 4 == 0) {
                BHT_HIP(hipEventRecord(b->ahead[slot], b->stream));
                b->ahead_rec[slot] = true;
            }
        }
        b->epochs_launched += k;
        for (bht_job *j : b->members)
            for (uint64_t e = 0; e < k; e++) {
                j->cur = 1 - j->cur;
                epoch_tail(j, j->epoch++);
            }
    }
    // retire completed runs
    std::vector<bht_job *> keep;
    bool synced = false;
    for (bht_job *j : b->members) {
        if (j->epoch < j->params.epochs) {
            keep.push_back(j);
            continue;
        }
        if (!synced) {
            BHT_HIP(hipStreamSynchronize(b->stream));
            synced = true;
        }
        j->in_batch = false;
        finish_run(j);
        b->finished.push_back(j);
        b->dirty = true;
    }
    b->members.swap(keep);
    return k;
}

}  // namespace

// ---- exported functions ----------------------------------------------------------------------

extern "C" {

void bht_params_default(bht_params *p) {  // lib.rs:241-266
    std::memset(p, 0, sizeof(*p));
    p->d_out = 2;
    p->bandwidth_mode = BHT_BANDWIDTH_FP64;
    p->epochs = 1000;
    p->momentum_switch_epoch = 250;
    p->stop_lying_epoch = 250;
    p->learning_rate = 0.0f;
    p->momentum = 0.5f;
    p->final_momentum = 0.8f;
    p->early_exaggeration = 12.0f;
    p->perplexity = 20.0f;
    p->min_boxes = 10;
    p->max_boxes = 1000;
}

const char *bht_last_error(void) { return g_last_error.c_str(); }

bht_status bht_ctx_create(int device, uint64_t hbm_budget_bytes, bht_ctx **out) {
    if (!out) return fail(BHT_ERR_INVALID, "null output pointer");
    *out = nullptr;
    return guarded([&] {
        int count = 0;
        BHT_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count)
            throw Error(ERR_INVALID, "device index out of range");
        auto ctx = std::make_unique<bht_ctx>();
        ctx->tune = std::make_unique<TuneCache>();
        ctx->device = device;
        ctx->budget = hbm_budget_bytes;
        BHT_HIP(hipSetDevice(device));
        BHT_HIP(hipStreamCreate(&ctx->stream));
        build_twiddles(ctx.get());
        exact_scratch_init(device);
        *out = ctx.release();
    });
}

namespace {

void ctx_free(bht_ctx *ctx) {
    hipSetDevice(ctx->device);
    ctx->release(ctx->tw.table, ctx->tw.total * sizeof(float2));
    ctx->release(ctx->tw.len_of_boxes, (MAX_BOXES + 1) * sizeof(int));
    ctx->release(ctx->tw.off_of_boxes, (MAX_BOXES + 1) * sizeof(int));
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    ctx->trim(0);
    delete ctx;
}

void ctx_acquire(bht_ctx *ctx) {
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->users++;
}

void ctx_drop(bht_ctx *ctx) {
    bool last;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        last = --ctx->users == 0 && ctx->closed;
    }
    if (last) ctx_free(ctx);
}

}  // namespace

void bht_ctx_destroy(bht_ctx *ctx) {
    if (!ctx) return;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        if (ctx->users > 0) {
            ctx->closed = true;
            return;
        }
    }
    ctx_free(ctx);
}

uint64_t bht_ctx_bytes_in_use(const bht_ctx *ctx) {
    return ctx ? const_cast<bht_ctx *>(ctx)->used() : 0;
}

bht_status bht_job_create(bht_ctx *ctx, uint64_t n_samples, const bht_params *params,
                          bht_job **out) {
    if (!ctx || !out) return fail(BHT_ERR_INVALID, "null argument");
    *out = nullptr;
    return guarded([&] {
        bht_params p;
        if (params) p = *params;
        else bht_params_default(&p);
        if (n_samples >= (1ull << 31)) throw Error(ERR_INVALID, "n_samples must be below 2^31");
        static std::atomic<uint64_t> next_serial{0};
        *out = new bht_job(ctx, n_samples, p);
        (*out)->serial = next_serial++;
        ctx_acquire(ctx);
    });
}

void bht_job_destroy(bht_job *job) {
    if (!job) return;
    bht_ctx *ctx = job->ctx;
    hipSetDevice(ctx->device);
    delete job;  // returns its buffers to the context
    ctx_drop(ctx);
}

bht_status bht_job_set_params(bht_job *job, const bht_params *params) {
    if (!job || !params) return fail(BHT_ERR_INVALID, "null argument");
    job->params = *params;
    return BHT_OK;
}

bht_status bht_job_get_params(const bht_job *job, bht_params *out) {
    if (!job || !out) return fail(BHT_ERR_INVALID, "null argument");
    *out = job->params;
    return BHT_OK;
}

bht_status bht_job_set_data_f32(bht_job *job, const float *x, uint64_t n, uint32_t d_in) {
    if (!job || !x) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (n != job->n) throw Error(ERR_INVALID, "data rows must equal n_samples");
        if (d_in == 0) throw Error(ERR_INVALID, "d_in must be positive");
        job->x.ensure(job->ctx, (size_t)n * d_in * 4);
        BHT_HIP(hipMemcpyAsync(job->x.p, x, (size_t)n * d_in * 4, hipMemcpyHostToDevice,
                               job->stream));
        BHT_HIP(hipStreamSynchronize(job->stream));
        job->d_in = d_in;
        job->have_data = true;
    });
}

bht_status bht_job_set_initial_embedding(bht_job *job, const float *y, uint64_t len) {
    if (!job || (!y && len)) return fail(BHT_ERR_INVALID, "null argument");
    job->init.assign(y, y + len);
    job->have_init = true;
    return BHT_OK;
}

bht_status bht_job_exact(bht_job *job, uint32_t metric, bht_epoch_cb cb, void *user) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (metric != BHT_METRIC_EUCLIDEAN && metric != BHT_METRIC_SQEUCLIDEAN)
            throw Error(ERR_INVALID, "unknown metric");
        if (job->dim() < 1 || job->dim() > 7)
            throw Error(ERR_INVALID, "the exact path supports d_out in 1..7");
        check_perplexity(job);
        validate_init(job);
        prepare_exact(job, (int)metric);
        run_exact_epochs(job, job->params.epochs, cb, user);
        finish_run(job);
    });
}

bht_status bht_job_fit_sne(bht_job *job, bht_epoch_cb cb, void *user) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        check_dims(job, FIT_FITSNE);
        check_perplexity(job);
        if (has_cached_affinities(job)) {
            fit(job, FIT_FITSNE, 0.0f, cb, user, true);
            return;
        }
        build_affinities_from_data(job);
        fit(job, FIT_FITSNE, 0.0f, cb, user, false);
    });
}

bht_status bht_job_fit_sne_with_neighbors(bht_job *job, const uint64_t *nbr_index,
                                          const float *nbr_distance, const uint64_t *row_lens,
                                          uint64_t n_rows, bht_epoch_cb cb, void *user) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        check_dims(job, FIT_FITSNE);
        fit_with_neighbors(job, FIT_FITSNE, 0.0f, nbr_index, nbr_distance, row_lens, n_rows, cb,
                           user);
    });
}

bht_status bht_job_barnes_hut(bht_job *job, float theta, bht_epoch_cb cb, void *user) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (!(theta > 0.0f)) throw Error(ERR_INVALID, MSG_THETA);  // lib.rs:1182-1188
        check_perplexity(job);
        check_dims(job, FIT_BH);
        if (has_cached_affinities(job)) {
            fit(job, FIT_BH, theta, cb, user, true);
            return;
        }
        build_affinities_from_data(job);
        fit(job, FIT_BH, theta, cb, user, false);
    });
}

bht_status bht_job_barnes_hut_with_neighbors(bht_job *job, float theta, const uint64_t *nbr_index,
                                             const float *nbr_distance, const uint64_t *row_lens,
                                             uint64_t n_rows, bht_epoch_cb cb, void *user) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (!(theta > 0.0f)) throw Error(ERR_INVALID, MSG_THETA);
        check_perplexity(job);
        check_dims(job, FIT_BH);
        fit_with_neighbors(job, FIT_BH, theta, nbr_index, nbr_distance, row_lens, n_rows, cb,
                           user);
    });
}

bht_status bht_job_get_embedding(const bht_job *job, float *out, uint64_t len) {
    if (!job || !out) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (len != job->n * job->dim()) throw Error(ERR_INVALID, "length must be n * d_out");
        if (!job->y[job->cur].p) throw Error(ERR_STATE, "no embedding yet");
        bht_job *j = const_cast<bht_job *>(job);
        download_points(j, j->ycur(), out);
    });
}

bht_status bht_job_get_affinities(const bht_job *job, uint64_t *rows, uint32_t *cols, float *vals,
                                  uint64_t *nnz, float *perplexity) {
    if (!job || !nnz) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (!job->have_p) throw Error(ERR_STATE, "no affinities");  // affinities() -> None
        if (perplexity) *perplexity = job->params.perplexity;
        if (!rows) {
            *nnz = job->nnz;
            return;
        }
        if (*nnz < job->nnz) throw Error(ERR_INVALID, "output arrays too small");
        std::vector<uint32_t> r, c;
        std::vector<float> v;
        download_csr(const_cast<bht_job *>(job), r, c, &v);
        for (size_t i = 0; i < r.size(); i++) rows[i] = r[i];
        if (cols && job->nnz) std::memcpy(cols, c.data(), job->nnz * 4);
        if (vals) {
            if (job->nnz) std::memcpy(vals, v.data(), job->nnz * 4);
            // lib.rs:1023-1029: affinities() divides the stored P by the exaggeration unless
            // stop-lying has fired, and stop_lying_epoch == 0 never sets the flag (the values it
            // rescaled up front are divided once more).  During a run the device CSR values keep
            // the exaggeration (the epochs rescale their tiled copy; the CSR follows at the end
            // of the run), and stop_lying_fired is set exactly when the stored P lost it, so a
            // prepared run always exports the device values divided by the exaggeration.
            const bool exaggerated = job->prepared || !job->stop_lying_fired;
            if (exaggerated) {
                const float scale = 1.0f / job->params.early_exaggeration;
                for (uint64_t i = 0; i < job->nnz; i++) vals[i] = vals[i] * scale;
            }
        }
        *nnz = job->nnz;
    });
}

bht_status bht_job_set_affinities(bht_job *job, const uint64_t *rows, const uint32_t *cols,
                                  const float *vals, uint64_t n_rows_plus_one, uint64_t nnz,
                                  float perplexity) {
    if (!job || !rows || (nnz && (!cols || !vals))) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        // The kernels index y by these offsets and columns without bounds checks, so the CSR is
        // checked on the host before anything is uploaded.
        check_csr(rows, cols, n_rows_plus_one, nnz);
        // A renumbered job holds its optimiser state in engine order: bring it back to sample
        // order first, since the new CSR is in sample order (the reference keeps y as it is).
        if (job->permuted) unpermute_state(job);
        std::vector<uint32_t> r(n_rows_plus_one);
        for (uint64_t i = 0; i < n_rows_plus_one; i++) r[i] = (uint32_t)rows[i];
        job->rows.ensure(job->ctx, n_rows_plus_one * 4);
        job->cols.ensure(job->ctx, nnz * 4);
        job->vals.ensure(job->ctx, nnz * 4);
        BHT_HIP(hipMemcpyAsync(job->rows.p, r.data(), r.size() * 4, hipMemcpyHostToDevice,
                               job->stream));
        BHT_HIP(hipMemcpyAsync(job->cols.p, cols, nnz * 4, hipMemcpyHostToDevice, job->stream));
        BHT_HIP(hipMemcpyAsync(job->vals.p, vals, nnz * 4, hipMemcpyHostToDevice, job->stream));
        job->nnz = nnz;
        // with_affinities (lib.rs:1043-1051): adopt the perplexity; the cache key is rows.len()
        job->params.perplexity = perplexity;
        job->cached_perplexity = perplexity;
        job->have_cached_perplexity = true;
        job->have_p = n_rows_plus_one == job->n + 1 && n_rows_plus_one > 1;
        if (job->have_p) reset_order(job);
        job->prepared = false;  // a run prepared on the old P is not continued
        l2_sync(job->stream);
        BHT_HIP(hipStreamSynchronize(job->stream));
    });
}

bht_status bht_check_affinities(const uint64_t *rows, const uint32_t *cols,
                                uint64_t n_rows_plus_one, uint64_t nnz) {
    if (!rows || (nnz && !cols)) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] { check_csr(rows, cols, n_rows_plus_one, nnz); });
}

bht_status bht_job_prepare(bht_job *job, uint32_t method, float theta) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (method != BHT_METHOD_FITSNE && method != BHT_METHOD_BARNES_HUT)
            throw Error(ERR_INVALID, "unknown method");
        int m = method == BHT_METHOD_FITSNE ? FIT_FITSNE : FIT_BH;
        if (m == FIT_BH && !(theta > 0.0f)) throw Error(ERR_INVALID, MSG_THETA);
        check_dims(job, m);
        if (!job->have_p) {
            check_perplexity(job);
            build_affinities_from_data(job);
        }
        validate_init(job);
        prepare_run(job, m, theta, true);
    });
}

bht_status bht_job_step(bht_job *job, uint64_t n_epochs, bht_epoch_cb cb, void *user) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] { run_epochs(job, n_epochs, cb, user); });
}

bht_status bht_job_finish(bht_job *job) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] { finish_run(job); });
}

bht_status bht_job_synchronize(bht_job *job) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] { BHT_HIP(hipStreamSynchronize(job->stream)); });
}

bht_status bht_job_forces(bht_job *job, float *attractive, float *repulsive, float *inverse_z) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (!job->prepared) throw Error(ERR_STATE, "bht_job_prepare must precede bht_job_forces");
        if (job->in_batch) throw Error(ERR_STATE, "the job is advanced by its batch");
        repulsion(job);
        attract_tiled(single_set(job), job->pos.as<float>(), false, job->stream);
        EpochScalars sc;
        BHT_HIP(hipMemcpyAsync(&sc, job->scal.p, sizeof(sc), hipMemcpyDeviceToHost, job->stream));
        if (attractive) download_points(job, job->pos.as<float>(), attractive);
        if (repulsive) download_points(job, job->frep.as<float>(), repulsive);
        BHT_HIP(hipStreamSynchronize(job->stream));
        if (inverse_z) *inverse_z = sc.inv_z;
    });
}

bht_status bht_job_get_state(const bht_job *job, float *y, float *uy, float *gains) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (!job->uy.p) throw Error(ERR_STATE, "no optimiser state");
        bht_job *j = const_cast<bht_job *>(job);
        recenter(j);
        if (y) download_points(j, j->ycur(), y);
        if (uy) download_points(j, j->uy.as<float>(), uy);
        if (gains) download_points(j, j->gains.as<float>(), gains);
    });
}

bht_status bht_job_set_state(bht_job *job, const float *y, const float *uy, const float *gains) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (!job->prepared) throw Error(ERR_STATE, "bht_job_prepare must precede set_state");
        if (y) upload_points(job, y, job->ycur());
        if (uy) upload_points(job, uy, job->uy.as<float>());
        if (gains) upload_points(job, gains, job->gains.as<float>());
        // the state is taken as given: no recentring of an earlier update is pending
        EpochScalars *sc = job->scal.as<EpochScalars>();
        BHT_HIP(hipMemsetAsync(&sc->mean_pending, 0, sizeof(int), job->stream));
        l2_sync(job->stream);
        BHT_HIP(hipStreamSynchronize(job->stream));
        job->lazy_ready = false;
    });
}

bht_status bht_job_epoch_form(const bht_job *job, int32_t *attraction_blocks) {
    if (!job || !attraction_blocks) return fail(BHT_ERR_INVALID, "null argument");
    *attraction_blocks = job->last_choice;
    return BHT_OK;
}

bht_status bht_job_run_state(const bht_job *job, uint64_t *next_epoch, int *exaggerated) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    if (next_epoch) *next_epoch = job->epoch;
    if (exaggerated) *exaggerated = device_p_exaggerated(job) ? 1 : 0;
    return BHT_OK;
}

bht_status bht_job_grid_info(const bht_job *job, uint32_t *boxes, uint32_t *fft_len, float *lo,
                             float *hi) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (!job->scal.p) throw Error(ERR_STATE, "no run");
        EpochScalars sc;
        BHT_HIP(hipStreamSynchronize(job->stream));
        BHT_HIP(hipMemcpy(&sc, job->scal.p, sizeof(sc), hipMemcpyDeviceToHost));
        if (boxes) *boxes = (uint32_t)sc.plan.boxes;
        if (fft_len) *fft_len = (uint32_t)sc.plan.len;
        if (lo) *lo = sc.plan.lo;
        if (hi) *hi = sc.plan.hi;
    });
}

bht_status bht_job_set_timing(bht_job *job, int enable) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        BHT_HIP(hipStreamSynchronize(job->stream));
        job->timing = enable != 0;
        job->events_used = 0;
    });
}

bht_status bht_job_get_timing(bht_job *job, double *out3) {
    if (!job || !out3) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        BHT_HIP(hipStreamSynchronize(job->stream));
        BHT_HIP(hipStreamSynchronize(job->side));
        double rep = 0.0, att = 0.0;
        size_t epochs = job->events_used / 4;
        for (size_t e = 0; e < epochs; e++) {
            float a = 0.f, b = 0.f;
            BHT_HIP(hipEventElapsedTime(&a, job->events[4 * e], job->events[4 * e + 1]));
            // forked: the attraction is the first launch on its stream after the fork mark;
            // in order: it follows the repulsion
            BHT_HIP(hipEventElapsedTime(&b, job->events[4 * e + (job->timed_fork ? 0 : 1)],
                                        job->events[4 * e + 2]));
            rep += a;
            att += b;
        }
        out3[0] = rep;
        out3[1] = att;
        out3[2] = (double)epochs;
    });
}

bht_status bht_job_kl_divergence(bht_job *job, double *reference_kl, double *true_kl) {
    if (!job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (job->fit_kind == FIT_NONE) throw Error(ERR_STATE, "no fit yet");  // -> None
        if (job->fit_kind == FIT_EXACT) {  // evaluate_error (tsne/mod.rs:579-618)
            const int saved = job->method;
            job->method = FIT_EXACT;
            const LaunchSet L = single_set(job);
            exact_kl(L.jobs, job->pdense.as<float>(), (int64_t)job->n, (int)job->dim(),
                     job->zrow.as<double>(), job->stream);
            EpochScalars h;
            BHT_HIP(hipMemcpyAsync(&h, job->scal.p, sizeof(h), hipMemcpyDeviceToHost,
                                   job->stream));
            BHT_HIP(hipStreamSynchronize(job->stream));
            job->method = saved;
            if (reference_kl) *reference_kl = h.kl_ref;
            if (true_kl) *true_kl = h.kl_true;
            return;
        }
        // the reference indexes p_values by the sample index (mod.rs:768-770)
        if (job->nnz < job->n)
            throw Error(ERR_INVALID, "kl_divergence needs at least one affinity per sample");
        // Z from the fit's own estimator on the current embedding (mod.rs:635-720)
        const int saved_method = job->method;
        const float saved_theta = job->theta;
        job->method = job->fit_kind;
        job->theta = job->fit_theta;
        if (job->method == FIT_FITSNE) ensure_fitsne_workspace(job);
        else ensure_bh_workspace(job);
        repulsion(job);
        EpochScalars *sc = job->scal.as<EpochScalars>();
        sparse_kl(job->ycur(), (int)job->n, (int)job->dim(), job->rows.as<uint32_t>(),
                  job->cols.as<uint32_t>(), job->vals.as<float>(), job->qv.as<float>(),
                  job->permuted ? job->perm.as<uint32_t>() : nullptr, sc,
                  job->dpart.as<double>(), job->stream);
        EpochScalars h;
        BHT_HIP(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, job->stream));
        BHT_HIP(hipStreamSynchronize(job->stream));
        job->method = saved_method;
        job->theta = saved_theta;
        if (reference_kl) *reference_kl = h.kl_ref;
        if (true_kl) *true_kl = h.kl_true;
    });
}

bht_status bht_knn_f32_ex(bht_ctx *ctx, const float *x, uint64_t n, uint32_t d_in, uint32_t k,
                          uint32_t mode, uint32_t *idx_out, float *dist_out,
                          uint64_t *exact_rows) {
    if (!ctx || !x || !idx_out || !dist_out) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (k == 0 || k >= n) throw Error(ERR_INVALID, "k must be in [1, n - 1]");
        if (mode != BHT_KNN_AUTO && mode != BHT_KNN_EXACT) throw Error(ERR_INVALID, "unknown mode");
        DevBuf dx(ctx), di(ctx), dd(ctx), temp(ctx);
        dx.ensure(ctx, (size_t)n * d_in * 4);
        di.ensure(ctx, (size_t)n * k * 4);
        dd.ensure(ctx, (size_t)n * k * 4);
        BHT_HIP(hipMemcpy(dx.p, x, (size_t)n * d_in * 4, hipMemcpyHostToDevice));
        const size_t tb = mode == BHT_KNN_AUTO ? knn_temp_bytes((int64_t)n, (int)d_in, (int)k) : 0;
        if (tb) temp.ensure(ctx, tb);
        uint64_t rows = n;
        if (tb)
            knn_fast(dx.as<float>(), (int64_t)n, (int)d_in, (int)k, di.as<uint32_t>(),
                     dd.as<float>(), temp.p, ctx->stream, &rows);
        else
            knn_exact(dx.as<float>(), (int64_t)n, (int)d_in, (int)k, di.as<uint32_t>(),
                      dd.as<float>(), ctx->stream);
        BHT_HIP(hipStreamSynchronize(ctx->stream));
        BHT_HIP(hipMemcpy(idx_out, di.p, (size_t)n * k * 4, hipMemcpyDeviceToHost));
        BHT_HIP(hipMemcpy(dist_out, dd.p, (size_t)n * k * 4, hipMemcpyDeviceToHost));
        if (exact_rows) *exact_rows = rows;
    });
}

bht_status bht_exact_repulsion_f64(bht_ctx *ctx, const float *y, uint64_t n, uint32_t d,
                                   double *neg_out, double *z_out) {
    if (!ctx || !y || !neg_out || !z_out) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (d != 1 && d != 2) throw Error(ERR_INVALID, "d must be 1 or 2");
        if (n == 0 || n >= (1ull << 31)) throw Error(ERR_INVALID, "n must be in [1, 2^31)");
        const int nb = exact_repulsion_blocks((int64_t)n);
        DevBuf dy(ctx), dn(ctx), dz(ctx);
        dy.ensure(ctx, (size_t)n * d * 4);
        dn.ensure(ctx, (size_t)n * d * 8);
        dz.ensure(ctx, (size_t)nb * 8);
        BHT_HIP(hipMemcpy(dy.p, y, (size_t)n * d * 4, hipMemcpyHostToDevice));
        exact_repulsion_f64(dy.as<float>(), (int64_t)n, (int)d, dn.as<double>(), dz.as<double>(),
                            ctx->stream);
        BHT_HIP(hipStreamSynchronize(ctx->stream));
        std::vector<double> zp(nb);
        BHT_HIP(hipMemcpy(zp.data(), dz.p, (size_t)nb * 8, hipMemcpyDeviceToHost));
        BHT_HIP(hipMemcpy(neg_out, dn.p, (size_t)n * d * 8, hipMemcpyDeviceToHost));
        double z = 0.0;
        for (double v : zp) z += v;
        *z_out = z;
    });
}

bht_status bht_exact_repulsion_f64_noalloc(bht_ctx *ctx, const float *y_dev, uint64_t n,
                                           uint32_t d, double *neg_dev, double *z_dev,
                                           void *stream) {
    if (!ctx || !y_dev || !neg_dev || !z_dev) return fail(BHT_ERR_INVALID, "null argument");
    if (d != 1 && d != 2) return fail(BHT_ERR_INVALID, "d must be 1 or 2");
    if (n == 0 || n > exact_noalloc_max_points())
        return fail(BHT_ERR_INVALID, "n must be in [1, 16777216] without allocation");
    return guarded([&] {
        BHT_HIP(hipSetDevice(ctx->device));
        hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
        exact_repulsion_f64_noalloc(ctx->device, y_dev, (int64_t)n, (int)d, neg_dev, z_dev, s);
    });
}

bht_status bht_knn_f32(bht_ctx *ctx, const float *x, uint64_t n, uint32_t d_in, uint32_t k,
                       uint32_t *idx_out, float *dist_out) {
    return bht_knn_f32_ex(ctx, x, n, d_in, k, BHT_KNN_AUTO, idx_out, dist_out, nullptr);
}

bht_status bht_search_beta_f32(bht_ctx *ctx, const float *dist, uint64_t n, uint32_t k,
                               float perplexity, uint32_t mode, float *p_out) {
    if (!ctx || !dist || !p_out) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        DevBuf dd(ctx), dp(ctx);
        dd.ensure(ctx, (size_t)n * k * 4);
        dp.ensure(ctx, (size_t)n * k * 4);
        BHT_HIP(hipMemcpy(dd.p, dist, (size_t)n * k * 4, hipMemcpyHostToDevice));
        search_beta(dd.as<float>(), dp.as<float>(), (int64_t)n, (int)k, perplexity, (int)mode,
                    ctx->stream);
        BHT_HIP(hipStreamSynchronize(ctx->stream));
        BHT_HIP(hipMemcpy(p_out, dp.p, (size_t)n * k * 4, hipMemcpyDeviceToHost));
    });
}

bht_status bht_symmetrize(bht_ctx *ctx, const uint32_t *cols, const float *vals, uint64_t n,
                          uint32_t k, uint64_t *rows_out, uint32_t *cols_out, float *vals_out,
                          uint64_t *nnz) {
    if (!ctx || !cols || !vals || !nnz) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        bht_params p;
        bht_params_default(&p);
        bht_job job(ctx, n, p);
        for (uint64_t e = 0; e < n * k; e++)
            if (cols[e] >= n) throw Error(ERR_INVALID, "column index out of range");
        DevBuf dc(ctx), dv(ctx);
        dc.ensure(ctx, n * k * 4);
        dv.ensure(ctx, n * k * 4);
        BHT_HIP(hipMemcpy(dc.p, cols, n * k * 4, hipMemcpyHostToDevice));
        BHT_HIP(hipMemcpy(dv.p, vals, n * k * 4, hipMemcpyHostToDevice));
        // reuse the job's affinity path without the bandwidth search: p is given
        const int64_t nn = (int64_t)n;
        const uint64_t ne = n * k;
        DevBuf pair(ctx), nm(ctx), nmpos(ctx), inc(ctx), rowc(ctx), inoff(ctx), dest(ctx),
            edge(ctx), dest2(ctx), edge2(ctx), tot(ctx), stemp(ctx), rtemp(ctx);
        pair.ensure(ctx, ne * 4);
        nm.ensure(ctx, ne * 4);
        nmpos.ensure(ctx, ne * 4);
        inc.ensure(ctx, n * 4);
        rowc.ensure(ctx, n * 4);
        inoff.ensure(ctx, n * 4);
        dest.ensure(ctx, ne * 4);
        edge.ensure(ctx, ne * 4);
        dest2.ensure(ctx, ne * 4);
        edge2.ensure(ctx, ne * 4);
        tot.ensure(ctx, 16);
        stemp.ensure(ctx, scan_temp_bytes(ne));
        rtemp.ensure(ctx, radix_sort_temp_bytes(ne));
        job.rows.ensure(ctx, (n + 1) * 4);
        SymWork wk;
        wk.pair = pair.as<float>();
        wk.nonmutual = nm.as<uint32_t>();
        wk.nm_pos = nmpos.as<uint32_t>();
        wk.in_count = inc.as<uint32_t>();
        wk.row_count = rowc.as<uint32_t>();
        wk.in_off = inoff.as<uint32_t>();
        wk.rows = job.rows.as<uint32_t>();
        wk.dest = dest.as<uint32_t>();
        wk.edge = edge.as<uint32_t>();
        wk.dest_alt = dest2.as<uint32_t>();
        wk.edge_alt = edge2.as<uint32_t>();
        wk.total = tot.as<uint32_t>();
        wk.scan_temp = stemp.p;
        wk.sort_temp = rtemp.p;
        wk.alloc_out = [&](uint64_t m) {
            job.cols.ensure(ctx, m * 4);
            job.vals.ensure(ctx, m * 4);
            wk.out_cols = job.cols.as<uint32_t>();
            wk.out_vals = job.vals.as<float>();
        };
        uint64_t m = symmetrize_device(dc.as<uint32_t>(), dv.as<float>(), nn, (int)k, wk,
                                       job.stream);
        BHT_HIP(hipStreamSynchronize(job.stream));
        if (!rows_out) {
            *nnz = m;
            return;
        }
        if (*nnz < m) throw Error(ERR_INVALID, "output arrays too small");
        std::vector<uint32_t> r(n + 1);
        BHT_HIP(hipMemcpy(r.data(), job.rows.p, (n + 1) * 4, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i <= n; i++) rows_out[i] = r[i];
        if (cols_out) BHT_HIP(hipMemcpy(cols_out, job.cols.p, m * 4, hipMemcpyDeviceToHost));
        if (vals_out) BHT_HIP(hipMemcpy(vals_out, job.vals.p, m * 4, hipMemcpyDeviceToHost));
        *nnz = m;
    });
}

// ---- batches ---------------------------------------------------------------------------------

bht_status bht_batch_create(bht_ctx *ctx, bht_batch **out) {
    if (!ctx || !out) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        *out = new bht_batch(ctx);
        ctx_acquire(ctx);
    });
}

void bht_batch_destroy(bht_batch *batch) {
    if (!batch) return;
    hipSetDevice(batch->ctx->device);
    hipStreamSynchronize(batch->stream);
    for (bht_job *j : batch->members) {
        j->in_batch = false;
        j->prepared = false;
    }
    bht_ctx *ctx = batch->ctx;
    delete batch;
    ctx_drop(ctx);
}

bht_status bht_batch_add(bht_batch *batch, bht_job *job, uint32_t method, float theta) {
    if (!batch || !job) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        if (method != BHT_METHOD_FITSNE && method != BHT_METHOD_BARNES_HUT)
            throw Error(ERR_INVALID, "unknown method");
        batch_admit(batch, job, method == BHT_METHOD_FITSNE ? FIT_FITSNE : FIT_BH, theta);
    });
}

bht_status bht_batch_step(bht_batch *batch, uint64_t max_epochs, uint64_t *epochs_advanced) {
    if (!batch) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        const uint64_t k = batch_step(batch, max_epochs);
        if (epochs_advanced) *epochs_advanced = k;
    });
}

bht_status bht_batch_take_finished(bht_batch *batch, bht_job **out, uint64_t capacity,
                                   uint64_t *count) {
    if (!batch || !count || (capacity && !out)) return fail(BHT_ERR_INVALID, "null argument");
    const uint64_t m = std::min<uint64_t>(capacity, batch->finished.size());
    for (uint64_t i = 0; i < m; i++) out[i] = batch->finished[i];
    batch->finished.erase(batch->finished.begin(), batch->finished.begin() + (ptrdiff_t)m);
    *count = m;
    return BHT_OK;
}

uint64_t bht_batch_size(const bht_batch *batch) { return batch ? batch->members.size() : 0; }

bht_status bht_batch_set_concurrent(bht_batch *batch, int enable) {
    if (!batch) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        BHT_HIP(hipStreamSynchronize(batch->stream));
        batch->drop_graph();
        batch->fork = enable != 0;
    });
}

bht_status bht_batch_set_graph(bht_batch *batch, int enable) {
    if (!batch) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] {
        BHT_HIP(hipStreamSynchronize(batch->stream));
        batch->drop_graph();
        batch->use_graph = enable != 0;
    });
}

bht_status bht_batch_synchronize(bht_batch *batch) {
    if (!batch) return fail(BHT_ERR_INVALID, "null argument");
    return guarded([&] { BHT_HIP(hipStreamSynchronize(batch->stream)); });
}

bht_status bht_job_footprint(uint64_t n, uint32_t d_in, const bht_params *params, uint32_t method,
                             uint64_t *bytes) {
    if (!bytes) return fail(BHT_ERR_INVALID, "null argument");
    bht_params p;
    if (params) p = *params;
    else bht_params_default(&p);
    const uint64_t D = p.d_out ? p.d_out : 2;
    const uint64_t k = p.perplexity > 0.0f ? (uint64_t)(3.0f * p.perplexity) : 0;
    const uint64_t ne = n * k, nnz = 2 * ne;
    uint64_t b = n * d_in * 4;                      // samples
    b += (n + 1) * 4 + nnz * 8 + n * 4;             // CSR + KL weights
    const TiledSizes tz = tiled_sizes((int)n, (int)D, nnz);
    b += tz.persistent;                             // tiled copy of P
    b += 6 * n * D * 4 + 4096 * 8 + 4096 * 16;      // optimiser state + partials
    b += 2 * (uint64_t)MAX_UPDATE_PARTS * 8;        // update partials
    b += n * D * 4;                                 // attraction sums
    // transients, one at a time: the affinity build, the tiled layout build
    b += std::max<uint64_t>(ne * 12 * 4 + radix_sort_temp_bytes(ne), tz.temp);
    if (!(p.flags & BHT_FLAG_NO_REORDER) && n >= REORDER_MIN_POINTS)
        b += 5 * n * 4 + n * D * 4 + scan_temp_bytes(n);  // renumbering; its transients fit above
    if (method == BHT_METHOD_BARNES_HUT) {
        b += n * (2 * 8 + 2 * 4) + group_sort_temp_bytes(n);  // Morton sort
        if (D >= 5) b += 3 * n * 8;                            // two-word codes
        const uint64_t node_bytes = D >= 3 && D <= 7 ? (uint64_t)nd_node_floats((int)D) * 4 : 16;
        const uint64_t nodes = (D == 2 ? 4 : 2) * n;
        b += nodes * (node_bytes + 8 + 8) + (D == 2 ? n * 4 : 0);  // tree
        b += pad_tile(n) * 24 + group_sort_temp_bytes(pad_tile(n));  // batch arena
    } else {
        uint32_t maxb = std::min<uint32_t>(std::max<uint32_t>(p.max_boxes, 1), MAX_BOXES);
        maxb = std::max(maxb, std::min<uint32_t>(std::max<uint32_t>(p.min_boxes, 1), MAX_BOXES));
        const uint64_t nodes = 3ull * maxb, len = (uint64_t)next_smooth(6 * (int64_t)maxb - 1);
        const uint64_t h = len / 2 + 1;
        b += n * 16 + group_sort_temp_bytes(n);                   // box sort
        b += (D == 2 ? nodes * nodes : nodes) * 16;               // node grid
        if (D == 2) b += 2 * nodes * len * 8;                     // packed spectra
        b += 2 * h * h * 4;                                       // kernel transforms
        b += 2 * ((n + 63) / 64) * 36 * 4;                        // spread partials
        if (D == 2) b += (uint64_t)FFT_FB_BLOCKS * 2 * len * 8;   // global FFT lines
        if (D == 2 && n >= DYN_FFT_MIN_POINTS) b += n * 8;       // sort output
        b += pad_tile(n) * 16 + group_sort_temp_bytes(pad_tile(n));  // batch arena
    }
    *bytes = b;
    return BHT_OK;
}

}  // extern "C"
