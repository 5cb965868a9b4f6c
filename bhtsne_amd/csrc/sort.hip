// Device-wide exclusive scan and stable LSD radix sort of (key, u32 value) pairs.
//
// Used to group points by FIt-SNE box (u32 keys), to order Barnes-Hut Morton codes (u64 keys) and
// to bucket the non-mutual kNN edges by destination row during symmetrisation.  Stability is what
// makes every downstream reduction order (and therefore every result) independent of scheduling.
#include "internal.hpp"

namespace bht {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 4096

// Per-tile sums.
template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tile_sums(const uint32_t *in, uint64_t n,
                                                                 T *sums) {
    __shared__ double red[SCAN_THREADS / 64];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; i++) {
        uint64_t idx = base + (uint64_t)i * SCAN_THREADS + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    // integer sum through doubles is exact below 2^53
    double t = block_sum_f64((double)s, red);
    if (threadIdx.x == 0) sums[blockIdx.x] = (T)t;
}

// Exclusive scan of the tile sums by one block (tiles <= a few hundred thousand).
template <typename T>
__global__ void __launch_bounds__(1024) k_scan_sums(T *sums, uint64_t nt) {
    __shared__ T part[1024];
    uint64_t per = (nt + 1023) / 1024;
    uint64_t b = threadIdx.x * per, e = b + per < nt ? b + per : nt;
    T s = 0;
    for (uint64_t i = b; i < e; i++) s += sums[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
        for (int i = 0; i < 1024; i++) {
            T v = part[i];
            part[i] = run;
            run += v;
        }
    }
    __syncthreads();
    T run = part[threadIdx.x];
    for (uint64_t i = b; i < e; i++) {
        T v = sums[i];
        sums[i] = run;
        run += v;
    }
}

// Rescan each tile with its offset: thread-sequential runs of SCAN_ITEMS, block scan of runs.
template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const uint32_t *in, T *out,
                                                             uint64_t n, const T *sums,
                                                             T *total_out) {
    __shared__ T part[SCAN_THREADS];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    T s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        uint64_t idx = base + i;
        v[i] = idx < n ? in[idx] : 0u;
        s += v[i];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over the 256 runs
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        T t = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    T run = sums[blockIdx.x] + (threadIdx.x ? part[threadIdx.x - 1] : 0);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        uint64_t idx = base + i;
        if (idx < n) out[idx] = run;
        run += v[i];
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1)
        *total_out = run;
}

}  // namespace

size_t scan_temp_bytes(uint64_t n) {
    return (size_t)(ceil_div(n, SCAN_TILE) + 1) * sizeof(uint64_t);
}

template <typename T>
void scan_exclusive(const uint32_t *in, T *out, uint64_t n, void *temp, T *total_out,
                    hipStream_t s) {
    if (n == 0) return;
    unsigned nt = ceil_div(n, SCAN_TILE);
    T *sums = (T *)temp;
    hipLaunchKernelGGL(k_scan_tile_sums<T>, dim3(nt), dim3(SCAN_THREADS), 0, s, in, n, sums);
    hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(1024), 0, s, sums, (uint64_t)nt);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3(nt), dim3(SCAN_THREADS), 0, s, in, out, n, sums,
                       total_out);
    BHT_LAUNCH_CHECK();
}

template void scan_exclusive<uint32_t>(const uint32_t *, uint32_t *, uint64_t, void *, uint32_t *,
                                       hipStream_t);
template void scan_exclusive<uint64_t>(const uint32_t *, uint64_t *, uint64_t, void *, uint64_t *,
                                       hipStream_t);

// ---- radix sort ------------------------------------------------------------------------------

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;  // 2048 keys per tile

static_assert(RS_TILE == RS_TILE_KEYS, "segment alignment must match the tile size");

// Where a tile's keys live and where its digit counts go.  Unsegmented: tile t covers
// [t * RS_TILE, ...) of n keys and its count of digit d sits at d * ntiles + t.  Segmented: the
// tile belongs to one segment and covers a slice of that segment's range; the counts of a segment
// form a block (digit-major) of their own, and the segment's digit totals sit at the block's start
// in a table of the same shape.  Each digit's counts are scanned within their segment, and the
// scatter adds the totals of the smaller digits and the segment's base.
struct TileCtx {
    uint64_t first, end;   // key range of the tile, clipped to the segment
    size_t h0;             // count index of digit 0
    unsigned hstride;      // distance between consecutive digits
    size_t t0;             // index of the segment's digit-0 total
    uint64_t out_base;     // first output position of the segment
};

template <int BINS>
__device__ __forceinline__ TileCtx tile_ctx(uint64_t n, unsigned ntiles, const SortSeg *segs,
                                            const uint32_t *tile_seg) {
    TileCtx c;
    if (segs == nullptr) {
        c.first = (uint64_t)blockIdx.x * RS_TILE;
        c.end = n;
        c.h0 = blockIdx.x;
        c.hstride = ntiles;
        c.t0 = 0;
        c.out_base = 0;
        return c;
    }
    const SortSeg g = segs[tile_seg[blockIdx.x]];
    const unsigned lt = blockIdx.x - g.tile0;
    c.first = g.base + (uint64_t)lt * RS_TILE;
    c.end = g.base + g.n;
    c.h0 = (size_t)BINS * g.tile0 + lt;
    c.hstride = g.ntiles;
    c.t0 = (size_t)BINS * g.tile0;
    c.out_base = g.base;
    return c;
}

// A sort whose pass count is decided on the device: passes from *npass on do nothing (every key's
// digit there is 0), and pass *npass - 1 writes the final buffers.
__device__ __forceinline__ bool pass_skipped(const int *npass, int pass) {
    return npass != nullptr && pass >= *npass;
}

// Digit counts of one tile (BITS-bit digits at `shift`).
template <typename K, int BITS>
__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const K *keys, uint64_t n, int shift,
                                                        uint32_t *hist, unsigned ntiles,
                                                        const SortSeg *segs,
                                                        const uint32_t *tile_seg,
                                                        const int *npass, int pass) {
    constexpr int BINS = 1 << BITS;
    __shared__ uint32_t h[BINS];
    if (pass_skipped(npass, pass)) return;
    const TileCtx tc = tile_ctx<BINS>(n, ntiles, segs, tile_seg);
    for (int b = threadIdx.x; b < BINS; b += RS_THREADS) h[b] = 0;
    __syncthreads();
    for (int r = 0; r < RS_ITEMS; r++) {
        uint64_t idx = tc.first + (uint64_t)r * RS_THREADS + threadIdx.x;
        if (idx < tc.end) atomicAdd(&h[(unsigned)(keys[idx] >> shift) & (BINS - 1u)], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < BINS; b += RS_THREADS)
        hist[tc.h0 + (size_t)b * tc.hstride] = h[b];
}

// In place: each digit's tile counts become their exclusive prefix within the segment (tile
// order), and the digit's total goes to totals.  A wave takes DPW consecutive digits of one
// segment (DPW = 1, or SEG_DPW for segments of at most 64 tiles: a batch's many small segments
// would otherwise launch a wave per (segment, digit)); their counts are loaded DIG_BATCH digits
// at a time.
constexpr int SEG_DPW = 16;
constexpr int DIG_BATCH = 4;

template <int BITS>
__global__ void __launch_bounds__(256) k_rs_digit_scan(uint32_t *hist, uint32_t *totals,
                                                       unsigned ntiles, const SortSeg *segs,
                                                       unsigned nsegs, const int *npass,
                                                       int pass, int dpw) {
    constexpr unsigned BINS = 1u << BITS;
    if (pass_skipped(npass, pass)) return;
    const unsigned lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
    const unsigned groups = BINS / (unsigned)dpw;  // digit groups per segment
    if (wave >= (uint64_t)nsegs * groups) return;
    const unsigned g = (unsigned)(wave / groups), d0 = (unsigned)(wave % groups) * (unsigned)dpw;
    size_t h0, t0;
    unsigned cnt;
    if (segs) {
        const SortSeg sg = segs[g];
        h0 = (size_t)BINS * sg.tile0;
        t0 = (size_t)BINS * sg.tile0;
        cnt = sg.ntiles;
    } else {
        h0 = 0;
        t0 = 0;
        cnt = ntiles;
    }
    if (cnt == 0) return;  // an empty segment: no tile reads its totals
    // each lane scans a contiguous run of a digit's counts; the lanes' sums are scanned across
    // the wave
    const unsigned per = (cnt + 63) / 64, b = lane * per, e = b + per < cnt ? b + per : cnt;
    for (unsigned dq = 0; dq < (unsigned)dpw; dq += DIG_BATCH) {
        const unsigned nd = (unsigned)dpw - dq < (unsigned)DIG_BATCH ? (unsigned)dpw - dq
                                                                     : (unsigned)DIG_BATCH;
        uint32_t sum[DIG_BATCH];
#pragma unroll
        for (int u = 0; u < DIG_BATCH; u++) {
            sum[u] = 0;
            if ((unsigned)u < nd) {
                const size_t hd = h0 + (size_t)(d0 + dq + u) * cnt;
                for (unsigned t = b; t < e; t++) sum[u] += hist[hd + t];
            }
        }
#pragma unroll
        for (int u = 0; u < DIG_BATCH; u++) {
            if ((unsigned)u >= nd) continue;
            uint32_t incl = sum[u];
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t v = __shfl_up(incl, o, 64);
                if ((int)lane >= o) incl += v;
            }
            const size_t hd = h0 + (size_t)(d0 + dq + u) * cnt;
            uint32_t run = incl - sum[u];
            for (unsigned t = b; t < e; t++) {
                const uint32_t v = hist[hd + t];
                hist[hd + t] = run;
                run += v;
            }
            if (lane == 63) totals[t0 + d0 + dq + u] = incl;
        }
    }
}

// Exclusive prefix of the BINS digit totals of a segment into base[] (all threads of the block).
template <int BINS>
__device__ void digit_prefix(const uint32_t *totals, uint32_t *base, uint32_t *wsum) {
    constexpr int PER = BINS >= RS_THREADS ? BINS / RS_THREADS : 1;
    static_assert(BINS < RS_THREADS || BINS % RS_THREADS == 0, "digits per thread");
    const bool active = BINS >= RS_THREADS || (int)threadIdx.x < BINS;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (int i = 0; i < PER; i++) {
        v[i] = active ? totals[threadIdx.x * PER + i] : 0u;
        sum += v[i];
    }
    uint32_t incl = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    for (int q = 0; q < w; q++) run += wsum[q];
    if (!active) return;
#pragma unroll
    for (int i = 0; i < PER; i++) {
        base[threadIdx.x * PER + i] = run;
        run += v[i];
    }
}

// Stable scatter of one tile: rounds of 256 keys; within a round a key's rank among the equal
// digits of its wave comes from one ballot per digit bit, and the waves before it add their
// counts.
template <typename K, int BITS>
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const K *keys, const uint32_t *vals,
                                                           K *keys_out, uint32_t *vals_out,
                                                           uint64_t n, int shift,
                                                           const uint32_t *offs,
                                                           const uint32_t *totals,
                                                           unsigned ntiles, const SortSeg *segs,
                                                           const uint32_t *tile_seg,
                                                           const int *npass, int pass, K *keys_fin,
                                                           uint32_t *vals_fin) {
    constexpr int BINS = 1 << BITS;
    __shared__ uint32_t base[BINS];
    if (pass_skipped(npass, pass)) return;
    if (npass != nullptr && pass == *npass - 1) {
        keys_out = keys_fin;
        vals_out = vals_fin;
    }
    __shared__ uint32_t wcnt[RS_THREADS / 64][BINS];
    __shared__ uint32_t wsum[RS_THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const TileCtx tc = tile_ctx<BINS>(n, ntiles, segs, tile_seg);
    // base[d] = segment base + keys of smaller digits + keys of digit d in earlier tiles
    // every key of the segment has digit 0 in this pass (a high digit of small keys): the stable
    // order is the input order, so the tile is copied
    if (totals[tc.t0] == (uint32_t)(tc.end - tc.out_base)) {
        for (int r = 0; r < RS_ITEMS; r++) {
            const uint64_t idx = tc.first + (uint64_t)r * RS_THREADS + threadIdx.x;
            if (idx < tc.end) {
                keys_out[idx] = keys[idx];
                vals_out[idx] = vals[idx];
            }
        }
        return;
    }
    digit_prefix<BINS>(totals + tc.t0, base, wsum);
    __syncthreads();
    for (int b = threadIdx.x; b < BINS; b += RS_THREADS)
        base[b] += offs[tc.h0 + (size_t)b * tc.hstride] + (uint32_t)tc.out_base;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint64_t tile0 = tc.first;
    n = tc.end;
    for (int r = 0; r < RS_ITEMS; r++) {
        for (int b = threadIdx.x; b < BINS; b += RS_THREADS)
#pragma unroll
            for (int q = 0; q < RS_THREADS / 64; q++) wcnt[q][b] = 0;
        __syncthreads();
        uint64_t idx = tile0 + (uint64_t)r * RS_THREADS + threadIdx.x;
        bool valid = idx < n;
        K key = valid ? keys[idx] : (K)0;
        unsigned d = (unsigned)(key >> shift) & (BINS - 1u);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; b++) {
            uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        unsigned rank = (unsigned)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + rank;
            for (int q = 0; q < w; q++) pos += wcnt[q][d];
            keys_out[pos] = key;
            vals_out[pos] = vals[idx];
        }
        __syncthreads();
        for (int b = threadIdx.x; b < BINS; b += RS_THREADS) {
            uint32_t add = 0;
#pragma unroll
            for (int q = 0; q < RS_THREADS / 64; q++) add += wcnt[q][b];
            base[b] += add;
        }
        __syncthreads();
    }
}

}  // namespace

static size_t temp_bytes_bins(uint64_t n, int bins) {
    uint64_t ntiles = ceil_div(n, RS_TILE);
    uint64_t hist = ntiles * (uint64_t)bins;
    return (size_t)hist * 2 * sizeof(uint32_t) + 256;  // tile counts, digit totals
}

size_t radix_sort_temp_bytes(uint64_t n) { return temp_bytes_bins(n, 1 << 8); }
size_t group_sort_temp_bytes(uint64_t n) { return temp_bytes_bins(n, 1 << GROUP_SORT_BITS); }

// Sorts (keys, vals) by bits [begin_bit, end_bit) of the key, stable, BITS bits per pass.
// Ping-pongs between the two buffer pairs; returns true when the result is in (keys_alt,
// vals_alt).  With npass (device), the passes from *npass on are skipped and the result is in
// (keys_fin, vals_fin).
// Three launches per pass: tile counts, their per-digit scan, the scatter.
template <typename K, int BITS>
static bool radix_passes(K *keys, uint32_t *vals, K *keys_alt, uint32_t *vals_alt, uint64_t n,
                         unsigned ntiles, const SortSeg *segs, unsigned nsegs,
                         const uint32_t *tile_seg, int begin_bit, int end_bit, void *temp,
                         hipStream_t s, const int *npass = nullptr, K *keys_fin = nullptr,
                         uint32_t *vals_fin = nullptr, uint32_t max_seg_tiles = ~0u) {
    const uint64_t nh = (uint64_t)ntiles << BITS;
    uint32_t *hist = (uint32_t *)temp;
    uint32_t *totals = hist + nh;
    // digits per scan wave: SEG_DPW when every segment has at most 64 tiles
    const int dpw = segs != nullptr && max_seg_tiles <= 64 && (1 << BITS) % SEG_DPW == 0
                        ? SEG_DPW : 1;
    const unsigned scan_blocks = (unsigned)ceil_div(((uint64_t)nsegs << BITS) / dpw, 4);
    K *ki = keys, *ko = keys_alt;
    uint32_t *vi = vals, *vo = vals_alt;
    bool alt = false;
    int pass = 0;
    for (int shift = begin_bit; shift < end_bit; shift += BITS, pass++) {
        hipLaunchKernelGGL((k_rs_hist<K, BITS>), dim3(ntiles), dim3(RS_THREADS), 0, s, ki, n,
                           shift, hist, ntiles, segs, tile_seg, npass, pass);
        hipLaunchKernelGGL((k_rs_digit_scan<BITS>), dim3(scan_blocks), dim3(256), 0, s, hist,
                           totals, ntiles, segs, nsegs, npass, pass, dpw);
        hipLaunchKernelGGL((k_rs_scatter<K, BITS>), dim3(ntiles), dim3(RS_THREADS), 0, s, ki, vi,
                           ko, vo, n, shift, hist, totals, ntiles, segs, tile_seg, npass, pass,
                           keys_fin, vals_fin);
        BHT_LAUNCH_CHECK();
        K *tk = ki; ki = ko; ko = tk;
        uint32_t *tv = vi; vi = vo; vo = tv;
        alt = !alt;
    }
    return alt;
}

template <typename K>
bool radix_sort_pairs(K *keys, uint32_t *vals, K *keys_alt, uint32_t *vals_alt, uint64_t n,
                      int begin_bit, int end_bit, void *temp, hipStream_t s) {
    if (n == 0) return false;
    return radix_passes<K, 8>(keys, vals, keys_alt, vals_alt, n, ceil_div(n, RS_TILE), nullptr,
                              1, nullptr, begin_bit, end_bit, temp, s);
}

// The per-epoch sorts, group_sort_bits(end_bit, large)-bit digits; one segment (a job alone) or one
// segment per member (a batch; the arena's tiles are ntiles * RS_TILE keys, temp
// group_sort_temp_bytes of that).
template <typename K, int BITS>
static bool sort_group_bits(const LaunchSet &L, int end_bit, hipStream_t s) {
    if (L.segs) {
        if (L.sort_tiles == 0) return false;
        return radix_passes<K, BITS>((K *)L.keys, L.vals, (K *)L.keys_alt, L.vals_alt,
                                     (uint64_t)L.sort_tiles * RS_TILE, L.sort_tiles, L.segs,
                                     (unsigned)L.njobs, L.tile_seg, 0, end_bit, L.sort_temp, s,
                                     nullptr, nullptr, nullptr, L.max_seg_tiles);
    }
    if (L.sort_n == 0) return false;
    return radix_passes<K, BITS>((K *)L.keys, L.vals, (K *)L.keys_alt, L.vals_alt, L.sort_n,
                                 ceil_div(L.sort_n, RS_TILE), nullptr, 1, nullptr, 0, end_bit,
                                 L.sort_temp, s, L.sort_npass, (K *)L.keys_fin, L.vals_fin);
}

template <typename K>
bool sort_group(const LaunchSet &L, int end_bit, hipStream_t s) {
    const bool large = L.segs == nullptr && L.sort_n >= SORT7_MIN_KEYS;
    if (group_sort_bits(end_bit, large) == 7) return sort_group_bits<K, 7>(L, end_bit, s);
    return sort_group_bits<K, GROUP_SORT_BITS>(L, end_bit, s);
}

template bool radix_sort_pairs<uint32_t>(uint32_t *, uint32_t *, uint32_t *, uint32_t *, uint64_t,
                                         int, int, void *, hipStream_t);
template bool radix_sort_pairs<uint64_t>(uint64_t *, uint32_t *, uint64_t *, uint32_t *, uint64_t,
                                         int, int, void *, hipStream_t);
template bool sort_group<uint32_t>(const LaunchSet &, int, hipStream_t);
template bool sort_group<uint64_t>(const LaunchSet &, int, hipStream_t);

}  // namespace bht
