// Column-tiled layout of the affinity CSR and the attraction kernel that reads it
// (src/tsne/arena.rs:398-427: the attractive half of the gradient, sum_j p_ij q_ij (y_i - y_j)).
//
// The attraction gathers y[col] for every edge.  With the points in the diffusion order of P, the
// columns of a few thousand consecutive rows fall in a handful of index ranges.  So the rows are cut
// into blocks of TILE_ROWS rows (one workgroup each) and the columns into tiles of WIN_FLOATS / D
// points, at most 8192 (one LDS window each), and a block's edges are stored tile by tile: for
// each tile the block copies that tile's positions into LDS once and then reads every gathered
// position from LDS.
//
// An edge takes 6 bytes: its value and a 16-bit key, the column within the tile (13 bits) and the
// step of its row from the previous edge's (3 bits).  Rows are rebuilt in the kernel from the
// group's first row and the steps, with a scan across the wave.
//
// Layout (built once per CSR layout by tiled_build):
//   tkey[p], tval[p]   the edges, sorted by (row block, tile, row), stable, in groups.  Each group
//                      starts at a multiple of 8 and is padded to a multiple of 8 with edges of its
//                      last row (step 0); a row more than 7 past the previous one is reached
//                      through bridge edges of step 7.  Padding and bridges have column 0 and
//                      value 0.
//   blk_seg[b]         first segment of row block b (b = 0 .. nblocks; sentinel = nsegs)
//   seg_tile[s]        tile of segment s (a segment = one block's edges in one tile)
//   seg_e0[s]          first edge of segment s in the sorted order before padding (build only)
//   seg_g0[s]          first group of segment s (sentinel = ngroups)
//   grp_e0[g]          first edge of group g (sentinel = the padded edge count)
//   grp_row[g]         row (within its block) of group g's first edge
// A group is a run of whole rows of a segment, as many as fit one wave pass of GROUP_SPAN edges
// with their bridges and padding, or a single row that is longer.  A row's edges in a segment
// belong to exactly one group, which one wave processes alone, and every lane's 8 edges belong to
// one group or none.
//
// Within a group a wave takes GROUP_SPAN edges per pass, 8 consecutive edges per lane.  Each lane
// sums its edges by row (runs of equal rows), rows complete inside the lane are added to the
// block's LDS accumulators at once, and the rest are joined across lanes by a segmented scan; the
// lane where a row ends adds its sum.  Every accumulator is written by one wave per segment, in a
// fixed order, so the sums are deterministic.  After the last segment the block writes its rows'
// sums out.
#include "internal.hpp"

namespace bht {

namespace {

constexpr int T_THREADS = 512;
constexpr int T_WAVES = T_THREADS / 64;
// At most 1024 rows per block: blocks that short retire often enough that the repulsion, which runs
// beside the attraction, gets CUs (2048-row blocks: 1432 epochs/s at 1M points, 1024: 1517,
// 512: 1473).
constexpr int TILE_ROWS_MAX = 1024;
constexpr int GROUP_SPAN = 512;  // edges per wave pass: 64 lanes x 8
// With the repulsion running beside it, a job's attraction launches at most this many blocks,
// which take row blocks from a work queue, so that the repulsion always finds CUs; 384 of the 512
// that fit the device at once (two per CU) is the best share at 1M points.  The best share moves
// with the two sides' speeds, so the driver's autotuner picks it per job size; this is the value
// it starts from.  Alone on the device the attraction keeps a block per row block.
constexpr int ATT_QUEUE_BLOCKS = 384;
// LDS per block: window + the block's row positions + the accumulators (64,100 bytes at D = 2),
// under 80,064 bytes so that two blocks fit a CU (a block asking for the whole 160 KiB of a CU ran
// at half speed)
constexpr int WIN_FLOATS = 11800;
constexpr uint32_t COL_BITS = 13;      // column within a tile: at most 8192 points per tile
constexpr uint32_t ROW_STEP_MAX = 7;   // 3-bit row steps
constexpr uint32_t NO_ROW = 0xffffffffu;
// the accumulators of a block's rows as (x, y) pairs (y unused for D = 1), then a dummy pair
// per lane (adds that change no row); byte offset of the first dummy pair
constexpr uint32_t ACC_DUMMY_B = TILE_ROWS_MAX * 8;

typedef uint32_t u4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));

int bits_for(uint64_t count) {  // bits to hold 0 .. count - 1
    int b = 0;
    while (b < 64 && (1ull << b) < count) b++;
    return b;
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// ---- layout build ---------------------------------------------------------------------------

struct TileGeom {
    uint32_t rows_per_block, tile_points;
    int row_bits, tile_bits;
};

// key = block << (tile_bits + row_bits) | tile << row_bits | row within block; one wave per row
__global__ void k_tile_keys(const uint32_t *rows, const uint32_t *cols, int n, TileGeom g,
                            uint64_t *keys, uint32_t *idx) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int i = wave; i < n; i += nwaves) {
        const uint64_t b = (uint32_t)i / g.rows_per_block, r = (uint32_t)i % g.rows_per_block;
        for (uint32_t e = rows[i] + lane; e < rows[i + 1]; e += 64) {
            const uint64_t t = cols[e] / g.tile_points;
            keys[e] = (b << (g.tile_bits + g.row_bits)) | (t << g.row_bits) | r;
            idx[e] = e;
        }
    }
}

// Packs the sorted edges' columns (within their tiles) and values in sorted order and flags the
// first edge of every segment and of every (segment, row) run.
__global__ void k_tile_pack(const uint64_t *keys, const uint32_t *idx, const uint32_t *cols,
                            const float *vals, uint64_t nnz, TileGeom g, uint16_t *ucol,
                            float *uval, uint32_t *seg_flag, uint32_t *run_flag) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < nnz; p += stride) {
        const uint64_t k = keys[p];
        const uint32_t e = idx[p];
        const uint32_t t = (uint32_t)((k >> g.row_bits) & ((1ull << g.tile_bits) - 1));
        ucol[p] = (uint16_t)(cols[e] - t * g.tile_points);
        uval[p] = vals[e];
        const uint64_t prev = p > 0 ? keys[p - 1] : ~0ull;
        seg_flag[p] = (p == 0 || (prev >> g.row_bits) != (k >> g.row_bits)) ? 1u : 0u;
        run_flag[p] = (p == 0 || prev != k) ? 1u : 0u;
    }
}

// Segment table from the flags and their exclusive scan; the runs compacted the same way, with
// the row (within its block) of each.
__global__ void k_tile_segments(const uint64_t *keys, const uint32_t *seg_flag,
                                const uint32_t *seg_pos, const uint32_t *run_flag,
                                const uint32_t *run_pos, uint64_t nnz, TileGeom g,
                                uint32_t *seg_e0, uint32_t *seg_tile, uint32_t *seg_blk,
                                uint32_t *runs, uint32_t *run_row) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < nnz; p += stride) {
        const uint64_t k = keys[p];
        if (seg_flag[p]) {
            const uint32_t s = seg_pos[p];
            seg_e0[s] = (uint32_t)p;
            seg_tile[s] = (uint32_t)((k >> g.row_bits) & ((1ull << g.tile_bits) - 1));
            seg_blk[s] = (uint32_t)(k >> (g.row_bits + g.tile_bits));
        }
        if (run_flag[p]) {
            const uint32_t r = run_pos[p];
            runs[r] = (uint32_t)p;
            run_row[r] = (uint32_t)(k & ((1ull << g.row_bits) - 1));
        }
    }
}

__host__ __device__ constexpr uint32_t round8(uint32_t v) { return (v + 7u) & ~7u; }

// Splits the runs [r0, r1) of one segment (its edges end at seg_end) into groups, greedily: a
// group takes whole runs while its edges, the bridge edges between its runs and the padding to a
// multiple of 8 fit one GROUP_SPAN-edge pass; a run longer than that is a group of its own.  A run
// whose row is more than ROW_STEP_MAX past the previous run's gets (gap - 1) / ROW_STEP_MAX bridge
// edges before it.  Calls run_fn(r, group, offset of the run's first edge in its group, bridges,
// row step of its first edge) for every run and group_fn(group, first run, edges with bridges)
// for every group; returns the group count.
template <typename RF, typename GF>
__device__ uint32_t walk_groups(const uint32_t *runs, const uint32_t *run_row, uint32_t r0,
                                uint32_t r1, uint32_t seg_end, RF run_fn, GF group_fn) {
    uint32_t ng = 0, first = r0, glen = 0;
    for (uint32_t r = r0; r < r1; r++) {
        const uint32_t len = (r + 1 < r1 ? runs[r + 1] : seg_end) - runs[r];
        uint32_t nb = 0, step = 0;
        if (r > first) {
            const uint32_t gap = run_row[r] - run_row[r - 1];
            nb = (gap - 1) / ROW_STEP_MAX;
            step = gap - nb * ROW_STEP_MAX;
            if (round8(glen + nb + len) > (uint32_t)GROUP_SPAN) {  // run r opens the next group
                group_fn(ng, first, glen);
                ng++;
                first = r;
                glen = 0;
                nb = 0;
                step = 0;
            }
        }
        run_fn(r, ng, glen + nb, nb, step);
        glen += nb + len;
    }
    group_fn(ng, first, glen);
    return ng + 1;
}

// Groups per segment, zero past the last segment (the counts are scanned over the whole table),
// and the segment sentinel.  The first edge of a segment starts a run: its run index is run_pos
// there.
__global__ void k_tile_seg_groups(uint32_t *seg_e0, const uint32_t *run_pos, const uint32_t *runs,
                                  const uint32_t *run_row, const uint32_t *counts, uint64_t nnz,
                                  uint32_t max_segs, uint32_t *seg_ngrp) {
    const uint32_t nsegs = counts[0], nruns = counts[1];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < max_segs; s += stride) {
        if (s >= nsegs) {
            if (s == nsegs) seg_e0[s] = (uint32_t)nnz;
            seg_ngrp[s] = 0;
            continue;
        }
        const bool last = s + 1 == nsegs;
        const uint32_t e1 = last ? (uint32_t)nnz : seg_e0[s + 1];
        const uint32_t r0 = run_pos[seg_e0[s]], r1 = last ? nruns : run_pos[e1];
        seg_ngrp[s] = walk_groups(runs, run_row, r0, r1, e1,
                                  [](uint32_t, uint32_t, uint32_t, uint32_t, uint32_t) {},
                                  [](uint32_t, uint32_t, uint32_t) {});
    }
}

// blk_seg[b] = first segment of block b or later; per run its group, offset in the group and
// (bridges << 3 | row step); per group its edges with bridges and its first row.
__global__ void k_tile_index(const uint32_t *seg_blk, const uint32_t *seg_e0,
                             const uint32_t *seg_g0, const uint32_t *run_pos, const uint32_t *runs,
                             const uint32_t *run_row, const uint32_t *counts, uint32_t nblocks,
                             uint32_t *blk_seg, uint32_t *run_grp, uint32_t *run_off,
                             uint32_t *run_bs, uint32_t *glen, uint32_t *grp_row) {
    const uint32_t nsegs = counts[0], nruns = counts[1];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t b = tid; b <= nblocks; b += stride)
        blk_seg[b] = b == nblocks ? nsegs : lower_bound_u32(seg_blk, nsegs, b);
    for (uint32_t s = tid; s < nsegs; s += stride) {
        const uint32_t g0 = seg_g0[s], e1 = seg_e0[s + 1];
        const uint32_t r0 = run_pos[seg_e0[s]], r1 = s + 1 == nsegs ? nruns : run_pos[e1];
        walk_groups(
            runs, run_row, r0, r1, e1,
            [&](uint32_t r, uint32_t gi, uint32_t off, uint32_t nb, uint32_t step) {
                run_grp[r] = g0 + gi;
                run_off[r] = off;
                run_bs[r] = (nb << 3) | step;
            },
            [&](uint32_t gi, uint32_t first, uint32_t len) {
                glen[g0 + gi] = len;
                grp_row[g0 + gi] = run_row[first];
            });
    }
}

// Padded length of every group (a multiple of 8); zero past the last group.
__global__ void k_tile_pad_len(const uint32_t *glen, const uint32_t *seg_g0,
                               const uint32_t *counts, uint32_t max_groups, uint32_t *plen) {
    const uint32_t ngroups = seg_g0[counts[0]];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < max_groups; g += stride)
        plen[g] = g < ngroups ? round8(glen[g]) : 0u;
}

// Every edge to its place: key = column within the tile | row step << 13 (the step is that of the
// run's first edge, 0 for the rest of the run).
__global__ void k_tile_scatter(const uint16_t *ucol, const float *uval, const uint32_t *run_flag,
                               const uint32_t *run_pos, const uint32_t *runs,
                               const uint32_t *run_grp, const uint32_t *run_off,
                               const uint32_t *run_bs, const uint32_t *grp_e0, uint64_t nnz,
                               uint16_t *tkey, float *tval) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < nnz; p += stride) {
        const uint32_t r = run_pos[p] + run_flag[p] - 1u;
        const uint32_t first = runs[r];
        const uint32_t out = grp_e0[run_grp[r]] + run_off[r] + ((uint32_t)p - first);
        const uint32_t step = (uint32_t)p == first ? (run_bs[r] & 7u) : 0u;
        tkey[out] = (uint16_t)(ucol[p] | (step << 13));
        tval[out] = uval[p];
    }
}

// The bridge edges before the runs that need them (row step ROW_STEP_MAX) and the padding after
// every group (row step 0, the group's last row): column 0, value 0.  They add +-0 to a row's
// sum, which changes no bit.
__global__ void k_tile_fill(const uint32_t *runs_grp, const uint32_t *run_off,
                            const uint32_t *run_bs, const uint32_t *glen, const uint32_t *seg_g0,
                            const uint32_t *grp_e0, const uint32_t *counts, uint16_t *tkey,
                            float *tval) {
    const uint32_t nruns = counts[1], ngroups = seg_g0[counts[0]];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t r = tid; r < nruns; r += stride) {
        const uint32_t nb = run_bs[r] >> 3;
        const uint32_t at = grp_e0[runs_grp[r]] + run_off[r] - nb;
        for (uint32_t i = 0; i < nb; i++) {
            tkey[at + i] = (uint16_t)(ROW_STEP_MAX << 13);
            tval[at + i] = 0.0f;
        }
    }
    for (uint32_t g = tid; g < ngroups; g += stride) {
        for (uint32_t q = grp_e0[g] + glen[g]; q < grp_e0[g + 1]; q++) {
            tkey[q] = 0;
            tval[q] = 0.0f;
        }
    }
}

// ---- attraction -------------------------------------------------------------------------------

// The position at byte offset `off` of an LDS array of points (D floats each).
template <int D>
__device__ __forceinline__ float2 lds_pos_at(const float *a, uint32_t off) {
    BHT_LDS const char *b = (BHT_LDS const char *)lds_view(a) + off;
    if (D == 2) {
        const f2v v = *(BHT_LDS const f2v *)b;
        return make_float2(v.x, v.y);
    }
    return make_float2(*(BHT_LDS const float *)b, 0.0f);
}

// Copies points [first, first + count) of y (D floats each) into LDS, less the epoch's pending
// recentring shift (s0, s1); 16-byte loads when the start is 16-byte aligned.  A thread issues
// STAGE_BATCH loads before its first LDS store, so that a window (at most 2,950 float4s, under
// six per thread) costs one wait for memory instead of one per float4: with one block per CU
// (the attraction beside the repulsion) nothing else of the block runs while it stages.
constexpr int STAGE_BATCH = 8;

template <int D>
__device__ void stage_points(BHT_GLOBAL const float *y, uint32_t first, uint32_t count,
                             float *dst, float s0, float s1) {
    const uint32_t nf = count * D;
    BHT_GLOBAL const float *src = y + (size_t)first * D;
    BHT_LDS float *out = lds_view(dst);
    // coordinate t of the copy is axis t % D (the start is a point boundary)
    const f4v sub = D == 2 ? f4v{s0, s1, s0, s1} : f4v{s0, s0, s0, s0};
    if (((first * D) & 3u) == 0) {
        const uint32_t nf4 = nf / 4;  // whole float4s; the last nf % 4 floats one by one
        for (uint32_t k0 = threadIdx.x; k0 < nf4; k0 += STAGE_BATCH * T_THREADS) {
            f4v v[STAGE_BATCH];
#pragma unroll
            for (int u = 0; u < STAGE_BATCH; u++) {
                const uint32_t k = k0 + u * T_THREADS;
                if (k < nf4) v[u] = ((BHT_GLOBAL const f4v *)src)[k];
            }
#pragma unroll
            for (int u = 0; u < STAGE_BATCH; u++) {
                const uint32_t k = k0 + u * T_THREADS;
                if (k < nf4) ((BHT_LDS f4v *)out)[k] = v[u] - sub;
            }
        }
        if (threadIdx.x < (nf & 3u)) {
            const uint32_t t = (nf & ~3u) + threadIdx.x;
            out[t] = src[t] - (D == 2 && (t & 1u) ? s1 : s0);
        }
    } else {
        for (uint32_t k = threadIdx.x; k < nf; k += T_THREADS)
            out[k] = src[k] - (D == 2 && (k & 1u) ? s1 : s0);
    }
}

// s_waitcnt vmcnt(0): waits for every vector memory access of the wave in flight.
__device__ __forceinline__ void vm_wait_all() { __builtin_amdgcn_s_waitcnt(0x0f70); }

// A lane's 8 consecutive edges: the 16-bit keys two to a word (edge 2j in the low half of kw[j])
// and the values.
struct EdgeLoad {
    uint32_t kw[4];
    float pv[8];
};

__device__ __forceinline__ EdgeLoad load_edges(BHT_GLOBAL const uint16_t *tkey,
                                               BHT_GLOBAL const float *tval, uint32_t e0,
                                               uint32_t nnz) {
    EdgeLoad r;
    if (e0 + 8u <= nnz) {
        // plain loads: nontemporal ones measured 3% slower over the whole epoch
        const u4v k = *(BHT_GLOBAL const u4v *)(tkey + e0);
        const f4v v0 = *(BHT_GLOBAL const f4v *)(tval + e0);
        const f4v v1 = *(BHT_GLOBAL const f4v *)(tval + e0 + 4);
        r.kw[0] = k.x; r.kw[1] = k.y; r.kw[2] = k.z; r.kw[3] = k.w;
        r.pv[0] = v0.x; r.pv[1] = v0.y; r.pv[2] = v0.z; r.pv[3] = v0.w;
        r.pv[4] = v1.x; r.pv[5] = v1.y; r.pv[6] = v1.z; r.pv[7] = v1.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t lo = e0 + 2 * j < nnz ? tkey[e0 + 2 * j] : 0u;
            const uint32_t hi = e0 + 2 * j + 1 < nnz ? tkey[e0 + 2 * j + 1] : 0u;
            r.kw[j] = lo | (hi << 16);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) r.pv[u] = e0 + u < nnz ? tval[e0 + u] : 0.0f;
    }
    return r;
}

// Inclusive sum of v over the wave.
__device__ __forceinline__ uint32_t wave_incl_add(uint32_t v);

// The pair at byte offset at[u] of acc += (v0[u], v1[u]) for the lane's eight adds: all reads,
// then all writes (the eight offsets are distinct rows or the lane's dummy pair).  LDS atomic adds
// instead (the same order per row) ran slower: 1.76 ms at 1M points, 0.27 ms with the dummy adds
// masked off, against 0.26 ms.
template <int D>
__device__ __forceinline__ void acc_add8(float *acc, const uint32_t *at, const float *v0,
                                         const float *v1) {
    BHT_LDS char *a = (BHT_LDS char *)lds_view(acc);
    f2v cur[8];
#pragma unroll
    for (int u = 0; u < 8; u++) cur[u] = *(BHT_LDS f2v *)(a + at[u]);
#pragma unroll
    for (int u = 0; u < 8; u++) {
        cur[u].x = cur[u].x + v0[u];
        if (D == 2) cur[u].y = cur[u].y + v1[u];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) *(BHT_LDS f2v *)(a + at[u]) = cur[u];
}

// DPP controls (wave64): shifts within rows of 16 lanes, broadcasts of a row's last lane to the
// rows after it, and whole-wave shifts by one lane.
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114,
              DPP_ROW_SHR8 = 0x118, DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143,
              DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f(float old, float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL,
                                                      ROW_MASK, 0xf, false));
}

// One step of a segmented inclusive scan (sums restart where flag is set): each lane combines
// the value DPP moves to it from the lane `CTRL` names (none: -0.0 and no flag, which change
// nothing) with its own.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void seg_step(float &s0, float &s1, int &flag) {
    const float t0 = dpp_f<CTRL, ROW_MASK>(-0.0f, s0);
    const float t1 = dpp_f<CTRL, ROW_MASK>(-0.0f, s1);
    const int tf = __builtin_amdgcn_update_dpp(0, flag, CTRL, ROW_MASK, 0xf, false);
    if (!flag) {
        s0 = t0 + s0;
        s1 = t1 + s1;
    }
    flag |= tf;
}

// Segmented inclusive scan of (s0, s1) over the wave: within rows of 16 lanes, then across rows.
// It stops as soon as every lane's segment is complete.
__device__ __forceinline__ void seg_scan(float &s0, float &s1, int &flag) {
    if (__builtin_amdgcn_ballot_w64(flag == 0) == 0) return;
    seg_step<DPP_ROW_SHR1, 0xf>(s0, s1, flag);
    if (__builtin_amdgcn_ballot_w64(flag == 0) == 0) return;
    seg_step<DPP_ROW_SHR2, 0xf>(s0, s1, flag);
    if (__builtin_amdgcn_ballot_w64(flag == 0) == 0) return;
    seg_step<DPP_ROW_SHR4, 0xf>(s0, s1, flag);
    if (__builtin_amdgcn_ballot_w64(flag == 0) == 0) return;
    seg_step<DPP_ROW_SHR8, 0xf>(s0, s1, flag);
    if (__builtin_amdgcn_ballot_w64(flag == 0) == 0) return;
    seg_step<DPP_ROW_BCAST15, 0xa>(s0, s1, flag);
    if (__builtin_amdgcn_ballot_w64(flag == 0) == 0) return;
    seg_step<DPP_ROW_BCAST31, 0xc>(s0, s1, flag);
}

__device__ __forceinline__ uint32_t wave_incl_add(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_SHR1, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_SHR2, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_SHR4, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_SHR8, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_BCAST15, 0xa, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_BCAST31, 0xc, 0xf, false);
    return v;
}

template <int D, bool EPOCH>
__global__ void __launch_bounds__(T_THREADS) k_attract_tiled(const JobDev *jobs,
                                                             const uint32_t *map,
                                                             float *pos_out, int queued,
                                                             int self_head) {
    // static LDS: the arrays sit at fixed addresses, which fold into the instructions' offsets
    __shared__ __align__(16) float win[WIN_FLOATS];
    __shared__ __align__(16) float yrow[TILE_ROWS_MAX * D];
    __shared__ __align__(16) float acc[TILE_ROWS_MAX * 2 + 64 * 2];  // (x, y) pairs
    __shared__ uint32_t next_lb;
    // The attraction is the longer of the epoch's two streams: its waves take the SIMDs' issue
    // slots first, and the repulsion's kernels, which have slack, fill the gaps.
    __builtin_amdgcn_s_setprio(3);
    const BlockCtx bc = block_ctx(map, jobs, F_ATT);
    const JobDev &J = jobs[bc.job];
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // self_head: the epoch's head (the recentring mean, the epoch index) runs beside this kernel
    // on the repulsion's stream, so the kernel derives both from the last update's data itself
    const int epoch = __builtin_amdgcn_readfirstlane(self_head ? J.scal->epoch_next : J.scal->epoch);
    const uint32_t n = (uint32_t)J.n;
    const uint32_t R = J.tile_rows, W = J.tile_points;
    const uint32_t nnz = J.tile_nnz;
    BHT_GLOBAL const float *y = global_view((const float *)J.ybuf[epoch & 1]);
    BHT_GLOBAL const uint16_t *tkey = global_view(J.tkey);
    BHT_GLOBAL float *tval = global_view(J.tval);
    BHT_GLOBAL float *out = global_view(EPOCH ? J.fattr : pos_out);
    BHT_GLOBAL const uint32_t *blk_seg = global_view(J.blk_seg), *seg_tile = global_view(J.seg_tile);
    BHT_GLOBAL const uint32_t *seg_g0 = global_view(J.seg_g0), *grp_e0 = global_view(J.grp_e0);
    BHT_GLOBAL const uint32_t *grp_row = global_view(J.grp_row);
    // stop_lying fires after epoch == stop_lying_epoch (lib.rs:1168-1171): the first epoch to
    // read the rescaled values rescales them in place as it reads
    const bool stop_now = EPOCH && J.stop_lying_epoch != 0 && epoch == J.stop_lying_epoch + 1;
    const float inv_exag = J.inv_exag;
    float ys0, ys1;
    if (self_head) {
        pending_mean2(J, ys0, ys1);
    } else {
        ys0 = J.scal->ysub[0];
        ys1 = J.scal->ysub[1];
    }

    // queued: the blocks take row blocks from the job's work queue (reset by the update kernel);
    // otherwise each block has its own row block
    const uint32_t nrb = (n + R - 1) / R;
    uint32_t lb = (uint32_t)bc.lb;
    if (queued) {
        if (threadIdx.x == 0) next_lb = atomicAdd(&J.scal->att_queue, 1u);
        __syncthreads();
        lb = next_lb;
    }
    while (lb < nrb) {
        const uint32_t row0 = lb * R;
        const uint32_t nrows = row0 < n ? (n - row0 < R ? n - row0 : R) : 0u;
        stage_points<D>(y, row0, nrows, yrow, ys0, ys1);
        for (uint32_t k = threadIdx.x; k < nrows * 2; k += T_THREADS) lds_view(acc)[k] = 0.0f;
        const uint32_t s_begin = blk_seg[lb], s_end = blk_seg[lb + 1];
        for (uint32_t s = s_begin; s < s_end; s++) {
            const uint32_t t = seg_tile[s];
            const uint32_t first = t * W, count = n - first < W ? n - first : W;
            const uint32_t g_end = seg_g0[s + 1];
            // The wave's groups of this segment, 64 at a time: lane k holds the edge range of the
            // wave's k-th group.  Each pass's edges are loaded while the pass before is summed,
            // and the first pass's while the tile is copied into LDS.
            uint32_t gc = seg_g0[s] + wid;
            uint32_t hb = 0, he = 0, hr = 0, ng = 0, k = 0, e_end = 0, base = 0, grow8 = 0;
            EdgeLoad cur = {};
            auto open_groups = [&]() {
                const uint32_t gk = gc + (uint32_t)lane * T_WAVES;
                hb = gk < g_end ? grp_e0[gk] : 0u;
                he = gk < g_end ? grp_e0[gk + 1] : 0u;
                hr = gk < g_end ? grp_row[gk] * 8u : 0u;
                const uint32_t left = (g_end - gc + T_WAVES - 1) / T_WAVES;
                ng = left < 64u ? left : 64u;
                k = 0;
                base = (uint32_t)__builtin_amdgcn_readlane((int)hb, 0);
                e_end = (uint32_t)__builtin_amdgcn_readlane((int)he, 0);
                grow8 = (uint32_t)__builtin_amdgcn_readlane((int)hr, 0);
                cur = load_edges(tkey, tval, base + 8u * lane, nnz);
            };
            if (gc < g_end) open_groups();
            __syncthreads();  // the previous tile's reads are done
            stage_points<D>(y, first, count, win, ys0, ys1);
            __syncthreads();
            // The first pass's edges have landed (the copy waited for its loads).  Said
            // explicitly, so that the compiler's wait at the top of the pass loop does not also
            // cover the next pass's loads, issued just before it.
            vm_wait_all();
            while (gc < g_end) {
                while (true) {
                    // the pass after this one: the group's next 512 edges or the next group
                    uint32_t n_end = e_end, n_base = base + GROUP_SPAN, nk = k, n_grow8 = grow8;
                    bool more = n_base < e_end;
                    if (!more && k + 1 < ng) {
                        nk = k + 1;
                        n_base = (uint32_t)__builtin_amdgcn_readlane((int)hb, (int)nk);
                        n_end = (uint32_t)__builtin_amdgcn_readlane((int)he, (int)nk);
                        n_grow8 = (uint32_t)__builtin_amdgcn_readlane((int)hr, (int)nk);
                        more = true;
                    }
                    EdgeLoad nxt = {};
                    if (more) nxt = load_edges(tkey, tval, n_base + 8u * lane, nnz);
                    // A group starts at a multiple of 8 and holds a multiple of 8 edges, so a
                    // lane's edges are all in the group or all past its end.  The lanes past the
                    // end (the last lanes of the wave) hold real edges of other groups (columns
                    // inside the window); their row steps are left out of the scan, so their rows
                    // stay within 56 of the group's last and their reads inside the LDS block, and
                    // what they compute reaches no row.
                    const uint32_t e0 = base + 8u * lane;
                    const bool lane_ok = e0 < e_end;
                    // rows and columns as LDS byte offsets: the row's accumulator pair (8 bytes
                    // per row) and the column's position in the window (4 D bytes per point)
                    uint32_t row[8], col[8];
                    {
                        uint32_t run = 0;  // rows past the lane's first
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            const uint32_t w = cur.kw[j];
                            col[2 * j] = (w & ((1u << COL_BITS) - 1)) * (4u * D);
                            col[2 * j + 1] = ((w >> 16) & ((1u << COL_BITS) - 1)) * (4u * D);
                            run += (w >> COL_BITS) & 7u;
                            row[2 * j] = run;
                            run += w >> (COL_BITS + 16);
                            row[2 * j + 1] = run;
                        }
                        const uint32_t mine = lane_ok ? run : 0u;
                        const uint32_t before = grow8 + ((wave_incl_add(mine) - mine) << 3);
#pragma unroll
                        for (int u = 0; u < 8; u++) row[u] = before + (row[u] << 3);
                    }
                    float *pv = cur.pv;
                    if (stop_now) {
#pragma unroll
                        for (int u = 0; u < 8; u++) pv[u] = pv[u] * inv_exag;
                    }
                    if (stop_now && lane_ok) {
                        *(BHT_GLOBAL f4v *)(tval + e0) = f4v{pv[0], pv[1], pv[2], pv[3]};
                        *(BHT_GLOBAL f4v *)(tval + e0 + 4) = f4v{pv[4], pv[5], pv[6], pv[7]};
                    }
                    // p q (y_i - y_j), q = 1 / (1 + |y_i - y_j|^2) (hardware reciprocal)
                    float f0[8], f1[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const float2 yi = lds_pos_at<D>(yrow, D == 2 ? row[u] : row[u] >> 1);
                        const float2 yj = lds_pos_at<D>(win, col[u]);
                        const float d0 = yi.x - yj.x;
                        const float d1 = D == 2 ? yi.y - yj.y : 0.0f;
                        float dist = d0 * d0;
                        if (D == 2) dist = dist + d1 * d1;
                        const float f = pv[u] * __builtin_amdgcn_rcpf(dist + 1.0f);
                        f0[u] = f * d0;
                        f1[u] = f * d1;
                    }
                    // running sums of the lane's runs of equal rows (rows never decrease)
#pragma unroll
                    for (int u = 1; u < 8; u++) {
                        const bool same = row[u] == row[u - 1];
                        f0[u] = same ? f0[u - 1] + f0[u] : f0[u];
                        f1[u] = same ? f1[u - 1] + f1[u] : f1[u];
                    }
                    // the lanes' tail runs joined by a segmented scan across the wave (a lane in
                    // the group follows lanes in the group only)
                    const uint32_t tail_row = row[7];
                    const uint32_t prev_tail = (uint32_t)__builtin_amdgcn_update_dpp(
                        (int)NO_ROW, (int)tail_row, DPP_WAVE_SHR1, 0xf, 0xf, false);
                    int flag = (lane == 0 || prev_tail != tail_row || !lane_ok) ? 1 : 0;
                    float s0 = f0[7], s1 = f1[7];
                    seg_scan(s0, s1, flag);
                    // what the lanes before hold of this lane's head row
                    const float p0 = dpp_f<DPP_WAVE_SHR1, 0xf>(-0.0f, s0);
                    const float p1 = dpp_f<DPP_WAVE_SHR1, 0xf>(-0.0f, s1);
                    const bool joins = lane > 0 && prev_tail == row[0];
                    const uint32_t next_head = (uint32_t)__builtin_amdgcn_update_dpp(
                        (int)NO_ROW, (int)(lane_ok ? row[0] : NO_ROW), DPP_WAVE_SHL1, 0xf, 0xf,
                        false);
                    // every run that ends in this lane is added to its row once; the lane's other
                    // seven adds go to its own dummy pair.  No two lanes of the wave touch the
                    // same row here, and no other wave touches these rows in this segment, so
                    // the read-add-write needs no atomics (what several waves leave in a dummy
                    // pair is never read).
                    const uint32_t own = ACC_DUMMY_B + (uint32_t)lane * 8u;
                    uint32_t at[8];
                    float v0[8], v1[8];
#pragma unroll
                    for (int u = 0; u < 7; u++) {
                        const bool ends = row[u] != row[u + 1];
                        const bool head = joins && row[u] == row[0];
                        at[u] = ends ? row[u] : own;
                        v0[u] = head ? p0 + f0[u] : f0[u];
                        v1[u] = head ? p1 + f1[u] : f1[u];
                    }
                    at[7] = next_head != tail_row ? tail_row : own;
                    v0[7] = s0;
                    v1[7] = s1;
                    if (lane_ok) acc_add8<D>(acc, at, v0, v1);
                    if (!more) break;
                    cur = nxt;
                    base = n_base;
                    e_end = n_end;
                    grow8 = n_grow8;
                    k = nk;
                }
                gc += 64u * T_WAVES;
                if (gc < g_end) {
                    open_groups();
                    vm_wait_all();
                }
            }
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < nrows * D; k += T_THREADS)
            out[(size_t)row0 * D + k] = lds_view(acc)[D == 2 ? k : 2 * k];
        __syncthreads();  // the sums are out before the next row block's rows overwrite them
        if (queued) {
            if (threadIdx.x == 0) next_lb = atomicAdd(&J.scal->att_queue, 1u);
            __syncthreads();
            lb = next_lb;
        } else {
            lb += (uint32_t)bc.nb;
        }
    }
}

}  // namespace

uint32_t tiled_rows_per_block(int n) {
    // about 128 blocks for a job of 8,192 points or more (a large job reaches TILE_ROWS_MAX rows
    // per block); a power of two.  Every block copies each column tile its rows touch into LDS,
    // so short blocks copy more than they read of P: a 50,000-point job in 128-row blocks copied
    // about 80 MB of positions for 45 MB of edges.
    uint32_t r = 64;
    while (r < TILE_ROWS_MAX && (uint64_t)r * 128 < (uint64_t)(n > 0 ? n : 1)) r <<= 1;
    return r;
}

uint32_t tiled_tile_points(int dim) {
    return (uint32_t)std::min(WIN_FLOATS / dim, 1 << COL_BITS);
}

int tiled_blocks(int n) {
    const uint32_t r = tiled_rows_per_block(n);
    return (int)ceil_div((uint64_t)(n > 0 ? n : 1), r);
}

TiledSizes tiled_sizes(int n, int dim, uint64_t nnz) {
    TiledSizes z;
    const uint64_t nb = (uint64_t)tiled_blocks(n);
    const uint64_t nt = ceil_div((uint64_t)(n > 0 ? n : 1), tiled_tile_points(dim));
    z.max_segs = std::min<uint64_t>(nnz, nb * nt) + 1;
    // The groups of a segment cover disjoint, increasing row ranges of its block, so its bridges
    // cover at most rows_per_block - 1 rows, ROW_STEP_MAX each.
    const uint64_t bridges = std::min(z.max_segs, nnz) *
                             ((tiled_rows_per_block(n) - 1) / ROW_STEP_MAX);
    // two consecutive groups of a segment hold more than GROUP_SPAN - 8 edges with bridges
    z.max_groups = 2 * z.max_segs + 2 * ((nnz + bridges) / (GROUP_SPAN - 8)) + 2;
    // a group pads at most 7 edges, and there are no more groups than edges
    z.padded_cap = std::max<uint64_t>(nnz + bridges + 7 * std::min(z.max_groups, nnz), 8);
    z.persistent = z.padded_cap * 6 + (nb + 1) * 4 + z.max_segs * 12 + (z.max_groups + 1) * 8;
    const uint64_t nn = std::max<uint64_t>(nnz, 8);
    z.temp = nn * 8 * 2 + nn * 4 * 2 + radix_sort_temp_bytes(nn) + nn * 4 * 4 +
             scan_temp_bytes(std::max(nn, z.max_groups)) + z.max_segs * 8 +
             z.max_groups * 8 + 64 + 14 * 256;
    return z;
}

uint64_t tiled_build(const uint32_t *rows, const uint32_t *cols, const float *vals, int n,
                     int dim, uint64_t nnz, const TiledOut &o, void *temp, hipStream_t s) {
    const TiledSizes z = tiled_sizes(n, dim, nnz);
    const uint64_t nn = std::max<uint64_t>(nnz, 8);
    TileGeom g;
    g.rows_per_block = tiled_rows_per_block(n);
    g.tile_points = tiled_tile_points(dim);
    const uint32_t nblocks = (uint32_t)tiled_blocks(n);
    const uint64_t ntiles = ceil_div((uint64_t)(n > 0 ? n : 1), g.tile_points);
    g.row_bits = bits_for(g.rows_per_block);
    g.tile_bits = std::max(1, bits_for(ntiles));
    const int key_bits = g.row_bits + g.tile_bits + bits_for(nblocks);
    if (key_bits > 64) throw Error(ERR_INVALID, "tiled layout: too many points");
    // edge positions are 32-bit
    if (z.padded_cap > 0xffffffffull) throw Error(ERR_INVALID, "tiled layout: too many affinities");
    // temp carve-up (256-byte aligned pieces)
    char *p = (char *)temp;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += (bytes + 255) & ~(size_t)255;
        return (void *)q;
    };
    uint64_t *keys = (uint64_t *)take(nn * 8), *keys_alt = (uint64_t *)take(nn * 8);
    uint32_t *idx = (uint32_t *)take(nn * 4), *idx_alt = (uint32_t *)take(nn * 4);
    void *sort_temp = take(radix_sort_temp_bytes(nn));
    uint32_t *seg_flag = (uint32_t *)take(nn * 4), *seg_pos = (uint32_t *)take(nn * 4);
    uint32_t *run_flag = (uint32_t *)take(nn * 4), *run_pos = (uint32_t *)take(nn * 4);
    void *scan_temp = take(scan_temp_bytes(std::max(nn, z.max_groups)));
    uint32_t *seg_blk = (uint32_t *)take(z.max_segs * 4);
    uint32_t *seg_ngrp = (uint32_t *)take(z.max_segs * 4);
    uint32_t *glen = (uint32_t *)take(z.max_groups * 4);  // group edges with bridges
    uint32_t *plen = (uint32_t *)take(z.max_groups * 4);  // and padded
    uint32_t *counts = (uint32_t *)take(64);              // nsegs, nruns
    if (nnz == 0) {
        BHT_HIP(hipMemsetAsync(o.blk_seg, 0, (nblocks + 1) * 4, s));
        BHT_HIP(hipMemsetAsync(o.seg_g0, 0, 4, s));
        BHT_HIP(hipMemsetAsync(o.grp_e0, 0, 4, s));
        BHT_HIP(hipStreamSynchronize(s));
        return 0;
    }
    const int nb_rows = (int)std::min<uint64_t>(ceil_div((uint64_t)n, 4), 8192);
    hipLaunchKernelGGL(k_tile_keys, dim3(nb_rows), dim3(256), 0, s, rows, cols, n, g, keys, idx);
    BHT_LAUNCH_CHECK();
    const bool alt =
        radix_sort_pairs<uint64_t>(keys, idx, keys_alt, idx_alt, nnz, 0, key_bits, sort_temp, s);
    uint64_t *sk = alt ? keys_alt : keys;
    const uint32_t *si = alt ? idx_alt : idx;
    // the columns (within their tiles) and values in sorted order, in the sort's other key
    // buffer (nn >= 8 leaves room for the 16-byte alignment of the values)
    uint16_t *ucol = (uint16_t *)(alt ? keys : keys_alt);
    float *uval = (float *)((char *)ucol + ((nn * 2 + 15) & ~(uint64_t)15));
    const int nb_e = (int)std::min<uint64_t>(ceil_div(nnz, 256), 16384);
    hipLaunchKernelGGL(k_tile_pack, dim3(nb_e), dim3(256), 0, s, sk, si, cols, vals, nnz, g,
                       ucol, uval, seg_flag, run_flag);
    BHT_LAUNCH_CHECK();
    scan_exclusive<uint32_t>(seg_flag, seg_pos, nnz, scan_temp, counts, s);
    scan_exclusive<uint32_t>(run_flag, run_pos, nnz, scan_temp, counts + 1, s);
    // the index arrays are free once the edges are packed
    uint32_t *runs = idx, *run_row = idx_alt;
    hipLaunchKernelGGL(k_tile_segments, dim3(nb_e), dim3(256), 0, s, sk, seg_flag, seg_pos,
                       run_flag, run_pos, nnz, g, o.seg_e0, o.seg_tile, seg_blk, runs, run_row);
    BHT_LAUNCH_CHECK();
    const int nb_s = (int)std::min<uint64_t>(ceil_div(z.max_segs, 256), 4096);
    hipLaunchKernelGGL(k_tile_seg_groups, dim3(nb_s), dim3(256), 0, s, o.seg_e0, run_pos, runs,
                       run_row, counts, nnz, (uint32_t)z.max_segs, seg_ngrp);
    BHT_LAUNCH_CHECK();
    // seg_g0 = exclusive scan of the group counts over the segments and the sentinel
    scan_exclusive<uint32_t>(seg_ngrp, o.seg_g0, z.max_segs, scan_temp, counts + 2, s);
    // per-run tables in arrays free by now: the segment flags and the sorted keys
    uint32_t *run_grp = seg_flag, *run_off = seg_pos, *run_bs = (uint32_t *)sk;
    const int nb_g = (int)std::min<uint64_t>(ceil_div(z.max_segs + nblocks + 1, 256), 8192);
    hipLaunchKernelGGL(k_tile_index, dim3(nb_g), dim3(256), 0, s, seg_blk, o.seg_e0, o.seg_g0,
                       run_pos, runs, run_row, counts, nblocks, o.blk_seg, run_grp, run_off,
                       run_bs, glen, o.grp_row);
    BHT_LAUNCH_CHECK();
    // the padded group starts: an exclusive scan of the padded lengths; its total (the padded
    // edge count) is the sentinel after every group
    const int nb_p = (int)std::min<uint64_t>(ceil_div(z.max_groups, 256), 8192);
    hipLaunchKernelGGL(k_tile_pad_len, dim3(nb_p), dim3(256), 0, s, glen, o.seg_g0, counts,
                       (uint32_t)z.max_groups, plen);
    BHT_LAUNCH_CHECK();
    scan_exclusive<uint32_t>(plen, o.grp_e0, z.max_groups, scan_temp, o.grp_e0 + z.max_groups,
                             s);
    hipLaunchKernelGGL(k_tile_scatter, dim3(nb_e), dim3(256), 0, s, ucol, uval, run_flag, run_pos,
                       runs, run_grp, run_off, run_bs, o.grp_e0, nnz, o.tkey, o.tval);
    const int nb_f = (int)std::min<uint64_t>(ceil_div(std::min<uint64_t>(nnz, z.max_groups), 256),
                                             8192);
    hipLaunchKernelGGL(k_tile_fill, dim3(nb_f), dim3(256), 0, s, run_grp, run_off, run_bs, glen,
                       o.seg_g0, o.grp_e0, counts, o.tkey, o.tval);
    BHT_LAUNCH_CHECK();
    uint32_t padded = 0;
    BHT_HIP(hipMemcpyAsync(&padded, o.grp_e0 + z.max_groups, 4, hipMemcpyDeviceToHost, s));
    BHT_HIP(hipStreamSynchronize(s));
    if (padded > z.padded_cap) throw Error(ERR_STATE, "tiled layout: padded edges beyond their bound");
    return padded;
}

int attract_default_queue_blocks() { return ATT_QUEUE_BLOCKS; }

void attract_tiled(const LaunchSet &L, float *pos, bool epoch, hipStream_t s, int queue_blocks,
                   bool self_head) {
    if (L.dim >= 3) {  // plain CSR rows (Barnes-Hut, D = 3..7)
        nd_attract(L, pos, epoch, s);
        return;
    }
    const int queued = epoch && queue_blocks > 0 && L.map[F_ATT] == nullptr ? 1 : 0;
    const int self = epoch && self_head && L.map[F_ATT] == nullptr ? 1 : 0;
    const dim3 g(queued ? std::min(L.nb[F_ATT], queue_blocks) : L.nb[F_ATT]);
    if (L.dim == 2) {
        if (epoch)
            hipLaunchKernelGGL((k_attract_tiled<2, true>), g, dim3(T_THREADS), 0, s, L.jobs,
                               L.map[F_ATT], nullptr, queued, self);
        else
            hipLaunchKernelGGL((k_attract_tiled<2, false>), g, dim3(T_THREADS), 0, s, L.jobs,
                               L.map[F_ATT], pos, 0, 0);
    } else {
        if (epoch)
            hipLaunchKernelGGL((k_attract_tiled<1, true>), g, dim3(T_THREADS), 0, s, L.jobs,
                               L.map[F_ATT], nullptr, queued, self);
        else
            hipLaunchKernelGGL((k_attract_tiled<1, false>), g, dim3(T_THREADS), 0, s, L.jobs,
                               L.map[F_ATT], pos, 0, 0);
    }
    BHT_LAUNCH_CHECK();
}

}  // namespace bht
