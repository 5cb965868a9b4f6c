// Hand-written FP32 FFT convolution of the FIt-SNE charge grids (replaces src/tsne/fft.rs and the
// convolution loop of src/tsne/interpolation.rs:198-240).
//
// Layout and algebra:
//  * The D + 2 = 4 real charge grids (D = 2) are packed into two complex grids a = t0 + i t1 and
//    b = t2 + i t3.  The kernel K is real and even in both axes, so its spectrum is real and
//    conv(x + i y, K) = conv(x, K) + i conv(y, K): one complex transform carries two terms and
//    needs no real/complex split.
//  * Only rows/columns [0, nodes) of the len x len zero-padded grid are non-zero on input and only
//    those are needed on output, so the row passes run on `nodes` rows and the column pass
//    (forward FFT, multiply by the kernel spectrum, inverse FFT) is fused into one kernel per
//    column.
//  * The kernel spectrum is real-even: its quarter (len/2+1)^2 is computed from two real-even
//    passes, two rows (then two columns) per complex transform, pre-scaled by 1/len^2 as the
//    reference folds the inverse normalisation into it (interpolation.rs:221-225).
//  * Each line is transformed in LDS by a Stockham auto-sort FFT with radices 4, 2, 3, 5, 7
//    (every length here is 7-smooth, interpolation.rs:364-378); twiddles come from a
//    per-length table computed in double precision on the host.
#include "internal.hpp"

namespace bht {

constexpr int FFT_THREADS = 256;
constexpr int FFT_MAX_LEN = 6144;  // >= 6 * MAX_BOXES

__device__ __forceinline__ float2 cmul(float2 a, float2 w) {
    return make_float2(fmaf(a.x, w.x, -a.y * w.y), fmaf(a.x, w.y, a.y * w.x));
}

template <int R>
struct RootsOfUnity;  // cos / sin of 2 pi j / R, j < R, rounded from double
template <>
struct RootsOfUnity<3> {
    static constexpr float c[3] = {1.0f, -0.5f, -0.5f};
    static constexpr float s[3] = {0.0f, 0.86602540378443865f, -0.86602540378443865f};
};
template <>
struct RootsOfUnity<5> {
    static constexpr float c[5] = {1.0f, 0.30901699437494742f, -0.80901699437494742f,
                                   -0.80901699437494742f, 0.30901699437494742f};
    static constexpr float s[5] = {0.0f, 0.95105651629515357f, 0.58778525229247313f,
                                   -0.58778525229247313f, -0.95105651629515357f};
};
template <>
struct RootsOfUnity<7> {
    static constexpr float c[7] = {1.0f, 0.62348980185873353f, -0.22252093395631440f,
                                   -0.90096886790241913f, -0.90096886790241913f,
                                   -0.22252093395631440f, 0.62348980185873353f};
    static constexpr float s[7] = {0.0f, 0.78183148246802981f, 0.97492791218182361f,
                                   0.43388373911755812f, -0.43388373911755812f,
                                   -0.97492791218182361f, -0.78183148246802981f};
};

// In-register DFT of length R; forward uses exp(-2 pi i jk / R), inverse exp(+...).
template <int R, bool INV>
__device__ __forceinline__ void dft(float2 *x) {
    if constexpr (R == 2) {
        float2 a = x[0], b = x[1];
        x[0] = make_float2(a.x + b.x, a.y + b.y);
        x[1] = make_float2(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        float2 t0 = make_float2(x[0].x + x[2].x, x[0].y + x[2].y);
        float2 t1 = make_float2(x[0].x - x[2].x, x[0].y - x[2].y);
        float2 t2 = make_float2(x[1].x + x[3].x, x[1].y + x[3].y);
        float2 t3 = make_float2(x[1].x - x[3].x, x[1].y - x[3].y);
        x[0] = make_float2(t0.x + t2.x, t0.y + t2.y);
        x[2] = make_float2(t0.x - t2.x, t0.y - t2.y);
        if (!INV) {
            x[1] = make_float2(t1.x + t3.y, t1.y - t3.x);
            x[3] = make_float2(t1.x - t3.y, t1.y + t3.x);
        } else {
            x[1] = make_float2(t1.x - t3.y, t1.y + t3.x);
            x[3] = make_float2(t1.x + t3.y, t1.y - t3.x);
        }
    } else {
        float2 in[R];
#pragma unroll
        for (int m = 0; m < R; m++) in[m] = x[m];
#pragma unroll
        for (int q = 0; q < R; q++) {
            float re = in[0].x, im = in[0].y;
#pragma unroll
            for (int m = 1; m < R; m++) {
                const int j = (m * q) % R;
                const float c = RootsOfUnity<R>::c[j];
                const float s = INV ? RootsOfUnity<R>::s[j] : -RootsOfUnity<R>::s[j];
                re = fmaf(in[m].x, c, fmaf(-in[m].y, s, re));
                im = fmaf(in[m].x, s, fmaf(in[m].y, c, im));
            }
            x[q] = make_float2(re, im);
        }
    }
}

// Butterfly j of a Stockham stage of radix R (M = L / R butterflies, Ns = product of the earlier
// radices): the R inputs src[j + k M], twiddled, then a length-R DFT.  Every transform below runs
// its stages through this function, so all of them give the same bits.
template <int R, bool INV>
__device__ __forceinline__ void butterfly(const float2 *src, int j, int M, int Ns, int tstep,
                                          const float2 *tw, float2 *x) {
    const int k0 = j % Ns;
#pragma unroll
    for (int k = 0; k < R; k++) {
        float2 a = src[j + k * M];
        if (k > 0 && k0 > 0) {
            float2 w = tw[k0 * k * tstep];
            if (INV) w.y = -w.y;
            a = cmul(a, w);
        }
        x[k] = a;
    }
    dft<R, INV>(x);
}

// Output q of butterfly j goes to (j - j % Ns) R + j % Ns + q Ns.
__device__ __forceinline__ int stage_out(int j, int Ns, int R) {
    const int k0 = j % Ns;
    return (j - k0) * R + k0;
}

// One Stockham stage of radix R over a length-L line in LDS (all inputs are read into registers
// before any output is written, so one LDS buffer suffices).
template <int R, bool INV>
__device__ __forceinline__ void stockham_stage(float2 *buf, int L, int Ns, const float2 *tw) {
    constexpr int NB = (FFT_MAX_LEN / R + FFT_THREADS - 1) / FFT_THREADS;
    const int M = L / R;
    const int tstep = L / (Ns * R);
    float2 v[NB][R];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = threadIdx.x + b * FFT_THREADS;
        if (j < M) butterfly<R, INV>(buf, j, M, Ns, tstep, tw, v[b]);
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = threadIdx.x + b * FFT_THREADS;
        if (j < M) {
            const int base = stage_out(j, Ns, R);
#pragma unroll
            for (int q = 0; q < R; q++) buf[base + q * Ns] = v[b][q];
        }
    }
    __syncthreads();
}

template <bool INV>
__device__ __forceinline__ void fft_lds(float2 *buf, int L, const float2 *tw) {
    int Ns = 1, rem = L;
    while (rem % 4 == 0) { stockham_stage<4, INV>(buf, L, Ns, tw); Ns *= 4; rem /= 4; }
    while (rem % 2 == 0) { stockham_stage<2, INV>(buf, L, Ns, tw); Ns *= 2; rem /= 2; }
    while (rem % 3 == 0) { stockham_stage<3, INV>(buf, L, Ns, tw); Ns *= 3; rem /= 3; }
    while (rem % 5 == 0) { stockham_stage<5, INV>(buf, L, Ns, tw); Ns *= 5; rem /= 5; }
    while (rem % 7 == 0) { stockham_stage<7, INV>(buf, L, Ns, tw); Ns *= 7; rem /= 7; }
}

// The same stages between two global-memory lines of the block (src -> dst, one butterfly at a
// time, so few registers); for lines longer than a launch's LDS.  Returns the line holding the
// result.  Not inlined: the kernels' register budget is set by their LDS paths.
template <int R, bool INV>
__device__ __forceinline__ void pingpong_stage(const float2 *src, float2 *dst, int L, int Ns,
                                               const float2 *tw) {
    const int M = L / R;
    const int tstep = L / (Ns * R);
    for (int j = threadIdx.x; j < M; j += FFT_THREADS) {
        float2 x[R];
        butterfly<R, INV>(src, j, M, Ns, tstep, tw, x);
        const int base = stage_out(j, Ns, R);
#pragma unroll
        for (int q = 0; q < R; q++) dst[base + q * Ns] = x[q];
    }
    __syncthreads();
}

template <bool INV>
__device__ __attribute__((noinline)) float2 *fft_pingpong(float2 *a, float2 *b, int L,
                                                          const float2 *tw) {
    int Ns = 1, rem = L;
    auto step = [&]() { float2 *t = a; a = b; b = t; };
    while (rem % 4 == 0) { pingpong_stage<4, INV>(a, b, L, Ns, tw); step(); Ns *= 4; rem /= 4; }
    while (rem % 2 == 0) { pingpong_stage<2, INV>(a, b, L, Ns, tw); step(); Ns *= 2; rem /= 2; }
    while (rem % 3 == 0) { pingpong_stage<3, INV>(a, b, L, Ns, tw); step(); Ns *= 3; rem /= 3; }
    while (rem % 5 == 0) { pingpong_stage<5, INV>(a, b, L, Ns, tw); step(); Ns *= 5; rem /= 5; }
    while (rem % 7 == 0) { pingpong_stage<7, INV>(a, b, L, Ns, tw); step(); Ns *= 7; rem /= 7; }
    return a;
}

// One wave per line (lines up to WAVE_FFT_MAX points): the same Stockham stages in the wave's own
// LDS line, each stage reading all of its inputs into registers before writing any output (a
// wave's LDS operations complete in order), so a stage needs no block barrier and a line needs
// one buffer.
constexpr int WAVE_FFT_MAX = 1024;

template <int R, bool INV>
__device__ __forceinline__ void wave_stage(float2 *buf, int L, int Ns, const float2 *tw,
                                           int lane) {
    constexpr int NB = (WAVE_FFT_MAX / R + 63) / 64;
    const int M = L / R;
    const int tstep = L / (Ns * R);
    float2 v[NB][R];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = lane + 64 * b;
        if (j < M) butterfly<R, INV>(buf, j, M, Ns, tstep, tw, v[b]);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = lane + 64 * b;
        if (j < M) {
            const int base = stage_out(j, Ns, R);
#pragma unroll
            for (int q = 0; q < R; q++) buf[base + q * Ns] = v[b][q];
        }
    }
    __builtin_amdgcn_wave_barrier();
}

template <bool INV>
__device__ __forceinline__ void fft_wave(float2 *buf, int L, const float2 *tw, int lane) {
    int Ns = 1, rem = L;
    while (rem % 4 == 0) { wave_stage<4, INV>(buf, L, Ns, tw, lane); Ns *= 4; rem /= 4; }
    while (rem % 2 == 0) { wave_stage<2, INV>(buf, L, Ns, tw, lane); Ns *= 2; rem /= 2; }
    while (rem % 3 == 0) { wave_stage<3, INV>(buf, L, Ns, tw, lane); Ns *= 3; rem /= 3; }
    while (rem % 5 == 0) { wave_stage<5, INV>(buf, L, Ns, tw, lane); Ns *= 5; rem /= 5; }
    while (rem % 7 == 0) { wave_stage<7, INV>(buf, L, Ns, tw, lane); Ns *= 7; rem /= 7; }
}

// Line buffers.  A launch's dynamic LDS holds `cap` complex points per block (the column pass:
// cols_lds_points(cap)), sized for the lengths the job's grid has reached rather than for its
// capacity, so that these kernels fit a CU beside the attraction's blocks.  A block whose line is
// longer than its LDS (the grid grew past what the launch was sized for) transforms it between two
// global-memory lines of the job's fallback buffer instead, with the same arithmetic; only the
// first FFT_FB_BLOCKS blocks of a launch do so, striding over the lines.
__host__ __device__ constexpr int cols_lds_points(int cap) {
    return cap <= WAVE_FFT_MAX ? 4 * cap : cap;
}
// A batch's column pass also holds its lines' twiddle table after the four lines (C4: 3.16 ->
// 3.09 s of epochs); a single job keeps the smaller launch, which shares the CUs with its
// attraction.
__host__ __device__ constexpr int cols_lds_points_tw(int cap) {
    return cap <= WAVE_FFT_MAX ? 5 * cap : cap + cap / 4;
}

struct LdsLine {  // the block's line in LDS, transformed in place
    float2 *buf;
    __device__ float2 *line() const { return buf; }
    template <bool INV>
    __device__ __forceinline__ float2 *fft(float2 *x, int L, const float2 *tw) const {
        fft_lds<INV>(x, L, tw);
        return x;
    }
};

struct GlobalLines {  // two global lines of the block, transformed from one into the other
    float2 *a, *b;
    __device__ float2 *line() const { return a; }
    template <bool INV>
    __device__ __forceinline__ float2 *fft(float2 *x, int L, const float2 *tw) const {
        return fft_pingpong<INV>(x, x == a ? b : a, L, tw);
    }
};

__device__ __forceinline__ int signed_lag(int c, int L) { return c <= L / 2 ? c : c - L; }

// Column passes read each line with a stride of a whole row, so neighbouring columns share every
// cache line.  Workgroups go to the 8 XCDs in turn; work item k of a launch (k = block + i * grid,
// the grid a multiple of 8) is given line (k % 8) * per + k / 8, so that each XCD takes a
// contiguous range of columns and reads their shared lines through its own L2.  Items past the
// last line (from rounding `per` up) are skipped.
__device__ __forceinline__ int xcd_line(int k, int per) { return (k & 7) * per + (k >> 3); }

// ---- kernel spectrum -------------------------------------------------------------------------

// Runs body(line engine, first line, line stride, twiddles) with the block's LDS when a line of L
// points fits `cap`, else with the block's two lines of the job's fallback buffer (the first
// FFT_FB_BLOCKS blocks only).  When the LDS holds two lines, the length's twiddle table is copied
// after the line once per block and every stage reads it there (the same values).
template <typename Body>
__device__ __forceinline__ void with_line_buffer(float2 *lds, int cap, int L, const JobDev &J,
                                                 const BlockCtx &bc, const float2 *tw,
                                                 Body body) {
    if (L <= cap) {
        const float2 *twp = tw;
        if (2 * L <= cap) {
            float2 *t = lds + L;
            for (int i = threadIdx.x; i < L; i += FFT_THREADS) t[i] = tw[i];
            __syncthreads();
            twp = t;
        }
        body(LdsLine{lds}, bc.lb, bc.nb, twp);
    } else if (bc.lb < FFT_FB_BLOCKS) {
        float2 *a = J.fft_fb + (size_t)bc.lb * 2 * J.len_cap;
        body(GlobalLines{a, a + J.len_cap}, bc.lb,
             bc.nb < FFT_FB_BLOCKS ? bc.nb : FFT_FB_BLOCKS, tw);
    }
}

// Rows r1 = 2q, r2 = 2q + 1 of the squared Cauchy kernel (interpolation.rs:198-217), transformed
// together; both transforms are real (even rows), so real and imaginary parts separate directly.
__global__ void __launch_bounds__(FFT_THREADS) k_kernel_rows(const JobDev *jobs,
                                                             const uint32_t *map,
                                                             const float2 *tw_all, int cap) {
    extern __shared__ float2 lds[];
    const BlockCtx bc = block_ctx(map, jobs, F_KSPEC);
    const JobDev &J = jobs[bc.job];
    const int dim = J.dim;
    float *rker = dim == 2 ? J.rker : J.kspec;
    const GridPlan p = J.scal->plan;
    const int L = p.len, h = L / 2 + 1;
    const float2 *tw = tw_all + p.tw_off;
    const int npairs = dim == 2 ? (h + 1) / 2 : 1;
    with_line_buffer(lds, cap, L, J, bc, tw, [&](const auto &eng, int lb, int nb,
                                                 const float2 *twp) {
        float2 *buf = eng.line();
        for (int q = lb; q < npairs; q += nb) {
            const int r1 = 2 * q, r2 = 2 * q + 1;
            for (int c = threadIdx.x; c < L; c += FFT_THREADS) {
                float dc = p.spacing * (float)signed_lag(c, L);
                float v[2];
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    float sq = 0.0f;
                    sq = sq + dc * dc;
                    if (dim == 2) {
                        float dr = p.spacing * (float)signed_lag(t == 0 ? r1 : r2, L);
                        sq = sq + dr * dr;
                    }
                    float cauchy = 1.0f / (1.0f + sq);
                    v[t] = cauchy * cauchy;
                }
                buf[c] = make_float2(v[0], dim == 2 ? v[1] : 0.0f);
            }
            __syncthreads();
            const float2 *f = eng.template fft<false>(buf, L, twp);
            if (dim == 2) {
                for (int c = threadIdx.x; c < h; c += FFT_THREADS) {
                    rker[(size_t)r1 * h + c] = f[c].x;
                    if (r2 < h) rker[(size_t)r2 * h + c] = f[c].y;
                }
            } else {
                const float inv_total = 1.0f / (float)L;
                for (int c = threadIdx.x; c < h; c += FFT_THREADS) rker[c] = f[c].x * inv_total;
            }
            __syncthreads();
        }
    });
}

// Columns v1 = 2q, v2 = 2q + 1 of the row transforms, extended evenly over r, transformed
// together; the quarter spectrum is stored pre-scaled by 1 / len^2.
__global__ void __launch_bounds__(FFT_THREADS) k_kernel_cols(const JobDev *jobs,
                                                             const uint32_t *map,
                                                             const float2 *tw_all, int cap) {
    extern __shared__ float2 lds[];
    const BlockCtx bc = block_ctx(map, jobs, F_KSPEC);
    const JobDev &J = jobs[bc.job];
    const float *rker = J.rker;
    float *kspec = J.kspec;
    const GridPlan p = J.scal->plan;
    const int L = p.len, h = L / 2 + 1;
    const float2 *tw = tw_all + p.tw_off;
    const float inv_total = 1.0f / (float)(L * L);
    const int npairs = (h + 1) / 2;
    const int per = (npairs + 7) / 8;
    with_line_buffer(lds, cap, L, J, bc, tw, [&](const auto &eng, int lb, int nb,
                                                 const float2 *twp) {
        float2 *buf = eng.line();
        for (int k = lb; k < 8 * per; k += nb) {
            const int q = xcd_line(k, per);
            if (q >= npairs) continue;
            const int v1 = 2 * q, v2 = 2 * q + 1;
            for (int r = threadIdx.x; r < L; r += FFT_THREADS) {
                int rr = r <= L / 2 ? r : L - r;
                float a = rker[(size_t)rr * h + v1];
                float b = v2 < h ? rker[(size_t)rr * h + v2] : 0.0f;
                buf[r] = make_float2(a, b);
            }
            __syncthreads();
            const float2 *f = eng.template fft<false>(buf, L, twp);
            // stored by columns (kspec[v * h + u]): each column's writes here, and the column
            // pass's reads along a column, are contiguous
            for (int u = threadIdx.x; u < h; u += FFT_THREADS) {
                kspec[(size_t)v1 * h + u] = f[u].x * inv_total;
                if (v2 < h) kspec[(size_t)v2 * h + u] = f[u].y * inv_total;
            }
            __syncthreads();
        }
    });
}

// ---- data passes -----------------------------------------------------------------------------

// Forward row transforms of the packed grids: line (g, r), r < nodes.
__global__ void __launch_bounds__(FFT_THREADS) k_rows_fwd(const JobDev *jobs, const uint32_t *map,
                                                          const float2 *tw_all, int cap) {
    extern __shared__ float2 lds[];
    const BlockCtx bc = block_ctx(map, jobs, F_ROWS);
    const JobDev &J = jobs[bc.job];
    float2 *spec = J.spec;
    const size_t gstride = (size_t)J.nodes_cap * J.len_cap;
    const GridPlan p = J.scal->plan;
    const int L = p.len, nodes = p.nodes;
    const float2 *tw = tw_all + p.tw_off;
    const float2 *g2 = reinterpret_cast<const float2 *>(J.grid);
    const float inv_q = 1.0f / p.qscale;
    with_line_buffer(lds, cap, L, J, bc, tw, [&](const auto &eng, int lb, int nb,
                                                 const float2 *twp) {
        float2 *buf = eng.line();
        for (int line = lb; line < 2 * nodes; line += nb) {
            const int g = line / nodes, r = line % nodes;
            // (t0, t1 / q) and (t2 / q, t3 / q^2)
            const float sa = g == 0 ? 1.0f : inv_q, sb = g == 0 ? inv_q : inv_q * inv_q;
            for (int c = threadIdx.x; c < L; c += FFT_THREADS) {
                float2 v = c < nodes ? g2[2 * ((size_t)r * nodes + c) + g] : make_float2(0.f, 0.f);
                buf[c] = make_float2(v.x * sa, v.y * sb);
            }
            __syncthreads();
            const float2 *f = eng.template fft<false>(buf, L, twp);
            float2 *out = spec + g * gstride + (size_t)r * L;
            for (int c = threadIdx.x; c < L; c += FFT_THREADS) out[c] = f[c];
            __syncthreads();
        }
    });
}

// Column pass: forward transform, multiply by the real kernel spectrum, inverse transform.  The
// launch's LDS holds `cap` points: a wave per line (four lines per work item) when four lines fit,
// else the block per line.
__global__ void __launch_bounds__(FFT_THREADS) k_cols_conv(const JobDev *jobs, const uint32_t *map,
                                                           const float2 *tw_all, int cap) {
    extern __shared__ float2 lds[];
    const BlockCtx bc = block_ctx(map, jobs, F_COLS);
    const JobDev &J = jobs[bc.job];
    const float *kspec = J.kspec;
    float2 *spec = J.spec;
    const size_t gstride = (size_t)J.nodes_cap * J.len_cap;
    const GridPlan p = J.scal->plan;
    const int L = p.len, nodes = p.nodes, h = L / 2 + 1;
    const float2 *tw = tw_all + p.tw_off;
    if (L <= WAVE_FFT_MAX && 4 * L <= cap) {
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        float2 *f = lds + (size_t)w * L;
        // when the launch holds a fifth line, the length's twiddles are copied there once per
        // block and every stage reads them from LDS instead of memory (the same values)
        const float2 *twl = tw;
        if (5 * L <= cap) {
            float2 *t = lds + 4 * (size_t)L;
            for (int i = threadIdx.x; i < L; i += FFT_THREADS) t[i] = tw[i];
            __syncthreads();
            twl = t;
        }
        const int nlines = 2 * L, nquads = (nlines + 3) / 4;
        const int per = (nquads + 7) / 8;
        for (int k = bc.lb; k < 8 * per; k += bc.nb) {
            const int quad = xcd_line(k, per), line = 4 * quad + w;
            if (quad >= nquads || line >= nlines) continue;
            const int g = line / L, c = line % L;
            float2 *col = spec + g * gstride + c;
            for (int r = lane; r < L; r += 64)
                f[r] = r < nodes ? col[(size_t)r * L] : make_float2(0.f, 0.f);
            __builtin_amdgcn_wave_barrier();
            fft_wave<false>(f, L, twl, lane);
            const int cc = c <= L / 2 ? c : L - c;
            for (int u = lane; u < L; u += 64) {
                const int uu = u <= L / 2 ? u : L - u;
                const float kv = kspec[(size_t)cc * h + uu];
                const float2 z = f[u];
                f[u] = make_float2(z.x * kv, z.y * kv);
            }
            __builtin_amdgcn_wave_barrier();
            fft_wave<true>(f, L, twl, lane);
            for (int r = lane; r < nodes; r += 64) col[(size_t)r * L] = f[r];
            __builtin_amdgcn_wave_barrier();
        }
        return;
    }
    const int per = (2 * L + 7) / 8;
    with_line_buffer(lds, cap, L, J, bc, tw, [&](const auto &eng, int lb, int nb,
                                                 const float2 *twp) {
        float2 *buf = eng.line();
        for (int k = lb; k < 8 * per; k += nb) {
            const int line = xcd_line(k, per);
            if (line >= 2 * L) continue;
            const int g = line / L, c = line % L;
            float2 *col = spec + g * gstride + c;
            for (int r = threadIdx.x; r < L; r += FFT_THREADS)
                buf[r] = r < nodes ? col[(size_t)r * L] : make_float2(0.f, 0.f);
            __syncthreads();
            float2 *f = eng.template fft<false>(buf, L, twp);
            const int cc = c <= L / 2 ? c : L - c;
            for (int u = threadIdx.x; u < L; u += FFT_THREADS) {
                const int uu = u <= L / 2 ? u : L - u;
                const float kv = kspec[(size_t)cc * h + uu];
                float2 z = f[u];
                f[u] = make_float2(z.x * kv, z.y * kv);
            }
            __syncthreads();
            const float2 *r2 = eng.template fft<true>(f, L, twp);
            for (int r = threadIdx.x; r < nodes; r += FFT_THREADS) col[(size_t)r * L] = r2[r];
            __syncthreads();
        }
    });
}

// Inverse row transforms; the potentials of terms (t0, t1) / (t2, t3) land in the two halves of
// the float4 node grid.
__global__ void __launch_bounds__(FFT_THREADS) k_rows_inv(const JobDev *jobs, const uint32_t *map,
                                                          const float2 *tw_all, int cap) {
    extern __shared__ float2 lds[];
    const BlockCtx bc = block_ctx(map, jobs, F_ROWS);
    const JobDev &J = jobs[bc.job];
    const float2 *spec = J.spec;
    const size_t gstride = (size_t)J.nodes_cap * J.len_cap;
    const GridPlan p = J.scal->plan;
    const int L = p.len, nodes = p.nodes;
    const float2 *tw = tw_all + p.tw_off;
    float2 *g2 = reinterpret_cast<float2 *>(J.grid);
    const float q = p.qscale;
    with_line_buffer(lds, cap, L, J, bc, tw, [&](const auto &eng, int lb, int nb,
                                                 const float2 *twp) {
        float2 *buf = eng.line();
        for (int line = lb; line < 2 * nodes; line += nb) {
            const int g = line / nodes, r = line % nodes;
            const float sa = g == 0 ? 1.0f : q, sb = g == 0 ? q : q * q;
            const float2 *in = spec + g * gstride + (size_t)r * L;
            for (int c = threadIdx.x; c < L; c += FFT_THREADS) buf[c] = in[c];
            __syncthreads();
            const float2 *f = eng.template fft<true>(buf, L, twp);
            for (int c = threadIdx.x; c < nodes; c += FFT_THREADS)
                g2[2 * ((size_t)r * nodes + c) + g] = make_float2(f[c].x * sa, f[c].y * sb);
            __syncthreads();
        }
    });
}

// D = 1: one line, both packed signals (t0 + i t1, t2 + i 0) transformed, multiplied, inverted.
// Its launches always hold the job's whole capacity.
__global__ void __launch_bounds__(FFT_THREADS) k_line_conv_1d(const JobDev *jobs,
                                                              const uint32_t *map,
                                                              const float2 *tw_all) {
    extern __shared__ float2 buf[];
    const BlockCtx bc = block_ctx(map, jobs, F_LINE1);
    const JobDev &J = jobs[bc.job];
    const float *kspec = J.kspec;
    float4 *grid = J.grid;
    const GridPlan p = J.scal->plan;
    const int L = p.len, nodes = p.nodes;
    const float2 *tw = tw_all + p.tw_off;
    const int g = bc.lb;  // 0: (t0, t1 / q), 1: (t2 / q^2, 0)
    if (g > 1) return;
    const float q = p.qscale, inv_q = 1.0f / q;
    for (int c = threadIdx.x; c < L; c += FFT_THREADS) {
        float4 v = c < nodes ? grid[c] : make_float4(0.f, 0.f, 0.f, 0.f);
        buf[c] = g == 0 ? make_float2(v.x, v.y * inv_q) : make_float2(v.z * inv_q * inv_q, 0.0f);
    }
    __syncthreads();
    fft_lds<false>(buf, L, tw);
    for (int u = threadIdx.x; u < L; u += FFT_THREADS) {
        const int uu = u <= L / 2 ? u : L - u;
        const float k = kspec[uu];
        float2 z = buf[u];
        buf[u] = make_float2(z.x * k, z.y * k);
    }
    __syncthreads();
    fft_lds<true>(buf, L, tw);
    // Block 0 reads (x, y) and block 1 reads z; block 1 writes its potential into the unused .w
    // lane so that neither block overwrites what the other still reads.  The gather reads the
    // third potential from .w.
    for (int c = threadIdx.x; c < nodes; c += FFT_THREADS) {
        if (g == 0) {
            grid[c].x = buf[c].x;
            grid[c].y = buf[c].y * q;
        } else {
            grid[c].w = buf[c].x * q * q;
        }
    }
}

void fft_kernel_spectrum(const LaunchSet &L, hipStream_t s) {
    const int cap = L.fft_cap;
    const size_t lds_bytes = (size_t)cap * sizeof(float2);
    hipLaunchKernelGGL(k_kernel_rows, dim3(L.nb[F_KSPEC]), dim3(FFT_THREADS), lds_bytes, s,
                       L.jobs, L.map[F_KSPEC], L.tw->table, cap);
    if (L.dim == 2)
        hipLaunchKernelGGL(k_kernel_cols, dim3(L.nb[F_KSPEC]), dim3(FFT_THREADS), lds_bytes, s,
                           L.jobs, L.map[F_KSPEC], L.tw->table, cap);
    BHT_LAUNCH_CHECK();
}

void fft_convolve(const LaunchSet &L, hipStream_t s) {
    const int cap = L.fft_cap;
    const size_t lds_bytes = (size_t)cap * sizeof(float2);
    if (L.dim == 1) {
        hipLaunchKernelGGL(k_line_conv_1d, dim3(L.nb[F_LINE1]), dim3(FFT_THREADS), lds_bytes, s,
                           L.jobs, L.map[F_LINE1], L.tw->table);
        BHT_LAUNCH_CHECK();
        return;
    }
    const int cols_cap = L.map[F_COLS] != nullptr ? cols_lds_points_tw(cap) : cols_lds_points(cap);
    hipLaunchKernelGGL(k_rows_fwd, dim3(L.nb[F_ROWS]), dim3(FFT_THREADS), lds_bytes, s, L.jobs,
                       L.map[F_ROWS], L.tw->table, cap);
    hipLaunchKernelGGL(k_cols_conv, dim3(L.nb[F_COLS]), dim3(FFT_THREADS),
                       (size_t)cols_cap * sizeof(float2), s, L.jobs, L.map[F_COLS], L.tw->table,
                       cols_cap);
    hipLaunchKernelGGL(k_rows_inv, dim3(L.nb[F_ROWS]), dim3(FFT_THREADS), lds_bytes, s, L.jobs,
                       L.map[F_ROWS], L.tw->table, cap);
    BHT_LAUNCH_CHECK();
}

}  // namespace bht
