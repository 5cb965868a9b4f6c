/*
 * bhtsne_oracle.c -- CPU restatement of the bhtsne 0.7.10 hot path (Rust crate, not buildable
 * here: no cargo/rustc in this image).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is compared against,
 * and the CPU baseline that bench.py times beside it.  Nothing in the product (bhtsne_amd/,
 * include/) links, loads or calls it.
 *
 * Arithmetic follows the reference in T = f32: the same operation order, no FMA contraction
 * (built with -ffp-contract=off), Rust `signum` semantics, sequential sums where the reference
 * reduces.  Where the reference reduces with rayon (order not fixed) this file reduces
 * sequentially, i.e. the behaviour of a one-thread rayon pool, which is what the reference's own
 * path-equivalence tests use.  Per-row work that the reference runs in parallel is run with
 * OpenMP here (rows are independent, so results do not depend on the thread count).
 *
 * Deliberate differences, each stated where it happens:
 *   - kNN: brute force with the reference metric and a (distance, index) tie order
 *     (src/test.rs:866-881) instead of the VP tree, whose tie order is undefined.
 *   - FFT: an exact mixed-radix DFT in double precision, rounded to f32 at every point where the
 *     reference stores f32 (rustfft/realfft's own rounding is not reproducible without them).
 *   - BH sort: (code, index) instead of rustfft's unstable sort by code (ties are unordered there).
 *
 * Parity of every function here is pinned by tests/test_oracle_*.py (known answers from the
 * reference's tests, closed forms, brute force on small inputs).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------------------------ */
/* Threads                                                                                     */
/* ------------------------------------------------------------------------------------------ */

EXPORT int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ------------------------------------------------------------------------------------------ */
/* kNN (src/tsne/vptree.rs:274-319 result semantics; metric as in src/test.rs:856-862)        */
/* ------------------------------------------------------------------------------------------ */

/* Euclidean metric of the reference's tests and README: sqrt(sum((a - b)^2)), sequential f32. */
static float euclid_f32(const float *a, const float *b, int d) {
    float s = 0.0f;
    for (int i = 0; i < d; i++) {
        float t = a[i] - b[i];
        s = s + t * t;
    }
    return sqrtf(s);
}

static int pair_less(float da, uint32_t ia, float db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

/* k nearest neighbours of row i, self excluded, ascending (distance, index), into idx / dist
 * (k each); hd / hi are k-entry scratch.  Equivalent to the reference's "search k+1, drop self"
 * (vptree.rs:305-318) under a (distance, index) order. */
static void knn_row(const float *x, int64_t n, int d, int k, int64_t i, float *hd, uint32_t *hi,
                    uint32_t *idx, float *dist) {
    int cnt = 0;
    /* max-heap on (dist, idx) of the best k so far */
    for (int64_t j = 0; j < n; j++) {
        if (j == i) continue;
        float dj = euclid_f32(x + (size_t)j * d, x + (size_t)i * d, d);
        if (cnt < k) {
            int c = cnt++;
            hd[c] = dj;
            hi[c] = (uint32_t)j;
            while (c > 0) {
                int p = (c - 1) / 2;
                if (pair_less(hd[p], hi[p], hd[c], hi[c])) {
                    float tf = hd[p]; hd[p] = hd[c]; hd[c] = tf;
                    uint32_t ti = hi[p]; hi[p] = hi[c]; hi[c] = ti;
                    c = p;
                } else break;
            }
        } else if (pair_less(dj, (uint32_t)j, hd[0], hi[0])) {
            hd[0] = dj;
            hi[0] = (uint32_t)j;
            int c = 0;
            for (;;) {
                int l = 2 * c + 1, r = l + 1, m = c;
                if (l < k && pair_less(hd[m], hi[m], hd[l], hi[l])) m = l;
                if (r < k && pair_less(hd[m], hi[m], hd[r], hi[r])) m = r;
                if (m == c) break;
                float tf = hd[m]; hd[m] = hd[c]; hd[c] = tf;
                uint32_t ti = hi[m]; hi[m] = hi[c]; hi[c] = ti;
                c = m;
            }
        }
    }
    /* heap-sort ascending */
    for (int end = cnt - 1; end > 0; end--) {
        float tf = hd[0]; hd[0] = hd[end]; hd[end] = tf;
        uint32_t ti = hi[0]; hi[0] = hi[end]; hi[end] = ti;
        int c = 0;
        for (;;) {
            int l = 2 * c + 1, r = l + 1, m = c;
            if (l < end && pair_less(hd[m], hi[m], hd[l], hi[l])) m = l;
            if (r < end && pair_less(hd[m], hi[m], hd[r], hi[r])) m = r;
            if (m == c) break;
            float t2 = hd[m]; hd[m] = hd[c]; hd[c] = t2;
            uint32_t u2 = hi[m]; hi[m] = hi[c]; hi[c] = u2;
            c = m;
        }
    }
    for (int t = 0; t < k; t++) {
        idx[t] = t < cnt ? hi[t] : 0u;
        dist[t] = t < cnt ? hd[t] : 0.0f;
    }
}

/* k nearest neighbours of every row (brute force). */
EXPORT void oracle_knn(const float *x, int64_t n, int d, int k, uint32_t *idx_out,
                       float *dist_out) {
#pragma omp parallel
    {
        float *hd = (float *)malloc(sizeof(float) * (size_t)k);
        uint32_t *hi = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)k);
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < n; i++)
            knn_row(x, n, d, k, i, hd, hi, idx_out + (size_t)i * k, dist_out + (size_t)i * k);
        free(hd);
        free(hi);
    }
}

/* The same for the m rows sel[0..m) only (sampled checks at sizes where all rows would take too
 * long); output row t is row sel[t]. */
EXPORT void oracle_knn_rows(const float *x, int64_t n, int d, int k, const int64_t *sel,
                            int64_t m, uint32_t *idx_out, float *dist_out) {
#pragma omp parallel
    {
        float *hd = (float *)malloc(sizeof(float) * (size_t)k);
        uint32_t *hi = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)k);
#pragma omp for schedule(dynamic, 1)
        for (int64_t t = 0; t < m; t++)
            knn_row(x, n, d, k, sel[t], hd, hi, idx_out + (size_t)t * k, dist_out + (size_t)t * k);
        free(hd);
        free(hi);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Bandwidth search (src/tsne/mod.rs:201-275)                                                  */
/* ------------------------------------------------------------------------------------------ */

/* One row, f32 exactly as written in the reference. Returns the number of bisection steps. */
/* exp and ln of the bandwidth search: the host libm's expf / logf (what the reference's f32 exp()
 * and ln() call), or, when cr_explog is set, the correctly rounded values ((float) of the double
 * routine), which differ from the libm's by one ulp on ~6 in 10,000 arguments.  The second form
 * is what the device computes, so it pins the device kernel bit for bit; the first is the
 * reference. */
static int g_cr_explog = 0;
EXPORT void oracle_set_cr_explog(int on) { g_cr_explog = on; }
static float sb_exp(float x) { return g_cr_explog ? (float)exp((double)x) : expf(x); }
static float sb_log(float x) { return g_cr_explog ? (float)log((double)x) : logf(x); }

EXPORT int oracle_search_beta_row(float *p, const float *d, int k, float perplexity) {
    int found = 0;
    float beta = 1.0f;
    float min_beta = -FLT_MAX;
    float max_beta = FLT_MAX;
    const float tolerance = 1e-5f;
    int iteration = 0;
    float sum = 0.0f;
    const float two = 2.0f;
    const float log_perp = sb_log(perplexity);

    while (!found && iteration < 200) {
        for (int j = 0; j < k; j++) p[j] = sb_exp(-beta * (d[j] * d[j]));
        sum = 0.0f;
        for (int j = 0; j < k; j++) sum = sum + p[j];
        sum = sum + FLT_MIN;
        float entropy = 0.0f;
        for (int j = 0; j < k; j++) entropy = entropy + beta * p[j] * (d[j] * d[j]);
        entropy = entropy / sum + sb_log(sum);
        float diff = entropy - log_perp;
        if (diff < tolerance && -diff < tolerance) {
            found = 1;
        } else {
            if (diff > 0.0f) {
                min_beta = beta;
                if (max_beta == FLT_MAX || max_beta == -FLT_MAX)
                    beta *= two;
                else
                    beta = (beta + max_beta) / two;
            } else {
                max_beta = beta;
                if (min_beta == -FLT_MAX || min_beta == FLT_MAX)
                    beta /= two;
                else
                    beta = (beta + min_beta) / two;
            }
            if (isinf(beta) && beta > 0.0f) beta = FLT_MAX;
            if (isinf(beta) && beta < 0.0f) beta = -FLT_MAX;
        }
        iteration += 1;
    }
    for (int j = 0; j < k; j++) p[j] /= sum + FLT_EPSILON;
    return iteration;
}

/* Same routine in f64 (the reference's T = f64 instantiation), for the known-answer test of
 * src/test.rs:1023-1049 which runs in f64. */
EXPORT int oracle_search_beta_row_f64(double *p, const double *d, int k, double perplexity) {
    int found = 0;
    double beta = 1.0, min_beta = -DBL_MAX, max_beta = DBL_MAX;
    const double tolerance = 1e-5;
    int iteration = 0;
    double sum = 0.0;
    const double log_perp = log(perplexity);
    while (!found && iteration < 200) {
        for (int j = 0; j < k; j++) p[j] = exp(-beta * (d[j] * d[j]));
        sum = 0.0;
        for (int j = 0; j < k; j++) sum = sum + p[j];
        sum = sum + DBL_MIN;
        double entropy = 0.0;
        for (int j = 0; j < k; j++) entropy = entropy + beta * p[j] * (d[j] * d[j]);
        entropy = entropy / sum + log(sum);
        double diff = entropy - log_perp;
        if (diff < tolerance && -diff < tolerance) {
            found = 1;
        } else {
            if (diff > 0.0) {
                min_beta = beta;
                if (max_beta == DBL_MAX || max_beta == -DBL_MAX) beta *= 2.0;
                else beta = (beta + max_beta) / 2.0;
            } else {
                max_beta = beta;
                if (min_beta == -DBL_MAX || min_beta == DBL_MAX) beta /= 2.0;
                else beta = (beta + min_beta) / 2.0;
            }
            if (isinf(beta) && beta > 0.0) beta = DBL_MAX;
            if (isinf(beta) && beta < 0.0) beta = -DBL_MAX;
        }
        iteration += 1;
    }
    for (int j = 0; j < k; j++) p[j] /= sum + DBL_EPSILON;
    return iteration;
}

/* All rows (src/lib.rs:1094-1107: per-row, independent). */
EXPORT void oracle_search_beta(float *p, const float *d, int64_t n, int k, float perplexity) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        oracle_search_beta_row(p + (size_t)i * k, d + (size_t)i * k, k, perplexity);
}

/* ------------------------------------------------------------------------------------------ */
/* Symmetrisation (src/tsne/mod.rs:306-406), line by line                                      */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    uint32_t c;
    float v;
} colval;

static int cmp_colval(const void *a, const void *b) {
    uint32_t x = ((const colval *)a)->c, y = ((const colval *)b)->c;
    return (x > y) - (x < y);
}

/* Rust `slice::binary_search` on a sorted row; returns position or -1. */
static int64_t bsearch_u32(const uint32_t *row, int k, uint32_t key) {
    int64_t lo = 0, hi = k;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (row[mid] < key) lo = mid + 1;
        else if (row[mid] > key) hi = mid;
        else return mid;
    }
    return -1;
}

/* Two-call API: with rows_out == NULL returns nnz only; otherwise fills rows (n+1), cols, vals.
 * cols_in/vals_in are the per-row kNN columns and conditional P values (n*k each). */
EXPORT int64_t oracle_symmetrize(const uint32_t *cols_in, const float *vals_in, int64_t n, int k,
                                 uint64_t *rows_out, uint32_t *cols_out, float *vals_out) {
    size_t nk = (size_t)n * k;
    uint32_t *pc = (uint32_t *)malloc(sizeof(uint32_t) * nk);
    float *pv = (float *)malloc(sizeof(float) * nk);
    /* 1. per-row sort by column (mod.rs:318-329) */
#pragma omp parallel
    {
        colval *row = (colval *)malloc(sizeof(colval) * (size_t)k);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            for (int j = 0; j < k; j++) {
                row[j].c = cols_in[(size_t)i * k + j];
                row[j].v = vals_in[(size_t)i * k + j];
            }
            qsort(row, (size_t)k, sizeof(colval), cmp_colval);
            for (int j = 0; j < k; j++) {
                pc[(size_t)i * k + j] = row[j].c;
                pv[(size_t)i * k + j] = row[j].v;
            }
        }
        free(row);
    }
    /* 2. counts (mod.rs:343-352) */
    uint64_t *row_counts = (uint64_t *)calloc((size_t)n, sizeof(uint64_t));
    for (int64_t nn = 0; nn < n; nn++) {
        for (size_t i = (size_t)nn * k; i < (size_t)(nn + 1) * k; i++) {
            row_counts[nn] += 1;
            uint32_t c = pc[i];
            if (bsearch_u32(pc + (size_t)c * k, k, (uint32_t)nn) < 0) row_counts[c] += 1;
        }
    }
    uint64_t total = 0;
    for (int64_t i = 0; i < n; i++) total += row_counts[i];
    if (!rows_out) {
        free(pc);
        free(pv);
        free(row_counts);
        return (int64_t)total;
    }
    /* 3. prefix (mod.rs:356-363) */
    rows_out[0] = 0;
    for (int64_t i = 0; i < n; i++) rows_out[i + 1] = rows_out[i] + row_counts[i];
    /* 4. fill (mod.rs:365-397) */
    uint64_t *offset = (uint64_t *)calloc((size_t)n, sizeof(uint64_t));
    for (int64_t nn = 0; nn < n; nn++) {
        for (size_t i = (size_t)nn * k; i < (size_t)(nn + 1) * k; i++) {
            uint32_t c = pc[i];
            int64_t present = bsearch_u32(pc + (size_t)c * k, k, (uint32_t)nn);
            if (present >= 0) {
                if ((uint64_t)nn <= c) {
                    size_t m = (size_t)c * k + (size_t)present;
                    cols_out[rows_out[nn] + offset[nn]] = c;
                    cols_out[rows_out[c] + offset[c]] = (uint32_t)nn;
                    vals_out[rows_out[nn] + offset[nn]] = pv[i] + pv[m];
                    vals_out[rows_out[c] + offset[c]] = pv[i] + pv[m];
                }
            } else {
                cols_out[rows_out[nn] + offset[nn]] = c;
                cols_out[rows_out[c] + offset[c]] = (uint32_t)nn;
                vals_out[rows_out[nn] + offset[nn]] = pv[i];
                vals_out[rows_out[c] + offset[c]] = pv[i];
            }
            if (present < 0 || (uint64_t)nn <= c) {
                offset[nn] += 1;
                if (c != (uint64_t)nn) offset[c] += 1;
            }
        }
    }
    /* 5. halve (mod.rs:400-401) */
    for (uint64_t i = 0; i < total; i++) vals_out[i] *= 0.5f;
    free(offset);
    free(pc);
    free(pv);
    free(row_counts);
    return (int64_t)total;
}

/* A rayon `par_iter().sum()` splits the slice in halves down to small pieces, sums each piece
 * left to right and adds the halves back: a pairwise tree.  The split points depend on the pool
 * size; this uses halving down to 256 elements.  (A one-thread pool would add left to right,
 * whose error at 10^6 terms is ~1e-5 relative, far from any real run of the reference.) */
static float pairwise_sum_f32(const float *v, int64_t n, int64_t stride) {
    if (n <= 256) {
        float s = 0.0f;
        for (int64_t i = 0; i < n; i++) s = s + v[i * stride];
        return s;
    }
    int64_t h = n / 2;
    return pairwise_sum_f32(v, h, stride) + pairwise_sum_f32(v + h * stride, n - h, stride);
}

/* P <- P * exag / (sum(P) + eps)  (src/tsne/mod.rs:284-293).  Returns scale. */
EXPORT float oracle_normalize_p(float *vals, int64_t nnz, float exag) {
    float sum = pairwise_sum_f32(vals, nnz, 1);
    float scale = exag / (sum + FLT_EPSILON);
    for (int64_t i = 0; i < nnz; i++) vals[i] *= scale;
    return scale;
}

/* P <- P * (1/exag)  (src/tsne/mod.rs:525-531). */
EXPORT void oracle_stop_lying(float *vals, int64_t nnz, float exag) {
    float scale = 1.0f / exag;
    for (int64_t i = 0; i < nnz; i++) vals[i] *= scale;
}

/* learning rate, src/lib.rs:652-659 */
EXPORT float oracle_auto_learning_rate(int64_t n, float exag) {
    float a = (float)n / exag / 4.0f;
    return a > 50.0f ? a : 50.0f;
}

/* ------------------------------------------------------------------------------------------ */
/* Attraction + update + recentring (src/tsne/arena.rs:398-427, mod.rs:470-516, 534-568)      */
/* ------------------------------------------------------------------------------------------ */

EXPORT void oracle_edge_forces(const float *y, int64_t n, int D, const uint64_t *rows,
                               const uint32_t *cols, const float *vals, float *pos) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        float acc[8] = {0};
        const float *yi = y + (size_t)i * D;
        for (uint64_t e = rows[i]; e < rows[i + 1]; e++) {
            const float *yj = y + (size_t)cols[e] * D;
            float disp[8];
            float dist = 0.0f;
            for (int a = 0; a < D; a++) {
                float delta = yi[a] - yj[a];
                disp[a] = delta;
                dist += delta * delta;
            }
            float factor = vals[e] / (dist + 1.0f);
            for (int a = 0; a < D; a++) acc[a] += factor * disp[a];
        }
        for (int a = 0; a < D; a++) pos[(size_t)i * D + a] = acc[a];
    }
}

/* Rust f32::signum: +1 for +0.0 and positives, -1 for -0.0 and negatives, NaN for NaN. */
static float rust_signum(float x) {
    if (isnan(x)) return x;
    return copysignf(1.0f, x);
}

EXPORT void oracle_gradient_step(float *y, const float *pos, const float *neg, float *uy,
                                 float *gains, int64_t len, float learning_rate, float momentum,
                                 float inverse_norm) {
    const float inc = 0.2f, decay = 0.8f, min_gain = 0.01f;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; i++) {
        float gradient = pos[i] - neg[i] * inverse_norm;
        float s1 = rust_signum(gradient), s2 = rust_signum(uy[i]);
        float g = (s1 != s2) ? gains[i] + inc : gains[i] * decay;
        if (g < min_gain) g = min_gain;
        gains[i] = g;
        uy[i] = momentum * uy[i] - learning_rate * g * gradient;
        y[i] += uy[i];
    }
}

/* zero_mean (mod.rs:534-568): parallel fold/reduce of the per-axis sums, pairwise here.  Writes
 * the mean used into mean_out. */
EXPORT void oracle_zero_mean(float *y, int64_t n, int D, float *mean_out) {
    float m[8] = {0};
    for (int a = 0; a < D; a++) m[a] = pairwise_sum_f32(y + a, n, D);
    float nf = (float)n;
    for (int a = 0; a < D; a++) m[a] /= nf;
    for (int64_t i = 0; i < n; i++)
        for (int a = 0; a < D; a++) y[(size_t)i * D + a] -= m[a];
    if (mean_out)
        for (int a = 0; a < D; a++) mean_out[a] = m[a];
}

/* ------------------------------------------------------------------------------------------ */
/* FFT: exact mixed-radix complex DFT in double (stands in for rustfft/realfft, fft.rs)        */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    double re, im;
} cplx;

typedef struct {
    int n;
    int nf;
    int factors[64];
    cplx *tw; /* exp(-2 pi i j / n), j < n */
} fftplan;

static void plan_init(fftplan *p, int n) {
    p->n = n;
    p->nf = 0;
    int m = n;
    static const int primes[4] = {4, 2, 3, 5};
    /* radix 4 first, then 2, 3, 5, 7, then any other prime */
    while (m % 4 == 0) { p->factors[p->nf++] = 4; m /= 4; }
    for (int i = 1; i < 4; i++)
        while (m % primes[i] == 0) { p->factors[p->nf++] = primes[i]; m /= primes[i]; }
    for (int q = 7; m > 1; q++)
        while (m % q == 0) { p->factors[p->nf++] = q; m /= q; }
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; j++) {
        /* exact reduction of the angle keeps twiddles accurate to ~1 ulp */
        double a = -2.0 * M_PI * (double)j / (double)n;
        p->tw[j].re = cos(a);
        p->tw[j].im = sin(a);
    }
}

static void plan_free(fftplan *p) { free(p->tw); }

/* Recursive decimation in time.  out[0..n) = DFT(in[0], in[s], ..., in[(n-1)s]), sign -1 forward
 * (exponent -2 pi i jk/n), +1 inverse (unnormalised). */
static void fft_rec(const fftplan *pl, cplx *out, const cplx *in, int n, int s, int fi, int sign) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = pl->factors[fi];
    int m = n / p;
    for (int r = 0; r < p; r++) fft_rec(pl, out + r * m, in + r * s, m, s * p, fi + 1, sign);
    /* combine: X[k + q m] = sum_r W_n^{r(k+qm)} Y_r[k] */
    int step = pl->n / n; /* W_n^j = tw[j*step] */
    cplx tmp[64];
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) tmp[r] = out[r * m + k];
        for (int q = 0; q < p; q++) {
            double re = 0.0, im = 0.0;
            int kk = k + q * m;
            for (int r = 0; r < p; r++) {
                long e = ((long)r * kk) % n;
                cplx w = pl->tw[e * step];
                if (sign > 0) w.im = -w.im;
                re += tmp[r].re * w.re - tmp[r].im * w.im;
                im += tmp[r].re * w.im + tmp[r].im * w.re;
            }
            out[q * m + k].re = re;
            out[q * m + k].im = im;
        }
    }
}

static void fft_exec(const fftplan *pl, cplx *out, const cplx *in, int sign) {
    fft_rec(pl, out, in, pl->n, 1, 0, sign);
}

/* The same decimation in f32 arithmetic with f32 twiddles: a representative single-precision
 * FFT library, standing in for rustfft/realfft's f32 transforms when the f32 pipeline runs in
 * "f32 FFT" mode (oracle_set_fft_f32).  Its rounding is not rustfft's; it shows the size of an
 * f32 FFT's noise next to the double transform's. */
typedef struct {
    float re, im;
} cplxf;

static int g_fft_f32 = 0;

EXPORT void oracle_set_fft_f32(int on) { g_fft_f32 = on; }

static void fft_rec_f(const fftplan *pl, cplxf *out, const cplxf *in, int n, int s, int fi,
                      int sign) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = pl->factors[fi];
    int m = n / p;
    for (int r = 0; r < p; r++) fft_rec_f(pl, out + r * m, in + r * s, m, s * p, fi + 1, sign);
    int step = pl->n / n;
    cplxf tmp[64];
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) tmp[r] = out[r * m + k];
        for (int q = 0; q < p; q++) {
            float re = 0.0f, im = 0.0f;
            int kk = k + q * m;
            for (int r = 0; r < p; r++) {
                long e = ((long)r * kk) % n;
                float wr = (float)pl->tw[e * step].re, wi = (float)pl->tw[e * step].im;
                if (sign > 0) wi = -wi;
                re += tmp[r].re * wr - tmp[r].im * wi;
                im += tmp[r].re * wi + tmp[r].im * wr;
            }
            out[q * m + k].re = re;
            out[q * m + k].im = im;
        }
    }
}

static void fft_exec_any(const fftplan *pl, cplx *out, const cplx *in, int sign) {
    if (!g_fft_f32) {
        fft_exec(pl, out, in, sign);
        return;
    }
    int n = pl->n;
    cplxf *a = (cplxf *)malloc(sizeof(cplxf) * n), *b = (cplxf *)malloc(sizeof(cplxf) * n);
    for (int i = 0; i < n; i++) {
        a[i].re = (float)in[i].re;
        a[i].im = (float)in[i].im;
    }
    fft_rec_f(pl, b, a, n, 1, 0, sign);
    for (int i = 0; i < n; i++) {
        out[i].re = b[i].re;
        out[i].im = b[i].im;
    }
    free(a);
    free(b);
}

/* Double-precision 2-D real forward transform of a len x len row-major grid into the half
 * spectrum len x (len/2+1) (interleaved re, im), fft.rs:101-111 semantics. */
EXPORT void oracle_fft2_r2c_f64(const double *grid, double *spec, int len) {
    int h = len / 2 + 1;
    fftplan pl;
    plan_init(&pl, len);
    cplx *in = (cplx *)malloc(sizeof(cplx) * len), *out = (cplx *)malloc(sizeof(cplx) * len);
    for (int r = 0; r < len; r++) {
        for (int c = 0; c < len; c++) { in[c].re = grid[(size_t)r * len + c]; in[c].im = 0.0; }
        fft_exec(&pl, out, in, -1);
        for (int c = 0; c < h; c++) {
            spec[2 * ((size_t)r * h + c)] = out[c].re;
            spec[2 * ((size_t)r * h + c) + 1] = out[c].im;
        }
    }
    for (int c = 0; c < h; c++) {
        for (int r = 0; r < len; r++) {
            in[r].re = spec[2 * ((size_t)r * h + c)];
            in[r].im = spec[2 * ((size_t)r * h + c) + 1];
        }
        fft_exec(&pl, out, in, -1);
        for (int r = 0; r < len; r++) {
            spec[2 * ((size_t)r * h + c)] = out[r].re;
            spec[2 * ((size_t)r * h + c) + 1] = out[r].im;
        }
    }
    free(in);
    free(out);
    plan_free(&pl);
}

/* Inverse of the above (unnormalised), fft.rs:119-130 semantics including the cleared imaginary
 * parts of bins 0 and len/2 before the row c2r (fft.rs:213-222).  spec is consumed. */
EXPORT void oracle_fft2_c2r_f64(double *spec, double *grid, int len) {
    int h = len / 2 + 1;
    fftplan pl;
    plan_init(&pl, len);
    cplx *in = (cplx *)malloc(sizeof(cplx) * len), *out = (cplx *)malloc(sizeof(cplx) * len);
    for (int c = 0; c < h; c++) {
        for (int r = 0; r < len; r++) {
            in[r].re = spec[2 * ((size_t)r * h + c)];
            in[r].im = spec[2 * ((size_t)r * h + c) + 1];
        }
        fft_exec(&pl, out, in, +1);
        for (int r = 0; r < len; r++) {
            spec[2 * ((size_t)r * h + c)] = out[r].re;
            spec[2 * ((size_t)r * h + c) + 1] = out[r].im;
        }
    }
    for (int r = 0; r < len; r++) {
        spec[2 * ((size_t)r * h) + 1] = 0.0;
        if (len % 2 == 0) spec[2 * ((size_t)r * h + h - 1) + 1] = 0.0;
        for (int c = 0; c < len; c++) {
            if (c < h) {
                in[c].re = spec[2 * ((size_t)r * h + c)];
                in[c].im = spec[2 * ((size_t)r * h + c) + 1];
            } else {
                in[c].re = spec[2 * ((size_t)r * h + (len - c))];
                in[c].im = -spec[2 * ((size_t)r * h + (len - c)) + 1];
            }
        }
        fft_exec(&pl, out, in, +1);
        for (int c = 0; c < len; c++) grid[(size_t)r * len + c] = out[c].re;
    }
    free(in);
    free(out);
    plan_free(&pl);
}

/* ------------------------------------------------------------------------------------------ */
/* FIt-SNE repulsion (src/tsne/interpolation.rs:100-481) in f32 with the FFT above             */
/* ------------------------------------------------------------------------------------------ */

#define IP 3 /* INTERPOLATION_POINTS, interpolation.rs:24 */
#define MAX_BOXES 1000 /* interpolation.rs:42 */

/* next 7-smooth integer >= n (interpolation.rs:364-378) */
EXPORT int64_t oracle_next_smooth(int64_t n) {
    int64_t c = n < 1 ? 1 : n;
    for (;;) {
        int64_t v = c;
        static const int ps[4] = {2, 3, 5, 7};
        for (int i = 0; i < 4; i++)
            while (v % ps[i] == 0) v /= ps[i];
        if (v == 1) return c;
        c++;
    }
}

/* boxes per axis for an extent (interpolation.rs:353-358), min_boxes = 10 in the reference */
EXPORT int oracle_box_count(float extent, int min_boxes) {
    float scaled = ceilf(extent / 1.0f);
    int64_t s = 0;
    if (scaled > 0.0f) s = scaled >= 9.2e18f ? INT64_MAX : (int64_t)scaled;
    if (s < min_boxes) s = min_boxes;
    if (s > MAX_BOXES) s = MAX_BOXES;
    return (int)s;
}

typedef struct {
    float nodes[IP];
    float inv_denom[IP];
} lagrange;

/* interpolation.rs:400-422 */
static void lagrange_init(lagrange *b) {
    float half = 0.5f;
    float p_recip = 1.0f / (float)IP;
    for (int k = 0; k < IP; k++) b->nodes[k] = ((float)k + half) * p_recip;
    for (int k = 0; k < IP; k++) {
        float denom = 1.0f;
        for (int m = 0; m < IP; m++)
            if (m != k) denom = denom * (b->nodes[k] - b->nodes[m]);
        b->inv_denom[k] = 1.0f / denom;
    }
}

/* interpolation.rs:430-440 */
static void lagrange_weights(const lagrange *b, float local, float *out) {
    for (int k = 0; k < IP; k++) {
        float num = 1.0f;
        for (int m = 0; m < IP; m++)
            if (m != k) num = num * (local - b->nodes[m]);
        out[k] = num * b->inv_denom[k];
    }
}

EXPORT void oracle_lagrange_weights(float local, float *out) {
    lagrange b;
    lagrange_init(&b);
    lagrange_weights(&b, local, out);
}

static int64_t signed_lag(int64_t coord, int64_t len) {
    return coord <= len / 2 ? coord : coord - len;
}

/* f32 forward transform of a real len^D grid (row-major, D in {1,2}) into a half spectrum
 * (interleaved f32 re/im) -- rounding to f32 at the reference's storage points. */
static void fwd_f32(const fftplan *pl, const float *grid, float *spec, int len, int D) {
    int h = len / 2 + 1;
    int rows = D == 2 ? len : 1;
#pragma omp parallel
    {
        cplx *in = (cplx *)malloc(sizeof(cplx) * len), *out = (cplx *)malloc(sizeof(cplx) * len);
#pragma omp for schedule(static)
        for (int r = 0; r < rows; r++) {
            for (int c = 0; c < len; c++) { in[c].re = grid[(size_t)r * len + c]; in[c].im = 0.0; }
            fft_exec_any(pl, out, in,-1);
            for (int c = 0; c < h; c++) {
                spec[2 * ((size_t)r * h + c)] = (float)out[c].re;
                spec[2 * ((size_t)r * h + c) + 1] = (float)out[c].im;
            }
        }
        if (D == 2) {
#pragma omp for schedule(static)
            for (int c = 0; c < h; c++) {
                for (int r = 0; r < len; r++) {
                    in[r].re = spec[2 * ((size_t)r * h + c)];
                    in[r].im = spec[2 * ((size_t)r * h + c) + 1];
                }
                fft_exec_any(pl, out, in,-1);
                for (int r = 0; r < len; r++) {
                    spec[2 * ((size_t)r * h + c)] = (float)out[r].re;
                    spec[2 * ((size_t)r * h + c) + 1] = (float)out[r].im;
                }
            }
        }
        free(in);
        free(out);
    }
}

static void inv_f32(const fftplan *pl, float *spec, float *grid, int len, int D) {
    int h = len / 2 + 1;
    int rows = D == 2 ? len : 1;
#pragma omp parallel
    {
        cplx *in = (cplx *)malloc(sizeof(cplx) * len), *out = (cplx *)malloc(sizeof(cplx) * len);
        if (D == 2) {
#pragma omp for schedule(static)
            for (int c = 0; c < h; c++) {
                for (int r = 0; r < len; r++) {
                    in[r].re = spec[2 * ((size_t)r * h + c)];
                    in[r].im = spec[2 * ((size_t)r * h + c) + 1];
                }
                fft_exec_any(pl, out, in,+1);
                for (int r = 0; r < len; r++) {
                    spec[2 * ((size_t)r * h + c)] = (float)out[r].re;
                    spec[2 * ((size_t)r * h + c) + 1] = (float)out[r].im;
                }
            }
        }
#pragma omp for schedule(static)
        for (int r = 0; r < rows; r++) {
            float *row = spec + 2 * (size_t)r * h;
            row[1] = 0.0f;
            if (len % 2 == 0) row[2 * (h - 1) + 1] = 0.0f;
            for (int c = 0; c < len; c++) {
                if (c < h) { in[c].re = row[2 * c]; in[c].im = row[2 * c + 1]; }
                else { in[c].re = row[2 * (len - c)]; in[c].im = -row[2 * (len - c) + 1]; }
            }
            fft_exec_any(pl, out, in,+1);
            for (int c = 0; c < len; c++) grid[(size_t)r * len + c] = (float)out[c].re;
        }
        free(in);
        free(out);
    }
}

/* charge of a point for a term (interpolation.rs:446-456) */
static float charge_value(const float *pt, int D, int term) {
    if (term == 0) return 1.0f;
    if (term >= 1 && term <= D) return pt[term - 1];
    float s = 0.0f;
    for (int a = 0; a < D; a++) s = s + pt[a] * pt[a];
    return s;
}

/* node_of_combo (interpolation.rs:462-481) */
static int64_t node_of_combo(int combo, const int64_t *box, const float *w, int64_t len, int D,
                             float *weight_out) {
    int rem = combo;
    int64_t flat = 0;
    float weight = 1.0f;
    for (int a = 0; a < D; a++) {
        int k = rem % IP;
        rem /= IP;
        int64_t node = box[a] * IP + k;
        flat = flat * len + node;
        weight = weight * w[a * IP + k];
    }
    *weight_out = weight;
    return flat;
}

/* Repulsive forces and Z (interpolation.rs:100-134).  min_boxes = 10 reproduces the reference;
 * larger values give a finer grid.  info_out (may be NULL): [boxes, fft_len, min, max]. */
EXPORT void oracle_fitsne_repulsion(const float *y, int64_t n, int D, int min_boxes, float *neg,
                                    float *z_out, double *info_out) {
    if (n == 0) {
        *z_out = 0.0f;
        return;
    }
    /* coordinate_bounds (interpolation.rs:337-348) */
    float mn = INFINITY, mx = -INFINITY;
    for (int64_t i = 0; i < n * D; i++) {
        mn = fminf(mn, y[i]);
        mx = fmaxf(mx, y[i]);
    }
    float extent = mx - mn;
    if (!(extent > FLT_MIN)) extent = fmaxf(extent, FLT_MIN);
    int boxes = oracle_box_count(extent, min_boxes);
    int nodes = boxes * IP;
    int len = (int)oracle_next_smooth(2 * (int64_t)nodes - 1);
    float box_width = extent / (float)boxes;
    float spacing = box_width / (float)IP;
    if (info_out) {
        info_out[0] = boxes;
        info_out[1] = len;
        info_out[2] = mn;
        info_out[3] = mx;
    }

    /* locate_points (interpolation.rs:162-187) */
    int64_t *box = (int64_t *)malloc(sizeof(int64_t) * (size_t)n * D);
    float *w = (float *)malloc(sizeof(float) * (size_t)n * D * IP);
    lagrange basis;
    lagrange_init(&basis);
    float inv_box_width = 1.0f / box_width;
    int64_t last_box = boxes - 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        for (int a = 0; a < D; a++) {
            float offset = (y[(size_t)i * D + a] - mn) * inv_box_width;
            float b = floorf(offset);
            int64_t bid;
            if (b > 0.0f) {
                bid = b >= 9.2e18f ? last_box : (int64_t)b;
                if (bid > last_box) bid = last_box;
            } else {
                bid = 0;
            }
            box[(size_t)i * D + a] = bid;
            float local = offset - (float)bid;
            lagrange_weights(&basis, local, w + (size_t)i * D * IP + (size_t)a * IP);
        }
    }

    fftplan pl;
    plan_init(&pl, len);
    size_t grid_len = D == 2 ? (size_t)len * len : (size_t)len;
    int h = len / 2 + 1;
    size_t spec_len = D == 2 ? (size_t)len * h : (size_t)h;
    float *grid = (float *)malloc(sizeof(float) * grid_len);
    float *kspec = (float *)malloc(sizeof(float) * 2 * spec_len);
    float *wspec = (float *)malloc(sizeof(float) * 2 * spec_len);
    int terms = D + 2;
    float *pot = (float *)malloc(sizeof(float) * (size_t)n * terms);

    /* build_kernel_spectrum (interpolation.rs:198-226) */
#pragma omp parallel for schedule(static)
    for (int64_t flat = 0; flat < (int64_t)grid_len; flat++) {
        float sq = 0.0f;
        int64_t rem = flat;
        for (int a = 0; a < D; a++) {
            int64_t coord = rem % len;
            rem /= len;
            int64_t lag = signed_lag(coord, len);
            float disp = spacing * (float)lag;
            sq = sq + disp * disp;
        }
        float cauchy = 1.0f / (1.0f + sq);
        grid[flat] = cauchy * cauchy;
    }
    fwd_f32(&pl, grid, kspec, len, D);
    {
        int64_t tot = 1;
        for (int a = 0; a < D; a++) tot *= len;
        float inv_total = 1.0f / (float)tot;
        for (size_t i = 0; i < 2 * spec_len; i++) kspec[i] = kspec[i] * inv_total;
    }

    int combos = D == 2 ? IP * IP : IP;
    for (int term = 0; term < terms; term++) {
        /* spread_charges (interpolation.rs:247-269), sequential */
        memset(grid, 0, sizeof(float) * grid_len);
        for (int64_t i = 0; i < n; i++) {
            float charge = charge_value(y + (size_t)i * D, D, term);
            for (int c = 0; c < combos; c++) {
                float wt;
                int64_t flat = node_of_combo(c, box + (size_t)i * D, w + (size_t)i * D * IP, len,
                                             D, &wt);
                grid[flat] = grid[flat] + wt * charge;
            }
        }
        fwd_f32(&pl, grid, wspec, len, D);
        /* multiply_by_kernel (interpolation.rs:273-278): complex product, f32 */
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)spec_len; i++) {
            float ar = wspec[2 * i], ai = wspec[2 * i + 1];
            float br = kspec[2 * i], bi = kspec[2 * i + 1];
            wspec[2 * i] = ar * br - ai * bi;
            wspec[2 * i + 1] = ar * bi + ai * br;
        }
        inv_f32(&pl, wspec, grid, len, D);
        /* gather_potentials (interpolation.rs:282-301) */
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            float phi = 0.0f;
            for (int c = 0; c < combos; c++) {
                float wt;
                int64_t flat = node_of_combo(c, box + (size_t)i * D, w + (size_t)i * D * IP, len,
                                             D, &wt);
                phi = phi + wt * grid[flat];
            }
            pot[(size_t)i * terms + term] = phi;
        }
    }

    /* reconstruct (interpolation.rs:305-332) */
    float *zterm = (float *)malloc(sizeof(float) * (size_t)n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const float *phi = pot + (size_t)i * terms;
        const float *pt = y + (size_t)i * D;
        float density = phi[0];
        float sqn = phi[D + 1];
        float norm_sq = 0.0f, cross = 0.0f;
        for (int a = 0; a < D; a++) {
            float fm = phi[1 + a];
            neg[(size_t)i * D + a] = pt[a] * density - fm;
            norm_sq = norm_sq + pt[a] * pt[a];
            cross = cross + pt[a] * fm;
        }
        zterm[i] = (1.0f + norm_sq) * density - 2.0f * cross + sqn;
    }
    /* par_iter ... .sum() over points (interpolation.rs:309-328): pairwise */
    float zs = pairwise_sum_f32(zterm, n, 1);
    *z_out = zs - (float)n;

    free(zterm);
    free(pot);
    free(wspec);
    free(kspec);
    free(grid);
    plan_free(&pl);
    free(w);
    free(box);
}

/* The same interpolated repulsion in T = f64 (the reference's f64 instantiation,
 * interpolation.rs:100-332 with DBL constants), every grid and transform in double.  This is the
 * interpolation scheme's answer without f32 rounding: the yardstick for the f32 noise of any
 * implementation (the reference's and this engine's alike).  The box plan is taken from the f32
 * coordinates so that both use the same grid. */
EXPORT void oracle_fitsne_repulsion_f64(const float *yf, int64_t n, int D, int min_boxes,
                                        double *neg, double *z_out) {
    if (n == 0) {
        *z_out = 0.0;
        return;
    }
    float mnf = INFINITY, mxf = -INFINITY;
    for (int64_t i = 0; i < n * D; i++) {
        mnf = fminf(mnf, yf[i]);
        mxf = fmaxf(mxf, yf[i]);
    }
    float extf = fmaxf(mxf - mnf, FLT_MIN);
    int boxes = oracle_box_count(extf, min_boxes);
    int nodes = boxes * IP;
    int len = (int)oracle_next_smooth(2 * (int64_t)nodes - 1);
    const double mn = mnf, extent = fmax((double)mxf - (double)mnf, DBL_MIN);
    const double box_width = extent / boxes, spacing = box_width / IP;
    size_t grid_len = D == 2 ? (size_t)len * len : (size_t)len;
    int terms = D + 2;
    double nodes_pos[IP], inv_denom[IP];
    for (int k = 0; k < IP; k++) nodes_pos[k] = (k + 0.5) / IP;
    for (int k = 0; k < IP; k++) {
        double d = 1.0;
        for (int m = 0; m < IP; m++)
            if (m != k) d *= nodes_pos[k] - nodes_pos[m];
        inv_denom[k] = 1.0 / d;
    }
    int64_t *box = (int64_t *)malloc(sizeof(int64_t) * (size_t)n * D);
    double *w = (double *)malloc(sizeof(double) * (size_t)n * D * IP);
    for (int64_t i = 0; i < n; i++)
        for (int a = 0; a < D; a++) {
            double off = ((double)yf[i * D + a] - mn) / box_width;
            double b = floor(off);
            int64_t bid = b > 0.0 ? (int64_t)b : 0;
            if (bid > boxes - 1) bid = boxes - 1;
            box[i * D + a] = bid;
            double local = off - (double)bid;
            for (int k = 0; k < IP; k++) {
                double num = 1.0;
                for (int m = 0; m < IP; m++)
                    if (m != k) num *= local - nodes_pos[m];
                w[(i * D + a) * IP + k] = num * inv_denom[k];
            }
        }
    /* kernel on the grid and its full complex 2-D DFT (double) */
    fftplan pl;
    plan_init(&pl, len);
    cplx *kern = (cplx *)malloc(sizeof(cplx) * grid_len);
    cplx *g = (cplx *)malloc(sizeof(cplx) * grid_len);
    cplx *tmp_in = (cplx *)malloc(sizeof(cplx) * len), *tmp_out = (cplx *)malloc(sizeof(cplx) * len);
    int rows = D == 2 ? len : 1;
    for (int64_t f = 0; f < (int64_t)grid_len; f++) {
        double sq = 0.0;
        int64_t rem = f;
        for (int a = 0; a < D; a++) {
            int64_t lag = signed_lag(rem % len, len);
            rem /= len;
            sq += (spacing * lag) * (spacing * lag);
        }
        double c = 1.0 / (1.0 + sq);
        kern[f].re = c * c;
        kern[f].im = 0.0;
    }
#define FFT2(buf, sign)                                                                       \
    do {                                                                                      \
        for (int r = 0; r < rows; r++) {                                                      \
            fft_exec(&pl, tmp_out, buf + (size_t)r * len, sign);                               \
            memcpy(buf + (size_t)r * len, tmp_out, sizeof(cplx) * len);                        \
        }                                                                                     \
        if (D == 2)                                                                           \
            for (int c = 0; c < len; c++) {                                                   \
                for (int r = 0; r < len; r++) tmp_in[r] = buf[(size_t)r * len + c];           \
                fft_exec(&pl, tmp_out, tmp_in, sign);                                          \
                for (int r = 0; r < len; r++) buf[(size_t)r * len + c] = tmp_out[r];          \
            }                                                                                 \
    } while (0)
    FFT2(kern, -1);
    double *pot = (double *)malloc(sizeof(double) * (size_t)n * terms);
    int combos = D == 2 ? IP * IP : IP;
    for (int term = 0; term < terms; term++) {
        memset(g, 0, sizeof(cplx) * grid_len);
        for (int64_t i = 0; i < n; i++) {
            double ch;
            if (term == 0) ch = 1.0;
            else if (term <= D) ch = yf[i * D + term - 1];
            else {
                ch = 0.0;
                for (int a = 0; a < D; a++) ch += (double)yf[i * D + a] * yf[i * D + a];
            }
            for (int c = 0; c < combos; c++) {
                int rem = c;
                int64_t flat = 0;
                double wt = 1.0;
                for (int a = 0; a < D; a++) {
                    int k = rem % IP;
                    rem /= IP;
                    flat = flat * len + box[i * D + a] * IP + k;
                    wt *= w[(i * D + a) * IP + k];
                }
                g[flat].re += wt * ch;
            }
        }
        FFT2(g, -1);
        for (size_t f = 0; f < grid_len; f++) {
            double ar = g[f].re, ai = g[f].im, br = kern[f].re, bi = kern[f].im;
            g[f].re = ar * br - ai * bi;
            g[f].im = ar * bi + ai * br;
        }
        FFT2(g, +1);
        double inv_total = 1.0 / (double)grid_len;
        for (int64_t i = 0; i < n; i++) {
            double phi = 0.0;
            for (int c = 0; c < combos; c++) {
                int rem = c;
                int64_t flat = 0;
                double wt = 1.0;
                for (int a = 0; a < D; a++) {
                    int k = rem % IP;
                    rem /= IP;
                    flat = flat * len + box[i * D + a] * IP + k;
                    wt *= w[(i * D + a) * IP + k];
                }
                phi += wt * g[flat].re * inv_total;
            }
            pot[i * terms + term] = phi;
        }
    }
#undef FFT2
    double zs = 0.0;
    for (int64_t i = 0; i < n; i++) {
        const double *phi = pot + i * terms;
        double ns = 0.0, cross = 0.0;
        for (int a = 0; a < D; a++) {
            double p = yf[i * D + a];
            neg[i * D + a] = p * phi[0] - phi[1 + a];
            ns += p * p;
            cross += p * phi[1 + a];
        }
        zs += (1.0 + ns) * phi[0] - 2.0 * cross + phi[D + 1];
    }
    *z_out = zs - (double)n;
    free(pot);
    free(tmp_in);
    free(tmp_out);
    free(g);
    free(kern);
    plan_free(&pl);
    free(w);
    free(box);
}

/* ------------------------------------------------------------------------------------------ */
/* Barnes-Hut Morton arena for D = 2..7 (src/tsne/arena.rs:137-392, morton.rs:1-120,            */
/* morton/d2.rs .. d7.rs)                                                                      */
/* ------------------------------------------------------------------------------------------ */

typedef unsigned __int128 u128;

/* Bits per axis of the Morton code (morton/d2.rs:32, d3.rs:31, d4.rs:29, d5.rs:90, d6.rs:98,
 * d7.rs:93): 32 / 21 / 16 / 25 / 21 / 18, so D * bits is at most 64 for D <= 4 (u64 words) and at
 * most 126 for D >= 5 (u128 words). */
static int morton_bits(int D) {
    static const int b[8] = {0, 0, 32, 21, 16, 25, 21, 18};
    return (D >= 2 && D <= 7) ? b[D] : 0;
}

/* Traversal stack length of the reference (Morton::Stack: 128 / 192 / 256 / 800 / 1344 / 2304). */
static int morton_stack(int D) {
    static const int s[8] = {0, 0, 128, 192, 256, 800, 1344, 2304};
    return (D >= 2 && D <= 7) ? s[D] : 0;
}

/* Z-order interleave: bit b of axis a lands at position b * D + a (the definition the reference's
 * per-D spreaders part_1by1 .. part_1by6 implement with masks and tables; pinned by their
 * known-answer tests, d2.rs:56-75 .. d7.rs:149-158). */
static u128 morton_encode(const uint32_t *c, int D) {
    const int bits = morton_bits(D);
    u128 code = 0;
    for (int b = 0; b < bits; b++)
        for (int a = 0; a < D; a++) code |= (u128)((c[a] >> b) & 1u) << (b * D + a);
    return code;
}

EXPORT uint64_t oracle_morton_encode2(uint32_t c0, uint32_t c1) {
    const uint32_t c[2] = {c0, c1};
    return (uint64_t)morton_encode(c, 2);
}

/* Generic encode: the code's low and high 64-bit words. */
EXPORT void oracle_morton_encode(const uint32_t *c, int D, uint64_t *lo, uint64_t *hi) {
    const u128 code = morton_encode(c, D);
    *lo = (uint64_t)code;
    *hi = (uint64_t)(code >> 64);
}

EXPORT int oracle_morton_bits(int D) { return morton_bits(D); }

/* quantize (morton.rs:103-120) for one axis */
EXPORT uint32_t oracle_quantize(float p, float mn, float inv_scale, uint32_t max_bucket) {
    float scaled = floorf((p - mn) * inv_scale);
    if (scaled > 0.0f) {
        if (scaled >= 18446744073709551616.0f) return max_bucket; /* to_u64() is None */
        uint64_t v = (uint64_t)scaled;
        return v < (uint64_t)max_bucket ? (uint32_t)v : max_bucket;
    }
    return 0;
}

typedef struct {
    float com[7];
    uint32_t count;
    uint32_t first_child;
    uint8_t child_count;
    uint8_t level;
} bhnode;

#define SENTINEL 0xffffffffu

typedef struct {
    u128 code;
    uint32_t idx;
} codeidx;

static int cmp_codeidx(const void *a, const void *b) {
    const codeidx *x = (const codeidx *)a, *y = (const codeidx *)b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return (x->idx > y->idx) - (x->idx < y->idx);
}

static int msb128(u128 x) {
    const uint64_t hi = (uint64_t)(x >> 64), lo = (uint64_t)x;
    return hi ? 127 - __builtin_clzll(hi) : 63 - __builtin_clzll(lo);
}

typedef struct {
    bhnode *nodes;
    int64_t n_nodes;
    int D, bits;
    float hw2[33];
    uint32_t (*ranges)[2]; /* kept only when keep_ranges is set */
    uint32_t *order;       /* sorted point indices, kept with the ranges */
    int keep_ranges;
} arena_t;

/* Arena::rebuild (arena.rs:137-328) for D = 2..7 */
static void arena_build(arena_t *ar, const float *y, int64_t n, int D) {
    const int bits = morton_bits(D);
    ar->nodes = NULL;
    ar->n_nodes = 0;
    ar->D = D;
    ar->bits = bits;
    ar->ranges = NULL;
    ar->order = NULL;
    if (n == 0) return;
    /* bounding_box (arena.rs:65-94) */
    float mn[7], mx[7];
    for (int a = 0; a < D; a++) {
        mn[a] = FLT_MAX;
        mx[a] = -FLT_MAX;
    }
    for (int64_t i = 0; i < n; i++)
        for (int a = 0; a < D; a++) {
            mn[a] = fminf(mn[a], y[i * D + a]);
            mx[a] = fmaxf(mx[a], y[i * D + a]);
        }
    float extent[7], inv_scale[7];
    const float scale = (float)(1ull << bits);
    const uint32_t max_bucket = (uint32_t)((1ull << bits) - 1);
    float max_extent = 0.0f;
    for (int a = 0; a < D; a++) {
        extent[a] = mx[a] - mn[a];
        inv_scale[a] = extent[a] > 0.0f ? scale / extent[a] : 0.0f;
        max_extent = fmaxf(max_extent, extent[a]);
    }
    for (int level = 0; level <= bits; level++) {
        float hw = max_extent / (float)(1ull << (level + 1));
        ar->hw2[level] = hw * hw;
    }
    codeidx *sorted = (codeidx *)malloc(sizeof(codeidx) * (size_t)n);
    for (int64_t i = 0; i < n; i++) {
        uint32_t q[7];
        for (int a = 0; a < D; a++)
            q[a] = oracle_quantize(y[i * D + a], mn[a], inv_scale[a], max_bucket);
        sorted[i].code = morton_encode(q, D);
        sorted[i].idx = (uint32_t)i;
    }
    qsort(sorted, (size_t)n, sizeof(codeidx), cmp_codeidx);

    int64_t cap = 2 * n - 1;
    bhnode *nodes = (bhnode *)calloc((size_t)cap, sizeof(bhnode));
    uint32_t(*ranges)[2] = malloc(sizeof(uint32_t[2]) * (size_t)cap);
    int64_t nn = 0;
    nodes[0].count = (uint32_t)n;
    nodes[0].first_child = SENTINEL;
    nodes[0].level = (uint8_t)bits;
    ranges[0][0] = 0;
    ranges[0][1] = (uint32_t)n;
    nn = 1;
    const u128 mask = ((u128)1 << D) - 1;
    for (int64_t node = 0; node < nn; node++) {
        uint32_t start = ranges[node][0], end = ranges[node][1];
        if (end - start <= 1 || sorted[start].code == sorted[end - 1].code) continue;
        int highest = msb128(sorted[start].code ^ sorted[end - 1].code);
        int level = (bits - 1) - highest / D;
        int shift = D * (bits - 1 - level);
        nodes[node].level = (uint8_t)level;
        uint32_t first_child = (uint32_t)nn;
        int child_count = 0;
        uint32_t cs = start;
        while (cs < end) {
            u128 g = (sorted[cs].code >> shift) & mask;
            uint32_t ce = cs + 1;
            while (ce < end && ((sorted[ce].code >> shift) & mask) == g) ce++;
            nodes[nn].count = ce - cs;
            nodes[nn].first_child = SENTINEL;
            nodes[nn].child_count = 0;
            nodes[nn].level = (uint8_t)bits;
            ranges[nn][0] = cs;
            ranges[nn][1] = ce;
            nn++;
            child_count++;
            cs = ce;
        }
        nodes[node].first_child = first_child;
        nodes[node].child_count = (uint8_t)child_count; /* at most 2^D <= 128 */
    }
    /* leaf COM (arena.rs:264-284) */
    for (int64_t node = 0; node < nn; node++) {
        if (nodes[node].first_child != SENTINEL) continue;
        float c[7] = {0};
        for (uint32_t s = ranges[node][0]; s < ranges[node][1]; s++) {
            const float *p = y + (size_t)sorted[s].idx * D;
            for (int a = 0; a < D; a++) c[a] += p[a];
        }
        float inv = 1.0f / (float)nodes[node].count;
        for (int a = 0; a < D; a++) nodes[node].com[a] = c[a] * inv;
    }
    /* internal COM bottom-up (arena.rs:289-309) */
    for (int64_t node = nn - 1; node >= 0; node--) {
        if (nodes[node].first_child == SENTINEL) continue;
        float total = (float)nodes[node].count;
        float c[7] = {0};
        for (uint32_t ch = nodes[node].first_child;
             ch < nodes[node].first_child + nodes[node].child_count; ch++) {
            float wgt = (float)nodes[ch].count;
            for (int a = 0; a < D; a++) c[a] += nodes[ch].com[a] * wgt;
        }
        float inv = 1.0f / total;
        for (int a = 0; a < D; a++) nodes[node].com[a] = c[a] * inv;
    }
    if (ar->keep_ranges) {
        ar->ranges = ranges;
        ar->order = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
        for (int64_t i = 0; i < n; i++) ar->order[i] = sorted[i].idx;
    } else {
        free(ranges);
    }
    free(sorted);
    ar->nodes = nodes;
    ar->n_nodes = nn;
}

/* compute_non_edge_forces (arena.rs:338-392) with the reference's per-D stack length */
static void arena_non_edge(const arena_t *ar, int64_t index, float theta_sq, const float *y,
                           float *neg, float *q_sum, uint32_t *stack) {
    if (ar->n_nodes == 0) return;
    const int D = ar->D, cap = morton_stack(D);
    const float *query = y + (size_t)index * D;
    int top = 0;
    stack[top++] = 0;
    while (top > 0) {
        top--;
        const bhnode *node = &ar->nodes[stack[top]];
        float disp[7];
        float distance = 0.0f;
        for (int a = 0; a < D; a++) {
            float delta = query[a] - node->com[a];
            disp[a] = delta;
            distance += delta * delta;
        }
        if (node->first_child == SENTINEL) {
            if (distance == 0.0f) continue;
        } else if (ar->hw2[node->level] >= theta_sq * distance) {
            if (top + node->child_count > cap) abort(); /* the reference's fixed stack overflows */
            for (uint32_t c = 0; c < node->child_count; c++) stack[top++] = node->first_child + c;
            continue;
        }
        float inverse = 1.0f / (1.0f + distance);
        float magnitude = (float)node->count * inverse;
        *q_sum += magnitude;
        magnitude = magnitude * inverse;
        for (int a = 0; a < D; a++) neg[a] += magnitude * disp[a];
    }
}

/* Build the arena and export its nodes (for invariant tests).  Arrays sized 2n-1 (com: D per
 * node); returns the node count (-1 for an unsupported D).  hw2_out has bits + 1 entries.
 * ranges_out (2 per node: start, end in the sorted order) and order_out (n sorted point
 * indices) may be NULL. */
EXPORT int64_t oracle_bh_build_ranges(const float *y, int64_t n, int D, float *com,
                                      uint32_t *count, uint32_t *first_child,
                                      uint8_t *child_count, uint8_t *level, float *hw2_out,
                                      uint32_t *ranges_out, uint32_t *order_out) {
    if (!morton_bits(D)) return -1;
    arena_t ar;
    ar.keep_ranges = ranges_out != NULL || order_out != NULL;
    arena_build(&ar, y, n, D);
    if (ar.ranges) {
        for (int64_t i = 0; i < ar.n_nodes && ranges_out; i++) {
            ranges_out[2 * i] = ar.ranges[i][0];
            ranges_out[2 * i + 1] = ar.ranges[i][1];
        }
        for (int64_t i = 0; i < n && order_out; i++) order_out[i] = ar.order[i];
        free(ar.ranges);
        free(ar.order);
    }
    for (int64_t i = 0; i < ar.n_nodes; i++) {
        for (int a = 0; a < D; a++) com[D * i + a] = ar.nodes[i].com[a];
        count[i] = ar.nodes[i].count;
        first_child[i] = ar.nodes[i].first_child;
        child_count[i] = ar.nodes[i].child_count;
        level[i] = ar.nodes[i].level;
    }
    if (hw2_out)
        for (int l = 0; l <= ar.bits; l++) hw2_out[l] = ar.hw2[l];
    int64_t nn = ar.n_nodes;
    free(ar.nodes);
    return nn;
}

EXPORT int64_t oracle_bh_build_d(const float *y, int64_t n, int D, float *com, uint32_t *count,
                                 uint32_t *first_child, uint8_t *child_count, uint8_t *level,
                                 float *hw2_out) {
    return oracle_bh_build_ranges(y, n, D, com, count, first_child, child_count, level, hw2_out,
                                  NULL, NULL);
}

EXPORT int64_t oracle_bh_build(const float *y, int64_t n, float *com, uint32_t *count,
                               uint32_t *first_child, uint8_t *child_count, uint8_t *level,
                               float *hw2_out) {
    return oracle_bh_build_d(y, n, 2, com, count, first_child, child_count, level, hw2_out);
}

/* Barnes-Hut repulsion for every point: neg (n*D) and per-point q sums; returns sum of q (f32,
 * sequential as repulsion.rs:155). */
EXPORT float oracle_bh_repulsion_d(const float *y, int64_t n, int D, float theta, float *neg,
                                   float *q_sums) {
    if (!morton_bits(D)) abort();
    arena_t ar;
    ar.keep_ranges = 0;
    arena_build(&ar, y, n, D);
    float theta_sq = theta * theta;
#pragma omp parallel
    {
        uint32_t *stack = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)morton_stack(D));
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < n; i++) {
            float f[7] = {0};
            float q = 0.0f;
            arena_non_edge(&ar, i, theta_sq, y, f, &q, stack);
            for (int a = 0; a < D; a++) neg[i * D + a] = f[a];
            q_sums[i] = q;
        }
        free(stack);
    }
    float s = 0.0f;
    for (int64_t i = 0; i < n; i++) s = s + q_sums[i];
    free(ar.nodes);
    return s;
}

EXPORT float oracle_bh_repulsion(const float *y, int64_t n, float theta, float *neg,
                                 float *q_sums) {
    return oracle_bh_repulsion_d(y, n, 2, theta, neg, q_sums);
}

/* ------------------------------------------------------------------------------------------ */
/* One approximate epoch and the run loop (src/lib.rs:961-1014, 1162-1179; repulsion.rs)       */
/* ------------------------------------------------------------------------------------------ */

/* Forces for the current state: pos, neg (n*D) and returns inverse_norm (1/Z).
 * method 0 = FIt-SNE (min_boxes, D = 1, 2), 1 = Barnes-Hut (theta, D = 2..7). */
EXPORT float oracle_forces(const float *y, int64_t n, int D, const uint64_t *rows,
                           const uint32_t *cols, const float *vals, int method, float theta,
                           int min_boxes, float *pos, float *neg) {
    oracle_edge_forces(y, n, D, rows, cols, vals, pos);
    if (method == 0) {
        float z = 0.0f;
        oracle_fitsne_repulsion(y, n, D, min_boxes, neg, &z, NULL);
        return 1.0f / z;
    }
    float *q = (float *)malloc(sizeof(float) * (size_t)n);
    float s = oracle_bh_repulsion_d(y, n, D, theta, neg, q);
    free(q);
    return 1.0f / s;
}

/* Run epochs [e0, e1) of the approximate loop.  vals holds the current (normalised, possibly
 * exaggerated) P and is rescaled in place by stop-lying exactly as the reference does.
 * sched: [learning_rate, momentum, final_momentum, early_exaggeration];
 * isched: [momentum_switch_epoch, stop_lying_epoch].  *momentum_io carries the momentum across
 * calls (the switch overwrites it permanently, lib.rs:1172-1174). */
EXPORT void oracle_run_epochs(float *y, float *uy, float *gains, int64_t n, int D,
                              const uint64_t *rows, const uint32_t *cols, float *vals,
                              int method, float theta, int min_boxes, const float *sched,
                              const int64_t *isched, int64_t e0, int64_t e1, float *momentum_io) {
    int64_t len = n * D;
    float *pos = (float *)malloc(sizeof(float) * (size_t)len);
    float *neg = (float *)malloc(sizeof(float) * (size_t)len);
    float lr = sched[0], final_m = sched[2], exag = sched[3];
    float momentum = *momentum_io;
    int64_t nnz = (int64_t)rows[n];
    for (int64_t e = e0; e < e1; e++) {
        float inv_norm = oracle_forces(y, n, D, rows, cols, vals, method, theta, min_boxes, pos,
                                       neg);
        oracle_gradient_step(y, pos, neg, uy, gains, len, lr, momentum, inv_norm);
        oracle_zero_mean(y, n, D, NULL);
        if (e == isched[1] && e != 0) oracle_stop_lying(vals, nnz, exag);
        if (e == isched[0]) momentum = final_m;
    }
    *momentum_io = momentum;
    free(pos);
    free(neg);
}

/* ------------------------------------------------------------------------------------------ */
/* KL divergence (src/tsne/mod.rs:635-775)                                                     */
/* ------------------------------------------------------------------------------------------ */

/* sparse_kl_divergence with the reference's indexing (p_values[index], the sample index, mod.rs
 * 768-770) in f32 (returned via ref_out), and the true KL (edge values) accumulated in double. */
EXPORT void oracle_sparse_kl(const float *y, int64_t n, int D, const uint64_t *rows,
                             const uint32_t *cols, const float *vals, float inverse_q_sum,
                             float *ref_out, double *true_out) {
    float *partials = (float *)malloc(sizeof(float) * (size_t)n);
    double tk = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : tk)
    for (int64_t i = 0; i < n; i++) {
        float cost = 0.0f;
        const float *a = y + (size_t)i * D;
        for (uint64_t e = rows[i]; e < rows[i + 1]; e++) {
            const float *b = y + (size_t)cols[e] * D;
            float q = 0.0f;
            for (int t = 0; t < D; t++) {
                float dd = a[t] - b[t];
                q = q + dd * dd;
            }
            q = (1.0f / (1.0f + q)) * inverse_q_sum;
            cost += vals[i] * logf((vals[i] + FLT_MIN) / (q + FLT_MIN));
            double pe = vals[e];
            if (pe > 0.0) tk += pe * log(pe / (double)q);
        }
        partials[i] = cost;
    }
    *ref_out = pairwise_sum_f32(partials, n, 1); /* par_iter().sum() (mod.rs:774) */
    *true_out = tk;
    free(partials);
}

/* kl_divergence() (lib.rs:448-477) for the approximate fits: Z from the fit's own estimator. */
EXPORT void oracle_kl_divergence(const float *y, int64_t n, int D, const uint64_t *rows,
                                 const uint32_t *cols, const float *vals, int method, float theta,
                                 int min_boxes, float *ref_out, double *true_out) {
    float *neg = (float *)malloc(sizeof(float) * (size_t)n * D);
    float z;
    if (method == 0) {
        oracle_fitsne_repulsion(y, n, D, min_boxes, neg, &z, NULL);
    } else {
        float *q = (float *)malloc(sizeof(float) * (size_t)n);
        z = oracle_bh_repulsion_d(y, n, D, theta, neg, q);
        free(q);
    }
    free(neg);
    oracle_sparse_kl(y, n, D, rows, cols, vals, 1.0f / z, ref_out, true_out);
}

/* ------------------------------------------------------------------------------------------ */
/* FP64 exact gradient (the mathematics the approximations target; no reference quirks)       */
/* ------------------------------------------------------------------------------------------ */

/* Exact repulsion and Z over all pairs in double.  neg (n*D), returns Z = sum_{i != j} w_ij. */
EXPORT double oracle_exact_repulsion_f64(const float *y, int64_t n, int D, double *neg) {
    double z = 0.0;
#pragma omp parallel for schedule(dynamic, 32) reduction(+ : z)
    for (int64_t i = 0; i < n; i++) {
        double f[8] = {0};
        double zi = 0.0;
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            double d2 = 0.0, dl[8];
            for (int a = 0; a < D; a++) {
                dl[a] = (double)y[i * D + a] - (double)y[j * D + a];
                d2 += dl[a] * dl[a];
            }
            double w = 1.0 / (1.0 + d2);
            zi += w;
            for (int a = 0; a < D; a++) f[a] += w * w * dl[a];
        }
        for (int a = 0; a < D; a++) neg[i * D + a] = f[a];
        z += zi;
    }
    return z;
}

/* Exact repulsion for a subset of query rows against all points; returns their partial Z. */
EXPORT double oracle_exact_repulsion_rows_f64(const float *y, int64_t n, int D,
                                              const int64_t *rows_sel, int64_t m, double *neg) {
    double z = 0.0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : z)
    for (int64_t t = 0; t < m; t++) {
        int64_t i = rows_sel[t];
        double f[8] = {0};
        double zi = 0.0;
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            double d2 = 0.0, dl[8];
            for (int a = 0; a < D; a++) {
                dl[a] = (double)y[i * D + a] - (double)y[j * D + a];
                d2 += dl[a] * dl[a];
            }
            double w = 1.0 / (1.0 + d2);
            zi += w;
            for (int a = 0; a < D; a++) f[a] += w * w * dl[a];
        }
        for (int a = 0; a < D; a++) neg[t * D + a] = f[a];
        z += zi;
    }
    return z;
}

/* Exact attraction in double for every row. */
EXPORT void oracle_exact_attraction_f64(const float *y, int64_t n, int D, const uint64_t *rows,
                                        const uint32_t *cols, const float *vals, double *pos) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double f[8] = {0};
        for (uint64_t e = rows[i]; e < rows[i + 1]; e++) {
            double d2 = 0.0, dl[8];
            for (int a = 0; a < D; a++) {
                dl[a] = (double)y[i * D + a] - (double)y[(size_t)cols[e] * D + a];
                d2 += dl[a] * dl[a];
            }
            double w = (double)vals[e] / (1.0 + d2);
            for (int a = 0; a < D; a++) f[a] += w * dl[a];
        }
        for (int a = 0; a < D; a++) pos[i * D + a] = f[a];
    }
}

/* ------------------------------------------------------------------------------------------ */
/* The exact fit (src/lib.rs:489-646), restated with the reference's own semantics: dense rows  */
/* include each point's own entry (d_ii = 0), the metric value is squared again inside          */
/* search_beta, and Z sums q over all n^2 ordered pairs, diagonal included.  (The FP64          */
/* yardstick above, oracle_exact_repulsion_f64, is the quirk-free O(n^2) repulsion.)            */
/* ------------------------------------------------------------------------------------------ */

/* Dense distances (compute_pairwise_distance_matrix, mod.rs:161-189): the upper triangle from
 * f(x_i, x_j), mirrored; the diagonal 0.  metric 0: sqrt of the left-to-right sum of squared
 * differences; 1: the sum itself. */
EXPORT void oracle_exact_distances(const float *x, int64_t n, int d, int metric, float *dist) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; i++) {
        dist[(size_t)i * n + i] = 0.0f;
        for (int64_t j = i + 1; j < n; j++) {
            float s = 0.0f;
            for (int t = 0; t < d; t++) {
                float diff = x[(size_t)i * d + t] - x[(size_t)j * d + t];
                s = s + diff * diff;
            }
            dist[(size_t)i * n + j] = metric == 0 ? sqrtf(s) : s;
        }
    }
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++) dist[(size_t)j * n + i] = dist[(size_t)i * n + j];
}

/* Dense P (lib.rs:529-561): search_beta on every full row, P_ij = P_ji = p_ij + p_ji for i < j,
 * then normalize_p_values with the exaggeration (and stop_lying at once when stop_lying_epoch is
 * 0, finalize_p_and_seed lib.rs:663-669). */
EXPORT void oracle_exact_affinities(const float *x, int64_t n, int d, int metric,
                                    float perplexity, float exag, int stop_lying_now, float *P) {
    float *dist = (float *)malloc(sizeof(float) * (size_t)(n * n));
    oracle_exact_distances(x, n, d, metric, dist);
    oracle_search_beta(P, dist, n, (int)n, perplexity);
    free(dist);
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++) {
            float s = P[(size_t)i * n + j] + P[(size_t)j * n + i];
            P[(size_t)i * n + j] = s;
            P[(size_t)j * n + i] = s;
        }
    oracle_normalize_p(P, n * n, exag);
    if (stop_lying_now) oracle_stop_lying(P, n * n, exag);
}

/* Exact gradient of one epoch (lib.rs:571-623): d_ij = sum of powi(2) differences, q = 1/(1+d),
 * Z = sum of all n^2 q (pairwise f32 sum, as rayon's), inverse 1/Z in f32, and per row
 * dy_i += (y_i - y_j) * ((p - q inv) * q) left to right over j.  Returns 1/Z. */
EXPORT float oracle_exact_gradient(const float *y, int64_t n, int D, const float *P, float *dy) {
    float *q = (float *)malloc(sizeof(float) * (size_t)(n * n));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) {
            float s = 0.0f;
            for (int a = 0; a < D; a++) {
                float diff = y[(size_t)i * D + a] - y[(size_t)j * D + a];
                s = s + diff * diff;
            }
            q[(size_t)i * n + j] = 1.0f / (1.0f + s);
        }
    const float inv = 1.0f / pairwise_sum_f32(q, n * n, 1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        float acc[8] = {0};
        for (int64_t j = 0; j < n; j++) {
            const float qq = q[(size_t)i * n + j];
            const float m = (P[(size_t)i * n + j] - qq * inv) * qq;
            for (int a = 0; a < D; a++)
                acc[a] += (y[(size_t)i * D + a] - y[(size_t)j * D + a]) * m;
        }
        for (int a = 0; a < D; a++) dy[(size_t)i * D + a] = acc[a];
    }
    free(q);
    return inv;
}

/* Epochs [e0, e1) of the exact loop: gradient, update_solution (the same step as the approximate
 * paths' with a zero repulsion), zero_mean, the schedule (lib.rs:1162-1179).  sched / isched /
 * momentum_io as in oracle_run_epochs. */
EXPORT void oracle_exact_epochs(float *y, float *uy, float *gains, int64_t n, int D, float *P,
                                const float *sched, const int64_t *isched, int64_t e0, int64_t e1,
                                float *momentum_io) {
    int64_t len = n * D;
    float *dy = (float *)malloc(sizeof(float) * (size_t)len);
    float *zero = (float *)calloc((size_t)len, sizeof(float));
    float lr = sched[0], final_m = sched[2], exag = sched[3];
    float momentum = *momentum_io;
    for (int64_t e = e0; e < e1; e++) {
        oracle_exact_gradient(y, n, D, P, dy);
        oracle_gradient_step(y, dy, zero, uy, gains, len, lr, momentum, 0.0f);
        oracle_zero_mean(y, n, D, NULL);
        if (e == isched[1] && e != 0) oracle_stop_lying(P, n * n, exag);
        if (e == isched[0]) momentum = final_m;
    }
    *momentum_io = momentum;
    free(dy);
    free(zero);
}

/* evaluate_error (mod.rs:579-618): sum over all n^2 pairs of p ln((p + min) / (q inv + min)), in
 * f32 (returned) and in double (true_out). */
EXPORT float oracle_exact_kl(const float *y, int64_t n, int D, const float *P, double *true_out) {
    float *q = (float *)malloc(sizeof(float) * (size_t)(n * n));
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) {
            float s = 0.0f;
            for (int a = 0; a < D; a++) {
                float diff = y[(size_t)i * D + a] - y[(size_t)j * D + a];
                s = s + diff * diff;
            }
            q[(size_t)i * n + j] = 1.0f / (1.0f + s);
        }
    const float inv = 1.0f / pairwise_sum_f32(q, n * n, 1);
    float kl = 0.0f;
    double t = 0.0;
    for (int64_t e = 0; e < n * n; e++) {
        const float p = P[e], qn = q[e] * inv;
        kl = kl + p * logf((p + FLT_MIN) / (qn + FLT_MIN));
        t += (double)p * log(((double)p + (double)FLT_MIN) / ((double)qn + (double)FLT_MIN));
    }
    free(q);
    if (true_out) *true_out = t;
    return kl;
}
