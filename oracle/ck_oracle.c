/*
 * ck_oracle.c - CPU ORACLE of the soft-decision receive tail.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the reference's
 * CPU algorithms (commkit 3.11.0, Numba kernels) so that tests can check the
 * HIP library against them and bench.py can time them as the CPU baseline.
 * Nothing in the product (commkit_amd/) links, loads or calls it.
 *
 * Built by oracle/Makefile with -O2 -fno-fast-math -ffp-contract=off and
 * OpenMP: every float32 product and sum rounds on its own, as in the
 * reference's decoder (commkit/coding/_numba_loops.py:109, "compiled
 * without fast-math so min-sum matches the kernel bit for bit").
 *
 * Functions and the reference lines each restates:
 *   oracle_ldpc_layered      commkit/coding/_numba_loops.py:31-147
 *                            (phi, min_sum_row, sum_product_row, satisfied,
 *                            ldpc_layered): rows in the given schedule order,
 *                            parallel over frames.
 *   oracle_ldpc_layered_probe  the layered sum-product with phi evaluated in
 *                            float64 and rounded once (phi_probe): not the
 *                            reference; it measures the float32 algorithm's
 *                            sensitivity to the rounding of phi.
 *   oracle_ldpc_flooding     no reference counterpart (the reference has only
 *                            the layered schedule).  Restates this library's
 *                            flooding definition: each check updates from
 *                            Q = APP_prev - R_old; then APP = L + R_1 + R_2 +
 *                            ... over the variable's checks in ascending
 *                            check index, left to right.
 *   oracle_llr_points        commkit/mapping/_numba_llr.py:35-124
 *   oracle_llr_separable     commkit/mapping/_numba_llr.py:84-92,127-176
 *   oracle_llr_separable_maxlog  commkit/mapping/_numba_llr.py:179-235
 *   oracle_ccdm_match        commkit/coding/_numba_loops.py:944-1004
 *                            (ccdm_shares, ccdm_match): the constant-
 *                            composition matcher's 64-bit range coder.
 *   oracle_ccdm_dematch      commkit/coding/_numba_loops.py:1010-1102
 *                            (low_bit, ccdm_dematch): its inverse, with the
 *                            interval start kept exactly in 32-bit limbs.
 * The reference LLR kernels are fast-math (reassociation allowed), so LLRs
 * match them to rounding, not bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXACT_FLOOR 1e-26f
#define LLR_TASK 256

/* ------------------------------------------------------------------------ */
/* LDPC decoding                                                            */
/* ------------------------------------------------------------------------ */

static float phi(float x) {
    /* -log tanh(x/2), x clipped to [1e-7, 30] (_numba_loops.py:36-40) */
    if (x < 1e-7f) x = 1e-7f;
    if (x > 30.0f) x = 30.0f;
    return -logf(tanhf(0.5f * x));
}

/* The conditioning probe's phi: the same clipped function evaluated in
 * float64 as log((1 + e^-x) / (1 - e^-x)) and rounded once to float32, so it
 * differs from phi() only by float32 rounding (a fraction of an ulp to a few
 * ulps).  Not the reference's arithmetic: decoding with it measures how far a
 * float32 sum-product decoder moves when phi is rounded differently. */
static float phi_probe(float x) {
    if (x < 1e-7f) x = 1e-7f;
    if (x > 30.0f) x = 30.0f;
    const double d = (double)x;
    return (float)log((1.0 + exp(-d)) / -expm1(-d));
}

typedef float (*phi_fn)(float);

static void min_sum_row(float* a, float* r, const int32_t* cols, int e0, int e1, float scale,
                        float offset) {
    /* _numba_loops.py:43-68 */
    float min1 = INFINITY, min2 = INFINITY;
    int idx = -1;
    int neg = 0;
    for (int e = e0; e < e1; ++e) {
        float q = a[cols[e]] - r[e];
        neg ^= (q < 0.0f);
        float mag = fabsf(q);
        if (mag < min1) {
            min2 = min1;
            min1 = mag;
            idx = e;
        } else if (mag < min2) {
            min2 = mag;
        }
    }
    float m1 = scale * min1;
    m1 = m1 - offset;
    m1 = m1 > 0.0f ? m1 : 0.0f;
    float m2 = scale * min2;
    m2 = m2 - offset;
    m2 = m2 > 0.0f ? m2 : 0.0f;
    for (int e = e0; e < e1; ++e) {
        int c = cols[e];
        float q = a[c] - r[e];
        float mag = (e == idx) ? m2 : m1;
        float rn = (neg ^ (q < 0.0f)) ? -mag : mag;
        a[c] = q + rn;
        r[e] = rn;
    }
}

static void sum_product_row(float* a, float* r, const int32_t* cols, int e0, int e1, phi_fn f) {
    /* _numba_loops.py:71-94: pass 1 keeps Q in a and phi(|Q|) in r */
    float total = 0.0f;
    int neg = 0;
    for (int e = e0; e < e1; ++e) {
        int c = cols[e];
        float q = a[c] - r[e];
        neg ^= (q < 0.0f);
        float p = f(fabsf(q));
        total += p;
        a[c] = q;
        r[e] = p;
    }
    for (int e = e0; e < e1; ++e) {
        int c = cols[e];
        float q = a[c];
        float mag = f(total - r[e]);
        float rn = (neg ^ (q < 0.0f)) ? -mag : mag;
        a[c] = q + rn;
        r[e] = rn;
    }
}

static int satisfied(const float* a, const int32_t* row_ptr, const int32_t* cols, int m) {
    /* _numba_loops.py:97-106 */
    for (int row = 0; row < m; ++row) {
        int parity = 0;
        for (int e = row_ptr[row]; e < row_ptr[row + 1]; ++e) parity ^= (a[cols[e]] < 0.0f);
        if (parity) return 0;
    }
    return 1;
}

static void layered(float* app, int64_t R, int32_t N, const int32_t* row_ptr, const int32_t* cols,
                    int32_t m, int32_t sum_product, float scale, float offset,
                    int32_t max_iterations, int32_t early_stop, int32_t* iters, uint8_t* valid,
                    phi_fn f);

/* Layered decoding of R frames of app (R, N) in place; rows in schedule
 * order (row_ptr, cols) as LDPCCode._schedule gives them. */
void oracle_ldpc_layered(float* app, int64_t R, int32_t N, const int32_t* row_ptr,
                         const int32_t* cols, int32_t m, int32_t sum_product, float scale,
                         float offset, int32_t max_iterations, int32_t early_stop,
                         int32_t* iters, uint8_t* valid) {
    layered(app, R, N, row_ptr, cols, m, sum_product, scale, offset, max_iterations, early_stop,
            iters, valid, phi);
}

/* Layered sum-product with phi_probe in place of phi: the conditioning
 * probe (no reference counterpart; test infrastructure). */
void oracle_ldpc_layered_probe(float* app, int64_t R, int32_t N, const int32_t* row_ptr,
                               const int32_t* cols, int32_t m, int32_t max_iterations,
                               int32_t early_stop, int32_t* iters, uint8_t* valid) {
    layered(app, R, N, row_ptr, cols, m, 1, 1.0f, 0.0f, max_iterations, early_stop, iters, valid,
            phi_probe);
}

static void layered(float* app, int64_t R, int32_t N, const int32_t* row_ptr, const int32_t* cols,
                    int32_t m, int32_t sum_product, float scale, float offset,
                    int32_t max_iterations, int32_t early_stop, int32_t* iters, uint8_t* valid,
                    phi_fn f) {
    const int E = row_ptr[m];
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < R; ++b) {
        float* a = app + b * (int64_t)N;
        float* r = (float*)calloc((size_t)E, sizeof(float));
        int it = 0;
        int ok = early_stop && satisfied(a, row_ptr, cols, m);
        while (!ok && it < max_iterations) {
            for (int row = 0; row < m; ++row) {
                if (sum_product)
                    sum_product_row(a, r, cols, row_ptr[row], row_ptr[row + 1], f);
                else
                    min_sum_row(a, r, cols, row_ptr[row], row_ptr[row + 1], scale, offset);
            }
            ++it;
            if (early_stop) ok = satisfied(a, row_ptr, cols, m);
        }
        if (!early_stop) ok = satisfied(a, row_ptr, cols, m);
        iters[b] = it;
        valid[b] = (uint8_t)ok;
        free(r);
    }
}

/* Flooding decoding (this library's definition).  var_ptr/var_edges list,
 * per variable, its edges (indices into cols) in ascending natural check
 * index; L is the channel LLRs (R, N) (punctured positions 0). */
void oracle_ldpc_flooding(float* app, int64_t R, int32_t N, const int32_t* row_ptr,
                          const int32_t* cols, int32_t m, const int32_t* var_ptr,
                          const int32_t* var_edges, int32_t sum_product, float scale,
                          float offset, int32_t max_iterations, int32_t early_stop,
                          int32_t* iters, uint8_t* valid) {
    const int E = row_ptr[m];
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < R; ++b) {
        float* a = app + b * (int64_t)N;
        float* L = (float*)malloc((size_t)N * sizeof(float));
        float* r = (float*)calloc((size_t)E, sizeof(float));
        float* rn_all = (float*)calloc((size_t)E, sizeof(float));
        memcpy(L, a, (size_t)N * sizeof(float));
        int it = 0;
        int ok = early_stop && satisfied(a, row_ptr, cols, m);
        while (!ok && it < max_iterations) {
            for (int row = 0; row < m; ++row) {
                int e0 = row_ptr[row], e1 = row_ptr[row + 1];
                int neg = 0;
                if (sum_product) {
                    float total = 0.0f;
                    for (int e = e0; e < e1; ++e) {
                        float q = a[cols[e]] - r[e];
                        neg ^= (q < 0.0f);
                        total += phi(fabsf(q));
                    }
                    for (int e = e0; e < e1; ++e) {
                        float q = a[cols[e]] - r[e];
                        float mag = phi(total - phi(fabsf(q)));
                        rn_all[e] = (neg ^ (q < 0.0f)) ? -mag : mag;
                    }
                } else {
                    float min1 = INFINITY, min2 = INFINITY;
                    int idx = -1;
                    for (int e = e0; e < e1; ++e) {
                        float q = a[cols[e]] - r[e];
                        neg ^= (q < 0.0f);
                        float mag = fabsf(q);
                        if (mag < min1) {
                            min2 = min1;
                            min1 = mag;
                            idx = e;
                        } else if (mag < min2) {
                            min2 = mag;
                        }
                    }
                    float m1 = scale * min1;
                    m1 = m1 - offset;
                    m1 = m1 > 0.0f ? m1 : 0.0f;
                    float m2 = scale * min2;
                    m2 = m2 - offset;
                    m2 = m2 > 0.0f ? m2 : 0.0f;
                    for (int e = e0; e < e1; ++e) {
                        float q = a[cols[e]] - r[e];
                        float mag = (e == idx) ? m2 : m1;
                        rn_all[e] = (neg ^ (q < 0.0f)) ? -mag : mag;
                    }
                }
            }
            memcpy(r, rn_all, (size_t)E * sizeof(float));
            for (int v = 0; v < N; ++v) {
                float sum = L[v];
                for (int t = var_ptr[v]; t < var_ptr[v + 1]; ++t) sum = sum + r[var_edges[t]];
                a[v] = sum;
            }
            ++it;
            if (early_stop) ok = satisfied(a, row_ptr, cols, m);
        }
        if (!early_stop) ok = satisfied(a, row_ptr, cols, m);
        iters[b] = it;
        valid[b] = (uint8_t)ok;
        free(L);
        free(r);
        free(rn_all);
    }
}

/* ------------------------------------------------------------------------ */
/* LLRs                                                                     */
/* ------------------------------------------------------------------------ */

static float prepare_exp(const float* metric, float* e, int M) {
    /* _numba_llr.py:35-44 */
    float peak = metric[0];
    for (int m = 1; m < M; ++m)
        if (metric[m] > peak) peak = metric[m];
    for (int m = 0; m < M; ++m) e[m] = expf(metric[m] - peak);
    return peak;
}

/* idx: (2, k, half) int32, the points of each bit with value v */
static float log_half_sum(const float* metric, const float* e, const int32_t* idx, int k,
                          int half, int v, int b, float peak) {
    /* _numba_llr.py:47-66 */
    const int32_t* h = idx + ((int64_t)v * k + b) * half;
    float s = 0.0f;
    for (int j = 0; j < half; ++j) s += e[h[j]];
    if (s >= EXACT_FLOOR) return logf(s);
    float p = metric[h[0]];
    for (int j = 1; j < half; ++j)
        if (metric[h[j]] > p) p = metric[h[j]];
    s = 0.0f;
    for (int j = 0; j < half; ++j) s += expf(metric[h[j]] - p);
    return logf(s) + (p - peak);
}

static float bit_llr(const float* metric, const float* e, const int32_t* idx, int k, int half,
                     int b, float peak, int exact) {
    /* _numba_llr.py:69-81 */
    if (exact)
        return log_half_sum(metric, e, idx, k, half, 0, b, peak) -
               log_half_sum(metric, e, idx, k, half, 1, b, peak);
    const int32_t* h0 = idx + (int64_t)b * half;
    const int32_t* h1 = idx + ((int64_t)k + b) * half;
    float p0 = metric[h0[0]], p1 = metric[h1[0]];
    for (int j = 1; j < half; ++j) {
        if (metric[h0[j]] > p0) p0 = metric[h0[j]];
        if (metric[h1[j]] > p1) p1 = metric[h1[j]];
    }
    return p0 - p1;
}

/* LLRs from the metrics of every point (_numba_llr.py:95-124).
 * x_re/x_im (n,), c_re/c_im/log_pmf (M,), idx (2, k, M/2), out (n, k). */
void oracle_llr_points(const float* x_re, const float* x_im, int64_t n, const float* c_re,
                       const float* c_im, const float* log_pmf, const int32_t* idx, int32_t M,
                       int32_t k, float inv_s2, int32_t exact, float* out) {
    const int64_t tasks = (n + LLR_TASK - 1) / LLR_TASK;
    const int half = M / 2;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < tasks; ++t) {
        float* metric = (float*)malloc((size_t)M * sizeof(float));
        float* e = (float*)malloc((size_t)M * sizeof(float));
        int64_t end = (t + 1) * LLR_TASK < n ? (t + 1) * LLR_TASK : n;
        for (int64_t i = t * LLR_TASK; i < end; ++i) {
            float xr = x_re[i], xi = x_im[i];
            for (int m = 0; m < M; ++m) {
                float dr = xr - c_re[m], di = xi - c_im[m];
                metric[m] = -(dr * dr + di * di) * inv_s2 + log_pmf[m];
            }
            float peak = exact ? prepare_exp(metric, e, M) : 0.0f;
            for (int b = 0; b < k; ++b)
                out[i * k + b] = bit_llr(metric, e, idx, k, half, b, peak, exact);
        }
        free(metric);
        free(e);
    }
}

static void axis_llrs(float x, const float* lev, const float* logp, const int32_t* idx,
                      const int32_t* bits, int L, int kb, float inv_s2, int exact,
                      float* metric, float* e, float* out_row) {
    /* _numba_llr.py:84-92 */
    for (int m = 0; m < L; ++m) {
        float d = x - lev[m];
        metric[m] = -(d * d) * inv_s2 + logp[m];
    }
    float peak = exact ? prepare_exp(metric, e, L) : 0.0f;
    for (int b = 0; b < kb; ++b)
        out_row[bits[b]] = bit_llr(metric, e, idx, kb, L / 2, b, peak, exact);
}

/* LLRs from the metrics of each axis's levels (_numba_llr.py:127-176).
 * idx_a (2, kb, L/2); bits_a (kb,) positions; out (n, k) with k = 2 kb. */
void oracle_llr_separable(const float* x_re, const float* x_im, int64_t n, const float* lev_re,
                          const float* lev_im, const float* logp_re, const float* logp_im,
                          const int32_t* idx_re, const int32_t* idx_im, const int32_t* bits_re,
                          const int32_t* bits_im, int32_t L, int32_t kb, float inv_s2,
                          int32_t exact, float* out) {
    const int64_t tasks = (n + LLR_TASK - 1) / LLR_TASK;
    const int k = 2 * kb;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < tasks; ++t) {
        float* m_re = (float*)malloc((size_t)L * sizeof(float));
        float* e_re = (float*)malloc((size_t)L * sizeof(float));
        int64_t end = (t + 1) * LLR_TASK < n ? (t + 1) * LLR_TASK : n;
        for (int64_t i = t * LLR_TASK; i < end; ++i) {
            axis_llrs(x_re[i], lev_re, logp_re, idx_re, bits_re, L, kb, inv_s2, exact, m_re,
                      e_re, out + i * k);
            axis_llrs(x_im[i], lev_im, logp_im, idx_im, bits_im, L, kb, inv_s2, exact, m_re,
                      e_re, out + i * k);
        }
        free(m_re);
        free(e_re);
    }
}

static void axis_maxlog(float x, const float* lev, const int32_t* lab, const int32_t* lo,
                        const int32_t* hi, const int32_t* pos, int L, int kb, float inv_s2,
                        float* out_row) {
    /* _numba_llr.py:179-210 */
    int s = (int)rintf((x - lev[0]) / (lev[1] - lev[0]));
    s = s < 0 ? 0 : (s > L - 1 ? L - 1 : s);
    float t = x - lev[s];
    int s2 = t > 0 ? s + 1 : s - 1;
    if (s2 >= 0 && s2 < L) {
        float t2 = x - lev[s2];
        if (t2 * t2 < t * t) {
            s = s2;
            t = t2;
        }
    }
    float m_own = -(t * t) * inv_s2;
    for (int b = 0; b < kb; ++b) {
        float m_other = -3.0e38f;
        int j = lo[b * L + s];
        if (j >= 0) {
            float u = x - lev[j];
            float v = -(u * u) * inv_s2;
            if (v > m_other) m_other = v;
        }
        j = hi[b * L + s];
        if (j >= 0) {
            float u = x - lev[j];
            float v = -(u * u) * inv_s2;
            if (v > m_other) m_other = v;
        }
        if ((lab[s] >> (kb - 1 - b)) & 1)
            out_row[pos[b]] = m_other - m_own;
        else
            out_row[pos[b]] = m_own - m_other;
    }
}

/* Closed-form max-log LLRs of a uniform separable constellation
 * (_numba_llr.py:213-235).  lab_a (L,), lo_a/hi_a (kb, L), pos_a (kb,). */
void oracle_llr_separable_maxlog(const float* x_re, const float* x_im, int64_t n,
                                 const float* lev_re, const int32_t* lab_re,
                                 const int32_t* lo_re, const int32_t* hi_re,
                                 const int32_t* pos_re, const float* lev_im,
                                 const int32_t* lab_im, const int32_t* lo_im,
                                 const int32_t* hi_im, const int32_t* pos_im, int32_t L,
                                 int32_t kb, float inv_s2, float* out) {
    const int64_t tasks = (n + LLR_TASK - 1) / LLR_TASK;
    const int k = 2 * kb;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < tasks; ++t) {
        int64_t end = (t + 1) * LLR_TASK < n ? (t + 1) * LLR_TASK : n;
        for (int64_t i = t * LLR_TASK; i < end; ++i) {
            axis_maxlog(x_re[i], lev_re, lab_re, lo_re, hi_re, pos_re, L, kb, inv_s2,
                        out + i * k);
            axis_maxlog(x_im[i], lev_im, lab_im, lo_im, hi_im, pos_im, L, kb, inv_s2,
                        out + i * k);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Constant-composition distribution matching                               */
/* ------------------------------------------------------------------------ */

#define CCDM_TOP ((int64_t)1 << 32)
#define CCDM_HALF ((int64_t)1 << 31)
#define CCDM_LIMB 0xFFFFFFFFull

/* rng split in proportion to counts / remaining, floor rounding; the
 * remainder goes to the largest count (lowest index among equals). */
static void ccdm_split(const int64_t* counts, int M, int64_t rng, int64_t remaining,
                       int64_t* shares) {
    int big = 0;
    int64_t rest = rng;
    for (int i = 0; i < M; ++i) {
        shares[i] = rng * counts[i] / remaining;
        rest -= shares[i];
        if (counts[i] > counts[big]) big = i;
    }
    shares[big] += rest;
}

/* bits (R, k) int8 -> out (R, n) point indices. */
void oracle_ccdm_match(const int8_t* bits, int64_t R, int64_t k, const int64_t* composition,
                       int32_t M, int64_t n, int32_t* out) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t b = 0; b < R; ++b) {
        int64_t counts[256], shares[256];
        memcpy(counts, composition, sizeof(int64_t) * (size_t)M);
        const int8_t* u = bits + b * k;
        int64_t pos = 0, d = 0, rng = CCDM_TOP, remaining = n;
        for (int j = 0; j < 32; ++j, ++pos) d = (d << 1) | (pos < k ? u[pos] : 0);
        for (int64_t t = 0; t < n; ++t) {
            ccdm_split(counts, M, rng, remaining, shares);
            int64_t cum = 0, share = 0;
            int s = 0;
            for (int i = 0; i < M; ++i) {
                share = shares[i];
                if (d < cum + share) {
                    s = i;
                    break;
                }
                cum += share;
            }
            out[b * n + t] = s;
            d -= cum;
            rng = share;
            --counts[s];
            --remaining;
            while (rng < CCDM_HALF) {
                rng <<= 1;
                d = (d << 1) | (pos < k ? u[pos] : 0);
                ++pos;
            }
        }
    }
}

/* Bit 2^-j (j >= 1) of the limbs (32 bits each, most significant first). */
static int ccdm_bit(const uint64_t* low, int64_t j) {
    return (int)((low[(j - 1) >> 5] >> (31 - ((j - 1) & 31))) & 1u);
}

/* symbols (R, n) point indices -> bits (R, k), valid (R,). */
void oracle_ccdm_dematch(const int32_t* symbols, int64_t R, int64_t n,
                         const int64_t* composition, int32_t M, int64_t k,
                         int64_t buffer_bits, int8_t* bits, uint8_t* valid) {
    const int64_t limbs = buffer_bits / 32 + 2;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t b = 0; b < R; ++b) {
        int64_t counts[256], shares[256];
        memcpy(counts, composition, sizeof(int64_t) * (size_t)M);
        uint64_t* low = (uint64_t*)calloc((size_t)limbs, sizeof(uint64_t));
        int64_t rng = CCDM_TOP, shift = 32, remaining = n;
        int ok = 1;
        for (int64_t t = 0; t < n; ++t) {
            const int32_t s = symbols[b * n + t];
            if (s < 0 || s >= M || counts[s] == 0) {
                ok = 0;
                break;
            }
            ccdm_split(counts, M, rng, remaining, shares);
            int64_t cum = 0;
            for (int i = 0; i < s; ++i) cum += shares[i];
            /* low += cum 2^-shift, the carry running towards 2^-1. */
            int64_t i = (shift - 1) >> 5;
            const uint64_t v = (uint64_t)cum << (31 - ((shift - 1) & 31));
            low[i] += v & CCDM_LIMB;
            uint64_t carry = (low[i] >> 32) + (v >> 32);
            low[i] &= CCDM_LIMB;
            while (carry && i > 0) {
                --i;
                low[i] += carry;
                carry = low[i] >> 32;
                low[i] &= CCDM_LIMB;
            }
            rng = shares[s];
            --counts[s];
            --remaining;
            while (rng < CCDM_HALF) {
                rng <<= 1;
                ++shift;
            }
            if (shift > buffer_bits) {
                ok = 0;
                break;
            }
        }
        /* v = ceil(low 2^k): the first k bits, plus one if any bit follows. */
        int8_t* out = bits + b * k;
        int beyond = 0;
        const int64_t first = k >> 5;
        if (low[first] & (CCDM_LIMB >> (k & 31))) beyond = 1;
        for (int64_t l = first + 1; l < limbs && !beyond; ++l)
            if (low[l]) beyond = 1;
        for (int64_t j = 1; j <= k; ++j) out[j - 1] = (int8_t)ccdm_bit(low, j);
        int up = beyond;
        for (int64_t j = k; up && j > 0; --j) {
            up = out[j - 1];
            out[j - 1] ^= 1;
        }
        if (ok && up) ok = 0; /* v = 2^k: no payload */
        if (ok && beyond) {
            /* v 2^-k must lie in [low, low + rng 2^-shift): every bit of low
             * after the k-th above its last 32 is one, and 2^w - (those 32
             * bits) < rng. */
            const int64_t lo = (k + 1 > shift - 31) ? k + 1 : shift - 31;
            for (int64_t j = k + 1; j < lo && ok; ++j)
                if (!ccdm_bit(low, j)) ok = 0;
            if (ok) {
                int64_t f = 0;
                for (int64_t j = lo; j <= shift; ++j) f = (f << 1) | ccdm_bit(low, j);
                ok = ((int64_t)1 << (shift - lo + 1)) - f < rng;
            }
        }
        valid[b] = (uint8_t)ok;
        free(low);
    }
}
