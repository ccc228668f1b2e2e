/*
 * chain_oracle.c -- CPU restatement of TAL's DP chaining (the AVX-512BW build
 * of dp_chain::mm_dp_vectorized, src/chaining/parallel_chaining_32_bit.h:
 * 216-412).  TEST INFRASTRUCTURE ONLY (see chain_oracle.h).
 *
 * For each anchor i (in order; anchors sorted by r):
 *   st advances past predecessors more than dr away on r (64-bit unsigned
 *   difference, :232) and to i - max_iter at most (:235).
 *   Candidates j in [st, i-1] pass when |ddr - ddq| <= bw, ddr != 0,
 *   ddq > 0, ddq <= dq and ddq <= dr (ddr = r_i - r_j, ddq = q_i - q_j in
 *   int32 of the 32-bit SoA copies, :262-271); their score is
 *   f[j] + min(ddr, ddq, l_i) - gap(|ddr - ddq|) in int32 wrap (:283-293).
 *   f[i] = the best score if above l_i, else l_i; p[i] = the highest j with
 *   that score, or -1 (:304-362 and :395-402); v[i] = v[p[i]] if larger,
 *   else f[i] (:408).
 * The gap cost depends on which path the reference takes:
 *   i-1-st <= 5: its scalar loop, (int)(dd * .01 * avg_qspan) in double,
 *     truncated (:386), plus ilog2(dd) >> 1;
 *   otherwise: its 16-lane vector loop, dd converted to float times
 *     (float)(0.01 * avg_qspan), rounded toward -inf (:181-199), plus
 *     (31 - lzcnt(dd)) >> 1 for dd != 0.
 * Both conversions are x86's: a value out of int32 range converts to
 * INT32_MIN (cvttsd2si in the scalar loop, vcvtps2dq in the vector one).
 * avg_qspan = (float)(sum of l) / n in float (:222).  Candidates the vector
 * loop masks out score 0 there and so never beat l_i >= 0; lengths above
 * INT32_MAX (where l_i would read as negative) are outside this restatement.
 * Evaluation follows the source as written with IEEE double and float
 * arithmetic (no reassociation).
 */
#include "chain_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>

static int g_mutation = 0;

void chain_oracle_set_mutation(int m) { g_mutation = m; }

/* x86's float/double to int32 conversions: INT32_MIN out of range */
static int32_t x86_cvt_f32(float x)
{
    if (g_mutation == 4 && x >= 2147483648.0f) return INT32_MAX; /* planted: saturate */
    return x >= -2147483648.0f && x < 2147483648.0f ? (int32_t)x : INT32_MIN;
}

static int32_t x86_cvt_f64(double x)
{
    if (g_mutation == 4 && x >= 2147483648.0) return INT32_MAX;
    return x > -2147483649.0 && x < 2147483648.0 ? (int32_t)x : INT32_MIN;
}

static int32_t ilog2_u32(uint32_t v)
{
    int32_t r = -1;
    while (v) {
        v >>= 1;
        r++;
    }
    return r;
}

void chain_oracle_dp(const chain_oracle_anchor *a, int64_t n, int32_t bw, int32_t max_iter,
                     int32_t dr, int32_t dq, uint32_t *f, int32_t *p, int32_t *v)
{
    if (n <= 0) return;
    uint64_t sum = 0;
    for (int64_t i = 0; i < n; i++) sum += (uint32_t)a[i].l;
    const float avg_qspan = (float)sum / (float)n;
    const float avg_val = (float)(0.01 * (double)avg_qspan);
    int64_t st = 0;
    for (int64_t i = 0; i < n; i++) {
        const int32_t li = (int32_t)(uint32_t)a[i].l;
        while (st < i && !(a[i].r - a[st].r <= (uint64_t)(uint32_t)dr)) ++st;
        if (i - st > max_iter) st = i - max_iter;
        const int vector_path = !(i - 1 - st <= 5);
        const int32_t ri = (int32_t)(uint32_t)a[i].r, qi = (int32_t)(uint32_t)a[i].q;
        int32_t max_f = li, max_j = -1;
        for (int64_t j = i - 1; j >= st; j--) {
            const int32_t rj = (int32_t)(uint32_t)a[j].r, qj = (int32_t)(uint32_t)a[j].q;
            const int32_t ddr = (int32_t)((uint32_t)ri - (uint32_t)rj);
            const int32_t ddq = (int32_t)((uint32_t)qi - (uint32_t)qj);
            const int32_t diff = (int32_t)((uint32_t)ddr - (uint32_t)ddq);
            const int32_t dd = diff < 0 ? (int32_t)(0u - (uint32_t)diff) : diff;
            if (dd > bw && g_mutation != 1) continue;
            if (ddr == 0 || ddq <= 0) continue;
            if (ddq > dq || ddq > dr) continue;
            int32_t oc = ddr < ddq ? ddr : ddq;
            if (oc > li) oc = li;
            int32_t gap;
            const int32_t log_dd = dd ? ilog2_u32((uint32_t)dd) : 0;
            if (vector_path && g_mutation != 2) {
                const float prod = (float)dd * avg_val;
                gap = (int32_t)((uint32_t)x86_cvt_f32(floorf(prod)) + (uint32_t)(log_dd >> 1));
            } else {
                gap = (int32_t)((uint32_t)x86_cvt_f64((double)dd * .01 * (double)avg_qspan) +
                                (uint32_t)(log_dd >> 1));
            }
            const int32_t sc = (int32_t)((uint32_t)f[j] + (uint32_t)oc - (uint32_t)gap);
            if (sc > max_f || (g_mutation == 3 && sc == max_f)) {
                max_f = sc;
                max_j = (int32_t)j;
            }
        }
        f[i] = (uint32_t)max_f;
        p[i] = max_j;
        v[i] = max_j >= 0 && v[max_j] > max_f ? v[max_j] : max_f;
    }
}

void chain_oracle_dp_sets(const chain_oracle_anchor *a, const int64_t *off, int32_t nsets,
                          int32_t bw, int32_t max_iter, int32_t dr, int32_t dq, uint32_t *f,
                          int32_t *p, int32_t *v, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (int32_t k = 0; k < nsets; k++)
        chain_oracle_dp(a + off[k], off[k + 1] - off[k], bw, max_iter, dr, dq, f + off[k],
                        p + off[k], v + off[k]);
}
