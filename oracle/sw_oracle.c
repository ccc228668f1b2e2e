/*
 * sw_oracle.c -- CPU restatement of TAL's full inter-task Smith-Waterman.
 *
 * TEST INFRASTRUCTURE ONLY (see sw_oracle.h).
 *
 * The vector restatement follows the AVX-512BW build of
 * src/alignment/Smith-Waterman/PairWiseSW_{16,8,32}.cpp:
 *   getScores                          _16 :136-147, _8 :134-145, _32 :136-152
 *   smithWatermanBatchWrapper          _16 :239-377 (padding pairs :247-252,
 *                                      group bounds and DUMMY fill :296-345)
 *   smithWatermanInterTaskVectorized   _16 :149-237 (init :170-174, four-row
 *                                      blocks :181-222, MAIN_CODE :62-76)
 * A group is SIMD_WIDTH consecutive pairs (64 / 32 / 16 for 8 / 16 / 32
 * bits).  Every lane runs the DP over the group's padded matrix: rows up to
 * the group's longest len1 and columns up to its longest len2, each rounded
 * up to a multiple of 4; past its own lengths a lane's sequences hold
 * DUMMY1 = 'B' (seq1) and DUMMY2 = 'D' (seq2).  The group bounds have the
 * lane width's type: int8 in the 8-bit class, so a group whose rounded bound
 * reaches 128 wraps to -128 and its loops do not run (scores 0).  Each lane
 * value is kept at the lane width and every add wraps the way the epi8 /
 * epi16 / epi32 adds do; max is signed.
 *
 * The four-row blocks of the reference visit the same cells with the same
 * operands as a row-by-row sweep: row i+1 at column j takes the F that row i
 * produced at j and the H row i produced at j-1.  The restatement therefore
 * sweeps row by row.
 *
 * The scalar restatement follows smithWatermanOnePair, PairWiseSW.cpp:63-114.
 */
#include "sw_oracle.h"

#include <stdlib.h>
#include <string.h>

#define DUMMY1 'B' /* PairWiseSW_16.cpp:50 */
#define DUMMY2 'D' /* :51 */
#define LOW_INIT (-63) /* LOW_INIT_VALUE, PairWiseSW.h:60 */

static int g_mutation = 0;

void sw_oracle_set_mutation(int m) { g_mutation = m; }

/* value of x in a lane of the given width (two's-complement wrap) */
static inline int32_t lane(int bits, int64_t x)
{
    if (bits == 8) return (int8_t)(uint8_t)(x & 0xFF);
    if (bits == 16) return (int16_t)(uint16_t)(x & 0xFFFF);
    return (int32_t)(uint32_t)(x & 0xFFFFFFFFu);
}

static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

/* group bound as the wrapper computes it: max over lanes, then rounded up to
 * a multiple of 4, both in the width's type (int8_t maxLen1 in _8 :293,
 * int16_t in _16 :296, int32_t in _32 :301) */
static int32_t group_bound(int bits, int32_t maxlen)
{
    const int32_t r = ((maxlen + 3) / 4) * 4;
    if (bits == 8) return (int8_t)(uint8_t)(r & 0xFF);
    if (bits == 16) return (int16_t)(uint16_t)(r & 0xFFFF);
    return r;
}

/* One lane: the DP of smithWatermanInterTaskVectorized over nrow x ncol. */
static int32_t lane_dp(int bits, int32_t match, int32_t mismatch, int32_t open, int32_t ext,
                       const uint8_t *s1, const uint8_t *s2, int32_t nrow, int32_t ncol,
                       int32_t *F, int32_t *H)
{
    for (int32_t j = 0; j <= ncol; j++) { /* :170-174 */
        F[j] = open;
        H[j] = 0;
    }
    if (g_mutation == 2)
        for (int32_t j = 0; j <= ncol; j++) F[j] = 0;
    int32_t maxs = 0;
    for (int32_t i = 0; i < nrow; i++) {
        int32_t e = g_mutation == 3 ? 0 : open; /* e11..e41 = gapOpenVec, :183-184 */
        int32_t hdiag = 0;                       /* h00..h30 = zeroVec, :191 */
        for (int32_t j = 0; j < ncol; j++) {
            const int32_t f = F[j];
            int32_t sbt = s1[i] == s2[j] ? match : mismatch;
            if (g_mutation == 4) sbt = s1[i] == s2[j] ? mismatch : match;
            const int32_t m = lane(bits, (int64_t)hdiag + sbt);
            int32_t h = imax(imax(m, e), f);
            if (g_mutation != 1) h = imax(h, 0);
            maxs = imax(maxs, h);
            const int32_t t = lane(bits, (int64_t)h + open);
            e = imax(t, lane(bits, (int64_t)e + ext));
            F[j] = imax(lane(bits, (int64_t)f + ext), g_mutation == 5 ? lane(bits, (int64_t)h + ext) : t);
            hdiag = H[j];
            H[j] = h;
        }
    }
    return maxs;
}

static void run_group(int bits, int32_t match, int32_t mismatch, int32_t open, int32_t ext,
                      sw_oracle_pair *pairs, const uint8_t *seq_buf, int32_t n, int32_t g0,
                      int W, int32_t S, uint8_t *s1, uint8_t *s2, int32_t *F, int32_t *H)
{
    int32_t max1 = 0, max2 = 0;
    for (int k = 0; k < W; k++) {
        const int32_t p = g0 + k;
        /* pairs past numPairs are padding with len1 = len2 = 0 (:247-252) */
        const int32_t l1 = p < n ? pairs[p].len1 : 0, l2 = p < n ? pairs[p].len2 : 0;
        if (l1 > max1) max1 = l1;
        if (l2 > max2) max2 = l2;
    }
    const int32_t nrow = group_bound(bits, max1), ncol = group_bound(bits, max2);
    for (int k = 0; k < W && g0 + k < n; k++) {
        sw_oracle_pair *sp = &pairs[g0 + k];
        int32_t score = 0;
        if (nrow > 0 && ncol > 0) {
            const uint8_t *a = seq_buf + 2 * (int64_t)sp->id * S;
            const uint8_t *b = seq_buf + (2 * (int64_t)sp->id + 1) * S;
            for (int32_t r = 0; r < nrow; r++) s1[r] = r < sp->len1 ? a[r] : DUMMY1;
            for (int32_t c = 0; c < ncol; c++)
                s2[c] = c < sp->len2 ? b[c] : (g_mutation == 6 ? DUMMY1 : DUMMY2);
            score = lane_dp(bits, match, mismatch, open, ext, s1, s2, nrow, ncol, F, H);
        }
        sp->score = score; /* int8/int16 maxScore widened to int32, :233 */
    }
}

int sw_oracle_get_scores(int bits, int32_t w_match, int32_t w_mismatch, int32_t w_open,
                         int32_t w_extend, sw_oracle_pair *pairs, const uint8_t *seq_buf,
                         int32_t n, int nthreads)
{
    if (bits != 8 && bits != 16 && bits != 32) return -1;
    const int W = bits == 8 ? 64 : bits == 16 ? 32 : 16;
    const int32_t S = bits == 8 ? 128 : 1024; /* MAX_SEQ_LEN_8/16/32 */
    /* _MM_SET1_EPI8/16/32 of the int32 weights keep the low bits */
    const int32_t match = lane(bits, w_match), mismatch = lane(bits, w_mismatch);
    const int32_t open = lane(bits, w_open), ext = lane(bits, w_extend);
    const int32_t ngroups = (n + W - 1) / W;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        uint8_t *s1 = malloc(1 << 16), *s2 = malloc(1 << 16);
        int32_t *F = malloc(sizeof(int32_t) * ((1 << 16) + 1));
        int32_t *H = malloc(sizeof(int32_t) * ((1 << 16) + 1));
#pragma omp for schedule(dynamic, 4)
        for (int32_t g = 0; g < ngroups; g++)
            run_group(bits, match, mismatch, open, ext, pairs, seq_buf, n, g * W, W, S, s1, s2,
                      F, H);
        free(s1);
        free(s2);
        free(F);
        free(H);
    }
    return 0;
}

void sw_oracle_scalar(int32_t w_match, int32_t w_mismatch, int32_t w_open, int32_t w_extend,
                      sw_oracle_pair *pairs, const uint8_t *seq_buf, int32_t n, int32_t slot)
{
    int32_t *F = malloc(sizeof(int32_t) * 70000), *H = malloc(sizeof(int32_t) * 70000);
    for (int32_t k = 0; k < n; k++) {
        sw_oracle_pair *p = &pairs[k];
        const uint8_t *seq1 = seq_buf + 2 * (int64_t)p->id * slot;
        const uint8_t *seq2 = seq_buf + (2 * (int64_t)p->id + 1) * slot;
        const int32_t nrow = p->len1, ncol = p->len2;
        for (int32_t j = 0; j <= ncol; j++) {
            F[j] = LOW_INIT;
            H[j] = 0;
        }
        int32_t maxs = 0;
        for (int32_t i = 1; i <= nrow; i++) {
            int32_t e = LOW_INIT, hdiag = 0;
            for (int32_t j = 1; j <= ncol; j++) {
                e = imax(e + w_extend, H[j - 1] + w_open);
                F[j] = imax(F[j] + w_extend, H[j] + w_open);
                const int32_t m = hdiag + (seq1[i - 1] == seq2[j - 1] ? w_match : w_mismatch);
                hdiag = H[j];
                int32_t h = 0;
                if (e > h) h = e;
                if (F[j] > h) h = F[j];
                if (m > h) h = m;
                H[j] = h;
                if (maxs < h) maxs = h;
            }
        }
        p->score = maxs;
    }
    free(F);
    free(H);
}
