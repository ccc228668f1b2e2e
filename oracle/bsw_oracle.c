/*
 * bsw_oracle.c -- CPU restatement of TAL's banded Smith-Waterman extension.
 *
 * TEST INFRASTRUCTURE ONLY (see bsw_oracle.h).  The product path never links
 * this file; it is built into liboracle.so for tests/, smoke() and the
 * cpu_baseline leg of bench.py.
 *
 * The vector restatement follows the AVX-512BW build of
 * src/alignment/BSW/bandedSWA.cpp, which is the reference's own SIMD path on
 * an AVX-512 Xeon:
 *   16-bit: getScores16 :2664-2688, smithWatermanBatchWrapper16 :2690-2960,
 *           smithWaterman512_16 :2962-3356, MAIN_CODE16/ZSCORE16 :1868-1906.
 *   8-bit : getScores8 :1970-1995, smithWatermanBatchWrapper8 :1997-2261,
 *           smithWaterman512_8 :2263-2661, MAIN_CODE8/ZSCORE8 :1826-1865.
 * A group of 32 (16-bit) or 64 (8-bit) consecutive pairs shares the row loop,
 * the column range [beg,end), the band narrowing and the all-lanes-zero exit,
 * so a pair's outputs can depend on the other pairs of its group.  Each lane
 * value is kept at the reference's width and every add/sub wraps the way the
 * epi16/epi8 instructions do; compares are signed or unsigned exactly where
 * the reference uses cmpgt/min_epu/max_epu.
 *
 * The scalar restatement follows scalarBandedSWA :116-237.
 */
#include "bsw_oracle.h"

#include <stdlib.h>
#include <string.h>

#define W16 32
#define W8 64

/* Observation points of the 16-bit group loop, for the row-pair model
 * (rowpair_model.c defines them before including this file); empty here. */
#ifndef OBS_GROUP
#define OBS_GROUP()
#define OBS_ROW(i)
#define OBS_STORE(i, j, l, hst, fst, h, e)
#define OBS_COLDONE(i, j)
#define OBS_STOP(i)
#define OBS_ROWEND(i)
#endif
#define MAXLEN16 32768 /* MAX_SEQ_LEN16, bandedSWA.h:83 */
#define MAXLEN8 128    /* MAX_SEQ_LEN8,  bandedSWA.h:82 */
#define CODE_N 4       /* AMBIG, bandedSWA.cpp:44 */
#define PAD_REF 99     /* DUMMY1, :45 */
#define PAD_QRY 100    /* DUMMY2, :46 */
#define ALL16 0xFFFFFFFFu
#define ALL8 0xFFFFFFFFFFFFFFFFull

struct bsw_oracle {
    int o_del, e_del, o_ins, e_ins, zdrop, end_bonus;
    int8_t w_match, w_mismatch, w_ambig; /* w_mismatch stored negated */
    /* Per-object scratch that lives across groups and calls, [col][lane]. */
    int16_t *F16, *Hh16, *Hv16;
    uint16_t *r16, *q16;
    int8_t *F8, *Hh8, *Hv8;
    uint8_t *r8, *q8;
    int stale_mode;
    int mutation; /* planted mistake for the pin tests; 0 = none */
    uint64_t n_groups, n_stale_decisive, n_stale_nonzero;
};

/* ---------------------------------------------------------------- helpers */
static inline int16_t w16(int32_t x) { return (int16_t)(uint16_t)(uint32_t)x; }
static inline int16_t smax16(int16_t a, int16_t b) { return a > b ? a : b; }
static inline int16_t smin16(int16_t a, int16_t b) { return a < b ? a : b; }
static inline int16_t umin16(int16_t a, int16_t b)
{
    return (uint16_t)a < (uint16_t)b ? a : b;
}
static inline int16_t abs16(int16_t a) { return a < 0 ? w16(-(int32_t)a) : a; }

static inline int8_t w8(int32_t x) { return (int8_t)(uint8_t)(uint32_t)x; }
static inline int8_t smax8(int8_t a, int8_t b) { return a > b ? a : b; }
static inline int8_t smin8(int8_t a, int8_t b) { return a < b ? a : b; }
static inline int8_t umin8(int8_t a, int8_t b)
{
    return (uint8_t)a < (uint8_t)b ? a : b;
}
static inline int8_t abs8(int8_t a) { return a < 0 ? w8(-(int32_t)a) : a; }

/* ------------------------------------------------------------ object life */
bsw_oracle *bsw_oracle_create(int o_del, int e_del, int o_ins, int e_ins,
                              int zdrop, int end_bonus, int8_t w_match,
                              int8_t w_mismatch)
{
    bsw_oracle *o = (bsw_oracle *)calloc(1, sizeof(*o));
    if (!o) return NULL;
    o->o_del = o_del;
    o->e_del = e_del;
    o->o_ins = o_ins;
    o->e_ins = e_ins;
    o->zdrop = zdrop;
    o->end_bonus = end_bonus;
    o->w_match = w_match;
    o->w_mismatch = (int8_t)(w_mismatch * -1); /* bandedSWA.cpp:65 */
    o->w_ambig = -1;                           /* DEFAULT_AMBIG, :68 */
    /* Fresh scratch reads as zero, like the page-fresh _mm_malloc blocks of
     * a newly constructed object (:79-86).  One spare column/row on top of
     * the reference's sizes keeps the init stores at index 1 in bounds when
     * a group's maximum length is 0. */
    size_t n16 = (size_t)(MAXLEN16 + 2) * W16;
    size_t n8 = (size_t)(MAXLEN8 + 2) * W8;
    o->F16 = (int16_t *)calloc(n16, sizeof(int16_t));
    o->Hh16 = (int16_t *)calloc(n16, sizeof(int16_t));
    o->Hv16 = (int16_t *)calloc(n16, sizeof(int16_t));
    o->r16 = (uint16_t *)calloc(n16, sizeof(uint16_t));
    o->q16 = (uint16_t *)calloc(n16, sizeof(uint16_t));
    o->F8 = (int8_t *)calloc(n8, 1);
    o->Hh8 = (int8_t *)calloc(n8, 1);
    o->Hv8 = (int8_t *)calloc(n8, 1);
    o->r8 = (uint8_t *)calloc(n8, 1);
    o->q8 = (uint8_t *)calloc(n8, 1);
    if (!o->F16 || !o->Hh16 || !o->Hv16 || !o->r16 || !o->q16 || !o->F8 ||
        !o->Hh8 || !o->Hv8 || !o->r8 || !o->q8) {
        bsw_oracle_destroy(o);
        return NULL;
    }
    return o;
}

void bsw_oracle_destroy(bsw_oracle *o)
{
    if (!o) return;
    free(o->F16);
    free(o->Hh16);
    free(o->Hv16);
    free(o->r16);
    free(o->q16);
    free(o->F8);
    free(o->Hh8);
    free(o->Hv8);
    free(o->r8);
    free(o->q8);
    free(o);
}

void bsw_oracle_set_stale_mode(bsw_oracle *o, int mode) { o->stale_mode = mode; }

void bsw_oracle_set_mutation(bsw_oracle *o, int m) { o->mutation = m; }

void bsw_oracle_stats(const bsw_oracle *o, uint64_t out[3])
{
    out[0] = o->n_groups;
    out[1] = o->n_stale_decisive;
    out[2] = o->n_stale_nonzero;
}

uint32_t bsw_oracle_f16_nonzero_mask(const bsw_oracle *o, int c)
{
    uint32_t m = 0;
    for (int l = 0; l < W16; l++)
        if (o->F16[(size_t)c * W16 + l] != 0) m |= 1u << l;
    return m;
}

/* Per-pair band, computed the way the 16-bit wrapper does it
 * (bandedSWA.cpp:2882-2904): uint16 wrap of len2*max + eb - o, integer
 * division, then +1.0 and truncation, floored at 1, capped by w. */
static int band16(const bsw_oracle *o, int len2, int32_t w)
{
    int smax = 0;
    if (smax < o->w_match) smax = o->w_match;
    if (smax < o->w_mismatch) smax = o->w_mismatch;
    if (smax < o->w_ambig) smax = o->w_ambig;
    uint16_t ql = (uint16_t)(len2 * smax);
    uint16_t t = (uint16_t)(ql + w16(o->end_bonus - o->o_ins));
    int b = (int)(t / o->e_ins + 1.0);
    if (b < 1) b = 1;
    int band = (uint16_t)(w < b ? w : b);
    t = (uint16_t)(ql + w16(o->end_bonus - o->o_del));
    b = (int)(t / o->e_del + 1.0);
    if (b < 1) b = 1;
    band = (uint16_t)(band < b ? band : b);
    return band;
}

uint64_t bsw_oracle_band_cells(const bsw_oracle *o, int len1, int len2,
                               int32_t w)
{
    int b = band16(o, len2, w);
    uint64_t c = 0;
    for (int i = 0; i < len1; i++) {
        int hi = i + b + 1 < len2 ? i + b + 1 : len2;
        int lo = i - b > 0 ? i - b : 0;
        if (hi > lo) c += (uint64_t)(hi - lo);
    }
    return c;
}

/* ================================================================ 16-bit */

/* Zero mask of scratch column c: bit l set when F|H_h == 0 in lane l
 * (the cmpeq(or(F,H), 0) tests of the narrowing scans, :3239-3258). */
static inline uint32_t col_zero16(const int16_t *F, const int16_t *Hh, int c)
{
    uint32_t m = 0;
    const int16_t *f = F + (size_t)c * W16, *h = Hh + (size_t)c * W16;
    for (int l = 0; l < W16; l++)
        if ((f[l] | h[l]) == 0) m |= 1u << l;
    return m;
}

/* smithWaterman512_16 restated, bandedSWA.cpp:2962-3356. */
static void dp_group16(bsw_oracle *o, uint16_t nrow, uint16_t ncol,
                       oracle_seqpair *p, int32_t w, const uint16_t *myband)
{
    int16_t *F = o->F16, *Hh = o->Hh16;
    const int16_t *Hv = o->Hv16;
    const uint16_t *rs = o->r16, *qs = o->q16;
    const int16_t match = o->w_match, mism = o->w_mismatch, ambig = o->w_ambig;
    const int16_t e_del = w16(o->e_del), oe_del = w16(o->o_del + o->e_del);
    const int16_t e_ins = w16(o->e_ins), oe_ins = w16(o->o_ins + o->e_ins);
    const int16_t zdrop = w16(o->zdrop);

    int16_t tlen[W16], qlen[W16], head[W16], tail[W16], mlen[W16], band[W16];
    int16_t maxScore[W16], x[W16], y[W16], gscore[W16], max_off[W16];
    int16_t max_ie[W16], exit0[W16];
    int16_t e11[W16], h10[W16], maxRS[W16], y1[W16], idx[W16];

    int32_t minq = 10000000; /* :3000-3005 */
    for (int l = 0; l < W16; l++) {
        tlen[l] = (int16_t)(uint16_t)p[l].len1;
        qlen[l] = (int16_t)(uint16_t)p[l].len2;
        if (p[l].len2 < minq) minq = p[l].len2;
    }
    minq -= 1;
    for (int l = 0; l < W16; l++) {
        band[l] = (int16_t)myband[l];
        tail[l] = qlen[l];
        head[l] = 0;
        mlen[l] = umin16(w16(qlen[l] + band[l]), tlen[l]); /* :3021-3022 */
        maxScore[l] = Hv[l];                                /* :3033 */
        x[l] = y[l] = 0;
        gscore[l] = -1;
        max_off[l] = 0;
        max_ie[l] = 0;
        exit0[l] = (int16_t)0xFFFF;
    }
    for (int j = 0; j < ncol; j++) /* :3034-3035 */
        memset(F + (size_t)j * W16, 0, W16 * sizeof(int16_t));
    if (o->stale_mode == 1)
        memset(F + (size_t)ncol * W16, 0, W16 * sizeof(int16_t));
    int fcol_written = 0; /* has this group stored F[ncol] yet? */

    int beg = 0, end = ncol, nbeg = 0, nend = ncol;
    OBS_GROUP();
    for (int16_t i = 0; i < (int)nrow; i++) {
        /* column range of this row, :3059-3063 */
        beg = nbeg;
        end = nend;
        if (beg < i - w) beg = i - w;
        if (end > i + w + 1) end = i + w + 1;
        if (end > ncol) end = ncol;
        const int16_t i1 = w16(i + 1);
        for (int l = 0; l < W16; l++) {
            e11[l] = 0;
            h10[l] = (beg == 0) ? Hv[(size_t)(i + (o->mutation == 4 ? 0 : 1)) * W16 + l] : 0; /* :3065 */
            maxRS[l] = 0;
            y1[l] = 0;
        }
        /* per-lane band head/tail, :3080-3088 */
        uint32_t same = 0;
        for (int l = 0; l < W16; l++) {
            int16_t ph = head[l], pt = tail[l];
            head[l] = smax16(head[l], w16(i - band[l]));
            tail[l] = umin16(tail[l], w16(i1 + band[l]));
            tail[l] = umin16(tail[l], qlen[l]);
            if (head[l] == ph && tail[l] == pt) same |= 1u << l;
        }
        /* trim columns that left a lane's band, :3092-3112 */
        for (int c = beg; c < end && same != ALL16; c++) {
            uint32_t below = 0;
            for (int l = 0; l < W16; l++)
                if (head[l] > w16(c)) below |= 1u << l;
            if (below == 0) break;
            for (int l = 0; l < W16; l++) {
                int clear = ((below >> l) & 1) || (w16(c + 1) > tail[l]);
                if (clear) {
                    F[(size_t)c * W16 + l] = 0;
                    Hh[(size_t)c * W16 + l] = 0;
                }
            }
        }
        /* lanes that stop contributing, :3118-3129 */
        for (int l = 0; l < W16; l++)
            if (i1 > mlen[l] || tail[l] == head[l] || head[l] > tail[l])
                exit0[l] = 0;
        OBS_ROW(i);

        /* the column loop, :3134-3196 */
        int16_t jj = w16(beg); /* j512: column index + 1 after the bump */
        int16_t j;
        for (j = (int16_t)beg; j < end; j++) {
            const int16_t pj = jj;
            jj = w16(jj + 1);
            const size_t cj = (size_t)j * W16;
            const uint16_t *rrow = rs + (size_t)i * W16;
            for (int l = 0; l < W16; l++) {
                const int16_t h00 = Hh[cj + l], f11 = F[cj + l];
                const uint16_t s1 = rrow[l], s2 = qs[cj + l];
                /* MAIN_CODE16, :1883-1906 */
                int16_t sbt = (s1 == s2) ? match : mism;
                uint16_t smx = s1 > s2 ? s1 : s2;
                if ((smx & 0x8000u) && o->mutation != 5) sbt = ambig;
                int16_t m = w16(h00 + sbt);
                if (h00 == 0) m = 0;
                int16_t h11 = smax16(m, e11[l]);
                h11 = smax16(h11, f11);
                const int swap = o->mutation == 3;
                int16_t v = smax16(w16(m - (swap ? oe_del : oe_ins)), 0);
                e11[l] = smax16(v, w16(e11[l] - (swap ? e_del : e_ins)));
                v = smax16(w16(m - (swap ? oe_ins : oe_del)), 0);
                int16_t f21 = smax16(v, w16(f11 - (swap ? e_ins : e_del)));
                /* masked stores, :3150-3155 */
                const int left = head[l] > pj;
                const int out1 = (pj > tail[l]) || left;
                const int16_t hst = out1 ? 0 : h10[l];
                if (out1) f21 = 0;
                /* mutation 8: the F of column tail (one past the lane's row-max
                 * band) is not stored, as the scalar path stores eh[end].e = 0
                 * there (:201) */
                if (o->mutation == 8 && pj == tail[l]) f21 = 0;
                /* row max and its column, :3158-3169 */
                const int16_t b = maxRS[l];
                const int16_t mr = smax16(b, h11);
                const int upd = o->mutation == 1 ? (h11 > b) : ((mr > b) || (mr == h11));
                const int out2 = (jj > tail[l]) || left;
                if (!out2) {
                    if (upd) y1[l] = jj;
                    maxRS[l] = mr;
                }
                F[cj + l] = f21;
                Hh[cj + l] = hst;
                OBS_STORE(i, j, l, hst, f21, h11, e11[l]);
                h10[l] = h11;
                /* global score at the last query column, :3176-3195 */
                if (j >= minq) {
                    const int at_end = (jj == qlen[l]);
                    const int live = exit0[l] < 0 || o->mutation == 6;
                    const int past = jj > tail[l];
                    if (at_end && live && !past) {
                        if (!(gscore[l] > h11)) max_ie[l] = i1;
                        gscore[l] = smax16(gscore[l], h11);
                    }
                }
            }
            OBS_COLDONE(i, j);
        }
        /* boundary store after the loop, at j (== end, or beg if the loop
         * did not run), :3197-3203 */
        for (int l = 0; l < W16; l++) {
            const int out1 = (jj > tail[l]) || (head[l] > jj);
            Hh[(size_t)j * W16 + l] = out1 ? 0 : h10[l];
            F[(size_t)j * W16 + l] = 0;
            OBS_STORE(i, j, l, Hh[(size_t)j * W16 + l], 0, 0, 0);
        }
        if (j == ncol) fcol_written = 1;

        /* all lanes scored zero in this row: stop, :3205-3209 */
        uint32_t zrow = 0;
        for (int l = 0; l < W16; l++)
            if (maxRS[l] == 0) zrow |= 1u << l;
        if (zrow == ALL16) {
            OBS_STOP(i);
            break;
        }
        for (int l = 0; l < W16; l++) {
            if ((zrow >> l) & 1) exit0[l] = 0;
            const int16_t before = maxScore[l];
            if (exit0[l] < 0) maxScore[l] = smax16(maxScore[l], maxRS[l]);
            if (maxScore[l] > before) { /* :3217-3226 */
                y[l] = y1[l];
                x[l] = i1;
                max_off[l] = smax16(max_off[l], abs16(w16(y1[l] - i1)));
            }
            /* ZSCORE16, :1868-1881 (no gap-extension scaling) */
            const int16_t di = w16(i1 - x[l]), dj = w16(y1[l] - y[l]);
            int16_t off = (di > dj) ? w16(di - dj) : w16(dj - di);
            if (o->mutation == 2) off = w16(-off);
            const int16_t drop = w16(w16(maxScore[l] - maxRS[l]) - off);
            if (drop > zdrop) exit0[l] = 0;
        }
        OBS_ROWEND(i);

        /* narrow the group's column range, :3239-3260 */
        int c;
        for (c = beg; c < end; c++) {
            if (col_zero16(F, Hh, c) == ALL16) nbeg = c;
            else break;
        }
        for (c = end; c >= beg; c--)
            if (col_zero16(F, Hh, c) != ALL16) break;
        nend = (c + 2 < ncol) ? c + 2 : ncol;

        /* AVX-512 16-bit only: the head/tail scans run over the new range
         * (:3267), so the tail scan can start at column ncol before this
         * group has stored it. */
        if (o->mutation != 7) {
            beg = nbeg;
            end = nend;
        }

        /* head scan, :3269-3291 */
        uint32_t still = ALL16;
        int16_t lc = w16(beg);
        for (c = beg; c < end; c++) {
            uint32_t z = 0;
            for (int l = 0; l < W16; l++) {
                int16_t v = F[(size_t)c * W16 + l] | Hh[(size_t)c * W16 + l];
                if (v == 0 && exit0[l] < 0) z |= 1u << l;
            }
            if (z == 0) break;
            z &= still;
            lc = w16(lc + 1);
            for (int l = 0; l < W16; l++)
                if ((z >> l) & 1) head[l] = lc;
            still = z;
        }

        /* Is this tail scan about to test a stale F[ncol]?  Only the zero
         * test of lanes that are live and have H_h[ncol] == 0 can depend on
         * it; each lane's scan result is independent of the other lanes, so
         * the stale bit decides lane l iff the two outcomes differ. */
        if (end == ncol && !fcol_written && end >= beg) {
            const int16_t *fc = F + (size_t)ncol * W16;
            const int16_t *hc = Hh + (size_t)ncol * W16;
            int decisive = 0, nonzero = 0;
            for (int l = 0; l < W16; l++) {
                if (!(exit0[l] < 0) || hc[l] != 0) continue;
                /* outcome if the lane is nonzero at ncol: index stays tail */
                int16_t t_nz = smin16(w16(tail[l] + 2), qlen[l]);
                /* outcome if zero at ncol: walk down while zero */
                int16_t ix = w16(ncol - 1);
                for (int cc = ncol - 1; cc >= beg; cc--) {
                    int16_t v = F[(size_t)cc * W16 + l] | Hh[(size_t)cc * W16 + l];
                    if (v != 0) break;
                    ix = w16(cc - 1);
                }
                int16_t t_z = smin16(w16(ix + 2), qlen[l]);
                if (t_nz != t_z) {
                    decisive = 1;
                    if (fc[l] != 0) nonzero = 1;
                }
            }
            o->n_stale_decisive += (uint64_t)decisive;
            o->n_stale_nonzero += (uint64_t)nonzero;
        }

        /* tail scan, :3293-3316 */
        for (int l = 0; l < W16; l++) idx[l] = tail[l];
        still = ALL16;
        lc = w16(end);
        for (c = end; c >= beg; c--) {
            uint32_t z = 0;
            for (int l = 0; l < W16; l++) {
                int16_t v = F[(size_t)c * W16 + l] | Hh[(size_t)c * W16 + l];
                if (v == 0 && exit0[l] < 0) z |= 1u << l;
            }
            if (z == 0) break;
            z &= still;
            lc = w16(lc - 1);
            for (int l = 0; l < W16; l++)
                if ((z >> l) & 1) idx[l] = lc;
            still = z;
        }
        for (int l = 0; l < W16; l++)
            tail[l] = smin16(w16(idx[l] + 2), qlen[l]);
    }

    for (int l = 0; l < W16; l++) { /* :3327-3353 */
        p[l].score = maxScore[l];
        p[l].tle = x[l];
        p[l].qle = y[l];
        p[l].max_off = max_off[l];
        p[l].gscore = gscore[l];
        p[l].gtle = max_ie[l];
    }
    o->n_groups++;
}

/* smithWatermanBatchWrapper16 restated for one group, :2787-2942. */
static void setup_group16(bsw_oracle *o, oracle_seqpair *p, const uint8_t *ref,
                          const uint8_t *qer, int32_t w)
{
    uint16_t *rs = o->r16, *qs = o->q16;
    int16_t *Hv = o->Hv16, *Hh = o->Hh16;
    int smax = 0;
    if (smax < o->w_match) smax = o->w_match;
    if (smax < o->w_mismatch) smax = o->w_mismatch;
    if (smax < o->w_ambig) smax = o->w_ambig;

    int16_t h0[W16];
    uint16_t qlen[W16], myband[W16];
    uint16_t maxLen1 = 0, maxLen2 = 0;
    for (int l = 0; l < W16; l++) {
        h0[l] = (int16_t)(uint16_t)p[l].h0;
        const uint8_t *s = ref + (int64_t)p[l].idr;
        for (int k = 0; k < p[l].len1; k++) {
            rs[(size_t)k * W16 + l] = (s[k] == CODE_N) ? 0xFFFF : s[k];
            Hv[(size_t)k * W16 + l] = 0;
        }
        qlen[l] = (uint16_t)(p[l].len2 * smax);
        if (maxLen1 < p[l].len1) maxLen1 = (uint16_t)p[l].len1;
    }
    for (int l = 0; l < W16; l++)
        for (int k = p[l].len1; k <= maxLen1; k++) {
            rs[(size_t)k * W16 + l] = PAD_REF;
            Hv[(size_t)k * W16 + l] = PAD_REF;
        }
    /* first column H(i,-1): h0 - o_del - k*e_del, clipped at 0 (:2828-2837) */
    {
        int16_t t[W16];
        for (int l = 0; l < W16; l++) {
            Hv[l] = h0[l];
            t[l] = w16(h0[l] - o->o_del);
        }
        for (int k = 1; k < maxLen1; k++)
            for (int l = 0; l < W16; l++) {
                t[l] = w16(t[l] - o->e_del);
                Hv[(size_t)k * W16 + l] = smax16(t[l], 0);
            }
    }
    for (int l = 0; l < W16; l++) {
        const uint8_t *s = qer + (int64_t)p[l].idq;
        for (int k = 0; k < p[l].len2; k++) {
            qs[(size_t)k * W16 + l] = (s[k] == CODE_N) ? 0xFFFF : s[k];
            Hh[(size_t)k * W16 + l] = 0;
        }
        if (maxLen2 < p[l].len2) maxLen2 = (uint16_t)p[l].len2;
    }
    for (int l = 0; l < W16; l++)
        for (int k = p[l].len2; k <= maxLen2; k++) {
            qs[(size_t)k * W16 + l] = PAD_QRY;
            Hh[(size_t)k * W16 + l] = 0;
        }
    /* row -1 H(-1,j): h0, then h0 - oe_ins, then -e_ins steps, clipped at 0
     * (:2867-2879).  Stored at index 1 even when maxLen2 < 2. */
    {
        const int16_t oe = w16(o->o_ins + o->e_ins);
        int16_t t[W16];
        for (int l = 0; l < W16; l++) {
            Hh[l] = h0[l];
            t[l] = (h0[l] > oe) ? w16(h0[l] - oe) : 0;
            Hh[W16 + l] = t[l];
        }
        for (int k = 2; k < maxLen2; k++)
            for (int l = 0; l < W16; l++) {
                t[l] = smax16(w16(t[l] - o->e_ins), 0);
                Hh[(size_t)k * W16 + l] = t[l];
            }
    }
    /* per-lane band (:2882-2904) */
    int32_t bsize = w;
    for (int l = 0; l < W16; l++) {
        uint16_t t = (uint16_t)(qlen[l] + w16(o->end_bonus - o->o_ins));
        int b = (int)(t / o->e_ins + 1.0);
        b = b > 1 ? b : 1;
        myband[l] = (uint16_t)(bsize < b ? bsize : b);
    }
    for (int l = 0; l < W16; l++) {
        uint16_t t = (uint16_t)(qlen[l] + w16(o->end_bonus - o->o_del));
        int b = (int)(t / o->e_del + 1.0);
        b = b > 1 ? b : 1;
        myband[l] = (uint16_t)(myband[l] < b ? myband[l] : b);
        bsize = bsize < myband[l] ? myband[l] : bsize;
    }
    dp_group16(o, maxLen1, maxLen2, p, bsize, myband);
}

void bsw_oracle_get_scores16(bsw_oracle *o, oracle_seqpair *pairs,
                             const uint8_t *ref, const uint8_t *qer,
                             int32_t n, int32_t w)
{
    oracle_seqpair g[W16];
    for (int32_t base = 0; base < n; base += W16) {
        int cnt = (n - base < W16) ? n - base : W16;
        memcpy(g, pairs + base, (size_t)cnt * sizeof(oracle_seqpair));
        /* padding lanes, :2710-2717 (h0/idr/idq defined as 0) */
        for (int l = cnt; l < W16; l++) {
            memset(&g[l], 0, sizeof(g[l]));
            g[l].id = base + l;
        }
        setup_group16(o, g, ref, qer, w);
        for (int l = 0; l < cnt; l++) {
            oracle_seqpair *d = pairs + base + l;
            d->score = g[l].score;
            d->tle = g[l].tle;
            d->gtle = g[l].gtle;
            d->qle = g[l].qle;
            d->gscore = g[l].gscore;
            d->max_off = g[l].max_off;
        }
    }
}

/* ================================================================= 8-bit */

static inline uint64_t col_zero8(const int8_t *F, const int8_t *Hh, int c)
{
    uint64_t m = 0;
    const int8_t *f = F + (size_t)c * W8, *h = Hh + (size_t)c * W8;
    for (int l = 0; l < W8; l++)
        if ((f[l] | h[l]) == 0) m |= 1ull << l;
    return m;
}

/* smithWaterman512_8 restated, bandedSWA.cpp:2263-2661. */
static void dp_group8(bsw_oracle *o, uint8_t nrow, uint8_t ncol,
                      oracle_seqpair *p, int32_t w, const uint8_t *myband)
{
    int8_t *F = o->F8, *Hh = o->Hh8;
    const int8_t *Hv = o->Hv8;
    const uint8_t *rs = o->r8, *qs = o->q8;
    const int8_t match = o->w_match, mism = o->w_mismatch, ambig = o->w_ambig;
    const int8_t e_del = w8(o->e_del), oe_del = w8(o->o_del + o->e_del);
    const int8_t e_ins = w8(o->e_ins), oe_ins = w8(o->o_ins + o->e_ins);
    const int8_t zdrop = w8(o->zdrop);

    int8_t tlen[W8], qlen[W8], head[W8], tail[W8], mlen[W8], band[W8];
    int8_t maxScore[W8], x[W8], y[W8], gscore[W8], max_off[W8];
    int8_t max_ie[W8], exit0[W8];
    int8_t e11[W8], h10[W8], maxRS[W8], y1[W8], idx[W8];

    int32_t minq = 10000000;
    for (int l = 0; l < W8; l++) {
        tlen[l] = (int8_t)(uint8_t)p[l].len1;
        qlen[l] = (int8_t)(uint8_t)p[l].len2;
        if (p[l].len2 < minq) minq = p[l].len2;
    }
    minq -= 1;
    for (int l = 0; l < W8; l++) {
        band[l] = (int8_t)myband[l];
        tail[l] = qlen[l];
        head[l] = 0;
        mlen[l] = umin8(w8(qlen[l] + band[l]), tlen[l]);
        maxScore[l] = Hv[l];
        x[l] = y[l] = 0;
        gscore[l] = -1;
        max_off[l] = 0;
        max_ie[l] = 0;
        exit0[l] = (int8_t)0xFF;
    }
    for (int j = 0; j < ncol; j++) memset(F + (size_t)j * W8, 0, W8);

    int beg = 0, end = ncol, nbeg = 0, nend = ncol;
    for (int16_t i = 0; i < (int)nrow; i++) {
        beg = nbeg;
        end = nend;
        if (beg < i - w) beg = i - w;
        if (end > i + w + 1) end = i + w + 1;
        if (end > ncol) end = ncol;
        const int8_t i1 = w8(i + 1);
        for (int l = 0; l < W8; l++) {
            e11[l] = 0;
            h10[l] = (beg == 0) ? Hv[(size_t)(i + 1) * W8 + l] : 0;
            maxRS[l] = 0;
            y1[l] = 0;
        }
        uint64_t same = 0;
        for (int l = 0; l < W8; l++) {
            int8_t ph = head[l], pt = tail[l];
            head[l] = smax8(head[l], w8(i - band[l]));
            tail[l] = umin8(tail[l], w8(i1 + band[l]));
            tail[l] = umin8(tail[l], qlen[l]);
            if (head[l] == ph && tail[l] == pt) same |= 1ull << l;
        }
        for (int c = beg; c < end && same != ALL8; c++) {
            uint64_t below = 0;
            for (int l = 0; l < W8; l++)
                if (head[l] > w8(c)) below |= 1ull << l;
            if (below == 0) break;
            for (int l = 0; l < W8; l++) {
                int clear = ((below >> l) & 1) || (w8(c + 1) > tail[l]);
                if (clear) {
                    F[(size_t)c * W8 + l] = 0;
                    Hh[(size_t)c * W8 + l] = 0;
                }
            }
        }
        for (int l = 0; l < W8; l++)
            if (i1 > mlen[l] || tail[l] == head[l] || head[l] > tail[l])
                exit0[l] = 0;

        int8_t jj = w8(beg);
        int16_t j;
        for (j = (int16_t)beg; j < end; j++) {
            const int8_t pj = jj;
            jj = w8(jj + 1);
            const size_t cj = (size_t)j * W8;
            const uint8_t *rrow = rs + (size_t)i * W8;
            for (int l = 0; l < W8; l++) {
                const int8_t h00 = Hh[cj + l], f11 = F[cj + l];
                const uint8_t s1 = rrow[l], s2 = qs[cj + l];
                /* MAIN_CODE8, :1842-1865 */
                int8_t sbt = (s1 == s2) ? match : mism;
                uint8_t smx = s1 > s2 ? s1 : s2;
                if (smx & 0x80u) sbt = ambig;
                int8_t m = w8(h00 + sbt);
                if (h00 == 0) m = 0;
                int8_t h11 = smax8(m, e11[l]);
                h11 = smax8(h11, f11);
                int8_t v = smax8(w8(m - oe_ins), 0);
                e11[l] = smax8(v, w8(e11[l] - e_ins));
                v = smax8(w8(m - oe_del), 0);
                int8_t f21 = smax8(v, w8(f11 - e_del));
                const int left = head[l] > pj;
                const int out1 = (pj > tail[l]) || left;
                const int8_t hst = out1 ? 0 : h10[l];
                if (out1) f21 = 0;
                const int8_t b = maxRS[l];
                const int8_t mr = smax8(b, h11);
                const int upd = (mr > b) || (mr == h11);
                const int out2 = (jj > tail[l]) || left;
                if (!out2) {
                    if (upd) y1[l] = jj;
                    maxRS[l] = mr;
                }
                F[cj + l] = f21;
                Hh[cj + l] = hst;
                h10[l] = h11;
                if (j >= minq) {
                    const int at_end = (jj == qlen[l]);
                    const int live = exit0[l] < 0;
                    const int past = jj > tail[l];
                    if (at_end && live && !past) {
                        if (!(gscore[l] > h11)) max_ie[l] = i1;
                        gscore[l] = smax8(gscore[l], h11);
                    }
                }
            }
        }
        for (int l = 0; l < W8; l++) {
            const int out1 = (jj > tail[l]) || (head[l] > jj);
            Hh[(size_t)j * W8 + l] = out1 ? 0 : h10[l];
            F[(size_t)j * W8 + l] = 0;
        }

        uint64_t zrow = 0;
        for (int l = 0; l < W8; l++)
            if (maxRS[l] == 0) zrow |= 1ull << l;
        if (zrow == ALL8) break;
        for (int l = 0; l < W8; l++) {
            if ((zrow >> l) & 1) exit0[l] = 0;
            const int8_t before = maxScore[l];
            if (exit0[l] < 0) maxScore[l] = smax8(maxScore[l], maxRS[l]);
            if (maxScore[l] > before) {
                y[l] = y1[l];
                x[l] = i1;
                max_off[l] = smax8(max_off[l], abs8(w8(y1[l] - i1)));
            }
            const int8_t di = w8(i1 - x[l]), dj = w8(y1[l] - y[l]);
            const int8_t off = (di > dj) ? w8(di - dj) : w8(dj - di);
            const int8_t drop = w8(w8(maxScore[l] - maxRS[l]) - off);
            if (drop > zdrop) exit0[l] = 0;
        }

        int c;
        for (c = beg; c < end; c++) {
            if (col_zero8(F, Hh, c) == ALL8) nbeg = c;
            else break;
        }
        for (c = end; c >= beg; c--)
            if (col_zero8(F, Hh, c) != ALL8) break;
        nend = (c + 2 < ncol) ? c + 2 : ncol;

        /* 8-bit: head/tail scans keep this row's beg/end (:2570-2621) */
        uint64_t still = ALL8;
        int8_t lc = w8(beg);
        for (c = beg; c < end; c++) {
            uint64_t z = 0;
            for (int l = 0; l < W8; l++) {
                int8_t v = F[(size_t)c * W8 + l] | Hh[(size_t)c * W8 + l];
                if (v == 0 && exit0[l] < 0) z |= 1ull << l;
            }
            if (z == 0) break;
            z &= still;
            lc = w8(lc + 1);
            for (int l = 0; l < W8; l++)
                if ((z >> l) & 1) head[l] = lc;
            still = z;
        }
        for (int l = 0; l < W8; l++) idx[l] = tail[l];
        still = ALL8;
        lc = w8(end);
        for (c = end; c >= beg; c--) {
            uint64_t z = 0;
            for (int l = 0; l < W8; l++) {
                int8_t v = F[(size_t)c * W8 + l] | Hh[(size_t)c * W8 + l];
                if (v == 0 && exit0[l] < 0) z |= 1ull << l;
            }
            if (z == 0) break;
            z &= still;
            lc = w8(lc - 1);
            for (int l = 0; l < W8; l++)
                if ((z >> l) & 1) idx[l] = lc;
            still = z;
        }
        for (int l = 0; l < W8; l++)
            tail[l] = smin8(w8(idx[l] + 2), qlen[l]);
    }

    for (int l = 0; l < W8; l++) {
        p[l].score = maxScore[l];
        p[l].tle = x[l];
        p[l].qle = y[l];
        p[l].max_off = max_off[l];
        p[l].gscore = gscore[l];
        p[l].gtle = max_ie[l];
    }
    o->n_groups++;
}

/* smithWatermanBatchWrapper8 restated for one group, :2086-2244. */
static void setup_group8(bsw_oracle *o, oracle_seqpair *p, const uint8_t *ref,
                         const uint8_t *qer, int32_t w)
{
    uint8_t *rs = o->r8, *qs = o->q8;
    int8_t *Hv = o->Hv8, *Hh = o->Hh8;
    int8_t smax = 0;
    if (smax < o->w_match) smax = o->w_match;
    if (smax < o->w_mismatch) smax = o->w_mismatch;
    if (smax < o->w_ambig) smax = o->w_ambig;

    int8_t h0[W8];
    uint8_t qlen[W8], myband[W8];
    uint16_t maxLen1 = 0;
    uint8_t maxLen2 = 0;
    for (int l = 0; l < W8; l++) {
        h0[l] = (int8_t)(uint8_t)p[l].h0;
        const uint8_t *s = ref + (int64_t)p[l].idr;
        for (int k = 0; k < p[l].len1; k++) {
            rs[(size_t)k * W8 + l] = (s[k] == CODE_N) ? 0xFF : s[k];
            Hv[(size_t)k * W8 + l] = 0;
        }
        qlen[l] = (uint8_t)(p[l].len2 * smax);
        if (maxLen1 < p[l].len1) maxLen1 = (uint16_t)p[l].len1;
    }
    for (int l = 0; l < W8; l++)
        for (int k = p[l].len1; k <= maxLen1; k++) {
            rs[(size_t)k * W8 + l] = PAD_REF;
            Hv[(size_t)k * W8 + l] = PAD_REF;
        }
    {
        int8_t t[W8];
        for (int l = 0; l < W8; l++) {
            Hv[l] = h0[l];
            t[l] = w8(h0[l] - o->o_del);
        }
        for (int k = 1; k < maxLen1; k++)
            for (int l = 0; l < W8; l++) {
                t[l] = w8(t[l] - o->e_del);
                Hv[(size_t)k * W8 + l] = smax8(t[l], 0);
            }
    }
    for (int l = 0; l < W8; l++) {
        const uint8_t *s = qer + (int64_t)p[l].idq;
        for (int k = 0; k < p[l].len2; k++) {
            qs[(size_t)k * W8 + l] = (s[k] == CODE_N) ? 0xFF : s[k];
            Hh[(size_t)k * W8 + l] = 0;
        }
        if (maxLen2 < p[l].len2) maxLen2 = (uint8_t)p[l].len2;
    }
    for (int l = 0; l < W8; l++)
        for (int k = p[l].len2; k <= maxLen2; k++) {
            qs[(size_t)k * W8 + l] = PAD_QRY;
            Hh[(size_t)k * W8 + l] = 0;
        }
    {
        const int8_t oe = w8(o->o_ins + o->e_ins);
        int8_t t[W8];
        for (int l = 0; l < W8; l++) {
            Hh[l] = h0[l];
            t[l] = (h0[l] > oe) ? w8(h0[l] - oe) : 0;
            Hh[W8 + l] = t[l];
        }
        for (int k = 2; k < maxLen2; k++)
            for (int l = 0; l < W8; l++) {
                t[l] = smax8(w8(t[l] - o->e_ins), 0);
                Hh[(size_t)k * W8 + l] = t[l];
            }
    }
    int32_t bsize = w;
    for (int l = 0; l < W8; l++) {
        uint8_t t = (uint8_t)(qlen[l] + w8(o->end_bonus - o->o_ins));
        int b = (int)(t / o->e_ins + 1.0);
        b = b > 1 ? b : 1;
        myband[l] = (uint8_t)(bsize < b ? bsize : b);
    }
    for (int l = 0; l < W8; l++) {
        uint8_t t = (uint8_t)(qlen[l] + w8(o->end_bonus - o->o_del));
        int b = (int)(t / o->e_del + 1.0);
        b = b > 1 ? b : 1;
        myband[l] = (uint8_t)(myband[l] < b ? myband[l] : b);
        bsize = bsize < myband[l] ? myband[l] : bsize;
    }
    dp_group8(o, (uint8_t)maxLen1, maxLen2, p, bsize, myband);
}

void bsw_oracle_get_scores8(bsw_oracle *o, oracle_seqpair *pairs,
                            const uint8_t *ref, const uint8_t *qer,
                            int32_t n, int32_t w)
{
    oracle_seqpair g[W8];
    if (n <= 0) return;
    const int32_t last_len2 = pairs[n - 1].len2;
    for (int32_t base = 0; base < n; base += W8) {
        int cnt = (n - base < W8) ? n - base : W8;
        memcpy(g, pairs + base, (size_t)cnt * sizeof(oracle_seqpair));
        /* padding lanes, :2013-2018: len1 = 0, len2 = last pair's len2 */
        for (int l = cnt; l < W8; l++) {
            memset(&g[l], 0, sizeof(g[l]));
            g[l].id = base + l;
            g[l].len2 = last_len2;
        }
        setup_group8(o, g, ref, qer, w);
        for (int l = 0; l < cnt; l++) {
            oracle_seqpair *d = pairs + base + l;
            d->score = g[l].score;
            d->tle = g[l].tle;
            d->gtle = g[l].gtle;
            d->qle = g[l].qle;
            d->gscore = g[l].gscore;
            d->max_off = g[l].max_off;
        }
    }
}

/* ================================================================ scalar */

/* Scoring matrix laid out as bwa_fill_scmat does (bench-bsw.cpp:56-64). */
static void fill_mat(const bsw_oracle *o, int8_t mat[25])
{
    int k = 0;
    for (int i = 0; i < 4; i++) {
        for (int j = 0; j < 4; j++) mat[k++] = (i == j) ? o->w_match : o->w_mismatch;
        mat[k++] = o->w_ambig;
    }
    for (int j = 0; j < 5; j++) mat[k++] = o->w_ambig;
}

/* scalarBandedSWA restated, bandedSWA.cpp:116-237, with the scoring matrix
 * mat (the one the reference object was constructed with). */
static int scalar_one(const bsw_oracle *o, const int8_t *mat, int qlen, const uint8_t *query,
                      int tlen, const uint8_t *target, int32_t w, int h0,
                      int *_qle, int *_tle, int *_gtle, int *_gscore,
                      int *_max_off)
{
    const int oe_del = o->o_del + o->e_del, oe_ins = o->o_ins + o->e_ins;
    int *eh_h = (int *)calloc((size_t)qlen + 1, sizeof(int));
    int *eh_e = (int *)calloc((size_t)qlen + 1, sizeof(int));
    int8_t *qp = (int8_t *)malloc((size_t)(qlen > 0 ? qlen : 1) * 5);
    for (int k = 0, i = 0; k < 5; k++)
        for (int j = 0; j < qlen; j++) qp[i++] = mat[k * 5 + query[j]];
    /* row -1 (:143-145) */
    eh_h[0] = h0;
    if (qlen >= 1) eh_h[1] = h0 > oe_ins ? h0 - oe_ins : 0;
    for (int j = 2; j <= qlen && eh_h[j - 1] > o->e_ins; j++)
        eh_h[j] = eh_h[j - 1] - o->e_ins;
    /* band clipping (:147-156) */
    int mx = 0;
    for (int i = 0; i < 25; i++) mx = mx > mat[i] ? mx : mat[i];
    int max_ins = (int)((double)(qlen * mx + o->end_bonus - o->o_ins) / o->e_ins + 1.);
    max_ins = max_ins > 1 ? max_ins : 1;
    w = w < max_ins ? w : max_ins;
    int max_del = (int)((double)(qlen * mx + o->end_bonus - o->o_del) / o->e_del + 1.);
    max_del = max_del > 1 ? max_del : 1;
    w = w < max_del ? w : max_del;

    int max = h0, max_i = -1, max_j = -1, max_ie = -1, gscore = -1, max_off = 0;
    int beg = 0, end = qlen;
    for (int i = 0; i < tlen; i++) {
        int f = 0, h1, m = 0, mj = -1, j;
        const int8_t *q = &qp[target[i] * qlen];
        if (beg < i - w) beg = i - w;
        if (end > i + w + 1) end = i + w + 1;
        if (end > qlen) end = qlen;
        if (beg == 0) {
            h1 = h0 - (o->o_del + o->e_del * (i + 1));
            if (h1 < 0) h1 = 0;
        } else
            h1 = 0;
        for (j = beg; j < end; j++) {
            int M = eh_h[j], e = eh_e[j], h, t;
            eh_h[j] = h1;
            M = M ? M + q[j] : 0;
            h = M > e ? M : e;
            h = h > f ? h : f;
            h1 = h;
            mj = m > h ? mj : j;
            m = m > h ? m : h;
            t = M - oe_del;
            t = t > 0 ? t : 0;
            e -= o->e_del;
            e = e > t ? e : t;
            eh_e[j] = e;
            t = M - oe_ins;
            t = t > 0 ? t : 0;
            f -= o->e_ins;
            f = f > t ? f : t;
        }
        eh_h[end] = h1;
        eh_e[end] = 0;
        if (j == qlen) {
            max_ie = gscore > h1 ? max_ie : i;
            gscore = gscore > h1 ? gscore : h1;
        }
        if (m == 0) break;
        if (m > max) {
            max = m, max_i = i, max_j = mj;
            max_off = max_off > abs(mj - i) ? max_off : abs(mj - i);
        } else if (o->zdrop > 0) {
            if (i - max_i > mj - max_j) {
                if (max - m - ((i - max_i) - (mj - max_j)) * o->e_del > o->zdrop) break;
            } else {
                if (max - m - ((mj - max_j) - (i - max_i)) * o->e_ins > o->zdrop) break;
            }
        }
        for (j = beg; j < end && eh_h[j] == 0 && eh_e[j] == 0; j++)
            ;
        beg = j;
        for (j = end; j >= beg && eh_h[j] == 0 && eh_e[j] == 0; j--)
            ;
        end = j + 2 < qlen ? j + 2 : qlen;
    }
    free(eh_h);
    free(eh_e);
    free(qp);
    *_qle = max_j + 1;
    *_tle = max_i + 1;
    *_gtle = max_ie + 1;
    *_gscore = gscore;
    *_max_off = max_off;
    return max;
}

void bsw_oracle_scalar_mat(bsw_oracle *o, const int8_t *mat, oracle_seqpair *pairs,
                           const uint8_t *ref, const uint8_t *qer, int32_t n, int32_t w)
{
    for (int32_t i = 0; i < n; i++) {
        oracle_seqpair *p = pairs + i;
        p->score = scalar_one(o, mat, p->len2, qer + p->idq, p->len1, ref + p->idr, w,
                              p->h0, &p->qle, &p->tle, &p->gtle, &p->gscore,
                              &p->max_off);
    }
}

void bsw_oracle_scalar(bsw_oracle *o, oracle_seqpair *pairs, const uint8_t *ref,
                       const uint8_t *qer, int32_t n, int32_t w)
{
    int8_t mat[25];
    fill_mat(o, mat);
    for (int32_t i = 0; i < n; i++) {
        oracle_seqpair *p = pairs + i;
        p->score = scalar_one(o, mat, p->len2, qer + p->idq, p->len1, ref + p->idr, w,
                              p->h0, &p->qle, &p->tle, &p->gtle, &p->gscore,
                              &p->max_off);
    }
}

/* Full-band DP: the scalar recurrence and exits with every in-band cell
 * evaluated (no narrowing to the nonzero region). */
int bsw_oracle_band_dp(const bsw_oracle *o, int qlen, const uint8_t *query,
                       int tlen, const uint8_t *target, int32_t w, int h0,
                       int *_qle, int *_tle, int *_gtle, int *_gscore,
                       int *_max_off)
{
    int8_t mat[25];
    fill_mat(o, mat);
    const int oe_del = o->o_del + o->e_del, oe_ins = o->o_ins + o->e_ins;
    int mx = 0;
    for (int i = 0; i < 25; i++) mx = mx > mat[i] ? mx : mat[i];
    int max_ins = (int)((double)(qlen * mx + o->end_bonus - o->o_ins) / o->e_ins + 1.);
    max_ins = max_ins > 1 ? max_ins : 1;
    w = w < max_ins ? w : max_ins;
    int max_del = (int)((double)(qlen * mx + o->end_bonus - o->o_del) / o->e_del + 1.);
    max_del = max_del > 1 ? max_del : 1;
    w = w < max_del ? w : max_del;

    /* Hprev[j+1] = H(i-1, j), Hprev[0] = H(i-1, -1); Eprev[j] = E(i, j). */
    int *Hprev = (int *)calloc((size_t)qlen + 2, sizeof(int));
    int *Hcur = (int *)calloc((size_t)qlen + 2, sizeof(int));
    int *E = (int *)calloc((size_t)qlen + 1, sizeof(int));
    Hprev[0] = h0;
    for (int j = 0; j < qlen; j++) {
        int v = (j == 0) ? (h0 > oe_ins ? h0 - oe_ins : 0) : Hprev[j] - o->e_ins;
        Hprev[j + 1] = v > 0 ? v : 0;
    }
    int max = h0, max_i = -1, max_j = -1, max_ie = -1, gscore = -1, max_off = 0;
    for (int i = 0; i < tlen; i++) {
        int lo = i - w > 0 ? i - w : 0;
        int hi = i + w + 1 < qlen ? i + w + 1 : qlen;
        int hb = h0 - (o->o_del + o->e_del * (i + 1));
        Hcur[0] = hb > 0 ? hb : 0;
        int f = 0, m = 0, mj = -1;
        for (int j = 0; j < qlen; j++) {
            if (j < lo || j >= hi) {
                Hcur[j + 1] = 0;
                E[j] = 0;
                f = 0;
                continue;
            }
            int M = Hprev[j];
            M = M ? M + mat[target[i] * 5 + query[j]] : 0;
            int e = E[j];
            int h = M > e ? M : e;
            h = h > f ? h : f;
            Hcur[j + 1] = h;
            if (h >= m) { mj = j; m = h; }
            int t = M - oe_del;
            t = t > 0 ? t : 0;
            e -= o->e_del;
            E[j] = e > t ? e : t;
            t = M - oe_ins;
            t = t > 0 ? t : 0;
            f -= o->e_ins;
            f = f > t ? f : t;
        }
        if (hi == qlen && qlen > 0 && lo < hi) {
            int hl = Hcur[qlen];
            max_ie = gscore > hl ? max_ie : i;
            gscore = gscore > hl ? gscore : hl;
        }
        int *tmp = Hprev;
        Hprev = Hcur;
        Hcur = tmp;
        if (m == 0) break;
        if (m > max) {
            max = m, max_i = i, max_j = mj;
            max_off = max_off > abs(mj - i) ? max_off : abs(mj - i);
        } else if (o->zdrop > 0) {
            if (i - max_i > mj - max_j) {
                if (max - m - ((i - max_i) - (mj - max_j)) * o->e_del > o->zdrop) break;
            } else {
                if (max - m - ((mj - max_j) - (i - max_i)) * o->e_ins > o->zdrop) break;
            }
        }
    }
    free(Hprev);
    free(Hcur);
    free(E);
    *_qle = max_j + 1;
    *_tle = max_i + 1;
    *_gtle = max_ie + 1;
    *_gscore = gscore;
    *_max_off = max_off;
    return max;
}
