/*
 * bsw_synth.c -- deterministic synthetic extension-pair workloads (see
 * bsw_synth.h).  Used by the benchmark and the tests to build inputs shaped
 * like the SeqPair batches bench-bsw.cpp feeds getScores16
 * (benchmarks/bench-bsw.cpp:117-169), without any input files.
 */
#include "bsw_synth.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t s[4];
} rng_t;

static inline uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t next64(rng_t *r) /* xoshiro256** */
{
    uint64_t *s = r->s;
    const uint64_t out = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return out;
}

static void rng_seed(rng_t *r, uint64_t seed, int64_t k)
{
    uint64_t x = seed + (uint64_t)k * 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline int32_t uni(rng_t *r, int32_t lo, int32_t hi) /* inclusive */
{
    return lo + (int32_t)(next64(r) % (uint64_t)(hi - lo + 1));
}

static inline int ppm(rng_t *r, uint32_t p) { return (next64(r) % 1000000u) < p; }

static inline double unit(rng_t *r) /* (0,1) */
{
    return ((double)(next64(r) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

typedef struct {
    int32_t len1, len2, h0, w;
    int cls; /* workload-5 class: 0 plain, 1 N, 2 repeat, 3 divergent */
} shape_t;

/* The shape draws come first in each pair's stream so that shape() and
 * fill() agree. */
static int draw_shape(int cfg, rng_t *r, shape_t *s)
{
    s->w = 100;
    s->cls = 0;
    switch (cfg) {
    case 1:
        s->len1 = uni(r, 151, 250);
        s->len2 = 151;
        s->h0 = uni(r, 10, 60);
        return 0;
    case 2:
        s->len2 = uni(r, 30, 250);
        s->len1 = s->len2 + uni(r, 0, 100);
        s->h0 = uni(r, 0, 120);
        return 0;
    case 3: {
        double u1 = unit(r), u2 = unit(r);
        double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        long L = lround(400.0 * exp(0.6 * z));
        if (L < 100) L = 100;
        if (L > 2000) L = 2000;
        s->len2 = (int32_t)L;
        s->len1 = s->len2 + uni(r, 0, 100);
        s->h0 = uni(r, 0, 120);
        static const int32_t ws[4] = {50, 100, 200, 500};
        s->w = ws[next64(r) & 3];
        return 0;
    }
    case 5: {
        s->len1 = uni(r, 151, 250);
        s->len2 = 151;
        int c = uni(r, 0, 99);
        s->cls = (c < 5) ? 1 : (c < 15) ? 2 : (c < 25) ? 3 : 0;
        double u = unit(r);
        s->h0 = (int32_t)lround(10.0 * exp(u * log(3000.0)));
        return 0;
    }
    default:
        return -1;
    }
}

int tal_synth_shape(int cfg, uint64_t seed, int64_t first, int32_t n,
                    int32_t *len1, int32_t *len2, int32_t *h0, int32_t *w)
{
    for (int32_t k = 0; k < n; k++) {
        rng_t r;
        shape_t s;
        rng_seed(&r, seed, first + k);
        if (draw_shape(cfg, &r, &s)) return -1;
        if (len1) len1[k] = s.len1;
        if (len2) len2[k] = s.len2;
        if (h0) h0[k] = s.h0;
        if (w) w[k] = s.w;
    }
    return 0;
}

/* Query of length lq derived from ref[0..lr): substitutions at sub_ppm, 1-3 bp
 * insertions/deletions at indel_ppm per emitted base; random bases once the
 * ref runs out. */
static void derive_query(rng_t *r, const uint8_t *ref, int32_t lr, uint8_t *q,
                         int32_t lq, uint32_t sub_ppm, uint32_t indel_ppm)
{
    int32_t pos = 0, k = 0;
    while (k < lq) {
        if (pos >= lr) {
            q[k++] = (uint8_t)(next64(r) & 3);
            continue;
        }
        uint32_t x = (uint32_t)(next64(r) % 1000000u);
        if (x < indel_ppm) {
            int32_t L = uni(r, 1, 3);
            if (next64(r) & 1) {
                for (int32_t t = 0; t < L && k < lq; t++) q[k++] = (uint8_t)(next64(r) & 3);
            } else {
                pos += L;
            }
        } else if (x < indel_ppm + sub_ppm) {
            q[k++] = (uint8_t)((ref[pos++] + 1 + (next64(r) % 3)) & 3);
        } else {
            q[k++] = ref[pos++];
        }
    }
}

typedef struct { /* SeqPair, bandedSWA.h:90-99 */
    int32_t idr, idq, id, len1, len2, h0, seqid, regid;
    int32_t score, tle, gtle, qle, gscore, max_off;
} synth_pair;

static void fill_one(int cfg, uint64_t seed, int64_t k, int64_t ro, int64_t qo,
                     synth_pair *p, uint8_t *ref, uint8_t *qer)
{
    rng_t r;
    shape_t s;
    rng_seed(&r, seed, k);
    draw_shape(cfg, &r, &s);
    uint8_t *R = ref + ro, *Q = qer + qo;
    if (cfg == 5 && s.cls == 2) {
        int32_t period = uni(&r, 1, 6);
        uint8_t unitb[6];
        for (int t = 0; t < period; t++) unitb[t] = (uint8_t)(next64(&r) & 3);
        for (int32_t t = 0; t < s.len1; t++) R[t] = unitb[t % period];
    } else {
        for (int32_t t = 0; t < s.len1; t++) R[t] = (uint8_t)(next64(&r) & 3);
    }
    uint32_t sub = 10000, indel = 2000;
    if (cfg == 2) sub = 20000, indel = 5000;
    if (cfg == 3) sub = 25000, indel = 5000;
    if (cfg == 5 && s.cls == 3) sub = 200000, indel = 2000;
    derive_query(&r, R, s.len1, Q, s.len2, sub, indel);
    if (cfg == 5 && s.cls == 1) {
        uint32_t rate = (uint32_t)uni(&r, 10000, 50000);
        for (int32_t t = 0; t < s.len1; t++)
            if (ppm(&r, rate)) R[t] = 4;
        for (int32_t t = 0; t < s.len2; t++)
            if (ppm(&r, rate)) Q[t] = 4;
    }
    p->idr = (int32_t)ro;
    p->idq = (int32_t)qo;
    p->id = (int32_t)k;
    p->len1 = s.len1;
    p->len2 = s.len2;
    p->h0 = s.h0;
    p->seqid = p->regid = -1;
    p->score = p->tle = p->gtle = p->qle = p->gscore = p->max_off = -1;
}

int tal_synth_fill(int cfg, uint64_t seed, int64_t first, int32_t n,
                   const int64_t *ref_off, const int64_t *qer_off, void *pairs,
                   uint8_t *ref, uint8_t *qer, int nthreads)
{
    if (cfg != 1 && cfg != 2 && cfg != 3 && cfg != 5) return -1;
    synth_pair *P = (synth_pair *)pairs;
    (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int32_t k = 0; k < n; k++)
        fill_one(cfg, seed, first + k, ref_off[k], qer_off[k], P + k, ref, qer);
    return 0;
}
