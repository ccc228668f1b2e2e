/*
 * bsw_simd.cpp -- vectorised CPU restatement of TAL's banded extension group
 * loop, used as the CPU baseline.  TEST INFRASTRUCTURE ONLY: nothing in the
 * product links it; bench.py's cpu_baseline leg times it and tests/ check it
 * bit-exact against the lane-by-lane restatement in bsw_oracle.c.
 *
 * It follows the same reference code as bsw_oracle.c -- getScores16 /
 * smithWatermanBatchWrapper16 / smithWaterman512_16
 * (src/alignment/BSW/bandedSWA.cpp:2664-3356) and the 8-bit path
 * (:1970-2661) -- but evaluates each group's 32 (16-bit) or 64 (8-bit) lanes
 * as one 512-bit vector per column, the way the reference's AVX-512BW build
 * does, so that its speed on a host CPU stands for that build's.  The vector
 * type is GCC's generic vector extension; the compiler emits AVX-512BW for
 * it when the translation unit is built for that target (a second copy is
 * built for AVX2 hosts, where each vector is two 256-bit halves).  Adds and
 * subtracts wrap (done on the unsigned type), compares and max/min are
 * signed or unsigned exactly where the reference's are.  The per-object
 * scratch ([column][lane] vectors) persists across groups and calls, as the
 * reference object's does.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#if defined(__AVX512BW__)
#include <immintrin.h>
#endif

#include "bsw_oracle.h"

#ifndef SIMD_NS
#define SIMD_NS simd_generic
#endif

namespace SIMD_NS {

constexpr int MAXLEN16 = 32768; /* MAX_SEQ_LEN16, bandedSWA.h:83 */
constexpr int MAXLEN8 = 128;    /* MAX_SEQ_LEN8, bandedSWA.h:82 */
constexpr int CODE_N = 4, PAD_REF = 99, PAD_QRY = 100;

template <int BITS> struct Lanes;
template <> struct Lanes<16> {
    typedef int16_t S;
    typedef uint16_t U;
    static constexpr int L = 32;
    static constexpr int MAXLEN = MAXLEN16;
    typedef S VS __attribute__((vector_size(64)));
    typedef U VU __attribute__((vector_size(64)));
    /* bit l set when lane l of the compare result m is true */
    static inline uint64_t bits(VS m)
    {
#if defined(__AVX512BW__)
        return (uint32_t)_mm512_movepi16_mask((__m512i)m);
#else
        uint64_t r = 0;
        for (int l = 0; l < L; l++) r |= (uint64_t)(m[l] < 0) << l;
        return r;
#endif
    }
    static constexpr uint64_t ALL = 0xFFFFFFFFull;
};
template <> struct Lanes<8> {
    typedef int8_t S;
    typedef uint8_t U;
    static constexpr int L = 64;
    static constexpr int MAXLEN = MAXLEN8;
    typedef S VS __attribute__((vector_size(64)));
    typedef U VU __attribute__((vector_size(64)));
    static inline uint64_t bits(VS m)
    {
#if defined(__AVX512BW__)
        return (uint64_t)_mm512_movepi8_mask((__m512i)m);
#else
        uint64_t r = 0;
        for (int l = 0; l < L; l++) r |= (uint64_t)(m[l] < 0) << l;
        return r;
#endif
    }
    static constexpr uint64_t ALL = ~0ull;
};

template <int BITS> struct Obj {
    typedef Lanes<BITS> Ln;
    typedef typename Ln::VS VS;
    typedef typename Ln::VU VU;
    VS *F, *Hh, *Hv;
    VU *rs, *qs;
};

struct Object {
    int o_del, e_del, o_ins, e_ins, zdrop, end_bonus;
    int8_t w_match, w_mismatch, w_ambig; /* w_mismatch stored negated (:65) */
    Obj<16> o16;
    Obj<8> o8;
};

template <int BITS> bool alloc(Obj<BITS> &o)
{
    typedef typename Obj<BITS>::VS VS;
    typedef typename Obj<BITS>::VU VU;
    const size_t n = (size_t)Lanes<BITS>::MAXLEN + 2;
    o.F = (VS *)aligned_alloc(64, n * sizeof(VS));
    o.Hh = (VS *)aligned_alloc(64, n * sizeof(VS));
    o.Hv = (VS *)aligned_alloc(64, n * sizeof(VS));
    o.rs = (VU *)aligned_alloc(64, n * sizeof(VU));
    o.qs = (VU *)aligned_alloc(64, n * sizeof(VU));
    if (!o.F || !o.Hh || !o.Hv || !o.rs || !o.qs) return false;
    /* fresh scratch reads as zero, like a newly constructed object's */
    memset(o.F, 0, n * sizeof(VS));
    memset(o.Hh, 0, n * sizeof(VS));
    memset(o.Hv, 0, n * sizeof(VS));
    memset(o.rs, 0, n * sizeof(VU));
    memset(o.qs, 0, n * sizeof(VU));
    return true;
}

template <int BITS> void release(Obj<BITS> &o)
{
    free(o.F);
    free(o.Hh);
    free(o.Hv);
    free(o.rs);
    free(o.qs);
}

/* The group loop (smithWaterman512_16, :2962-3356; smithWaterman512_8,
 * :2263-2661) on one group whose SoA inputs the setup has written. */
template <int BITS>
void dp_group(const Object &ob, Obj<BITS> &o, int nrow, int ncol, oracle_seqpair *p, int32_t w,
              const typename Lanes<BITS>::S *myband)
{
    typedef Lanes<BITS> Ln;
    typedef typename Ln::S S;
    typedef typename Ln::U U;
    typedef typename Ln::VS VS;
    typedef typename Ln::VU VU;
    constexpr int L = Ln::L;
    auto wr = [](int32_t x) -> S { return (S)(U)(uint32_t)x; };
    auto bc = [](S x) -> VS { return VS{} + x; };
    auto add = [](VS a, VS b) -> VS { return (VS)((VU)a + (VU)b); };
    auto sub = [](VS a, VS b) -> VS { return (VS)((VU)a - (VU)b); };
    auto vmax = [](VS a, VS b) -> VS { return a > b ? a : b; };
    auto vmin = [](VS a, VS b) -> VS { return a < b ? a : b; };
    auto umin = [](VS a, VS b) -> VS { return (VS)((VU)a < (VU)b ? (VU)a : (VU)b); };
    VS *F = o.F, *Hh = o.Hh;
    const VS *Hv = o.Hv;
    const VU *rs = o.rs, *qs = o.qs;
    const VS zero = bc(0), match = bc(ob.w_match), mism = bc(ob.w_mismatch),
             ambig = bc(ob.w_ambig);
    const VS e_del = bc(wr(ob.e_del)), oe_del = bc(wr(ob.o_del + ob.e_del));
    const VS e_ins = bc(wr(ob.e_ins)), oe_ins = bc(wr(ob.o_ins + ob.e_ins));
    const VS zdrop = bc(wr(ob.zdrop));

    VS tlen, qlen, band;
    int32_t minq = 10000000; /* :3000-3005 */
    for (int l = 0; l < L; l++) {
        tlen[l] = (S)(U)p[l].len1;
        qlen[l] = (S)(U)p[l].len2;
        band[l] = myband[l];
        if (p[l].len2 < minq) minq = p[l].len2;
    }
    minq -= 1;
    VS tail = qlen, head = zero;
    const VS mlen = umin(add(qlen, band), tlen); /* :3021-3022 */
    VS maxScore = Hv[0], x = zero, y = zero, gscore = bc(-1), max_off = zero, max_ie = zero;
    VS exit0 = bc(-1);
    for (int j = 0; j < ncol; j++) F[j] = zero; /* :3034-3035 */

    int beg = 0, end = ncol, nbeg = 0, nend = ncol;
    for (int i = 0; i < nrow; i++) {
        beg = nbeg; /* :3059-3063 */
        end = nend;
        if (beg < i - w) beg = i - w;
        if (end > i + w + 1) end = i + w + 1;
        if (end > ncol) end = ncol;
        const VS i1 = bc(wr(i + 1));
        VS e11 = zero, maxRS = zero, y1 = zero;
        VS h10 = beg == 0 ? Hv[i + 1] : zero; /* :3065-3067 */
        /* per-lane band, :3080-3088 */
        const VS ph = head, pt = tail;
        head = vmax(head, sub(bc(wr(i)), band));
        tail = umin(tail, add(i1, band));
        tail = umin(tail, qlen);
        const bool same = Ln::bits((head == ph) & (tail == pt)) == Ln::ALL;
        /* trim, :3092-3112 */
        if (!same) {
            for (int c = beg; c < end; c++) {
                const VS below = head > bc(wr(c));
                if (Ln::bits(below) == 0) break;
                const VS clear = below | (bc(wr(c + 1)) > tail);
                F[c] = clear ? zero : F[c];
                Hh[c] = clear ? zero : Hh[c];
            }
        }
        /* :3118-3129 */
        exit0 = ((i1 > mlen) | (tail == head) | (head > tail)) ? zero : exit0;

        /* the column loop, :3134-3196 */
        S jj = wr(beg);
        const VU s1 = rs[i];
        int j;
        for (j = beg; j < end; j++) {
            const VS pj = bc(jj);
            jj = wr(jj + 1);
            const VS vj = bc(jj);
            const VS h00 = Hh[j], f11 = F[j];
            const VU s2 = qs[j];
            VS sbt = (s1 == s2) ? match : mism;
            const VU smx = s1 > s2 ? s1 : s2;
            sbt = ((VS)smx < zero) ? ambig : sbt; /* top bit: code 4 or a pad/odd byte */
            VS m = add(h00, sbt);
            m = (h00 == zero) ? zero : m;
            const VS h11 = vmax(vmax(m, e11), f11);
            e11 = vmax(vmax(sub(m, oe_ins), zero), sub(e11, e_ins));
            VS f21 = vmax(vmax(sub(m, oe_del), zero), sub(f11, e_del));
            const VS left = head > pj;
            const VS out1 = (pj > tail) | left;
            const VS hst = out1 ? zero : h10;
            f21 = out1 ? zero : f21;
            const VS mr = vmax(maxRS, h11);
            const VS upd = (mr > maxRS) | (mr == h11);
            const VS out2 = (vj > tail) | left;
            y1 = (~out2 & upd) ? vj : y1;
            maxRS = out2 ? maxRS : mr;
            F[j] = f21;
            Hh[j] = hst;
            h10 = h11;
            if (j >= minq) { /* :3176-3195 */
                const VS c = (vj == qlen) & (exit0 < zero) & ~(vj > tail);
                max_ie = (c & ~(gscore > h11)) ? i1 : max_ie;
                gscore = c ? vmax(gscore, h11) : gscore;
            }
        }
        { /* :3197-3203 */
            const VS vj = bc(jj);
            const VS out1 = (vj > tail) | (head > vj);
            Hh[j] = out1 ? zero : h10;
            F[j] = zero;
        }
        /* :3205-3211 */
        const VS zrow = maxRS == zero;
        if (Ln::bits(zrow) == Ln::ALL) break;
        exit0 = zrow ? zero : exit0;
        const VS before = maxScore;
        maxScore = (exit0 < zero) ? vmax(maxScore, maxRS) : maxScore;
        const VS inc = maxScore > before; /* :3217-3226 */
        y = inc ? y1 : y;
        x = inc ? i1 : x;
        {
            const VS d = sub(y1, i1);
            const VS ad = d < zero ? sub(zero, d) : d;
            max_off = inc ? vmax(max_off, ad) : max_off;
        }
        { /* z-drop without gap-extension scaling, :1868-1881 */
            const VS di = sub(i1, x), dj = sub(y1, y);
            const VS off = di > dj ? sub(di, dj) : sub(dj, di);
            const VS drop = sub(sub(maxScore, maxRS), off);
            exit0 = (drop > zdrop) ? zero : exit0;
        }
        /* narrowing, :3239-3260 */
        int c;
        for (c = beg; c < end; c++) {
            if (Ln::bits((F[c] | Hh[c]) == zero) == Ln::ALL)
                nbeg = c;
            else
                break;
        }
        for (c = end; c >= beg; c--)
            if (Ln::bits((F[c] | Hh[c]) == zero) != Ln::ALL) break;
        nend = (c + 2 < ncol) ? c + 2 : ncol;
        if (BITS == 16) { /* :3267 (the AVX-512 16-bit kernel only) */
            beg = nbeg;
            end = nend;
        }
        const VS live = exit0 < zero;
        /* head scan, :3269-3291 */
        VS still = bc(-1);
        S lc = wr(beg);
        for (c = beg; c < end; c++) {
            VS z = ((F[c] | Hh[c]) == zero) & live;
            if (Ln::bits(z) == 0) break;
            z &= still;
            lc = wr(lc + 1);
            head = z ? bc(lc) : head;
            still = z;
        }
        /* tail scan, :3293-3316 */
        VS idx = tail;
        still = bc(-1);
        lc = wr(end);
        for (c = end; c >= beg; c--) {
            VS z = ((F[c] | Hh[c]) == zero) & live;
            if (Ln::bits(z) == 0) break;
            z &= still;
            lc = wr(lc - 1);
            idx = z ? bc(lc) : idx;
            still = z;
        }
        tail = vmin(add(idx, bc(2)), qlen);
    }
    for (int l = 0; l < L; l++) { /* :3327-3353 */
        p[l].score = maxScore[l];
        p[l].tle = x[l];
        p[l].qle = y[l];
        p[l].max_off = max_off[l];
        p[l].gscore = gscore[l];
        p[l].gtle = max_ie[l];
    }
}

/* The batch wrappers' per-group setup (:2787-2942, :2086-2244): SoA
 * sequences with the ambiguity and pad codes, the first column H_v, the row
 * -1 H_h, the per-lane band. */
template <int BITS>
void setup_group(const Object &ob, Obj<BITS> &o, oracle_seqpair *p, const uint8_t *ref,
                 const uint8_t *qer, int32_t w)
{
    typedef Lanes<BITS> Ln;
    typedef typename Ln::S S;
    typedef typename Ln::U U;
    constexpr int L = Ln::L;
    auto wr = [](int32_t x) -> S { return (S)(U)(uint32_t)x; };
    U *rs = (U *)o.rs, *qs = (U *)o.qs;
    S *Hv = (S *)o.Hv, *Hh = (S *)o.Hh;
    const U amb = (U)~(U)0; /* code 4 as the reference maps it */
    int smax = 0;
    if (smax < ob.w_match) smax = ob.w_match;
    if (smax < ob.w_mismatch) smax = ob.w_mismatch;
    if (smax < ob.w_ambig) smax = ob.w_ambig;
    S h0[L];
    U qlen[L];
    S myband[L];
    int maxLen1 = 0, maxLen2 = 0;
    for (int l = 0; l < L; l++) {
        h0[l] = (S)(U)p[l].h0;
        const uint8_t *s = ref + (int64_t)p[l].idr;
        for (int k = 0; k < p[l].len1; k++) {
            rs[(size_t)k * L + l] = (s[k] == CODE_N) ? amb : s[k];
            Hv[(size_t)k * L + l] = 0;
        }
        qlen[l] = (U)(p[l].len2 * smax);
        if (maxLen1 < p[l].len1) maxLen1 = p[l].len1;
    }
    for (int l = 0; l < L; l++)
        for (int k = p[l].len1; k <= maxLen1; k++) {
            rs[(size_t)k * L + l] = PAD_REF;
            Hv[(size_t)k * L + l] = PAD_REF;
        }
    {
        S t[L];
        for (int l = 0; l < L; l++) {
            Hv[l] = h0[l];
            t[l] = wr(h0[l] - ob.o_del);
        }
        for (int k = 1; k < maxLen1; k++)
            for (int l = 0; l < L; l++) {
                t[l] = wr(t[l] - ob.e_del);
                Hv[(size_t)k * L + l] = t[l] > 0 ? t[l] : 0;
            }
    }
    for (int l = 0; l < L; l++) {
        const uint8_t *s = qer + (int64_t)p[l].idq;
        for (int k = 0; k < p[l].len2; k++) {
            qs[(size_t)k * L + l] = (s[k] == CODE_N) ? amb : s[k];
            Hh[(size_t)k * L + l] = 0;
        }
        if (maxLen2 < p[l].len2) maxLen2 = p[l].len2;
    }
    for (int l = 0; l < L; l++)
        for (int k = p[l].len2; k <= maxLen2; k++) {
            qs[(size_t)k * L + l] = PAD_QRY;
            Hh[(size_t)k * L + l] = 0;
        }
    {
        const S oe = wr(ob.o_ins + ob.e_ins);
        S t[L];
        for (int l = 0; l < L; l++) {
            Hh[l] = h0[l];
            t[l] = (h0[l] > oe) ? wr(h0[l] - oe) : 0;
            Hh[L + l] = t[l];
        }
        for (int k = 2; k < maxLen2; k++)
            for (int l = 0; l < L; l++) {
                const S v = wr(t[l] - ob.e_ins);
                t[l] = v > 0 ? v : 0;
                Hh[(size_t)k * L + l] = t[l];
            }
    }
    int32_t bsize = w;
    for (int l = 0; l < L; l++) { /* :2882-2904 */
        U t = (U)(qlen[l] + wr(ob.end_bonus - ob.o_ins));
        int b = (int)(t / ob.e_ins + 1.0);
        b = b > 1 ? b : 1;
        myband[l] = (S)(U)(bsize < b ? bsize : b);
    }
    for (int l = 0; l < L; l++) {
        U t = (U)(qlen[l] + wr(ob.end_bonus - ob.o_del));
        int b = (int)(t / ob.e_del + 1.0);
        b = b > 1 ? b : 1;
        const int mb = (U)myband[l];
        myband[l] = (S)(U)(mb < b ? mb : b);
        bsize = bsize < (int)(U)myband[l] ? (int)(U)myband[l] : bsize;
    }
    dp_group<BITS>(ob, o, (U)maxLen1, (U)maxLen2, p, bsize, myband);
}

template <int BITS>
void get_scores(Object &ob, Obj<BITS> &o, oracle_seqpair *pairs, const uint8_t *ref,
                const uint8_t *qer, int32_t n, int32_t w)
{
    constexpr int L = Lanes<BITS>::L;
    oracle_seqpair g[L];
    if (n <= 0) return;
    const int32_t last_len2 = pairs[n - 1].len2;
    for (int32_t base = 0; base < n; base += L) {
        const int cnt = (n - base < L) ? n - base : L;
        memcpy(g, pairs + base, (size_t)cnt * sizeof(oracle_seqpair));
        for (int l = cnt; l < L; l++) { /* padding lanes, :2710-2717 / :2013-2018 */
            memset(&g[l], 0, sizeof(g[l]));
            g[l].id = base + l;
            if (BITS == 8) g[l].len2 = last_len2;
        }
        setup_group<BITS>(ob, o, g, ref, qer, w);
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

} // namespace SIMD_NS

extern "C" {

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define ENTRY(name) CAT(name, SIMD_SUFFIX)

/* One thread-owned object per thread, contiguous chunks of `chunk` pairs
 * (bench-bsw's 20,000), as bsw_oracle_run_mt does for the lane oracle. */
int ENTRY(bsw_simd_run_mt)(int o_del, int e_del, int o_ins, int e_ins, int zdrop, int end_bonus,
                           int8_t w_match, int8_t w_mismatch, int bits, oracle_seqpair *pairs,
                           const uint8_t *ref, const uint8_t *qer, int32_t n, int32_t w,
                           int32_t chunk, int nthreads)
{
    using namespace SIMD_NS;
    int err = 0;
    if (chunk <= 0) chunk = 20000;
#pragma omp parallel num_threads(nthreads)
    {
        Object *ob = (Object *)calloc(1, sizeof(Object));
        bool ok = ob != nullptr;
        if (ok) {
            ob->o_del = o_del;
            ob->e_del = e_del;
            ob->o_ins = o_ins;
            ob->e_ins = e_ins;
            ob->zdrop = zdrop;
            ob->end_bonus = end_bonus;
            ob->w_match = w_match;
            ob->w_mismatch = (int8_t)(w_mismatch * -1); /* :65 */
            ob->w_ambig = -1;                           /* :68 */
            ok = bits == 8 ? alloc<8>(ob->o8) : alloc<16>(ob->o16);
        }
        if (!ok) {
#pragma omp atomic write
            err = -1;
        }
#pragma omp for schedule(dynamic, 1)
        for (int32_t base = 0; base < n; base += chunk) {
            if (!ok) continue;
            const int32_t cnt = n - base < chunk ? n - base : chunk;
            if (bits == 8)
                get_scores<8>(*ob, ob->o8, pairs + base, ref, qer, cnt, w);
            else
                get_scores<16>(*ob, ob->o16, pairs + base, ref, qer, cnt, w);
        }
        if (ob) {
            if (bits == 8)
                release<8>(ob->o8);
            else
                release<16>(ob->o16);
            free(ob);
        }
    }
    return err;
}

} /* extern "C" */
