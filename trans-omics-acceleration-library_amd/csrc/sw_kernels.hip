/*
 * sw_kernels.hip -- gfx950 kernels of the full inter-task Smith-Waterman
 * engine (include/tal_sw.h), the drop-in for PairWiseSW_{8,16,32}::getScores
 * (src/alignment/Smith-Waterman/PairWiseSW_16.cpp:136-377 and its 8/32-bit
 * twins).
 *
 * Semantics (the reference's AVX-512BW build): a group is 64 / 32 / 16
 * consecutive pairs; every pair of a group is scored over the group's padded
 * matrix (rows = longest len1, columns = longest len2, both rounded up to a
 * multiple of 4, in the class's integer type), with 'B' past a pair's own
 * seq1 and 'D' past its seq2 (_16.cpp:296-345).  The recurrence of MAIN_CODE
 * (_16.cpp:62-76), per cell, with every add wrapping at the lane width:
 *     m = Hdiag + (s1 == s2 ? match : mismatch)
 *     H = max(m, E, F, 0);  best = max(best, H)
 *     t = H + open;  E' = max(E + extend, t);  F' = max(F + extend, t)
 * with F = open and H = 0 above the first row (:170-174), E = open and
 * Hdiag = 0 left of the first column (:183-191).
 *
 * Layout on the GPU:
 *  - a unit is 32 pairs (8/16-bit) or 16 pairs (32-bit); a 16-lane quarter
 *    of a wave scores one unit, so a wave scores four.  In the 8/16-bit
 *    classes each lane holds two pairs of its unit (k and k+16) in the two
 *    16-bit halves of every register and the cell update runs on packed
 *    16-bit instructions; the 8-bit class holds value << 8 in each half, so
 *    that 16-bit adds wrap and 16-bit signed max orders exactly as the epi8
 *    instructions do.  The 32-bit class holds one pair per lane.
 *  - units are sorted by their bounds (widest first), so the four units of a
 *    wave have equal or near bounds; cells past a unit's own bounds are
 *    computed but masked out of its maximum.
 *  - rows are swept in strips of 16/8/4 rows held in registers (E, the
 *    previous column's H and the seq1 codes of each row); between strips the
 *    bottom row's H and F per column go to a per-wave slab in global memory,
 *    4 columns per 32-byte record per lane.  The seq1/seq2 codes are staged
 *    in the same slab, packed per lane.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sw_internal.h"

namespace talsw {

namespace {

typedef short s2v __attribute__((ext_vector_type(2)));
typedef unsigned short u2v __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t pmax(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s2v, a),
                                                                  __builtin_bit_cast(s2v, b)));
}
__device__ __forceinline__ uint32_t padd(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u2v, a) + __builtin_bit_cast(u2v, b));
}
/* min(a ^ b, 1) per 16-bit half: 0 where the halves are equal, else 1.  An
 * instruction of its own for the same reason as pmad below. */
__device__ __forceinline__ uint32_t pneq(uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_pk_min_u16 %0, %1, 1 op_sel_hi:[1,0]" : "=v"(r) : "v"(a ^ b));
    return r;
}
/* a * b + c per 16-bit half (low 16 bits).  Written as the instruction
 * itself: with a in {0, 1} the compiler would otherwise turn it into two
 * compares, two selects and a byte permute. */
__device__ __forceinline__ uint32_t pmad(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_pk_mad_u16 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ int32_t smax32(int32_t a, int32_t b) { return a > b ? a : b; }

/* One cell of MAIN_CODE for the lane's pairs.  PACK: two 16-bit lanes per
 * register (8/16-bit classes); otherwise one int32 lane. */
template <bool PACK> struct Cell;

template <> struct Cell<true> {
    uint32_t match, delta, open, ext;
    __device__ __forceinline__ uint32_t step(uint32_t a, uint32_t b, uint32_t d, uint32_t &e,
                                             uint32_t &f) const
    {
        const uint32_t neq = pneq(a, b);
        const uint32_t m = padd(d, pmad(neq, delta, match));
        const uint32_t h = pmax(pmax(pmax(m, e), f), 0u);
        const uint32_t t = padd(h, open);
        e = pmax(padd(e, ext), t);
        f = pmax(padd(f, ext), t);
        return h;
    }
    __device__ __forceinline__ uint32_t best(uint32_t acc, uint32_t h) const { return pmax(acc, h); }
};

template <> struct Cell<false> {
    uint32_t match, mismatch, open, ext;
    __device__ __forceinline__ uint32_t step(uint32_t a, uint32_t b, uint32_t d, uint32_t &e,
                                             uint32_t &f) const
    {
        const uint32_t m = d + (a == b ? match : mismatch);
        const int32_t h = smax32(smax32(smax32((int32_t)m, (int32_t)e), (int32_t)f), 0);
        const uint32_t t = (uint32_t)h + open;
        e = (uint32_t)smax32((int32_t)(e + ext), (int32_t)t);
        f = (uint32_t)smax32((int32_t)(f + ext), (int32_t)t);
        return (uint32_t)h;
    }
    __device__ __forceinline__ uint32_t best(uint32_t acc, uint32_t h) const
    {
        return (uint32_t)smax32((int32_t)acc, (int32_t)h);
    }
};

/* Slab of one wave, in 32-bit words:
 *   codes1 [rows_cap/4][64][4]   seq1 codes of row r at [r>>2][lane][r&3]
 *   codes2 [cols_cap/4][64][4]   seq2 codes, same shape over columns
 *   state  [cols_cap/4][64][8]   H (4 columns) then F (4 columns) of the row
 *                                below the last finished strip */
struct Slab {
    uint32_t *codes1, *codes2, *state;
};

/*
 * One 4-column tile of an H-row strip, the cells in anti-diagonal order:
 * step k visits (k - cc, cc) for the four columns, which do not depend on
 * each other, so dependent packed instructions are four cells apart.
 * Column cc carries its F down and the H it takes as the next row's
 * diagonal; row r carries E right (e[r]), and hp[r] hands H(r, cc) to
 * column cc + 1.  hu/fu: H and F of the row above the strip for the tile's
 * columns; hup: H above the column left of the tile; b: the tile's seq2
 * codes.  Outputs the strip's bottom-row H (ho) and the F below it (fo).
 * MASK: cells in rows >= rv or in a tile with cm == 0 stay out of the
 * maxima.
 */
template <bool PACK, int H, bool MASK>
__device__ __forceinline__ void tile_cells(const Cell<PACK> &C, const uint32_t (&s1)[H],
                                           uint32_t (&e)[H], uint32_t (&hp)[H],
                                           uint32_t (&acc)[4], uint32_t &hup, const u32x4 &hu,
                                           const u32x4 &fu, const u32x4 &b, int rv, uint32_t cm,
                                           uint32_t (&ho)[4], uint32_t (&fo)[4])
{
    uint32_t f[4] = {fu.x, fu.y, fu.z, fu.w};
    uint32_t d[4] = {hup, hu.x, hu.y, hu.z};
    hup = hu.w;
    const uint32_t bc[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
    for (int k = 0; k < H + 3; k++) {
#pragma unroll
        for (int cc = 0; cc < 4; cc++) {
            const int r = k - cc;
            if (r < 0 || r >= H) continue;
            const uint32_t h = C.step(s1[r], bc[cc], d[cc], e[r], f[cc]);
            acc[cc] = C.best(acc[cc], MASK ? (r < rv ? (h & cm) : 0u) : h);
            d[cc] = hp[r];
            hp[r] = h;
            if (r == H - 1) ho[cc] = h;
        }
    }
#pragma unroll
    for (int cc = 0; cc < 4; cc++) fo[cc] = f[cc];
}

/*
 * One strip of H rows starting at r0, over columns [0, NC).  MASK: some
 * lane's unit ends inside the strip or before NC, so cells past the unit's
 * own bounds (rows >= nrow, columns >= ncol) are kept out of its maximum.
 */
template <bool PACK, int H, bool MASK>
__device__ __forceinline__ uint32_t strip(const Cell<PACK> &C, const Slab &sl, int lane, int r0,
                                          int NC, bool first, bool last, int nrow, int ncol,
                                          uint32_t best)
{
    uint32_t s1[H], e[H], hp[H];
    const u32x4 *c1 = (const u32x4 *)sl.codes1 + (size_t)(r0 >> 2) * 64 + lane;
#pragma unroll
    for (int k = 0; k < H / 4; k++) {
        const u32x4 v = c1[(size_t)k * 64];
        s1[4 * k] = v.x;
        s1[4 * k + 1] = v.y;
        s1[4 * k + 2] = v.z;
        s1[4 * k + 3] = v.w;
    }
#pragma unroll
    for (int r = 0; r < H; r++) {
        e[r] = C.open;
        hp[r] = 0;
    }
    const int rv = nrow - r0; /* rows of this strip inside the lane's unit */
    const u32x4 *c2 = (const u32x4 *)sl.codes2 + lane;
    u32x4 *st = (u32x4 *)sl.state + 2 * lane;
    const u32x4 open4 = {C.open, C.open, C.open, C.open};
    const u32x4 zero4 = {0u, 0u, 0u, 0u};
    uint32_t hup = 0;
    uint32_t acc[4] = {best, best, best, best};
    u32x4 nb = c2[0];
    u32x4 nh = first ? zero4 : st[0];
    u32x4 nf = first ? open4 : st[1];
    for (int c0 = 0; c0 < NC; c0 += 4) {
        const int t = c0 >> 2;
        const u32x4 b = nb, hu = nh, fu = nf;
        if (c0 + 4 < NC) { /* prefetch the next 4-column record */
            nb = c2[(size_t)(t + 1) * 64];
            if (!first) {
                nh = st[(size_t)(t + 1) * 128];
                nf = st[(size_t)(t + 1) * 128 + 1];
            }
        }
        uint32_t ho[4], fo[4];
        tile_cells<PACK, H, MASK>(C, s1, e, hp, acc, hup, hu, fu, b, rv,
                                  c0 < ncol ? 0xFFFFFFFFu : 0u, ho, fo);
        if (!last) {
            st[(size_t)t * 128] = u32x4{ho[0], ho[1], ho[2], ho[3]};
            st[(size_t)t * 128 + 1] = u32x4{fo[0], fo[1], fo[2], fo[3]};
        }
    }
    return C.best(C.best(acc[0], acc[1]), C.best(acc[2], acc[3]));
}

__device__ __forceinline__ uint32_t from_lane(int src, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}

/*
 * One round of a unit split over K quarters of the wave (K = 2 or 4).
 * Strips of 16 rows go round-robin to the unit's quarters: in round j,
 * quarter s sweeps strip jK + s, one 4-column tile behind quarter s - 1,
 * whose strip lies directly above.  After each tile the bottom row of every
 * quarter's tile (H and F of four columns) is handed to the quarter below
 * across lanes, so only quarter 0 reads the state slab (the bottom row of the
 * previous round's last strip) and only quarter K - 1 writes it.
 * act: this quarter has a strip in this round; rv: rows of it inside the
 * lane's unit.  `ul` indexes the unit's lanes in the slab's state rows.
 */
template <bool PACK, int K, bool MASK>
__device__ __forceinline__ void round_split(const Cell<PACK> &C, const Slab &sl, int lane, int s,
                                            int ul, int r0, bool act, int rv, int NC, int ncol,
                                            bool first, bool last, uint32_t (&acc)[4])
{
    constexpr int H = 16;
    uint32_t s1[H], e[H], hp[H];
    const u32x4 *c1 = (const u32x4 *)sl.codes1 + (size_t)(r0 >> 2) * 64 + lane;
#pragma unroll
    for (int k = 0; k < H / 4; k++) {
        const u32x4 v = c1[(size_t)k * 64];
        s1[4 * k] = v.x;
        s1[4 * k + 1] = v.y;
        s1[4 * k + 2] = v.z;
        s1[4 * k + 3] = v.w;
    }
#pragma unroll
    for (int r = 0; r < H; r++) {
        e[r] = C.open;
        hp[r] = 0;
    }
    const int ntiles = NC >> 2;
    const u32x4 *c2 = (const u32x4 *)sl.codes2 + lane;
    u32x4 *st = (u32x4 *)sl.state + 2 * ul;
    const u32x4 open4 = {C.open, C.open, C.open, C.open};
    const u32x4 zero4 = {0u, 0u, 0u, 0u};
    const int src = s > 0 ? lane - 16 : lane;
    uint32_t hup = 0;
    u32x4 nb = c2[0];
    u32x4 nh = zero4, nf = open4; /* quarter 0: the slab row above, prefetched */
    if (s == 0 && !first) {
        nh = st[0];
        nf = st[1];
    }
    u32x4 xh = zero4, xf = open4; /* quarters > 0: handed down by quarter s - 1 */
    for (int tau = 0; tau < ntiles + K - 1; tau++) {
        const int c = tau - s;
        const bool inr = act && c >= 0 && c < ntiles;
        const u32x4 b = nb;
        const u32x4 hu = s == 0 ? nh : xh, fu = s == 0 ? nf : xf;
        const int cn = c + 1;
        if (cn >= 0 && cn < ntiles) { /* prefetch the tile of the next step */
            nb = c2[(size_t)cn * 64];
            if (s == 0 && !first) {
                nh = st[(size_t)cn * 128];
                nf = st[(size_t)cn * 128 + 1];
            }
        }
        uint32_t ho[4] = {0u, 0u, 0u, 0u}, fo[4] = {0u, 0u, 0u, 0u};
        if (inr) {
            tile_cells<PACK, H, MASK>(C, s1, e, hp, acc, hup, hu, fu, b, rv,
                                      4 * c < ncol ? 0xFFFFFFFFu : 0u, ho, fo);
            if (s == K - 1 && !last) {
                st[(size_t)c * 128] = u32x4{ho[0], ho[1], ho[2], ho[3]};
                st[(size_t)c * 128 + 1] = u32x4{fo[0], fo[1], fo[2], fo[3]};
            }
        }
        xh = u32x4{from_lane(src, ho[0]), from_lane(src, ho[1]), from_lane(src, ho[2]),
                   from_lane(src, ho[3])};
        xf = u32x4{from_lane(src, fo[0]), from_lane(src, fo[1]), from_lane(src, fo[2]),
                   from_lane(src, fo[3])};
    }
}

/* All rounds of a split unit; returns the lane's best over its strips. */
template <bool PACK, int K>
__device__ __forceinline__ uint32_t sweep_split(const Cell<PACK> &C, const Slab &sl, int lane,
                                                int s, int ul, int NR, int NC, int nrow, int ncol,
                                                int min_nrow, int min_ncol)
{
    uint32_t acc[4] = {0u, 0u, 0u, 0u};
    const int nstrips = (NR + 15) >> 4;
    const int nrounds = (nstrips + K - 1) / K;
    for (int j = 0; j < nrounds; j++) {
        const int si = j * K + s;
        const bool act = si < nstrips;
        const int r0 = (act ? si : nstrips - 1) * 16; /* inactive quarters idle */
        const bool first = j == 0, last = j == nrounds - 1;
        /* masked where a unit ends inside this round's rows or before NC */
        const bool m = (j + 1) * K * 16 > min_nrow || min_ncol < NC;
        if (m)
            round_split<PACK, K, true>(C, sl, lane, s, ul, r0, act, nrow - r0, NC, ncol, first,
                                       last, acc);
        else
            round_split<PACK, K, false>(C, sl, lane, s, ul, r0, act, nrow - r0, NC, ncol, first,
                                        last, acc);
        /* the next round's quarter 0 reads what quarter K - 1 stored */
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    return C.best(C.best(acc[0], acc[1]), C.best(acc[2], acc[3]));
}

/* All rows in strips of 16, then 8 and 4 for the remainder (the bounds
 * are multiples of 4).  A strip is masked only where some unit of the wave
 * ends before the strip's last row or before the widest unit's last
 * column; masked strips take at most 8 rows. */
template <bool PACK>
__device__ __forceinline__ uint32_t sweep(const Cell<PACK> &C, const Slab &sl, int lane, int NR,
                                          int NC, int nrow, int ncol, int min_nrow, int min_ncol)
{
    uint32_t best = 0;
    int r0 = 0;
    while (r0 < NR) {
        const int rem = NR - r0;
        int h = rem >= 16 ? 16 : rem >= 8 ? 8 : 4;
        const bool m = r0 + h > min_nrow || min_ncol < NC;
        /* masked strips are at most 8 rows: the row masks take registers */
        if (m && h == 16) h = 8;
        const bool first = r0 == 0, last = r0 + h >= NR;
        if (h == 16) {
            best = strip<PACK, 16, false>(C, sl, lane, r0, NC, first, last, nrow, ncol, best);
        } else if (h == 8) {
            best = m ? strip<PACK, 8, true>(C, sl, lane, r0, NC, first, last, nrow, ncol, best)
                     : strip<PACK, 8, false>(C, sl, lane, r0, NC, first, last, nrow, ncol, best);
        } else {
            best = m ? strip<PACK, 4, true>(C, sl, lane, r0, NC, first, last, nrow, ncol, best)
                     : strip<PACK, 4, false>(C, sl, lane, r0, NC, first, last, nrow, ncol, best);
        }
        r0 += h;
    }
    return best;
}

__device__ __forceinline__ int wave_max4(int v)
{
    const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    const int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return max(max(a, b), max(c, d));
}
__device__ __forceinline__ int wave_min4(int v)
{
    const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    const int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return min(min(a, b), min(c, d));
}

/* Bytes k..k+3 of a sequence as one little-endian word, the padding code in
 * the bytes at or past len.  Whole words are read as one aligned load when
 * the sequences are 4-byte aligned (their slots are multiples of 128 bytes,
 * so that is the buffer's own alignment); otherwise and for the last partial
 * word, byte loads at clamped offsets, which every lane issues without
 * branching.  `s` must be readable at offset 0 even when len is 0. */
__device__ __forceinline__ uint32_t word_at(const uint8_t *s, int len, int k, uint32_t pad,
                                            bool aligned)
{
    if (aligned && k + 4 <= len) return *(const uint32_t *)(s + k);
    const int last = max(len - 1, 0);
    uint32_t w = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t c = s[min(k + i, last)];
        w |= (k + i < len ? c : pad) << (8 * i);
    }
    return w;
}

} // namespace

/*
 * Group bounds and validation, one thread per group: the padded bounds of
 * the wrapper (_16.cpp:296-345; int8 bounds in the 8-bit class, _8.cpp:293,
 * 308, which wrap to -128 when they round up to 128), each unit's bounds and
 * sort key, the batch maxima and the error word.
 */
__device__ __forceinline__ uint32_t spread_bits(uint32_t x)
{
    x &= 0xFFFF;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

__global__ void sw_shapes(const tal_sw_pair *pairs, int32_t n, int32_t bits, size_t seq_len,
                          int32_t *unit_dims, uint32_t *keys, int32_t *ids, int32_t *ctl,
                          SwInert I)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int GL = bits == 8 ? 64 : bits == 16 ? 32 : 16;
    const int UL = bits == 32 ? 16 : 32;
    const int ngroups = (n + GL - 1) / GL;
    if (g >= ngroups) return;
    const int64_t S = bits == 8 ? 128 : 1024;
    const int limit = bits == 8 ? 127 : 1024;
    int max1 = 0, max2 = 0, err = 0;
    for (int k = 0; k < GL; k++) {
        const int p = g * GL + k;
        if (p >= n) break;
        const tal_sw_pair sp = pairs[p];
        if (sp.len1 < 0 || sp.len2 < 0 || sp.len1 > limit || sp.len2 > limit) {
            err = max(err, SW_ERR_RANGE);
            continue;
        }
        if (sp.id < 0 || (2 * (int64_t)sp.id + 1) * S + sp.len2 > (int64_t)seq_len ||
            2 * (int64_t)sp.id * S + sp.len1 > (int64_t)seq_len)
            err = max(err, SW_ERR_BOUNDS);
        max1 = max(max1, (int)sp.len1);
        max2 = max(max2, (int)sp.len2);
    }
    if (err) atomicMax(&ctl[SW_CTL_ERR], err);
    int nrow = ((max1 + 3) / 4) * 4, ncol = ((max2 + 3) / 4) * 4;
    if (bits == 8) {
        nrow = (int8_t)(uint8_t)(nrow & 0xFF);
        ncol = (int8_t)(uint8_t)(ncol & 0xFF);
    }
    /* a bound <= 0 leaves the reference's loops empty: every score is 0 */
    if (nrow <= 0 || ncol <= 0) nrow = ncol = 0;
    const int upg = GL / UL;
    for (int u = 0; u < upg; u++) {
        const int id = g * upg + u;
        unit_dims[id] = (nrow << 16) | ncol;
        keys[id] = ((uint32_t)(ncol >> 2) << 10) | (uint32_t)(nrow >> 2);
        ids[id] = id;
    }
    atomicMax(&ctl[SW_CTL_MAXROW], nrow);
    atomicMax(&ctl[SW_CTL_MAXCOL], ncol);
    if (!I.on) return;
    /* Regrouping keys: pairs of a zero-bound group keep score 0 and sort
     * last (key 0); the others sort by the Morton order of (len1, len2), so
     * that runs of the sorted order have near lengths in both dimensions. */
    int elig = 0, bad = 0;
    for (int k = 0; k < GL; k++) {
        const int p = g * GL + k;
        if (p >= n) break;
        const tal_sw_pair sp = pairs[p];
        uint32_t key = 0;
        if (nrow > 0 && sp.len1 >= 0 && sp.len2 >= 0) {
            key = (spread_bits(sp.len1) | (spread_bits(sp.len2) << 1)) + 1;
            elig++;
            /* the best local score is at most match * min(len1, len2) */
            if ((int64_t)max(I.match, 0) * min((int)sp.len1, (int)sp.len2) > I.lane_max) bad = 1;
        }
        I.pair_keys[p] = key;
        I.pair_ids[p] = p;
    }
    if (elig) atomicAdd(&ctl[SW_CTL_ELIGIBLE], elig);
    if (bad) atomicMax(&ctl[SW_CTL_NOT_INERT], 1);
}

/* One wave per pair: does seq1 hold the seq2 padding code 'D' or seq2 the
 * seq1 padding code 'B'?  Then a padded cell could score a match and the
 * reference's groups must be kept.  Pairs that fail validation are skipped
 * (the bounds kernel reports them). */
__global__ void sw_scan_codes(const tal_sw_pair *pairs, int32_t n, int32_t bits,
                              const uint8_t *seq, size_t seq_len, int32_t *ctl)
{
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= n) return;
    const tal_sw_pair sp = pairs[p];
    const int64_t S = bits == 8 ? 128 : 1024;
    const int limit = bits == 8 ? 127 : 1024;
    if (sp.len1 < 0 || sp.len2 < 0 || sp.len1 > limit || sp.len2 > limit || sp.id < 0 ||
        (2 * (int64_t)sp.id + 1) * S + sp.len2 > (int64_t)seq_len ||
        2 * (int64_t)sp.id * S + sp.len1 > (int64_t)seq_len)
        return;
    const uint8_t *a = seq + 2 * (int64_t)sp.id * S, *b = a + S;
    bool hit = false;
    for (int k = lane; k < sp.len1; k += 64) hit |= a[k] == 'D';
    for (int k = lane; k < sp.len2; k += 64) hit |= b[k] == 'B';
    if (hit) ctl[SW_CTL_NOT_INERT] = 1;
}

/* Bounds and sort key of each regrouped unit (UL consecutive entries of the
 * sorted pairs), in the layout of the bounds kernel. */
__global__ void sw_unit_dims(const tal_sw_pair *pairs, const int32_t *perm, int32_t m,
                             int32_t bits, int32_t *unit_dims, uint32_t *keys, int32_t *ids)
{
    const int UL = bits == 32 ? 16 : 32;
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= (m + UL - 1) / UL) return;
    int max1 = 0, max2 = 0;
    for (int k = 0; k < UL; k++) {
        const int i = u * UL + k;
        if (i >= m) break;
        const tal_sw_pair sp = pairs[perm[i]];
        max1 = max(max1, (int)sp.len1);
        max2 = max(max2, (int)sp.len2);
    }
    int nrow = ((max1 + 3) / 4) * 4, ncol = ((max2 + 3) / 4) * 4;
    if (nrow <= 0 || ncol <= 0) nrow = ncol = 0;
    unit_dims[u] = (nrow << 16) | ncol;
    keys[u] = ((uint32_t)(ncol >> 2) << 10) | (uint32_t)(nrow >> 2);
    ids[u] = u;
}

__global__ void sw_zero_scores(tal_sw_pair *pairs, const int32_t *idx, int32_t count)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) pairs[idx[k]].score = 0;
}

/*
 * The DP.  Persistent grid of one-wave blocks; a wave takes the next 4 / K
 * units of the sorted order from the task counter until none is left.  K = 1:
 * each quarter sweeps a unit of its own; K = 2 or 4: each unit is split over
 * K quarters (round_split), for batches with too few units to fill the GPU.
 */
template <int BITS, int K>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) sw_main(SwParams P, SwBuffers B)
{
    constexpr bool PACK = BITS != 32;
    constexpr int UPW = 4 / K; /* units per wave */
    const int lane = threadIdx.x;
    const int q = lane >> 4, l16 = lane & 15;
    const int j = q / K, s = q % K; /* unit slot and the quarter's part of it */
    Slab sl;
    {
        uint32_t *base = B.slabs + (size_t)blockIdx.x * P.slab_words;
        sl.codes1 = base;
        sl.codes2 = base + (size_t)P.rows_cap * 64;
        sl.state = sl.codes2 + (size_t)P.cols_cap * 64;
    }
    Cell<PACK> C;
    if constexpr (PACK) {
        C = Cell<true>{P.k_match, P.k_delta, P.k_open, P.k_ext};
    } else {
        C = Cell<false>{P.k_match, P.k_delta, P.k_open, P.k_ext};
    }
    const int ntasks = (P.nunits + UPW - 1) / UPW;
    for (;;) {
        int task = 0;
        if (lane == 0) task = atomicAdd(&B.ctl[SW_CTL_COUNTER], 1);
        task = __builtin_amdgcn_readfirstlane(task);
        if (task >= ntasks) break;
        const int ui = task * UPW + j;
        const int unit = ui < P.nunits ? B.order[ui] : -1;
        const int dims = unit >= 0 ? B.unit_dims[unit] : 0;
        const int nrow = dims >> 16, ncol = dims & 0xFFFF;
        const int NR = wave_max4(nrow), NC = wave_max4(ncol);
        const int min_nrow = wave_min4(nrow), min_ncol = wave_min4(ncol);
        /* the lane's pairs: k and k+16 of the unit (8/16-bit), k (32-bit) */
        const int ul = PACK ? 32 : 16;
        int pa, pb;
        if (B.perm) { /* regrouped units */
            const int i = unit * ul + l16;
            pa = unit >= 0 && i < P.m ? B.perm[i] : P.n;
            pb = PACK && unit >= 0 && i + 16 < P.m ? B.perm[i + 16] : P.n;
        } else {
            pa = unit >= 0 ? unit * ul + l16 : P.n;
            pb = PACK ? pa + 16 : P.n;
        }
        const bool va = pa < P.n, vb = PACK && unit >= 0 && pb < P.n;
        int la1 = 0, la2 = 0, lb1 = 0, lb2 = 0;
        /* an always-readable address for empty or absent sequences */
        const uint8_t *safe = (const uint8_t *)B.pairs;
        const uint8_t *a1 = safe, *a2 = safe, *b1 = safe, *b2 = safe;
        if (va) {
            const tal_sw_pair sp = B.pairs[pa];
            la1 = sp.len1;
            la2 = sp.len2;
            if (la1 > 0) a1 = B.seq + 2 * (int64_t)sp.id * P.slot;
            if (la2 > 0) a2 = B.seq + (2 * (int64_t)sp.id + 1) * P.slot;
        }
        if (vb) {
            const tal_sw_pair sp = B.pairs[pb];
            lb1 = sp.len1;
            lb2 = sp.len2;
            if (lb1 > 0) b1 = B.seq + 2 * (int64_t)sp.id * P.slot;
            if (lb2 > 0) b2 = B.seq + (2 * (int64_t)sp.id + 1) * P.slot;
        }
        /* stage the codes; rows/columns past a pair's length take the
         * padding codes 'B' / 'D' up to the unit's bounds.  Packed: row k
         * of pair a in the low half and of pair b in the high half. */
        const bool al = P.aligned4 != 0;
        /* split units sweep whole 16-row strips: stage rows up to that */
        const int NRs = K > 1 ? (NR + 15) & ~15 : NR;
        for (int r = 0; r < NRs; r += 4) {
            const uint32_t wa = word_at(a1, la1, r, 'B', al);
            u32x4 v;
            if (PACK) {
                const uint32_t wb = word_at(b1, lb1, r, 'B', al);
                v = u32x4{__builtin_amdgcn_perm(wb, wa, 0x0C040C00u),
                          __builtin_amdgcn_perm(wb, wa, 0x0C050C01u),
                          __builtin_amdgcn_perm(wb, wa, 0x0C060C02u),
                          __builtin_amdgcn_perm(wb, wa, 0x0C070C03u)};
            } else {
                v = u32x4{wa & 0xFF, (wa >> 8) & 0xFF, (wa >> 16) & 0xFF, wa >> 24};
            }
            ((u32x4 *)sl.codes1)[(size_t)(r >> 2) * 64 + lane] = v;
        }
        for (int c = 0; c < NC; c += 4) {
            const uint32_t wa = word_at(a2, la2, c, 'D', al);
            u32x4 v;
            if (PACK) {
                const uint32_t wb = word_at(b2, lb2, c, 'D', al);
                v = u32x4{__builtin_amdgcn_perm(wb, wa, 0x0C040C00u),
                          __builtin_amdgcn_perm(wb, wa, 0x0C050C01u),
                          __builtin_amdgcn_perm(wb, wa, 0x0C060C02u),
                          __builtin_amdgcn_perm(wb, wa, 0x0C070C03u)};
            } else {
                v = u32x4{wa & 0xFF, (wa >> 8) & 0xFF, (wa >> 16) & 0xFF, wa >> 24};
            }
            ((u32x4 *)sl.codes2)[(size_t)(c >> 2) * 64 + lane] = v;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        uint32_t best;
        if constexpr (K == 1) {
            best = sweep<PACK>(C, sl, lane, NR, NC, nrow, ncol, min_nrow, min_ncol);
        } else {
            best = sweep_split<PACK, K>(C, sl, lane, s, j * 16 + l16, NR, NC, nrow, ncol,
                                        min_nrow, min_ncol);
            /* the unit's best over its K quarters */
            best = C.best(best, from_lane(lane ^ 16, best));
            if (K == 4) best = C.best(best, from_lane(lane ^ 32, best));
        }
        if (s != 0) {
            /* scores are written by the unit's first quarter */
        } else if (PACK) {
            const int16_t lo = (int16_t)(best & 0xFFFF), hi = (int16_t)(best >> 16);
            if (va) B.pairs[pa].score = BITS == 8 ? (lo >> 8) : lo;
            if (vb) B.pairs[pb].score = BITS == 8 ? (hi >> 8) : hi;
        } else if (va) {
            B.pairs[pa].score = (int32_t)best;
        }
        /* the slab is rewritten by the next task: finish this one's reads */
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
}

hipError_t launch_sw_shapes(const tal_sw_pair *pairs, int32_t n, int32_t bits, size_t seq_len,
                            int32_t *unit_dims, uint32_t *keys, int32_t *ids, int32_t *ctl,
                            const SwInert &I, hipStream_t s)
{
    const int GL = bits == 8 ? 64 : bits == 16 ? 32 : 16;
    const int G = (n + GL - 1) / GL;
    if (G == 0) return hipSuccess;
    hipLaunchKernelGGL(sw_shapes, dim3((G + 127) / 128), dim3(128), 0, s, pairs, n, bits, seq_len,
                       unit_dims, keys, ids, ctl, I);
    return hipGetLastError();
}

hipError_t launch_sw_scan_codes(const tal_sw_pair *pairs, int32_t n, int32_t bits,
                                const uint8_t *seq, size_t seq_len, int32_t *ctl, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(sw_scan_codes, dim3((n + 3) / 4), dim3(256), 0, s, pairs, n, bits, seq,
                       seq_len, ctl);
    return hipGetLastError();
}

hipError_t launch_sw_unit_dims(const tal_sw_pair *pairs, const int32_t *perm, int32_t m,
                               int32_t bits, int32_t *unit_dims, uint32_t *keys, int32_t *ids,
                               hipStream_t s)
{
    const int UL = bits == 32 ? 16 : 32;
    const int U = (m + UL - 1) / UL;
    if (U == 0) return hipSuccess;
    hipLaunchKernelGGL(sw_unit_dims, dim3((U + 127) / 128), dim3(128), 0, s, pairs, perm, m, bits,
                       unit_dims, keys, ids);
    return hipGetLastError();
}

hipError_t launch_sw_zero_scores(tal_sw_pair *pairs, const int32_t *idx, int32_t count,
                                 hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(sw_zero_scores, dim3((count + 255) / 256), dim3(256), 0, s, pairs, idx,
                       count);
    return hipGetLastError();
}

template <int BITS>
void launch_width(int split, int grid, const SwParams &P, const SwBuffers &B, hipStream_t s)
{
    if (split == 4)
        hipLaunchKernelGGL((sw_main<BITS, 4>), dim3(grid), dim3(64), 0, s, P, B);
    else if (split == 2)
        hipLaunchKernelGGL((sw_main<BITS, 2>), dim3(grid), dim3(64), 0, s, P, B);
    else
        hipLaunchKernelGGL((sw_main<BITS, 1>), dim3(grid), dim3(64), 0, s, P, B);
}

hipError_t launch_sw_main(int bits, int split, int grid, const SwParams &P, const SwBuffers &B,
                          hipStream_t s)
{
    if (bits == 8)
        launch_width<8>(split, grid, P, B, s);
    else if (bits == 16)
        launch_width<16>(split, grid, P, B, s);
    else
        launch_width<32>(split, grid, P, B, s);
    return hipGetLastError();
}

} // namespace talsw
