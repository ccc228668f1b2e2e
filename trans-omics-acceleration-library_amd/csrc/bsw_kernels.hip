/*
 * bsw_kernels.hip -- gfx950 kernels of the banded Smith-Waterman extension.
 *
 * Mapping.  The reference vectorises across pairs: one AVX-512 lane per pair,
 * 32 pairs (16-bit) or 64 pairs (8-bit) per group sharing one row loop
 * (src/alignment/BSW/bandedSWA.cpp:2962-3356 and :2263-2661).  Here one
 * wavefront lane is one pair; a 64-lane wave runs two 16-bit groups (one per
 * half-wave) or one 8-bit group.  Group-wide decisions of the reference
 * (column range, band narrowing, all-lanes-zero exit, head/tail scans) become
 * ballots over the group's lanes.  Control flow is uniform inside a group, so
 * the two groups of a 16-bit wave diverge only as whole half-waves.
 *
 * State.  The reference's per-group H_h and F rows ([col][lane] int16/int8)
 * live in one dword per (column, lane): low half H_h, high half F.  They sit
 * in LDS when the wave's widest group fits, else in a per-block slab of
 * global memory.  H_v (the first-column boundary) is recomputed from its
 * closed form instead of being stored.
 *
 * Widths.  16-bit lanes use int16 arithmetic, which wraps exactly like the
 * epi16 instructions.  8-bit lanes keep value << 8 in an int16: add, sub,
 * signed max/compare and unsigned min then behave exactly like epi8 on the
 * high byte, so one code path serves both widths.
 */
#include <type_traits>

#include "bsw_internal.h"

namespace talbsw {

template <int BITS> struct Width {
    static constexpr int GL = BITS == 16 ? 32 : 64; /* lanes per group */
    static constexpr int GPW = 64 / GL;             /* groups per wave */
    static constexpr int SH = BITS == 16 ? 0 : 8;   /* value scale shift */
    /* Wrap an integer to the lane width and scale it. */
    static __device__ __forceinline__ int16_t v(int x)
    {
        return (int16_t)(uint16_t)((uint32_t)x << SH);
    }
    static __device__ __forceinline__ int out(int16_t s) { return (int)s >> SH; }
};

__device__ __forceinline__ int16_t smax16(int16_t a, int16_t b) { return a > b ? a : b; }
__device__ __forceinline__ int16_t smin16(int16_t a, int16_t b) { return a < b ? a : b; }
__device__ __forceinline__ int16_t umin16(int16_t a, int16_t b)
{
    return (uint16_t)a < (uint16_t)b ? a : b;
}
__device__ __forceinline__ int16_t add16(int16_t a, int16_t b) { return (int16_t)(a + b); }
__device__ __forceinline__ int16_t sub16(int16_t a, int16_t b) { return (int16_t)(a - b); }
__device__ __forceinline__ int16_t abs16(int16_t a) { return a < 0 ? (int16_t)(-(int)a) : a; }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

/* Two int16 lanes in one dword: packed add/sub wrap per half and max is
 * signed per half, exactly as the scalar int16 forms. */
typedef short s2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ s2v as_s2(uint32_t x) { return __builtin_bit_cast(s2v, x); }
__device__ __forceinline__ uint32_t as_u32(s2v x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ s2v max2(s2v a, s2v b) { return __builtin_elementwise_max(a, b); }
typedef unsigned short u2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u2v as_u2(uint32_t x) { return __builtin_bit_cast(u2v, x); }
__device__ __forceinline__ uint32_t u2_u32(u2v x) { return __builtin_bit_cast(uint32_t, x); }

/* Reductions over the lanes of one group, result in every lane of it: DPP
 * steps make each 16-lane row uniform (swap neighbours, swap pairs, mirror
 * the half-rows, mirror the row), then the row-pair swap (32-lane groups)
 * and the half-wave swap (64-lane groups) bring in the other rows.  Every
 * lane of a group is active or none is. */
template <int GL, typename Op> __device__ __forceinline__ int group_reduce(int x, Op op)
{
    /* (every source lane of these patterns is in the row, so bound_ctrl
     * never applies; it lets the move fold into the operation) */
    x = op(x, __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true));  /* quad_perm 1,0,3,2 */
    x = op(x, __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, true));  /* quad_perm 2,3,0,1 */
    x = op(x, __builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, true)); /* row_half_mirror */
    x = op(x, __builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, true)); /* row_mirror */
    {
        const auto r = __builtin_amdgcn_permlane16_swap((unsigned)x, (unsigned)x, false, false);
        x = op((int)r[0], (int)r[1]);
    }
    if (GL == 64) {
        const auto r = __builtin_amdgcn_permlane32_swap((unsigned)x, (unsigned)x, false, false);
        x = op((int)r[0], (int)r[1]);
    }
    return x;
}
template <int GL> __device__ __forceinline__ int group_max(int x)
{
    return group_reduce<GL>(x, [](int a, int b) { return max(a, b); });
}
template <int GL> __device__ __forceinline__ int group_min(int x)
{
    return group_reduce<GL>(x, [](int a, int b) { return min(a, b); });
}

/* Score of ref code s1 against query code s2 (raw bytes; pads are 99/100).
 * 16-bit: code 4 maps to 0xFFFF, so only code 4 has the top bit.
 * 8-bit : code 4 maps to 0xFF, and any byte >= 128 also has the top bit. */
template <int BITS>
__device__ __forceinline__ int16_t score_of(int s1, int s2, const KParams &K)
{
    bool amb;
    if (BITS == 16)
        amb = (s1 == 4) | (s2 == 4);
    else
        amb = (s1 == 4) | (s2 == 4) | (s1 >= 128) | (s2 >= 128);
    int16_t s = (s1 == s2) ? K.match : K.mism;
    return amb ? K.ambig : s;
}

typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(1))) uint32_t glb_u32;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) u32x4 glb_u32x4;
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;

/* Buffer resource of one wave's slab.  Slab accesses take 32-bit byte
 * offsets; an offset of OOB lies past every slab, where the hardware's range
 * check turns a load into zeros and drops a store, so a lane with nothing to
 * read or write in a block gives that offset instead of branching. */
typedef __amdgpu_buffer_rsrc_t rsrc_t;
constexpr uint32_t OOB = 0x80000000u;
constexpr int RSRC_WORD3 = 0x00020000; /* raw dword buffer (gfx9 family) */

__device__ __forceinline__ uint32_t bld32(rsrc_t r, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
}
__device__ __forceinline__ void bst32(rsrc_t r, uint32_t off, uint32_t v)
{
    __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)off, 0, 0);
}
__device__ __forceinline__ u32x4 bld128(rsrc_t r, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
}
__device__ __forceinline__ void bst128(rsrc_t r, uint32_t off, u32x4 v)
{
    __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)off, 0, 0);
}

/* One dword of state per (column, lane), in 8-column blocks of each lane:
 * [block][64 lanes][8 columns], so that a lane's block is 32 contiguous
 * bytes.  Every block has a home in the wave's global slab; a window of `nb`
 * consecutive blocks starting at block `lb` lives in LDS instead, as a ring
 * (block B in slot (B - lb + r0) mod nb).  The window follows the start of
 * the row's column range: columns below it are never stored again, so the
 * LDS holds the part of every row's range that fits, instead of a fixed set
 * of columns that later rows leave behind.  The row sweep runs over the LDS
 * blocks and then over the global blocks, each with its own kind of memory
 * instruction.  In the global part a lane that skips a block reads zeros and
 * stores nothing (offset OOB), so lanes outside their band fetch no state
 * lines and dirty none. */
struct StateRef {
    lds_u32 *lds;   /* this lane's 8 dwords of LDS slot 0 */
    rsrc_t slab;    /* the wave's slab */
    uint32_t gbase; /* byte offset of this lane's 8 dwords of global block 0 */
    int nb;         /* LDS slots (blocks) */
    int nblk;       /* state blocks in the slab */
    int lb = 0;     /* first block of the LDS window */
    int r0 = 0;     /* its slot */
    __device__ __forceinline__ bool in_lds(int blk) const
    {
        return (unsigned)(blk - lb) < (unsigned)nb;
    }
    __device__ __forceinline__ int slot(int blk) const
    {
        const int s = blk - lb + r0;
        return s >= nb ? s - nb : s;
    }
    /* byte offset of the global block at column b (a multiple of 8) */
    __device__ __forceinline__ uint32_t goff(int b) const { return gbase + ((uint32_t)b << 8); }
    /* single columns outside the sweep */
    __device__ __forceinline__ uint32_t ld(int c) const
    {
        const int blk = c >> 3;
        if (in_lds(blk)) return lds[((uint32_t)slot(blk) << 9) + (c & 7)];
        return bld32(slab, goff(c & ~7) + ((uint32_t)(c & 7) << 2));
    }
    __device__ __forceinline__ void st(int c, uint32_t v) const
    {
        const int blk = c >> 3;
        if (in_lds(blk))
            lds[((uint32_t)slot(blk) << 9) + (c & 7)] = v;
        else
            bst32(slab, goff(c & ~7) + ((uint32_t)(c & 7) << 2), v);
    }
    /* the block at column b (a multiple of 8) in LDS */
    __device__ __forceinline__ lds_u32 *lblk(int b) const
    {
        return lds + ((uint32_t)slot(b >> 3) << 9);
    }
    /* whole blocks wherever they live (outside the sweep) */
    __device__ __forceinline__ void ld8any(int b, uint32_t *v) const
    {
        u32x4 x, y;
        if (in_lds(b >> 3)) {
            const lds_u32x4 *p = (const lds_u32x4 *)lblk(b);
            x = p[0];
            y = p[1];
        } else {
            x = bld128(slab, goff(b));
            y = bld128(slab, goff(b) + 16);
        }
        v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
        v[4] = y.x; v[5] = y.y; v[6] = y.z; v[7] = y.w;
    }
    __device__ __forceinline__ void st8any(int b, const uint32_t *v) const
    {
        const u32x4 x = {v[0], v[1], v[2], v[3]}, y = {v[4], v[5], v[6], v[7]};
        if (in_lds(b >> 3)) {
            lds_u32x4 *p = (lds_u32x4 *)lblk(b);
            p[0] = x;
            p[1] = y;
        } else {
            bst128(slab, goff(b), x);
            bst128(slab, goff(b) + 16, y);
        }
    }
    /* column where the LDS window ends */
    __device__ __forceinline__ int wend() const { return (lb + nb) << 3; }
    /* the window back to block 0 (a new group) */
    __device__ __forceinline__ void reset()
    {
        lb = 0;
        r0 = 0;
    }
    /* Moves the window start up to block `to`: each block leaving the
     * window is written to its home in the slab, and its slot takes the
     * block entering at the window end. */
    __device__ void advance(int to)
    {
        while (lb < to) {
            lds_u32x4 *p = (lds_u32x4 *)(lds + ((uint32_t)r0 << 9));
            const uint32_t o = goff(lb << 3);
            bst128(slab, o, p[0]);
            bst128(slab, o + 16, p[1]);
            const int in = lb + nb;
            if (in < nblk) {
                const uint32_t oi = goff(in << 3);
                const u32x4 a = bld128(slab, oi), c = bld128(slab, oi + 16);
                p[0] = a;
                p[1] = c;
            }
            lb++;
            r0 = r0 + 1 == nb ? 0 : r0 + 1;
        }
    }
    /* the 8 columns of the global block at b (a multiple of 8), read only
     * where `need` (zeros elsewhere) */
    __device__ __forceinline__ void gld8(int b, uint32_t *v, bool need) const
    {
        const uint32_t o = need ? goff(b) : OOB;
        const u32x4 x = bld128(slab, o), y = bld128(slab, o + 16);
        v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
        v[4] = y.x; v[5] = y.y; v[6] = y.z; v[7] = y.w;
    }
    /* written only where `need`: lanes with nothing to store issue a
     * dropped store, so their lines are never dirtied */
    __device__ __forceinline__ void gst8(int b, const uint32_t *v, bool need) const
    {
        const uint32_t o = need ? goff(b) : OOB;
        bst128(slab, o, (u32x4){v[0], v[1], v[2], v[3]});
        bst128(slab, o + 16, (u32x4){v[4], v[5], v[6], v[7]});
    }
};

__device__ __forceinline__ uint32_t pack(int16_t h, int16_t f)
{
    return (uint32_t)(uint16_t)h | ((uint32_t)(uint16_t)f << 16);
}

constexpr int NONE_COL = 0x7FFFFFFF;

/* Query codes of the fast path: 0..4 as given, 5 past the end of the query
 * (the reference's DUMMY2 pad), eight 4-bit codes per dword. */
constexpr uint32_t QPAD = 5;

/* Score of ref byte s1 against query code c (fast path) for building the
 * per-row table. */
template <int BITS> __device__ __forceinline__ int table_score(int s1, int c, int match, int mism)
{
    const int qbyte = c == (int)QPAD ? 100 : c;
    const bool amb = (BITS == 16) ? (s1 == 4 || qbyte == 4)
                                  : (s1 == 4 || s1 >= 128 || qbyte == 4);
    return amb ? -1 : (s1 == qbyte ? match : mism);
}

/* Loads and stores of the caller's data.  SYS: the data may have been
 * written by another agent (DMA, the host) or another wave since this wave
 * last cached it, and results are read by them: system-scope accesses, which
 * bypass the caches for these lines only. */
template <bool SYS, class T> __device__ __forceinline__ T dld(const T *p)
{
    if (SYS) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    return *p;
}
/* Sequence bytes through a buffer resource over a packed sequence buffer:
 * offsets are the pair's offset plus the position (both below 2^31 + 2^15),
 * and SEQ_OOB for a position past the sequence, which reads 0 with no memory
 * access.  (OOB itself lies inside this range: it is past the end of the
 * state slabs only.) */
/* s_getreg operands (id | offset << 6 | (size - 1) << 11): the wave's
 * hardware ids (CU in bits 8-11, SH 12, SE 13-15) and its XCD */
constexpr int HWREG_HW_ID = 4 | (31 << 11);
constexpr int HWREG_XCC_ID = 20 | (31 << 11);
constexpr uint32_t SEQ_RANGE = 0x80008000u;
constexpr uint32_t SEQ_OOB = 0xFFFFFF00u;
static_assert(SEQ_OOB >= SEQ_RANGE, "the sequence pad offset must be out of range");
__device__ __forceinline__ rsrc_t seq_rsrc(const uint8_t *base)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, 0, (int)SEQ_RANGE, RSRC_WORD3);
}
template <bool SYS> __device__ __forceinline__ uint32_t seq_byte(rsrc_t r, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b8(r, (int)off, 0, SYS ? 0x11 : 0); /* 0x11: sc0 sc1 */
}
template <bool SYS, class T> __device__ __forceinline__ void dst(T *p, T v)
{
    if (SYS)
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else
        *p = v;
}

/*
 * Runs group g (or nothing when g < 0) on this half-wave / wave.  The body is
 * the reference's smithWaterman512_{16,8} row loop with its batch wrapper's
 * per-group setup folded in.
 *
 * qb is the byte offset of this lane's first query-code word in the slab
 * (words [word][64 lanes], at the head of the slab).
 */
template <int BITS, bool SYS = false>
__device__ void run_group(const KParams &Kref, const KBuffers &B, int g, StateRef &S, uint32_t qb,
                          int64_t stale_word = -1)
{
    typedef Width<BITS> Wd;
    constexpr int GL = Wd::GL;
    constexpr int SH = Wd::SH;
    const int lane = threadIdx.x;
    const int half = lane / GL;
    const int gl = lane % GL;
    const uint64_t gmask = GL == 64 ? ~0ull : (0xFFFFFFFFull << (32 * half));
    const int16_t ONE = Wd::v(1), TWO = Wd::v(2);

    /* parameters in registers (never indexed through the argument block) */
    const int n_pairs = Kref.n, wband = Kref.w;
    const int o_del = Kref.o_del, e_del = Kref.e_del, o_ins = Kref.o_ins, e_ins = Kref.e_ins;
    const int end_bonus = Kref.end_bonus, smaxv = Kref.smax;
    const int16_t k_match = Kref.match, k_mism = Kref.mism, k_ambig = Kref.ambig;
    const int16_t oe_ins_s = Kref.oe_ins_s, e_ins_s = Kref.e_ins_s;
    const int16_t oe_del_s = Kref.oe_del_s, e_del_s = Kref.e_del_s, zdrop_s = Kref.zdrop_s;
    const int raw_match = (int)k_match >> SH, raw_mism = (int)k_mism >> SH;

    /* ---- per-lane pair (wrapper :2787-2826 / :2092-2120) ---- */
    const bool gvalid = g >= 0;
    int len1 = 0, len2 = 0, h0 = 0, idr = 0, idq = 0;
    bool real = false;
    {
        const int p = gvalid ? g * GL + gl : 0;
        real = gvalid && p < n_pairs;
        if (real) {
            const tal_seqpair *sp = B.pairs + p;
            idr = dld<SYS>(&sp->idr);
            idq = dld<SYS>(&sp->idq);
            len1 = dld<SYS>(&sp->len1);
            len2 = dld<SYS>(&sp->len2);
            h0 = dld<SYS>(&sp->h0);
        } else if (gvalid && BITS == 8) {
            len2 = B.last_len2_dev ? *B.last_len2_dev : Kref.last_len2; /* :2017 */
        }
    }
    /* Lanes sorted by query length, longest first.  Everything the group
     * shares (rows, columns, range, narrowing, scans, exits) is a max, min,
     * any or all over its lanes, so the order of pairs across lanes changes
     * nothing; per-lane bit masks (stale column, dependence, final F) are
     * kept in the reference's lane order through `orig`/`inv`.  Sorted, the
     * lanes inside their band at a column are a prefix of the group, so the
     * state blocks of lanes outside it can be skipped whole. */
    int orig = gl; /* the pair's lane in the reference's group */
    int inv = lane; /* the lane now holding the pair of reference lane (lane % GL) */
    {
        int rank = 0;
        for (int o = 0; o < GL; o++) {
            const int ol = __shfl(len2, half * GL + o, 64);
            rank += (ol > len2 || (ol == len2 && o < gl)) ? 1 : 0;
        }
        const int to = (half * GL + rank) * 4;
        idr = __builtin_amdgcn_ds_permute(to, idr);
        idq = __builtin_amdgcn_ds_permute(to, idq);
        len1 = __builtin_amdgcn_ds_permute(to, len1);
        len2 = __builtin_amdgcn_ds_permute(to, len2);
        h0 = __builtin_amdgcn_ds_permute(to, h0);
        real = __builtin_amdgcn_ds_permute(to, real ? 1 : 0) != 0;
        orig = __builtin_amdgcn_ds_permute(to, gl);
        inv = __builtin_amdgcn_ds_permute((half * GL + orig) * 4, lane);
    }
    const int p = gvalid ? g * GL + orig : 0;
    const int nrow = group_max<GL>(len1);
    const int ncol = group_max<GL>(len2);

    /* per-lane band (:2882-2904): integer division then +1, floor 1, cap w */
    int qlm = BITS == 16 ? (int)(uint16_t)(len2 * smaxv) : (int)(uint8_t)(len2 * smaxv);
    int tb = BITS == 16 ? (int)(uint16_t)(qlm + end_bonus - o_ins)
                        : (int)(uint8_t)(qlm + end_bonus - o_ins);
    int b1 = tb / e_ins + 1;
    b1 = b1 > 1 ? b1 : 1;
    int band = wband < b1 ? wband : b1;
    band = BITS == 16 ? (int)(uint16_t)band : (int)(uint8_t)band;
    tb = BITS == 16 ? (int)(uint16_t)(qlm + end_bonus - o_del)
                    : (int)(uint8_t)(qlm + end_bonus - o_del);
    int b2 = tb / e_del + 1;
    b2 = b2 > 1 ? b2 : 1;
    band = band < b2 ? band : b2;

    const int16_t h0w = Wd::v(h0);
    const int16_t tlen = Wd::v(len1), qlen = Wd::v(len2), bandv = Wd::v(band);
    const int qcol = (int)qlen >> SH; /* the query length as the lane sees it */
    int16_t head = 0, tail = qlen;
    const int16_t mlen = umin16(add16(qlen, bandv), tlen);
    int16_t maxScore = h0w, x = 0, y = 0, gscore = Wd::v(-1), max_off = 0, max_ie = 0;
    bool live = true; /* exit0 */

    const rsrc_t rbuf = seq_rsrc(B.ref), qbuf = seq_rsrc(B.qer);

    S.reset();
    /* ---- H_h row -1 and F (:2867-2879, :3034) ---- */
    /* (stale_word >= 0: the group's mask, given by the caller) */
    const bool stale =
        BITS == 16 && gvalid &&
        (stale_word >= 0 ? (((uint32_t)stale_word >> orig) & 1u) != 0
                         : B.stale_in && ((dld<SYS>(B.stale_in + g) >> orig) & 1u));
    /* nzhi: the highest state column of this lane that may hold a nonzero
     * value; blocks above it and above the band need no store */
    int nzhi = -1;
    {
        /* whole 8-column blocks (columns past ncol get zeros; nothing reads
         * them as values) */
        int16_t hv = h0w;
        for (int b = 0; b <= ncol; b += 8) {
            uint32_t v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int c = b + k;
                if (c == 1)
                    hv = h0w > oe_ins_s ? sub16(h0w, oe_ins_s) : (int16_t)0;
                else if (c >= 2)
                    hv = (c < ncol) ? smax16(sub16(hv, e_ins_s), 0) : (int16_t)0;
                const int16_t f = (c < ncol) ? (int16_t)0 : (int16_t)(stale ? 1 : 0);
                v[k] = c <= ncol ? pack(hv, f) : 0u;
                if (v[k] != 0u) nzhi = c;
            }
            S.st8any(b, v);
        }
    }
    /* ---- query codes for the fast path ---- */
    /* Every int8 score fits the per-row byte table.  The fast path needs
     * query bytes 0-4, nonnegative gap open+extend penalties (a run of zero
     * cells keeps E and F at zero) and nonnegative gap extensions (E and F
     * are nonnegative, so e - e_ins and f - e_del then never wrap and their
     * max with 0 is an unsigned saturating subtraction). */
    bool fast = oe_ins_s >= 0 && oe_del_s >= 0 && e_ins_s >= 0 && e_del_s >= 0;
    {
        bool ok = true;
        const int nwords = (ncol + 1 + 7) / 8 + 1;
        /* 16 byte loads in flight at a time; columns past the query load
         * nothing and become the pad */
        for (int wd = 0; wd < nwords; wd += 2) {
            uint32_t by[16];
#pragma unroll
            for (int k = 0; k < 16; k++) {
                const int c = wd * 8 + k;
                by[k] = seq_byte<SYS>(qbuf, c < len2 ? (uint32_t)idq + (uint32_t)c : SEQ_OOB);
            }
#pragma unroll
            for (int h = 0; h < 2; h++) {
                uint32_t word = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int c = (wd + h) * 8 + k;
                    const bool in = c < len2;
                    ok = ok && (!in || by[8 * h + k] <= 4u);
                    word |= ((in ? by[8 * h + k] : QPAD) & 15u) << (4 * k);
                }
                if (wd + h < nwords) bst32(S.slab, qb + ((uint32_t)(wd + h) << 8), word);
            }
        }
        fast = (ballot(!ok) & gmask) == 0;
    }

    int beg = 0, end = ncol, nbeg = 0, nend = ncol;
    /* end of the previous row's column range [pbeg, pend] (its start pbeg is
     * at most this row's beg, so only pend is kept) and this lane's store
     * band in it; before row 0 no column has been stored (pend = -1) */
    int pend = -1, phd = 0, ptl = -1;
    /* stale stretch [st_lo, st_hi) below the range start and its first/last
     * nonzero column for this lane */
    int st_lo = -1, st_hi = -1, st_first = NONE_COL, st_last = -1;
    bool gdone = !gvalid;
    bool fwritten = false;
    uint64_t dep = 0;
    int s1_next = 0 < len1 ? (int)seq_byte<SYS>(rbuf, (uint32_t)idr) : 99;

    for (int i = 0;; i++) {
        const bool ract = !gdone && i < nrow;
        if (ballot(ract) == 0) break;
        if (!ract) {
            gdone = true;
            continue;
        }
        const int s1 = s1_next; /* ref code of this row; next row's load in flight */
        {
            const bool more = i + 1 < len1;
            const uint32_t b = seq_byte<SYS>(rbuf, more ? (uint32_t)idr + (uint32_t)(i + 1) : SEQ_OOB);
            s1_next = more ? (int)b : 99;
        }
        /* column range (:3059-3063) */
        beg = nbeg;
        end = nend;
        if (beg < i - wband) beg = i - wband;
        if (end > i + wband + 1) end = i + wband + 1;
        if (end > ncol) end = ncol;
        S.advance(beg >> 3); /* the LDS window starts at the range */
        const int16_t i1 = Wd::v(i + 1);
        int16_t e11 = 0, maxRS = 0, y1 = 0;
        int16_t h10 = 0;
        if (beg == 0) { /* H_v[i+1] (:2828-2837, :3065-3067) */
            if (i + 1 < nrow)
                h10 = smax16(Wd::v(h0 - o_del - (i + 1) * e_del), 0);
            else
                h10 = Wd::v(99); /* DUMMY1 left at H_v[maxLen1] (:2821-2825) */
        }
        /* per-lane band head/tail (:3080-3088) */
        bool same;
        {
            const int16_t ph = head, pt = tail;
            head = smax16(head, sub16(Wd::v(i), bandv));
            tail = umin16(tail, add16(i1, bandv));
            tail = umin16(tail, qlen);
            same = (ballot(head == ph && tail == pt) & gmask) == gmask;
        }
        const int hd = (int)head >> SH, tl = (int)tail >> SH;
        if (!same) {
            /* Trim (:3092-3112).  The loop runs while some lane's head is
             * beyond the column, i.e. over [beg, min(end, max head)), and
             * zeroes a column in lanes whose band excludes it (c < head or
             * c >= tail).  Every column of the previous row's range
             * [pbeg, pend] was stored then and is already zero outside that
             * row's [phd, ptl]; beg >= pbeg.  So only columns inside
             * [phd, ptl] or above pend can still be nonzero. */
            const int cstop = min(end, group_max<GL>(hd));
            /* head side and tail side inside the previous band */
            for (int c = max(beg, phd); c < min(min(hd, ptl + 1), cstop); c++) S.st(c, 0u);
            for (int c = max(max(beg, phd), tl); c < min(ptl + 1, cstop); c++) S.st(c, 0u);
            /* columns above the previous row's range */
            for (int c = max(beg, pend + 1); c < cstop; c++)
                if ((c < hd || c >= tl) && !(c >= phd && c <= ptl)) S.st(c, 0u);
        }
        if (i1 > mlen || tail == head || head > tail) live = false; /* :3118-3129 */

        /* Column mask as an unsigned range test on j - head: stores keep
         * column j iff head <= j <= tail (:3150-3157). */
        const uint32_t lim1 = hd <= tl ? (uint32_t)(tl - hd + 1) : 0u;

        /* the column loop (:3134-3196); fz/lz collect the first and last
         * nonzero store of [beg, end) */
        int fz = NONE_COL, lz = -1;
        int j = beg;
        int16_t jj = Wd::v(beg);
        if (fast) {
            /* Per-row score table: one signed byte per query code, codes 0-3
             * in tlo and 4-5 in thi, looked up for four columns at a time by
             * a byte permute. */
            uint32_t tlo = 0, thi = 0;
#pragma unroll
            for (int c = 0; c < 4; c++)
                tlo |= ((uint32_t)table_score<BITS>(s1, c, raw_match, raw_mism) & 255u) << (8 * c);
#pragma unroll
            for (int c = 4; c < 6; c++)
                thi |= ((uint32_t)table_score<BITS>(s1, c, raw_match, raw_mism) & 255u)
                       << (8 * (c - 4));
            /* E and F:  e' = max(m - oe_ins, 0, e - e_ins),
             *           f' = max(m - oe_del, 0, f - e_del),
             * the max with 0 folded into an unsigned saturating subtraction
             * (e, f >= 0). */
            const s2v OEd2 = {oe_del_s, oe_del_s}, OEi2 = {oe_ins_s, oe_ins_s};
            const u2v EDu = {(uint16_t)e_del_s, (uint16_t)e_del_s};
            const s2v OE = {oe_ins_s, oe_del_s};
            const s2v EE = {e_ins_s, e_del_s};
            const s2v ZERO2 = {0, 0};
            uint32_t ez = 0;                       /* low half: E entering the column */
            uint32_t hprev = (uint32_t)(uint16_t)h10; /* low half: H of the previous column */
            /* one column of the edge blocks; sbt: the score in the low half
             * of a dword (sign-extended); only low halves are meaningful in
             * m, h11 and the adds below.  E and F run as one packed pair: ez
             * holds (e, f'), the state dword (H_h, F). */
            auto cell = [&](uint32_t sv, uint32_t sbt) -> uint32_t {
                const uint32_t m = (sv & 0xFFFFu) == 0u ? 0u : sv + sbt;
                /* (e, f) of this column: e from ez, f from the state */
                const uint32_t ef = __builtin_amdgcn_perm(sv, ez, 0x07060100u);
                const s2v mm = __builtin_shufflevector(as_s2(m), as_s2(m), 0, 0);
                const s2v z = max2(max2(mm - OE, as_s2(ef) - EE), ZERO2);
                /* h11 = max(m, e, f) in the low half; the high half is
                 * never read (the permute below takes bytes 0-1 only) */
                uint32_t h11;
                asm("v_max_i16_sdwa %0, %1, %1 dst_sel:WORD_0 dst_unused:UNUSED_PAD "
                    "src0_sel:WORD_0 src1_sel:WORD_1\n\t"
                    "v_max_i16_e32 %0, %2, %0"
                    : "=&v"(h11)
                    : "v"(ef), "v"(m));
                /* the store: H of the previous column, F of this one */
                const uint32_t st = __builtin_amdgcn_perm(as_u32(z), hprev, 0x07060100u);
                ez = as_u32(z);
                hprev = h11;
                return st;
            };
            auto score = [&](uint32_t sc, int byte) -> uint32_t {
                const uint32_t b8 = (sc >> (8 * byte)) & 255u;
                return BITS == 16 ? (uint32_t)(int32_t)(int8_t)b8 : b8 << 8;
            };
            /* Row max and its column as one key, read off the stores: the
             * store at column j holds H of column j-1 (masked to the lane's
             * band [head, tail]), so max over j in (beg, end] of
             * (store_j << 16 | j) is max over columns j-1 of (h11 << 16 |
             * column+1).  H >= 0, so ties keep the latest column -- the
             * reference's rule (maxRS' >= maxRS takes the new column,
             * :3158-3169).  The row max counts column c iff head <= c <=
             * tail-1; the stores add only c = head-1, whose H is 0 (the
             * columns from beg to head-1 were trimmed to zero), and masked
             * zeros.  Zero-score keys never win over a positive row max, and
             * when the row max is 0 its column is not used. */
            uint32_t mkey = 0;
            /* first and last 8-column block holding a nonzero store */
            int bfirst = NONE_COL, blast = -1;
            /* 8-column blocks aligned to multiples of 8 (one query word,
             * static code positions). */
            /* This lane reads a block iff it meets [lo, tl] and stores it iff
             * it meets [lo, max(tl, nzhi)], lo = min(hd, pend + 1).  Below
             * lo the state is zero: the previous row stored [pbeg, pend] with
             * its band mask and the trim zeroed [beg, hd) where head moved, so
             * only column pend + 1 (never stored last row) can still hold an
             * older value there.  Above tl the inputs only feed masked cells,
             * and memory above nzhi already holds the masked zeros. */
            const int lo = min(hd, pend + 1);
            const int sthi = max(tl, nzhi);
            /* the row's blocks: those in the LDS window, then the global ones */
            uint32_t pr[8];
            /* the stores of the last edge block swept (zeros outside [beg,
             * end)); with pr, the row's last block stays in registers */
            uint32_t ev[8];
            int pb = -1;
            bool pneed = false;
            auto sweep = [&](auto gtag, int from, int to) {
                constexpr bool G = decltype(gtag)::value;
                /* one block: compute and store (a middle block's stores are
                 * left pending in pr/pb), fold its row-max key and nonzero
                 * flag into the row's */
                auto do_block = [&](int b, const uint32_t *sv, uint32_t qw, lds_u32 *lp) {
                    /* scores of columns b+0,2,4,6 (sce) and b+1,3,5,7 (sco) */
                    const uint32_t sce = __builtin_amdgcn_perm(thi, tlo, qw & 0x0F0F0F0Fu);
                    const uint32_t sco =
                        __builtin_amdgcn_perm(thi, tlo, (qw >> 4) & 0x0F0F0F0Fu);
                    const uint32_t rel0 = (uint32_t)(b - hd);
                    uint32_t any = 0, kb = 0;
                    if (b > beg && b + 8 <= end) {
                        /* Two columns at a time in packed 16-bit halves:
                         * the diagonal term m and F' of a column depend only
                         * on the state above it, so columns k and k+1 share
                         * every instruction; only E runs column by column.
                         * m = H + s, or 0 when H == 0, as t*s + H with
                         * t = min(H, 1). */
                        /* The band mask.  Inside [beg, end) the state
                         * left of a lane's head is zero, except at column
                         * pend+1 (not stored last row; end <= pend + 2, so it
                         * can only be column end-1), and zero inputs give
                         * zero stores here: E and F stay zero (oe_ins,
                         * oe_del >= 0) and the block is not the row's first,
                         * so no boundary H enters it.  Columns b..b+6 are
                         * thus masked at the tail only, column b+7 at both
                         * ends. */
                        const int tb = tl - b;
                        /* 16-bit: the scores shifted up a byte, so that
                         * every score has an odd byte position in one of
                         * the two sources (the permute's sign selectors
                         * 8-11 replicate the top bit of bytes 1, 3, 5, 7) */
                        const uint32_t sce1 = sce << 8, sco1 = sco << 8;
                        uint32_t e = ez, hp = hprev;
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const int k = 2 * q;
                            /* scores of columns k, k+1 in the two halves:
                             * 16-bit sign-extended (each byte with its sign
                             * byte), 8-bit as value << 8 (each byte above a
                             * zero byte) */
                            s2v S;
                            if (BITS == 16) {
                                const uint32_t sel = q < 2 ? 0x0A050801u : 0x0B070903u;
                                S = as_s2((q & 1) ? __builtin_amdgcn_perm(sco, sce, sel)
                                                  : __builtin_amdgcn_perm(sco1, sce1, sel));
                            } else {
                                const uint32_t sel = 0x000C000Cu | ((uint32_t)q << 8) |
                                                     ((uint32_t)(4 + q) << 24);
                                S = as_s2(__builtin_amdgcn_perm(sco, sce, sel));
                            }
                            const u2v H = as_u2(__builtin_amdgcn_perm(sv[k + 1], sv[k], 0x05040100u));
                            const u2v F = as_u2(__builtin_amdgcn_perm(sv[k + 1], sv[k], 0x07060302u));
                            uint32_t t, mw;
                            asm("v_pk_min_u16 %0, %1, %2" : "=v"(t) : "v"(u2_u32(H)), "v"(0x00010001u));
                            asm("v_pk_mad_u16 %0, %1, %2, %3"
                                : "=v"(mw)
                                : "v"(t), "v"(as_u32(S)), "v"(u2_u32(H)));
                            const s2v m = as_s2(mw);
                            const s2v fn = max2(m - OEd2, as_s2(u2_u32(__builtin_elementwise_sub_sat(F, EDu))));
                            const s2v am = m - OEi2;
                            const uint16_t e0 = (uint16_t)e;
                            const uint16_t e1 = (uint16_t)smax16(am.x, (int16_t)__builtin_elementwise_sub_sat(e0, (uint16_t)e_ins_s));
                            const uint16_t e2 = (uint16_t)smax16(am.y, (int16_t)__builtin_elementwise_sub_sat(e1, (uint16_t)e_ins_s));
                            const s2v E2 = as_s2(__builtin_amdgcn_perm((uint32_t)e1, (uint32_t)e, 0x05040100u));
                            const uint32_t h2 = as_u32(max2(max2(m, as_s2(u2_u32(F))), E2));
                            /* H of column k-1: the low half of hp entering the
                             * block, then the high half of the previous h2 */
                            const uint32_t st0 = __builtin_amdgcn_perm(as_u32(fn), hp, q == 0 ? 0x05040100u : 0x05040302u);
                            const uint32_t st1 = __builtin_amdgcn_perm(as_u32(fn), h2, 0x07060100u);
                            pr[k] = k <= tb ? st0 : 0u;
                            pr[k + 1] = (q < 3 ? k + 1 <= tb : rel0 + 7 < lim1) ? st1 : 0u;
                            hp = h2;
                            e = e2;
                        }
                        ez = e;
                        hprev = hp >> 16;
                        /* row-max key (store << 16 | column) and nonzero
                         * flag of the block, three operands at a time */
                        uint32_t kk[8];
#pragma unroll
                        for (int k = 0; k < 8; k++) kk[k] = (pr[k] << 16) | (uint32_t)k;
                        kb = max(max(max(kk[0], kk[1]), max(kk[2], kk[3])),
                                 max(max(kk[4], kk[5]), max(kk[6], kk[7])));
                        any = (pr[0] | pr[1] | pr[2]) | (pr[3] | pr[4] | pr[5]) | (pr[6] | pr[7]);
                        pb = b;
                        pneed = b <= sthi && b + 7 >= lo;
                    } else {
                        /* The row's first and last block: column by column,
                         * only the columns in [beg, end), each stored at
                         * once.  A lane that skipped loading a global block
                         * has it outside [lo, tl] and stores masked zeros. */
#pragma unroll
                        for (int k = 0; k < 8; k++) {
                            ev[k] = 0u;
                            if (b + k >= beg && b + k < end) {
                                const uint32_t st =
                                    cell(sv[k], score((k & 1) ? sco : sce, k >> 1));
                                const uint32_t r = rel0 + k < lim1 ? st : 0u;
                                if (G)
                                    bst32(S.slab, S.goff(b) + 4u * k, r);
                                else
                                    lp[k] = r;
                                ev[k] = r;
                                any |= r;
                                if (b + k > beg) kb = max(kb, (r << 16) | (uint32_t)k);
                            }
                        }
                    }
                    mkey = max(mkey, kb + (uint32_t)b);
                    if (any != 0u) {
                        bfirst = min(bfirst, b);
                        blast = b;
                    }
                };
                /* a global block's loads, issued for every lane: a lane
                 * reads only a block it may need (zeros elsewhere, and past
                 * the sweep) */
                auto load = [&](int b, uint32_t *v, uint32_t &q) {
                    S.gld8(b, v, b < to && b + 7 >= lo && b <= tl);
                    q = bld32(S.slab, qb + ((uint32_t)b << 5));
                };
                /* LDS part: the blocks from the window start on sit in
                 * consecutive ring slots, so their addresses step by one
                 * slot (a block one past the window reads the window's first
                 * slot, a valid address) */
                lds_u32 *pp = nullptr; /* LDS address of the pending block */
                /* the pending stores, if any (a global block only by the
                 * lanes that need it) */
                auto flush = [&]() {
                    if (G) {
                        S.gst8(pb, pr, pb >= 0 && pneed);
                    } else if (pb >= 0) {
                        lds_u32x4 *p = (lds_u32x4 *)pp;
                        p[0] = (u32x4){pr[0], pr[1], pr[2], pr[3]};
                        p[1] = (u32x4){pr[4], pr[5], pr[6], pr[7]};
                    }
                    pb = -1;
                };
                /* Two buffers used in turn: a block's loads land straight in
                 * the registers its computation reads, so the wait for them
                 * falls just before that computation, a whole block after
                 * they were issued, and never covers later stores. */
                if (from < to && !G) {
                    uint32_t A[8], Bv[8];
                    uint32_t qa = 0, qbv = 0;
                    int sl = S.r0; /* slot of the next block to load (from = the window start) */
                    auto lload = [&](int b, uint32_t *v, uint32_t &q) -> lds_u32 * {
                        lds_u32 *p = S.lds + ((uint32_t)sl << 9);
                        sl = sl + 1 == S.nb ? 0 : sl + 1;
                        const lds_u32x4 *p4 = (const lds_u32x4 *)p;
                        const u32x4 x = p4[0], y = p4[1];
                        v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
                        v[4] = y.x; v[5] = y.y; v[6] = y.z; v[7] = y.w;
                        q = bld32(S.slab, qb + ((uint32_t)b << 5));
                        return p;
                    };
                    lds_u32 *pa = lload(from, A, qa), *pbv;
                    for (int b = from;;) {
                        pbv = lload(b + 8, Bv, qbv);
                        flush();
                        do_block(b, A, qa, pa);
                        pp = pa;
                        b += 8;
                        if (b >= to) break;
                        pa = lload(b + 8, A, qa);
                        flush();
                        do_block(b, Bv, qbv, pbv);
                        pp = pbv;
                        b += 8;
                        if (b >= to) break;
                    }
                } else if (from < to) {
                    /* global blocks: loads two blocks ahead (three buffers),
                     * so that the memory latency is covered by two blocks of
                     * computation */
                    uint32_t A[8], Bv[8], C[8];
                    uint32_t qa = 0, qbv = 0, qc = 0;
                    load(from, A, qa);
                    load(from + 8, Bv, qbv);
                    for (int b = from;;) {
                        load(b + 16, C, qc);
                        flush();
                        do_block(b, A, qa, nullptr);
                        b += 8;
                        if (b >= to) break;
                        load(b + 16, A, qa);
                        flush();
                        do_block(b, Bv, qbv, nullptr);
                        b += 8;
                        if (b >= to) break;
                        load(b + 16, Bv, qbv);
                        flush();
                        do_block(b, C, qc, nullptr);
                        b += 8;
                        if (b >= to) break;
                    }
                }
                flush();
            };
            if (beg < end) {
                const int b0 = beg & ~7;
                const int we = S.wend();
                sweep(std::false_type{}, b0, min(end, we));
                sweep(std::true_type{}, max(b0, we), end);
                j = end;
            }
            h10 = (int16_t)hprev;
            jj = Wd::v(j);
            /* the key of column end-1 comes from the boundary store at end */
            if (beg < end) {
                const bool out1 = (jj > tail) || (head > jj);
                const uint32_t rb = out1 ? 0u : (uint32_t)(uint16_t)h10;
                mkey = max(mkey, (rb << 16) | (uint32_t)end);
            }
            maxRS = (int16_t)(mkey >> 16);
            y1 = Wd::v((int)(mkey & 0xFFFFu));
            /* exact first/last nonzero column of [beg, end) from the two
             * flagged blocks (one 8-column read each) */
            if (bfirst != NONE_COL) {
                uint32_t vf[8], vl[8];
                S.ld8any(bfirst, vf);
                /* the row's last block is still in registers: pr if it was
                 * a middle block, ev if an edge block */
                const int bl = (end - 1) & ~7;
                const bool lmid = bl > beg && bl + 8 <= end;
                if (blast == bl) {
#pragma unroll
                    for (int k = 0; k < 8; k++) vl[k] = lmid ? pr[k] : ev[k];
                } else {
                    S.ld8any(blast, vl);
                }
                uint32_t nzf = 0, nzl = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    nzf |= min(vf[k], 1u) << k;
                    nzl |= min(vl[k], 1u) << k;
                }
                /* columns [beg, end) of each block as a bit range */
                auto range = [&](int bb) -> uint32_t {
                    const int l = min(max(beg - bb, 0), 8), h = min(max(end - bb, 0), 8);
                    return ((1u << h) - 1u) & ~((1u << l) - 1u);
                };
                nzf &= range(bfirst);
                nzl &= range(blast);
                fz = bfirst + __builtin_ctz(nzf);
                lz = blast + 31 - __builtin_clz(nzl);
            }
        } else {
            const int s1b = s1;
            for (; j < end; j++) {
                const int16_t pj = jj;
                jj = add16(jj, ONE);
                const uint32_t sv = S.ld(j);
                const int16_t h00 = (int16_t)(sv & 0xFFFFu), f11 = (int16_t)(sv >> 16);
                const int s2 = j < len2 ? (int)seq_byte<SYS>(qbuf, (uint32_t)idq + (uint32_t)j) : 100;
                bool amb;
                if (BITS == 16)
                    amb = (s1b == 4) | (s2 == 4);
                else
                    amb = (s1b == 4) | (s2 == 4) | (s1b >= 128) | (s2 >= 128);
                const int16_t sbt = amb ? k_ambig : (s1b == s2 ? k_match : k_mism);
                const int16_t m = h00 == 0 ? (int16_t)0 : add16(h00, sbt);
                const int16_t h11 = smax16(smax16(m, e11), f11);
                e11 = smax16(smax16(sub16(m, oe_ins_s), 0), sub16(e11, e_ins_s));
                int16_t f21 = smax16(smax16(sub16(m, oe_del_s), 0), sub16(f11, e_del_s));
                const bool left = head > pj;
                const bool out1 = (pj > tail) || left;
                const int16_t hst = out1 ? (int16_t)0 : h10;
                if (out1) f21 = 0;
                const bool out2 = (jj > tail) || left;
                if (!out2 && h11 >= maxRS) {
                    y1 = jj;
                    maxRS = h11;
                }
                const uint32_t r = pack(hst, f21);
                S.st(j, r);
                h10 = h11;
                if (r != 0u) {
                    fz = min(fz, j);
                    lz = j;
                }
            }
        }
        uint32_t rj; /* the boundary store */
        { /* boundary store at j (end, or beg when the loop was empty) */
            const bool out1 = (jj > tail) || (head > jj);
            rj = pack(out1 ? (int16_t)0 : h10, 0);
            S.st(j, rj);
            if (j == ncol) fwritten = true;
            /* fz: first nonzero in [beg, end) (end if none);
             * lz: last nonzero in [beg, end] (beg - 1 if none) */
            fz = fz == NONE_COL ? end : fz;
            lz = (rj != 0u && j >= beg) ? j : (lz >= beg ? lz : beg - 1);
        }
        /* Loaded now, so that their latency overlaps the reductions below:
         * columns j+1 and j+2 (the head and tail scans look at most two
         * columns past the row's end, es <= end + 2) and the last query
         * column for gscore. */
        const bool gq = live && tail == qlen && qcol - 1 >= beg && qcol - 1 < end;
        const uint32_t vj1 = S.ld(j + 1), vj2 = S.ld(j + 2), vq = S.ld(gq ? qcol : j);
        auto colval = [&](int c) -> uint32_t {
            return c == j ? rj : c == j + 1 ? vj1 : c == j + 2 ? vj2 : S.ld(c);
        };
        pend = j;
        phd = hd;
        ptl = tl;
        nzhi = max(nzhi, min(tl, j)); /* this row's nonzero stores lie in [hd, tl] */
        /* gscore (:3176-3195): the cell of the last query column is updated
         * when it was computed this row, the lane is live and tail == qlen;
         * then H_h[qlen] holds it (head <= qlen, or the lane would not be
         * live, so the store at column qlen was not masked).  Nothing else
         * in the row reads gscore, so it is folded in at the row's end, after
         * the load has had the reductions to arrive. */
        auto fold_gscore = [&]() {
            if (gq) {
                const int16_t hq = (int16_t)(vq & 0xFFFFu);
                if (!(gscore > hq)) max_ie = i1;
                gscore = smax16(gscore, hq);
            }
        };

        /* all lanes zero in this row: the group stops (:3205-3209) */
        if ((ballot(maxRS == 0) & gmask) == gmask) {
            fold_gscore();
            gdone = true;
            continue;
        }
        if (maxRS == 0) live = false;
        {
            const int16_t before = maxScore;
            if (live) maxScore = smax16(maxScore, maxRS);
            if (maxScore > before) { /* :3217-3226 */
                y = y1;
                x = i1;
                max_off = smax16(max_off, abs16(sub16(y1, i1)));
            }
            /* z-drop (ZSCORE, :1868-1881) */
            const int16_t di = sub16(i1, x), dj = sub16(y1, y);
            const int16_t off = di > dj ? sub16(di, dj) : sub16(dj, di);
            const int16_t drop = sub16(sub16(maxScore, maxRS), off);
            if (drop > zdrop_s) live = false;
        }

        /*
         * Narrowing (:3239-3260) and the head/tail scans (:3269-3316) in
         * closed form.  Every column of [beg, end] was stored this row, and a
         * lane's stores are zero outside [head, tail], so the first and last
         * nonzero columns of the lane in that range are found by searching
         * from its own band edges:
         *   fz = first nonzero in [beg, end)  (end if none)
         *   lz = last nonzero in [beg, end]   (beg - 1 if none)
         * The forward narrowing stops at the first column where some lane is
         * nonzero (group min of fz); the backward one at the last (group max
         * of lz).  A head scan from column bs gives a live lane that is zero
         * at bs its first nonzero column in [bs, es) (es if none): the scan
         * only breaks at a column where no live lane is zero, which is past
         * every such lane's zero run.  The tail scan from es likewise gives a
         * live lane that is zero at es the last nonzero column in [bs, es]
         * (bs - 1 if none); other lanes keep tail.
         */
        const int c1 = group_min<GL>(fz);
        if (c1 > beg) nbeg = c1 - 1;
        const int lzmax = group_max<GL>(lz);
        const int lmax = beg <= end ? lzmax : end;
        nend = (lmax + 2 < ncol) ? lmax + 2 : ncol;
        int bs = beg, es = end;
        if (BITS == 16) { /* :3267, 16-bit AVX-512 only */
            bs = nbeg;
            es = nend;
        }
        /* The scans start at bs <= beg when the narrowing left the range
         * start behind (16-bit).  Columns below beg are never stored again
         * (beg only grows), so the first/last nonzero column of the stale
         * stretch [bs, beg) is kept per lane and extended as beg moves. */
        if (bs < beg) {
            if (bs != st_lo) { /* a new stretch */
                st_lo = bs;
                st_hi = bs;
                st_first = NONE_COL;
                st_last = -1;
            }
            for (int c = st_hi; c < beg; c++)
                if (S.ld(c) != 0u) {
                    st_first = min(st_first, c);
                    st_last = c;
                }
            st_hi = beg;
        }
        /* first nonzero >= bs in [bs, es) (es if none) */
        int fzs;
        if (bs < beg && st_first != NONE_COL && st_first < es) {
            fzs = st_first;
        } else if (fz < end && fz >= bs && fz < es) {
            fzs = fz;
        } else {
            int c = max(max(bs, beg), end);
            while (c < es && colval(c) == 0u) c++;
            fzs = c < es ? c : es;
        }
        /* last nonzero <= es in [bs, es] (bs - 1 if none) */
        int lzs;
        {
            int c = es;
            while (c > end && c >= bs && colval(c) == 0u) c--;
            if (c > end && c >= bs)
                lzs = c;
            else if (lz >= beg && lz >= bs && lz <= es)
                lzs = lz;
            else if (bs < beg && st_last >= bs && st_last <= es)
                lzs = st_last;
            else
                lzs = bs - 1;
        }
        if (live && bs < es && fzs > bs) head = Wd::v(fzs);

        /* 16-bit: is the tail scan about to read a stale F[ncol]?  A live
         * lane with H_h[ncol] == 0 then depends on it iff "nonzero at ncol"
         * (tail + 2) and "zero at ncol" (walk down) give different tails. */
        if (BITS == 16 && es == ncol && !fwritten && es >= bs) {
            bool dec = false;
            if (live && (S.ld(ncol) & 0xFFFFu) == 0u) {
                const int16_t tnz = smin16(add16(tail, TWO), qlen);
                /* the last nonzero column of [bs, ncol) ("zero at ncol"):
                 * here end < ncol (F[ncol] not yet stored) and bs <= end.
                 * Above the range only columns up to nzhi can be nonzero;
                 * in [beg, end] it is lz, and below beg the stale stretch's
                 * last nonzero. */
                int c = min(ncol - 1, nzhi);
                while (c > end && S.ld(c) == 0u) c--;
                if (c <= end) {
                    if (lz >= max(beg, bs))
                        c = lz;
                    else if (bs < beg && st_last >= bs)
                        c = st_last;
                    else
                        c = bs - 1;
                }
                const int16_t tz = smin16(add16(Wd::v(c), TWO), qlen);
                dec = tnz != tz;
            }
            /* bit of the reference's lane: read `dec` from the lane holding it */
            const int dec_o = __builtin_amdgcn_ds_bpermute(inv * 4, dec ? 1 : 0);
            dep |= ballot(dec_o != 0) & gmask;
        }

        /* tail scan result */
        {
            const int16_t idx = (live && es >= bs && lzs < es) ? Wd::v(lzs) : tail;
            tail = smin16(add16(idx, TWO), qlen);
        }
        fold_gscore();
    }

    if (real) { /* outputs (:3327-3353) */
        tal_seqpair *sp = B.pairs + p;
        dst<SYS>(&sp->score, Wd::out(maxScore));
        dst<SYS>(&sp->tle, Wd::out(x));
        dst<SYS>(&sp->qle, Wd::out(y));
        dst<SYS>(&sp->max_off, Wd::out(max_off));
        dst<SYS>(&sp->gscore, Wd::out(gscore));
        dst<SYS>(&sp->gtle, Wd::out(max_ie));
    }
    if (BITS == 16 && gvalid && B.dep_out) {
        const uint32_t d = (uint32_t)(dep >> (32 * half));
        const bool wr = fwritten;
        /* (system-scope, written through: waves of the same launch may
         * settle later groups on them) */
        if (gl == 0) {
            dst<true>(B.dep_out + g, d);
            dst<true>(B.wrote_out + g, (uint8_t)(wr ? 1 : 0));
        }
        uint32_t *fm = B.fmask + B.fmask_off[g];
        for (int b = 0; b < ncol; b += 8) {
            uint32_t v[8];
            S.ld8any(b, v);
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int nz_o = __builtin_amdgcn_ds_bpermute(inv * 4, (v[k] >> 16) != 0u ? 1 : 0);
                const uint64_t nz = ballot(nz_o != 0);
                if (gl == 0 && b + k < ncol) dst<true>(fm + b + k, (uint32_t)(nz >> (32 * half)));
            }
        }
    }
}

/* Groups scanned back at most when a finished group is settled in the
 * launch; past it the group is left to the resolution after the launch. */
constexpr int SETTLE_SCAN = 64;

/*
 * Settles the wave's two 16-bit groups right after they ran, inside the
 * launch, when that is already possible -- the stale-column resolution of
 * bsw_resolve for one group, done early.  A group whose run did not depend
 * on its F[ncol] (dep == 0) is final.  Otherwise its true mask comes from
 * the last earlier group that stored column c = ncol: if that group and the
 * groups on the way with ncol == c (whose `wrote` decides) are final, the
 * truth is known; equal to the assumed mask on the dep lanes, the group is
 * final, else it runs again at once with the true mask and is then final.
 * A group that cannot be settled yet stays pending, for the resolution and
 * rerun rounds after the launch, which read the same records (a settled
 * group's stale mask holds its truth, so they leave it alone).  Final
 * groups are flagged in gstate after all their system-scope stores; the
 * flags and the records they guard are read with system-scope loads.
 * Returns the half-wave's verdict: 0 final, 1 run again with `truth`,
 * 2 pending.
 */
__device__ int settle16(const KBuffers &B, int g, uint32_t &truth)
{
    const int lane = threadIdx.x, half = lane >> 5, gl = lane & 31;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    int verdict = 0; /* 0 final, 1 run again with `truth`, 2 pending */
    uint32_t tv = 0;
    const uint32_t d = gl == 0 && g >= 0 ? dld<true>(B.dep_out + g) : 0u;
    /* first the groups that did not depend on F[ncol]: final now, and
     * flagged before the other half looks (it may be its predecessor) */
    if (gl == 0 && g >= 0 && d == 0u) dst<true>(B.gstate + g, 1);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (gl == 0 && g >= 0) {
        if (d != 0u) {
            verdict = 2;
            const int c = B.shapes[g].ncol;
            uint32_t tr = B.image[c];
            bool known = true;
            for (int h = g - 1, steps = 0; h >= 0; h--, steps++) {
                if (steps == SETTLE_SCAN) {
                    known = false;
                    break;
                }
                const int nh = B.shapes[h].ncol;
                if (nh < c) continue;
                if (dld<true>(B.gstate + h) == 0) {
                    known = false;
                    break;
                }
                if (nh > c) {
                    tr = dld<true>(B.fmask + B.fmask_off[h] + c);
                    break;
                }
                if (dld<true>(B.wrote_out + h) != 0) {
                    tr = 0u;
                    break;
                }
            }
            if (known) {
                if ((tr & d) != (dld<true>(B.stale_in + g) & d)) {
                    verdict = 1;
                    tv = tr;
                } else {
                    verdict = 0;
                }
            }
        }
    }
    verdict = __shfl(verdict, half << 5, 64);
    truth = (uint32_t)__shfl((int)tv, half << 5, 64);
    return verdict;
}

/* Persistent grid: each 64-thread block is one wave that pulls tasks. */
template <int BITS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void bsw_main(KParams K, KBuffers B)
{
    extern __shared__ uint32_t lds_state[];
    typedef Width<BITS> Wd;
    const int lane = threadIdx.x;
    const int half = lane / Wd::GL;
    /* LDS: a window of lds_cols state columns.  The block's global slab
     * holds the query-code words ([word][64 lanes]) and the home of every
     * state block. */
    uint32_t *slab = B.slabs + (size_t)blockIdx.x * K.slab_cols * 64;
    StateRef S;
    S.lds = (lds_u32 *)lds_state + lane * 8;
    S.slab = __builtin_amdgcn_make_buffer_rsrc(slab, 0, K.slab_cols * 64 * (int)sizeof(uint32_t),
                                               RSRC_WORD3);
    S.gbase = ((uint32_t)K.query_words * 64u + (uint32_t)lane * 8u) * 4u;
    S.nb = K.lds_cols >> 3;
    S.nblk = (K.state_cols + 7) >> 3;
    const uint32_t qb = (uint32_t)lane * 4u;
    if (B.abort_flag && *B.abort_flag != 0) return;
    const int ntasks = B.ntasks_dev ? *B.ntasks_dev : K.ntasks;
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(B.task_counter, 1);
        t = __shfl(t, 0, 64);
        int g;
        bool run = true; /* false: a pending group, settled without running first */
        if (t < ntasks) {
            if (B.task_groups)
                g = B.task_groups[t * Wd::GPW + half];
            else {
                g = t * Wd::GPW + half;
                if (g >= K.ngroups) g = -1;
            }
        } else {
            /* Out of tasks: take a group left pending by some wave, in the
             * first half-wave, and try again now that more groups are final
             * (one pass; what stays pending goes to the resolution after the
             * launch, as do groups listed after this wave has left). */
            if (BITS != 16 || !B.pending) break;
            int k = 0;
            if (lane == 0) k = atomicAdd(B.pending + 1, 1);
            k = __shfl(k, 0, 64);
            if (k >= dld<true>(B.pending)) break;
            /* (a slot whose group is not visible yet still reads -1) */
            const int pg = dld<true>(B.pending + 2 + k);
            if (pg < 0 || pg >= K.ngroups) continue;
            g = half == 0 ? pg : -1;
            run = false;
        }
        /* One call site of run_group (a second would be a second inlined
         * copy): a group settled as "run again" comes round once more.  The
         * pass number and `run` are the same in every lane. */
        int gg = g;
        int64_t sw = -1;
        bool do_run = run;
        for (int pass = 0;; pass++) {
            if (do_run) run_group<BITS>(K, B, gg, S, qb, sw);
            if (BITS != 16 || !B.gstate) break;
            const int gl = lane & 31;
            if (pass == 1) { /* the reruns, with their true masks, are final */
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                if (gl == 0 && gg >= 0) dst<true>(B.gstate + gg, 1);
                break;
            }
            uint32_t truth;
            const int verdict = settle16(B, g, truth);
            if (gl == 0 && g >= 0 && verdict == 0) dst<true>(B.gstate + g, 1);
            if (run && B.pending && gl == 0 && g >= 0 && verdict == 2)
                dst<true>(B.pending + 2 + atomicAdd(B.pending, 1), g);
            if (ballot(verdict == 1) == 0) break;
            if (gl == 0 && verdict == 1) {
                dst<true>(B.stale_in + g, truth);
                if (B.early) atomicAdd(B.early, 1);
            }
            gg = verdict == 1 ? g : -1;
            sw = verdict == 1 ? (int64_t)truth : -1;
            do_run = true;
        }
    }
}

/* ------------------------------------------------ device-resident serving */

/* A value the host wrote into pinned memory: a system-scope load, which
 * always reads memory (never a cached copy). */
template <class T> __device__ __forceinline__ T hload(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T> __device__ __forceinline__ T *hload_ptr(T *const *p)
{
    return (T *)(uintptr_t)hload((const uint64_t *)p);
}

/*
 * The persistent serving kernel.  Each 64-thread block is one wave that
 * claims tasks [head, tail) of the host's ring in order, runs them and marks
 * them done.  It exits when the host sets stop or after idle_ticks of the
 * 100 MHz real-time clock in which no wave claimed or finished a task; the
 * host relaunches it when work arrives while it is gone, clearing `closing`
 * first.
 *
 * Claims are tickets: a wave that sees head below tail_seen (the device's
 * copy of the host's tail) takes the next ticket with a fetch-and-add on
 * head, which always succeeds (a compare-and-swap loop would make each claim
 * wait out a round of failed attempts by every idle wave).  A ticket past
 * the tail is held until its task is published; a wave that leaves while
 * holding one hands it back (the minimum goes to `handback`, from which the
 * next launch resumes, skipping tasks whose done flag is set).  So no task
 * is run twice or lost.  The blocks below
 * `pollers` are watchers: the only waves that read host memory while idle,
 * they copy the tail and the stop flag into tail_seen and stop_seen and run
 * no tasks.  Idle workers watch those device words, sleeping longer on each
 * idle turn (up to about 55 us), so that idle waves load neither the bus nor
 * the busy waves' memory.
 *
 * The kernel's own words (dev) are changed with system-scope atomics, which
 * pass every cache.  Idle waves poll them with agent-scope loads, served by
 * their XCD's L2 (so that polling puts no load on memory), and every few
 * turns with a system-scope load, in case that copy is stale.
 *
 * Coherence.  The host writes a task's descriptor, its request's pinned
 * records and (by DMA) its sequence bytes before it raises tail.  Every read
 * of these, and of the bookkeeping that earlier tasks of the request wrote,
 * is a system-scope load, and every result a system-scope store: these pass
 * the caches, so no wave ever reads a line another agent has changed since,
 * and nothing invalidates or writes back the caches as a whole (they hold
 * the other waves' state slabs).  The done flag is written once the task's
 * stores are complete.
 */
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void bsw_serve(
    KParams K16, KParams K8, const ServeTask *ring, int32_t ring_mask, const ServeReq *reqs,
    const ServeCtl *ctl, int32_t *done, int64_t *dev, uint32_t *slabs, int64_t idle_ticks,
    int32_t pollers)
{
    int64_t *head = dev, *tail_seen = dev + 1, *stop_seen = dev + 2, *alive = dev + 3;
    int64_t *last_active = dev + 4, *closing = dev + 5, *handback = dev + 6;
    int32_t *cu_rank = (int32_t *)(dev + SERVE_DEV_WORDS);
    extern __shared__ uint32_t lds_state[];
    const int lane = threadIdx.x;
    uint32_t *slab = slabs + (size_t)blockIdx.x * K16.slab_cols * 64;
    StateRef S;
    S.lds = (lds_u32 *)lds_state + lane * 8;
    S.slab = __builtin_amdgcn_make_buffer_rsrc(slab, 0, K16.slab_cols * 64 * (int)sizeof(uint32_t),
                                               RSRC_WORD3);
    S.gbase = ((uint32_t)K16.query_words * 64u + (uint32_t)lane * 8u) * 4u;
    S.nb = K16.lds_cols >> 3;
    S.nblk = (K16.state_cols + 7) >> 3;
    const uint32_t qb = (uint32_t)lane * 4u;
    if ((int)blockIdx.x < pollers) {
        /* A watcher: copies the host's tail and stop flag into the device
         * words the workers watch, and runs no tasks, so that it is never
         * busy when work arrives.  It stays while any worker stays (workers
         * leave after idle_ticks without a task). */
        if (lane == 0) {
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            for (;;) {
                const int64_t tl = hload(&ctl->tail);
                if (tl > __hip_atomic_load(tail_seen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    __hip_atomic_fetch_max(tail_seen, tl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                if (hload(&ctl->stop) != 0) {
                    __hip_atomic_store(stop_seen, (int64_t)1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_SYSTEM);
                    break;
                }
                /* (after a start-up grace period, in which workers count in) */
                if (__hip_atomic_load(alive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) == 0 &&
                    (int64_t)(__builtin_amdgcn_s_memrealtime() - t0) > idle_ticks)
                    break;
                __builtin_amdgcn_s_sleep(127);
            }
        }
        return;
    }
    if (lane == 0) __hip_atomic_fetch_add(alive, (int64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const int64_t started = (int64_t)__builtin_amdgcn_s_memrealtime();
    /* This wave's rank among the workers on its CU (by start order): a
     * rank-r wave that sees work waits about r * 7 us before it draws a
     * ticket, so that a small request spreads over the CUs before it doubles
     * up on one. */
    int rank = 0;
    if (lane == 0) {
        const uint32_t hw = __builtin_amdgcn_s_getreg(HWREG_HW_ID);
        const uint32_t xcc = __builtin_amdgcn_s_getreg(HWREG_XCC_ID);
        const uint32_t key = ((xcc & 7u) << 8) | ((hw >> 8) & 0xFFu); /* XCD, SE, SH, CU */
        rank = __hip_atomic_fetch_add(cu_rank + key, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    for (;;) {
        int64_t t = -1; /* the claimed task; -1: leave */
        if (lane == 0) {
            /* (relaxed throughout: what a task reads is read with
             * system-scope loads, so no ordering is needed here, and
             * acquire/release would invalidate or write back the caches
             * that hold the busy waves' slabs) */
            int64_t my = -1; /* a ticket this wave holds */
            int naps = 1;    /* idle turns sleep longer and longer */
            int turn = 0;
            bool primed = false; /* work seen; the rank's wait is done */
            for (;;) {
                /* polls read the XCD's L2; every 8th turn reads memory, so
                 * a stale line there delays a wave by at most a few turns */
                const bool fresh = (++turn & 7) == 0;
                auto ld = [&](const int64_t *p) {
                    return fresh ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)
                                 : __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                };
                const int64_t ts = ld(tail_seen);
                if (my >= 0 && my < ts) { /* the ticket's task is published */
                    t = my;
                    break;
                }
                const bool leave = ld(closing) != 0 || ld(stop_seen) != 0;
                if (!leave && my < 0 && ld(head) < ts) {
                    if (rank > 0 && !primed) {
                        primed = true;
                        for (int k = 0; k < 2 * min(rank, 16); k++) __builtin_amdgcn_s_sleep(127);
                        continue;
                    }
                    my = __hip_atomic_fetch_add(head, (int64_t)1, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_SYSTEM);
                    naps = 1;
                    continue;
                }
                primed = false;
                /* The grid leaves as a whole once no wave has claimed or
                 * finished a task for idle_ticks: a wave that sees that
                 * raises `closing` and every wave leaves at its next idle
                 * turn (a wave that leaves alone would leave the kernel
                 * running with fewer waves, and the host relaunches only a
                 * kernel that has ended). */
                bool go = leave;
                if (!go && (int64_t)__builtin_amdgcn_s_memrealtime() - max(ld(last_active), started) >
                               idle_ticks) {
                    __hip_atomic_store(closing, (int64_t)1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_SYSTEM);
                    go = true;
                }
                if (go) {
                    /* a ticket past the tail goes back: the next launch
                     * starts its claims there (bsw_serve_reset) */
                    if (my >= 0)
                        __hip_atomic_fetch_min(handback, my, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_SYSTEM);
                    break;
                }
                /* (a ticket holder is next in line: it naps briefly) */
                for (int k = 0; k < naps; k++) __builtin_amdgcn_s_sleep(127);
                naps = min(naps + 1, my >= 0 ? 4 : 16);
            }
        }
        {
            const int lo = __shfl((int)(uint32_t)t, 0, 64), hi = __shfl((int)(t >> 32), 0, 64);
            t = (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
        }
        if (t < 0) {
            if (lane == 0)
                __hip_atomic_fetch_add(alive, (int64_t)-1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            break;
        }
        /* a task below a hand-back may have run already (by its ticket's
         * holder, after the hand-back of an earlier ticket) */
        if (hload(done + (t & ring_mask)) == (int32_t)(t + 1)) continue;
        if (lane == 0)
            __hip_atomic_fetch_max(last_active, (int64_t)__builtin_amdgcn_s_memrealtime(),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        const int slot = (int)(t & ring_mask);
        const ServeTask *tk = ring + slot;
        const int bits = hload(&tk->bits), rq = hload(&tk->req);
        const int g0 = hload(&tk->g0), g1 = hload(&tk->g1);
        const ServeReq *R = reqs + rq;
        const int32_t n16 = hload(&R->n16), n8 = hload(&R->n8), w = hload(&R->w);
        if (bits == SERVE_COPY_IN || bits == SERVE_COPY_OUT) {
            /* One chunk of a request's copy, in 16-byte pieces, 8 KiB of
             * loads in flight per turn; system-scope on both sides (the
             * pinned source was written by the host, device stores are read
             * by other waves' system-scope loads).  Sizes and addresses are
             * multiples of 256 bytes, so no piece crosses the range. */
            const bool in = bits == SERVE_COPY_IN;
            const uint8_t *src = hload_ptr(in ? &R->cin_src : &R->cout_src);
            uint8_t *dstp = hload_ptr(in ? &R->cin_dst : &R->cout_dst);
            const int64_t total = hload(in ? &R->cin_bytes : &R->cout_bytes);
            const int64_t lo = (int64_t)g0 * SERVE_COPY_CHUNK;
            const int n = (int)min(total - lo, SERVE_COPY_CHUNK);
            if (n > 0) {
                const rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void *)(src + lo), 0, n,
                                                                    RSRC_WORD3);
                const rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(dstp + lo, 0, n, RSRC_WORD3);
                for (int base = 0; base < n; base += 8192) {
                    u32x4 v[8];
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        v[k] = __builtin_amdgcn_raw_buffer_load_b128(
                            rs, base + k * 1024 + lane * 16, 0, 0x11);
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        __builtin_amdgcn_raw_buffer_store_b128(v[k], rd, base + k * 1024 + lane * 16,
                                                               0, 0x11);
                }
            }
        } else if (bits == 0) {
            /* Stale-column resolution of the request's 16-bit groups from the
             * zero image of a fresh object (bsw_resolve with one segment):
             * lanes take groups, the groups to rerun go to the rerun list. */
            const uint32_t *dep = hload_ptr(&R->dep), *fmask = hload_ptr(&R->fmask);
            const uint8_t *wrote = hload_ptr(&R->wrote);
            uint32_t *stale = hload_ptr(&R->stale);
            const int64_t *foff = hload_ptr(&R->fmask_off);
            const int32_t *ncol = hload_ptr(&R->ncol16);
            int32_t *reruns = hload_ptr(&R->reruns);
            const int G = n16 / 32;
            int count = 0;
            for (int base = 0; base < G; base += 64) {
                const int g = base + lane;
                bool rr = false;
                uint32_t truth = 0;
                if (g < G) {
                    const uint32_t d = hload(dep + g);
                    if (d != 0u) {
                        const int c = hload(ncol + g);
                        for (int h = g - 1; h >= 0; h--) {
                            const int nh = hload(ncol + h);
                            if (nh > c) {
                                truth = hload(fmask + hload(foff + h) + c);
                                break;
                            }
                            if (nh == c && hload(wrote + h)) break;
                        }
                        rr = (truth & d) != (hload(stale + g) & d);
                    }
                }
                if (rr) dst<true>(stale + g, truth);
                const uint64_t m = ballot(rr);
                if (rr) dst<true>(reruns + 1 + count + __popcll(m & ((1ull << lane) - 1)), g);
                count += __popcll(m);
            }
            if (lane == 0) dst<true>(reruns, count);
        } else if (bits == 16) {
            KParams K = K16;
            K.n = n16;
            K.ngroups = n16 / 32;
            K.w = w;
            KBuffers B{};
            B.pairs = hload_ptr(&R->pairs16);
            B.ref = B.qer = hload_ptr(&R->seq);
            B.stale_in = hload_ptr(&R->stale);
            B.dep_out = hload_ptr(&R->dep);
            B.wrote_out = hload_ptr(&R->wrote);
            B.fmask = hload_ptr(&R->fmask);
            B.fmask_off = hload_ptr(&R->fmask_off);
            run_group<16, true>(K, B, lane < 32 ? g0 : g1, S, qb);
        } else {
            KParams K = K8;
            K.n = n8;
            K.ngroups = n8 / 64;
            K.w = w;
            KBuffers B{};
            B.pairs = hload_ptr(&R->pairs8);
            B.ref = B.qer = hload_ptr(&R->seq);
            run_group<8, true>(K, B, g0, S, qb);
        }
        /* every result store of the task (all system-scope, past the
         * caches) is complete before the done flag is written */
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (lane == 0) {
            __hip_atomic_store(done + slot, (int32_t)(t + 1), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_fetch_max(last_active, (int64_t)__builtin_amdgcn_s_memrealtime(),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

/* Before a launch of bsw_serve: claims resume from the lowest handed-back
 * ticket, and closing is cleared. */
__global__ __launch_bounds__(256) void bsw_serve_reset(int64_t *dev)
{
    int32_t *cu_rank = (int32_t *)(dev + SERVE_DEV_WORDS);
    for (int k = threadIdx.x; k < SERVE_CU_KEYS; k += blockDim.x) cu_rank[k] = 0;
    if (threadIdx.x == 0) {
        if (dev[6] < dev[0]) dev[0] = dev[6];
        dev[5] = 0;
        dev[6] = INT64_MAX;
    }
}

void serve_dev_init(int64_t words[SERVE_DEV_WORDS])
{
    for (int k = 0; k < SERVE_DEV_WORDS; k++) words[k] = 0;
    words[6] = INT64_MAX;
}

hipError_t launch_serve(int grid, size_t lds_bytes, const KParams &K16, const KParams &K8,
                        const ServeTask *ring, int32_t ring_mask, const ServeReq *reqs,
                        const ServeCtl *ctl, int32_t *done, int64_t *dev, uint32_t *slabs,
                        int64_t idle_ticks, int32_t pollers, hipStream_t s)
{
    hipLaunchKernelGGL(bsw_serve_reset, dim3(1), dim3(256), 0, s, dev);
    hipLaunchKernelGGL(bsw_serve, dim3(grid), dim3(64), lds_bytes, s, K16, K8, ring, ring_mask,
                       reqs, ctl, done, dev, slabs, idle_ticks, pollers);
    return hipGetLastError();
}

hipError_t set_serve_lds_limit(size_t bytes)
{
    return hipFuncSetAttribute((const void *)bsw_serve, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

/* Group shapes: rows/columns of each group and the 8-bit padding length. */
template <int GL>
__global__ void bsw_shapes(const tal_seqpair *pairs, int32_t n, int32_t ngroups,
                           int32_t pad_len2, const int32_t *pad_len2_dev, GroupShape *out,
                           int64_t *ncol_plus1)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    if (pad_len2_dev) pad_len2 = *pad_len2_dev;
    int nr = 0, nc = 0;
    for (int l = 0; l < GL; l++) {
        const int p = g * GL + l;
        int l1 = 0, l2 = pad_len2;
        if (p < n) {
            l1 = pairs[p].len1;
            l2 = pairs[p].len2;
            /* a negative length is reported as an oversized group */
            if (l1 < 0 || l2 < 0) l1 = l2 = 0x7FFFFFFF;
        }
        nr = max(nr, l1);
        nc = max(nc, l2);
    }
    out[g].nrow = nr;
    out[g].ncol = nc;
    if (ncol_plus1) ncol_plus1[g] = nc + 1;
}

/* Stale-column resolution for the 16-bit path.  For each group that depended
 * on F[ncol], find the last earlier group that stored that column (a group h
 * stores columns [0, ncol_h) at its start and column ncol_h iff wrote_h), and
 * compare its final F zero-mask there with the mask the group assumed. */
__global__ void bsw_resolve(int32_t ngroups, const GroupShape *shape, const uint32_t *dep,
                            const uint8_t *wrote, const uint32_t *fmask,
                            const int64_t *fmask_off, const uint32_t *image,
                            const int32_t *seg_first, uint32_t *stale, int32_t *rerun_list,
                            int32_t *rerun_count, const int32_t *abort_flag)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups || (abort_flag && *abort_flag != 0)) return;
    const uint32_t d = dep[g];
    if (d == 0u) return;
    const int c = shape[g].ncol;
    /* With segments (independent calls batched together), the search stops
     * at the segment's first group and the image is the zero image of a
     * fresh object. */
    const int lo = seg_first ? seg_first[g] : 0;
    uint32_t truth = image[c];
    for (int h = g - 1; h >= lo; h--) {
        const int nh = shape[h].ncol;
        if (nh > c) {
            truth = fmask[fmask_off[h] + c];
            break;
        }
        if (nh == c && wrote[h]) {
            truth = 0u;
            break;
        }
    }
    if ((truth & d) != (stale[g] & d)) {
        stale[g] = truth;
        const int k = atomicAdd(rerun_count, 1);
        rerun_list[k] = g;
    }
}

/* Highest column each group stores: ncol (if wrote) or ncol-1. */
__global__ void bsw_last_writer_scatter(int32_t ngroups, const GroupShape *shape,
                                        const uint8_t *wrote, int32_t *last, const int32_t *skip)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups || (skip && *skip != 0)) return;
    const int top = wrote[g] ? shape[g].ncol : shape[g].ncol - 1;
    if (top >= 0) atomicMax(last + top, g);
}

/* One block: suffix max of `last` over columns, then fold each column's last
 * writer into the handle's image of the F scratch zero-masks. */
__global__ __launch_bounds__(1024) void bsw_image_update(int32_t ncols, const GroupShape *shape,
                                                         const int32_t *last, const uint32_t *fmask,
                                                         const int64_t *fmask_off, uint32_t *image,
                                                         const int32_t *skip)
{
    if (skip && *skip != 0) return; /* reruns still pending: the caller updates later */
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = -1;
    __syncthreads();
    /* process columns from the top down in chunks of blockDim */
    for (int top = ncols - 1; top >= 0; top -= blockDim.x) {
        const int c = top - (int)threadIdx.x;
        int v = c >= 0 ? last[c] : -1;
        /* inclusive suffix max within the chunk via shared memory scan */
        __shared__ int32_t buf[1024];
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {
            int t = threadIdx.x >= (unsigned)o ? buf[threadIdx.x - o] : -1;
            __syncthreads();
            buf[threadIdx.x] = max(buf[threadIdx.x], t);
            __syncthreads();
        }
        const int lw = max(buf[threadIdx.x], carry);
        if (c >= 0 && lw >= 0) {
            const int nh = shape[lw].ncol;
            image[c] = nh > c ? fmask[fmask_off[lw] + c] : 0u;
        }
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1) carry = max(carry, buf[threadIdx.x]);
        __syncthreads();
    }
}

/* 8/16-bit split of the dispatcher: flag[k] = 1 for 8-bit pairs and
 * flag[n+k] = 1 for 16-bit pairs. */
__global__ void bsw_flag8(const tal_seqpair *pairs, int32_t n, int32_t match_pos, uint8_t *flag)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const tal_seqpair sp = pairs[k];
    const int mn = sp.len1 < sp.len2 ? sp.len1 : sp.len2;
    const long long bound = (long long)sp.h0 + (long long)match_pos * mn;
    const uint8_t f = (sp.len1 >= 0 && sp.len2 >= 0 && sp.len1 <= MAX_LEN8 &&
                       sp.len2 <= MAX_LEN8 && sp.h0 >= 0 && bound <= 127)
                          ? 1
                          : 0;
    flag[k] = f;         /* 8-bit path */
    flag[n + k] = 1 - f; /* 16-bit path */
}

/* Gather pairs by index list, and scatter outputs back. */
__global__ void bsw_gather(const tal_seqpair *src, const int32_t *idx, int32_t m, tal_seqpair *dst)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) dst[k] = src[idx[k]];
}

__global__ void bsw_scatter_out(const tal_seqpair *src, const int32_t *idx, int32_t m,
                                tal_seqpair *dst)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const tal_seqpair s = src[k];
    tal_seqpair *d = dst + idx[k];
    d->score = s.score;
    d->tle = s.tle;
    d->gtle = s.gtle;
    d->qle = s.qle;
    d->gscore = s.gscore;
    d->max_off = s.max_off;
}

/* ------------------------------------------- graph-replayed batch helpers */

/* Stable 8/16-bit partition without host reads, for captured graphs:
 * part_count counts the 8-bit pairs of each 256-pair block, part_offsets
 * turns the counts into exclusive offsets (one block) and checks the total
 * against the planned split, part_scatter writes each pair's index to its
 * side in input order. */
__device__ __forceinline__ bool flag8_of(const tal_seqpair &p, int match_pos)
{
    const int mn = min(p.len1, p.len2);
    return p.len1 >= 0 && p.len2 >= 0 && p.len1 <= MAX_LEN8 && p.len2 <= MAX_LEN8 &&
           p.h0 >= 0 && (long long)p.h0 + (long long)match_pos * mn <= 127;
}

__global__ __launch_bounds__(256) void part_count(const tal_seqpair *pairs, int32_t n,
                                                  int32_t match_pos, int32_t *blk)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    const bool f = k < n && flag8_of(pairs[k], match_pos);
    const int c = __popcll(__ballot(f));
    __shared__ int32_t wsum[4];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(1024) void part_offsets(int32_t *blk, int32_t nblk, int32_t planned8,
                                                     int32_t *err)
{
    __shared__ int32_t buf[1024];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nblk; base += 1024) {
        const int k = base + (int)threadIdx.x;
        const int v = k < nblk ? blk[k] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const int t = threadIdx.x >= (unsigned)o ? buf[threadIdx.x - o] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (k < nblk) blk[k] = carry + buf[threadIdx.x] - v; /* exclusive */
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0 && carry != planned8) atomicOr(err, 1);
}

__global__ __launch_bounds__(256) void part_scatter(const tal_seqpair *pairs, int32_t n,
                                                    int32_t match_pos, const int32_t *blk,
                                                    int32_t n8, int32_t *sel8, int32_t *sel16)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    const bool f = k < n && flag8_of(pairs[k], match_pos);
    const uint64_t bal = __ballot(f);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __shared__ int32_t wsum[4];
    if (lane == 0) wsum[wv] = __popcll(bal);
    __syncthreads();
    int before = 0;
    for (int t = 0; t < wv; t++) before += wsum[t];
    const int r8 = before + __popcll(bal & ((1ull << lane) - 1ull)); /* 8-bit rank in block */
    const int o8 = blk[blockIdx.x];
    if (k >= n) return;
    if (f) {
        if (o8 + r8 < n8) sel8[o8 + r8] = k;
    } else {
        const int o16 = blockIdx.x * 256 - o8;
        const int r16 = (int)threadIdx.x - r8;
        if (o16 + r16 < n - n8) sel16[o16 + r16] = k;
    }
}

/* Bounds of a planned 16-bit batch: every group within the planned columns
 * and the length limit. */
__global__ void plan_check16(const GroupShape *shape, int32_t ngroups, int32_t max_ncol,
                             int32_t *err)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    if (shape[g].ncol > max_ncol || shape[g].nrow > MAX_LEN16) atomicOr(err, 1);
}

/* The 8-bit padding length: len2 of the last 8-bit pair (bounded). */
__global__ void plan_pad8(const tal_seqpair *pairs8, int32_t n8, int32_t *out, int32_t *err)
{
    const int v = pairs8[n8 - 1].len2;
    *out = v;
    if (v < 0 || v > MAX_LEN8) atomicOr(err, 1);
}

__global__ void copy_i32(const int32_t *src, int32_t *dst) { *dst = *src; }

/* ---------------------------------------------------------------- launchers */

hipError_t launch_main(int bits, bool lds, int grid, size_t lds_bytes, const KParams &K,
                       const KBuffers &B, hipStream_t s)
{
    (void)lds;
    if (bits == 16)
        hipLaunchKernelGGL((bsw_main<16>), dim3(grid), dim3(64), lds_bytes, s, K, B);
    else
        hipLaunchKernelGGL((bsw_main<8>), dim3(grid), dim3(64), lds_bytes, s, K, B);
    return hipGetLastError();
}

hipError_t set_main_lds_limit(size_t bytes)
{
    hipError_t e = hipFuncSetAttribute((const void *)bsw_main<16>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void *)bsw_main<8>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t launch_shapes(int bits, const tal_seqpair *pairs, int32_t n, int32_t ngroups,
                         int32_t pad_len2, GroupShape *out, int64_t *ncol_plus1, hipStream_t s,
                         const int32_t *pad_len2_dev)
{
    const int tpb = 256, grid = (ngroups + tpb - 1) / tpb;
    if (grid == 0) return hipSuccess;
    if (bits == 16)
        hipLaunchKernelGGL(bsw_shapes<32>, dim3(grid), dim3(tpb), 0, s, pairs, n, ngroups, 0,
                           (const int32_t *)nullptr, out, ncol_plus1);
    else
        hipLaunchKernelGGL(bsw_shapes<64>, dim3(grid), dim3(tpb), 0, s, pairs, n, ngroups,
                           pad_len2, pad_len2_dev, out, ncol_plus1);
    return hipGetLastError();
}

hipError_t launch_resolve(int32_t ngroups, const GroupShape *shape, const uint32_t *dep,
                          const uint8_t *wrote, const uint32_t *fmask, const int64_t *fmask_off,
                          const uint32_t *image, const int32_t *seg_first, uint32_t *stale,
                          int32_t *rerun_list, int32_t *rerun_count, hipStream_t s,
                          const int32_t *abort_flag)
{
    const int tpb = 256, grid = (ngroups + tpb - 1) / tpb;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(bsw_resolve, dim3(grid), dim3(tpb), 0, s, ngroups, shape, dep, wrote,
                       fmask, fmask_off, image, seg_first, stale, rerun_list, rerun_count,
                       abort_flag);
    return hipGetLastError();
}

hipError_t launch_image_update(int32_t ngroups, int32_t ncols, const GroupShape *shape,
                               const uint8_t *wrote, int32_t *last, const uint32_t *fmask,
                               const int64_t *fmask_off, uint32_t *image, hipStream_t s,
                               const int32_t *skip)
{
    const int tpb = 256, grid = (ngroups + tpb - 1) / tpb;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(bsw_last_writer_scatter, dim3(grid), dim3(tpb), 0, s, ngroups, shape,
                       wrote, last, skip);
    hipLaunchKernelGGL(bsw_image_update, dim3(1), dim3(1024), 0, s, ncols, shape, last, fmask,
                       fmask_off, image, skip);
    return hipGetLastError();
}

hipError_t launch_partition(const tal_seqpair *pairs, int32_t n, int32_t match_pos,
                            int32_t planned8, int32_t *blk, int32_t *sel8, int32_t *sel16,
                            int32_t *err, hipStream_t s)
{
    const int nblk = (n + 255) / 256;
    if (nblk == 0) return hipSuccess;
    hipLaunchKernelGGL(part_count, dim3(nblk), dim3(256), 0, s, pairs, n, match_pos, blk);
    hipLaunchKernelGGL(part_offsets, dim3(1), dim3(1024), 0, s, blk, nblk, planned8, err);
    hipLaunchKernelGGL(part_scatter, dim3(nblk), dim3(256), 0, s, pairs, n, match_pos, blk,
                       planned8, sel8, sel16);
    return hipGetLastError();
}

hipError_t launch_plan_check16(const GroupShape *shape, int32_t ngroups, int32_t max_ncol,
                               int32_t *err, hipStream_t s)
{
    const int tpb = 256, grid = (ngroups + tpb - 1) / tpb;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(plan_check16, dim3(grid), dim3(tpb), 0, s, shape, ngroups, max_ncol, err);
    return hipGetLastError();
}

hipError_t launch_plan_pad8(const tal_seqpair *pairs8, int32_t n8, int32_t *out, int32_t *err,
                            hipStream_t s)
{
    if (n8 <= 0) return hipSuccess;
    hipLaunchKernelGGL(plan_pad8, dim3(1), dim3(1), 0, s, pairs8, n8, out, err);
    return hipGetLastError();
}

hipError_t launch_copy_i32(const int32_t *src, int32_t *dst, hipStream_t s)
{
    hipLaunchKernelGGL(copy_i32, dim3(1), dim3(1), 0, s, src, dst);
    return hipGetLastError();
}

hipError_t launch_flag8(const tal_seqpair *pairs, int32_t n, int32_t match_pos, uint8_t *flag,
                        hipStream_t s)
{
    const int tpb = 256, grid = (n + tpb - 1) / tpb;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(bsw_flag8, dim3(grid), dim3(tpb), 0, s, pairs, n, match_pos, flag);
    return hipGetLastError();
}

hipError_t launch_gather(const tal_seqpair *src, const int32_t *idx, int32_t m,
                         tal_seqpair *dst, hipStream_t s)
{
    const int tpb = 256, grid = (m + tpb - 1) / tpb;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(bsw_gather, dim3(grid), dim3(tpb), 0, s, src, idx, m, dst);
    return hipGetLastError();
}

hipError_t launch_scatter_out(const tal_seqpair *src, const int32_t *idx, int32_t m,
                              tal_seqpair *dst, hipStream_t s)
{
    const int tpb = 256, grid = (m + tpb - 1) / tpb;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(bsw_scatter_out, dim3(grid), dim3(tpb), 0, s, src, idx, m, dst);
    return hipGetLastError();
}

} // namespace talbsw
