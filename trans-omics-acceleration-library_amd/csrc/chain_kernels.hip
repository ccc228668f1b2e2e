/*
 * chain_kernels.hip -- gfx950 kernels of the DP chaining engine
 * (include/tal_chain.h), the drop-in for dp_chain::mm_dp_vectorized
 * (src/chaining/parallel_chaining_32_bit.h:216-412).
 *
 * One wave per anchor set (sets sorted longest first, persistent grid),
 * 64 anchors at a time: lanes = anchors score every candidate older than
 * the 64 together, then the 64 are finished in order against the candidates
 * among themselves, whose f and v are held one per lane.  r, q and f of the
 * latest anchors live in an LDS ring (1024 by default, 12 bytes each);
 * older candidates are read from HBM.  Ties go to the highest j, as in the
 * reference.
 *
 * The gap cost follows the path the reference takes for anchor i: its
 * scalar loop when i-1-st <= 5 ((int)(dd * .01 * avg_qspan) in double,
 * truncated, :386) and its 16-lane vector loop otherwise (dd as float times
 * (float)(0.01 * avg_qspan), rounded toward -inf, :181-199); both add
 * ilog2(dd) >> 1 for dd != 0.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "chain_internal.h"

namespace talchain {

namespace {

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t x, int m)
{
    const int lo = __shfl_xor((int)(uint32_t)x, m);
    const int hi = __shfl_xor((int)(uint32_t)(x >> 32), m);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

/* Wave max of a signed 32-bit value, the same in every lane: DPP row
 * shifts by 1, 2, 4 and 8 leave each 16-lane row's max in its lane 15
 * (lanes shifted in from outside the row take INT32_MIN), and the four
 * row maxima are read as scalars. */
__device__ __forceinline__ int32_t wave_max_i32(int32_t x)
{
    /* INT32_MIN, the identity of max, for lanes shifted in from outside the
     * row, lets the compiler fold each move into the max (v_max_i32_dpp) */
    x = max(x, __builtin_amdgcn_update_dpp(INT32_MIN, x, 0x111, 0xF, 0xF, false)); /* row_shr:1 */
    x = max(x, __builtin_amdgcn_update_dpp(INT32_MIN, x, 0x112, 0xF, 0xF, false)); /* row_shr:2 */
    x = max(x, __builtin_amdgcn_update_dpp(INT32_MIN, x, 0x114, 0xF, 0xF, false)); /* row_shr:4 */
    x = max(x, __builtin_amdgcn_update_dpp(INT32_MIN, x, 0x118, 0xF, 0xF, false)); /* row_shr:8 */
    const int32_t a = __builtin_amdgcn_readlane(x, 15), b = __builtin_amdgcn_readlane(x, 31);
    const int32_t c = __builtin_amdgcn_readlane(x, 47), d = __builtin_amdgcn_readlane(x, 63);
    return max(max(a, b), max(c, d));
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x += shfl_xor_u64(x, m);
    return x;
}

/* floor(log2(x)) for x > 0 */
__device__ __forceinline__ int32_t ilog2_u32(uint32_t x) { return 31 - __clz((int)x); }

} // namespace

/* Validation and sort keys, one wave per set: offsets must be non-negative
 * and non-decreasing, a set at most INT32_MAX anchors, every seed length
 * >= 0.  (No atomic per set: with 10^5 sets one contended word would
 * serialize them.) */
__global__ void chain_check(const tal_chain_anchor *a, const int64_t *off, int32_t nsets,
                            uint32_t *keys, int32_t *ids, int32_t *ctl)
{
    const int k = blockIdx.x;
    const int lane = threadIdx.x;
    if (k >= nsets) return;
    const int64_t b = off[k], e = off[k + 1];
    int err = 0;
    if (b < 0 || e < b) err = CHAIN_ERR_OFFSETS;
    else if (e - b > INT32_MAX) err = CHAIN_ERR_RANGE;
    bool neg = false;
    if (!err)
        for (int64_t i = b + lane; i < e; i += 64) neg |= a[i].l < 0;
    if (__ballot(neg)) err = CHAIN_ERR_RANGE;
    if (lane != 0) return;
    if (err) atomicMax(&ctl[CHAIN_CTL_ERR], err);
    if (k == 0) { /* the batch's anchors, for the host's choice of block shape */
        const int64_t t = off[nsets] - off[0];
        ctl[CHAIN_CTL_ANCHORS] = (int32_t)(uint32_t)t;
        ctl[CHAIN_CTL_ANCHORS + 1] = (int32_t)(t >> 32);
    }
    const int64_t n = err ? 0 : e - b;
    keys[k] = (uint32_t)min(n, (int64_t)0xFFFFFFFF);
    ids[k] = k;
}

/* x86 converts an out-of-range float or double to INT32_MIN, and the
 * reference's gap costs are x86 conversions; gfx950 saturates instead, so a
 * value at or above 2^31 is mapped to INT32_MIN by hand (below -2^31 both give
 * INT32_MIN). */
__device__ __forceinline__ int32_t cvt_floor_f32(float x)
{
    const float y = floorf(x);
    return !(y < 2147483648.0f) ? INT32_MIN : (int32_t)y;
}

__device__ __forceinline__ int32_t cvt_trunc_f64(double x)
{
    return !(x < 2147483648.0) ? INT32_MIN : (int32_t)x;
}

/* ilog2(dd) >> 1, with 0 for dd = 0: ilog2(dd | 1) */
__device__ __forceinline__ int32_t log_half(int32_t dd)
{
    return ilog2_u32((uint32_t)dd | 1u) >> 1;
}

/* The vector loop's gap cost (:129-150): dd as float times (float)(0.01 *
 * avg_qspan), floored, plus ilog2(dd) >> 1, in int32 wrap. */
__device__ __forceinline__ int32_t gap_vec(int32_t dd, float avg_val)
{
    return (int32_t)((uint32_t)cvt_floor_f32((float)dd * avg_val) + (uint32_t)log_half(dd));
}

/* The scalar loop's (:420): dd * .01 * avg_qspan in double, truncated. */
__device__ __forceinline__ int32_t gap_scalar(int32_t dd, double avg_d)
{
    return (int32_t)((uint32_t)cvt_trunc_f64((double)dd * .01 * avg_d) + (uint32_t)log_half(dd));
}

/* The set-wide constants of a candidate's filters and gap. */
struct SetConst {
    int32_t bw;
    uint32_t qlim;    /* max(min(dq, dr), 0): 0 < ddq <= min(dq, dr) as one test */
    float avg_val;    /* (float)(0.01 * avg_qspan) */
    double avg_d;     /* avg_qspan */
    const int32_t *gt; /* table mode: 1 - gap cost of the vector loop by dd (LDS) */
};

/* How a candidate is scored: the reference's scalar loop (windows of at
 * most 6 candidates), its vector loop computed (bw too large for a table,
 * or bw + min(dq, dr) >= 2^31), or its vector loop from the set's table. */
enum GapMode { GAP_SCALAR, GAP_VEC, GAP_TABLE };

/* Anchor i as a candidate's score needs it.  In table mode the offsets are
 * r_i - rbase - 1, q_i - 1 and l_i - 1 (see score()). */
struct Target {
    uint32_t r;
    int32_t q, l;
};

template <int MODE> __device__ __forceinline__ Target target(uint32_t r, int32_t q, int32_t l)
{
    return MODE == GAP_TABLE ? Target{r - 1u, (int32_t)((uint32_t)q - 1u), l - 1} : Target{r, q, l};
}

/*
 * Candidate j (rj = r_j - rbase, qj, f[j]) of anchor i: the reference's
 * filters and score (:262-293 for the vector loop, :386-420 for the scalar
 * one): |ddr - ddq| <= bw, ddr != 0, 0 < ddq <= min(dq, dr), score f[j] +
 * min(ddr, ddq, l_i) - gap in int32 wrap.  Returns whether it passes; the
 * score in sc.
 *
 * Table mode works on a = ddr - 1 and u = ddq - 1 (the target holds r_i - 1
 * and q_i - 1): ddr != 0 is a != -1, 0 < ddq <= qlim is u < qlim unsigned,
 * ddr - ddq = a - u, and min(ddr, ddq, l_i) = min(a, u, l_i - 1) + 1 with the
 * + 1 in the table (1 - gap).  That min is exact unless ddr = INT32_MIN (a
 * wraps); such a candidate has |ddr - ddq| >= 2^31 - qlim > bw when bw +
 * qlim < 2^31, which the host requires of table mode, so it fails anyway.
 */
template <int MODE>
__device__ __forceinline__ bool score(const SetConst &C, const Target &t, uint32_t rj, int32_t qj,
                                      int32_t fj, int32_t &sc)
{
    if (MODE == GAP_TABLE) {
        const int32_t a = (int32_t)(t.r - rj);
        const int32_t u = (int32_t)((uint32_t)t.q - (uint32_t)qj);
        const int32_t diff = (int32_t)((uint32_t)a - (uint32_t)u);
        const int32_t dd = diff < 0 ? (int32_t)(0u - (uint32_t)diff) : diff;
        /* dd in [0, bw] at its own entry; INT32_MIN (and dd > bw) at bw + 1 */
        const int32_t g = C.gt[min((uint32_t)dd, (uint32_t)C.bw + 1u)];
        sc = (int32_t)((uint32_t)min(min(a, u), t.l) + (uint32_t)fj + (uint32_t)g);
        /* (bitwise: one mask, no branches) */
        return !(dd > C.bw) & (a != -1) & ((uint32_t)u < C.qlim);
    }
    const int32_t ddr = (int32_t)(t.r - rj);
    const int32_t ddq = (int32_t)((uint32_t)t.q - (uint32_t)qj);
    const int32_t diff = (int32_t)((uint32_t)ddr - (uint32_t)ddq);
    const int32_t dd = diff < 0 ? (int32_t)(0u - (uint32_t)diff) : diff;
    const int32_t oc = min(min(ddr, ddq), t.l);
    const int32_t gap = MODE == GAP_VEC ? gap_vec(dd, C.avg_val) : gap_scalar(dd, C.avg_d);
    sc = (int32_t)((uint32_t)fj + (uint32_t)oc - (uint32_t)gap);
    return !(dd > C.bw) & (ddr != 0) & ((uint32_t)ddq - 1u < C.qlim);
}

/* One phase-A candidate j: keep it if it passes, is in the lane's window and
 * beats the lane's best (strict: j decreases, ties stay on the higher j). */
template <int MODE>
__device__ __forceinline__ void take_a(const SetConst &C, const Target &t, int32_t st, int32_t j,
                                       uint32_t rj, int32_t qj, int32_t fj, int32_t &bsc,
                                       int32_t &bj)
{
    int32_t sc;
    const bool ok = score<MODE>(C, t, rj, qj, fj, sc);
    if (ok & (j >= st) & (sc > bsc)) {
        bsc = sc;
        bj = j;
    }
}

/*
 * Phase A of a chunk: lane k takes anchor c0 + k (target t; its window
 * starts at st, INT32_MAX for lanes past the set) and scores the candidates
 * j in [max(st, jlo), jhi], all older than the chunk, j decreasing: every
 * lane reads the same j (broadcast reads), so the loop has no divergence
 * and its iterations are independent but for the running best.  From the
 * ring (RING) four slots at a time where they are contiguous (j = 3 mod 4:
 * two 16-byte reads of rq and one of fr), else from the anchor array and the
 * stored f outputs.
 */
template <int MODE, bool RING>
__device__ __forceinline__ void phase_a(const SetConst &C, const Target &t, int32_t st,
                                        int32_t jhi, int32_t jlo, int mask, const uint2 *rq,
                                        const int32_t *fr, const tal_chain_anchor *A,
                                        const uint32_t *F, uint32_t rbase, int32_t &bsc,
                                        int32_t &bj)
{
    int32_t j = jhi;
    if (RING) {
        for (; j >= jlo && (j & 3) != 3; j--) {
            const uint2 e = rq[j & mask];
            take_a<MODE>(C, t, st, j, e.x, (int32_t)e.y, fr[j & mask], bsc, bj);
        }
        /* groups of four slots j-3..j, j = 3 mod 4 (R is a multiple of 4:
         * no wrap inside); two register sets used in turn, so a group's
         * reads are issued while the previous group is scored (the last
         * group re-reads itself instead of reading past jlo) */
        auto ld4 = [&](int32_t jj, uint4 &x0, uint4 &x1, int4 &f4) {
            const int s0 = (jj & mask) - 3;
            x0 = *(const uint4 *)(rq + s0);
            x1 = *(const uint4 *)(rq + s0 + 2);
            f4 = *(const int4 *)(fr + s0);
        };
        auto take4 = [&](int32_t jj, const uint4 &x0, const uint4 &x1, const int4 &f4) {
            take_a<MODE>(C, t, st, jj, x1.z, (int32_t)x1.w, f4.w, bsc, bj);
            take_a<MODE>(C, t, st, jj - 1, x1.x, (int32_t)x1.y, f4.z, bsc, bj);
            take_a<MODE>(C, t, st, jj - 2, x0.z, (int32_t)x0.w, f4.y, bsc, bj);
            take_a<MODE>(C, t, st, jj - 3, x0.x, (int32_t)x0.y, f4.x, bsc, bj);
        };
        if (j - 3 >= jlo) {
            uint4 a0, a1, b0, b1;
            int4 af, bf;
            ld4(j, a0, a1, af);
            for (;;) {
                const bool m1 = j - 7 >= jlo;
                ld4(m1 ? j - 4 : j, b0, b1, bf);
                take4(j, a0, a1, af);
                j -= 4;
                if (!m1) break;
                const bool m2 = j - 7 >= jlo;
                ld4(m2 ? j - 4 : j, a0, a1, af);
                take4(j, b0, b1, bf);
                j -= 4;
                if (!m2) break;
            }
        }
        for (; j >= jlo; j--) {
            const uint2 e = rq[j & mask];
            take_a<MODE>(C, t, st, j, e.x, (int32_t)e.y, fr[j & mask], bsc, bj);
        }
    } else {
        for (; j - 3 >= jlo; j -= 4) {
            tal_chain_anchor x[4];
            int32_t g[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                x[u] = A[j - u];
                g[u] = (int32_t)F[j - u];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                take_a<MODE>(C, t, st, j - u, (uint32_t)x[u].r - rbase, x[u].q, g[u], bsc, bj);
        }
        for (; j >= jlo; j--) {
            const tal_chain_anchor x = A[j];
            take_a<MODE>(C, t, st, j, (uint32_t)x.r - rbase, x.q, (int32_t)F[j], bsc, bj);
        }
    }
}

/*
 * Phase B of a chunk: the anchors c0 .. c0 + kend - 1 in order.  Lane k
 * keeps the best candidate inside the chunk seen so far for anchor c0 + k
 * (isc, ij, iv).  Anchor m is finished from its lane's intra-chunk best and
 * its phase-A best (osc, oj, ov; the intra-chunk j are higher, so they win
 * ties): f, p and v (:304-408).  Then its f is final, and every later lane
 * k > m scores candidate m for its own anchor; m increases, so >= keeps
 * the highest j of equal scores.  No wave reduction is needed.  SCAL: some
 * anchor of the chunk (scal) takes the reference's scalar loop and its
 * double gap.
 */
template <int MODE, bool SCAL>
__device__ __forceinline__ void phase_b(const SetConst &C, int32_t c0, int32_t kend, int lane,
                                        uint32_t rr, int32_t qq, int32_t ll, int32_t st_lane,
                                        bool scal, int32_t osc, int32_t oj, int32_t ov,
                                        int32_t &fk, int32_t &pk, int32_t &vk)
{
    const Target t = target<MODE>(rr, qq, ll);
    const Target ts = target<GAP_SCALAR>(rr, qq, ll);
    int32_t isc = INT32_MIN, ij = -1, iv = 0;
    for (int32_t m = 0; m < kend; m++) {
        /* candidate m's own data comes first: its score terms do not wait
         * for its f */
        const uint32_t rj = __builtin_amdgcn_readlane(rr, m);
        const int32_t qj = __builtin_amdgcn_readlane(qq, m);
        const int32_t li = __builtin_amdgcn_readlane(ll, m);
        const int32_t t_isc = __builtin_amdgcn_readlane(isc, m);
        const int32_t t_osc = __builtin_amdgcn_readlane(osc, m);
        const int32_t t_ij = __builtin_amdgcn_readlane(ij, m);
        const int32_t t_oj = __builtin_amdgcn_readlane(oj, m);
        const int32_t t_iv = __builtin_amdgcn_readlane(iv, m);
        const int32_t t_ov = __builtin_amdgcn_readlane(ov, m);
        const bool intra = t_isc >= t_osc;
        const int32_t top = intra ? t_isc : t_osc;
        const bool take = top > li; /* top > l_i >= 0: a candidate, not a sentinel */
        const int32_t vp = intra ? t_iv : t_ov;
        const int32_t f = take ? top : li;
        const int32_t pj = take ? (intra ? t_ij : t_oj) : -1;
        const int32_t v = take && vp > f ? vp : f;
        fk = lane == m ? f : fk;
        pk = lane == m ? pj : pk;
        vk = lane == m ? v : vk;
        /* candidate c0 + m for the later anchors of the chunk */
        int32_t sc;
        bool ok = score<MODE>(C, t, rj, qj, f, sc);
        if (SCAL) {
            int32_t sc2;
            const bool ok2 = score<GAP_SCALAR>(C, ts, rj, qj, f, sc2);
            sc = scal ? sc2 : sc;
            ok = scal ? ok2 : ok;
        }
        if (ok & (lane > m) & (c0 + m >= st_lane) & (sc >= isc)) {
            isc = sc;
            ij = c0 + m;
            iv = v;
        }
    }
}

/*
 * The DP: one set per block at a time (persistent grid, sets longest
 * first); the WPB waves of a block share the set's LDS ring and split its
 * phase A.  64 anchors (a chunk) at a time:
 *
 *  1. st of each anchor of the chunk (:232-235), by every wave (the same
 *     result).  Sets whose r are sorted and span less than 2^32
 *     (minimap2's anchors are sorted by r): lane k binary-searches the first
 *     p >= st_prev with r_p >= r_i - dr in the ring (the first p within dr,
 *     which is where the reference's forward scan stops), then applies the
 *     max_iter clamp; for sorted r the reference's running st is max(that,
 *     i - max_iter).  Other sets (and windows past the ring): in order, from
 *     a window of 64 r values held one per lane (reloaded when st passes
 *     it), on 64-bit r.
 *  2. Phase A, lanes = anchors: the best candidate older than the chunk,
 *     the candidates split in WPB contiguous ranges, one per wave; wave 0
 *     takes the best of the ranges (ties to the highest j), then v of it
 *     from the v outputs.
 *  3. Phase B (wave 0), in order: anchor c0 + k takes the best of its
 *     phase-A best and of the candidates inside the chunk (f, p, v); its f
 *     is then final and the later anchors score it as a candidate.
 *  4. Wave 0 writes f to the ring and the outputs to HBM.
 *
 * LDS ring of `ring` (R) slots per block: rq[slot] = {r - rbase (low 32
 * bits), q}, written when a chunk's anchors are loaded, and fr[slot] = f,
 * written when the chunk ends; at phase A of chunk c0 the ring holds every j
 * >= c0 + 64 - R.  Older candidates (windows longer than the ring), and v,
 * are read from the anchor array and from the outputs, which wave 0 stored
 * when their chunk ended.  The candidates use the low 32 bits of r, as the
 * reference's uint32 SoA does.  Two barriers a chunk: after wave 0's ring
 * writes, and after the phase-A ranges.
 */
template <int WPB, bool TABLE>
__global__ void __launch_bounds__(64 * WPB) chain_main(ChainParams P, ChainBuffers B)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int R = P.ring, mask = R - 1;
    /* (readfirstlane: the compiler then knows the wave index is uniform,
     * and keeps what derives from it in scalar registers) */
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    uint2 *rq = (uint2 *)lds_raw;
    int32_t *fr = (int32_t *)(rq + R);
    int32_t *gt = fr + R;                  /* P.table entries */
    int2 *part = (int2 *)(gt + P.table);  /* WPB x 64 phase-A bests {score, j} */
    int32_t *task_word = (int32_t *)(part + 64 * WPB);
    const uint64_t dr64 = (uint64_t)(uint32_t)P.dr;
    constexpr int AMODE = TABLE ? GAP_TABLE : GAP_VEC;
    /* Every loop here is block-uniform: the task comes through LDS after a
     * barrier, and the lane-0 updates (the counts) come after the loop.  A
     * lane-divergent branch at the end of a loop body once let the
     * compiler's loop structure send lanes 1..63 back to the loop head
     * without lane 0. */
    uint64_t examined = 0, far_anchors = 0;
    for (;;) {
        if (threadIdx.x == 0) *task_word = atomicAdd(&B.ctl[CHAIN_CTL_COUNTER], 1);
        __syncthreads();
        const int task = __builtin_amdgcn_readfirstlane(*task_word);
        __syncthreads(); /* everyone has read the word before it is rewritten */
        if (task >= P.nsets) break;
        const int set = B.order[task];
        const int64_t a0 = B.off[set];
        const int32_t n = max((int32_t)(B.off[set + 1] - a0), 0);
        const tal_chain_anchor *A = B.a + a0;
        const uint4 *A4 = (const uint4 *)A;
        const uint32_t *F = B.f + a0;
        const int32_t *V = B.v + a0;
        /* avg_qspan = (float)sum(l) / n (:220-222); whether r is sorted */
        uint64_t sum = 0;
        bool unsorted = false;
        for (int i = lane; i < n; i += 64) {
            const tal_chain_anchor x = A[i];
            sum += (uint32_t)x.l;
            if (i + 1 < n) unsorted |= x.r > A[i + 1].r;
        }
        sum = wave_sum_u64(sum);
        const float avg_qspan = (float)sum / (float)n;
        const uint64_t r_first = n > 0 ? A[0].r : 0, r_last = n > 0 ? A[n - 1].r : 0;
        const bool sorted = !__ballot(unsorted) && r_last - r_first <= 0xFFFFFFFFull;
        SetConst C;
        C.bw = P.bw;
        C.qlim = (uint32_t)max(min(P.dq, P.dr), 0);
        C.avg_val = (float)(0.01 * (double)avg_qspan);
        C.avg_d = (double)avg_qspan;
        C.gt = gt;
        /* 1 - the vector loop's gap cost for dd = 0 .. bw, and for dd =
         * INT32_MIN (|INT32_MIN| wraps to itself) at bw + 1; read after the
         * first chunk's barrier */
        if (TABLE)
            for (int d = threadIdx.x; d <= P.bw + 1; d += 64 * WPB)
                gt[d] = (int32_t)(1u - (uint32_t)gap_vec(d <= P.bw ? d : INT32_MIN, C.avg_val));
        /* any base works for the low 32 bits of r; with the set's first r,
         * r - rbase is exact for sorted sets */
        const uint32_t rbase = (uint32_t)r_first;
        int32_t st = 0;
        /* the st window of unsorted sets: r of anchors wb .. wb + 63 */
        int32_t wb = -64;
        uint64_t wr = 0;
        uint4 nxt = make_uint4(0, 0, 0, 0);
        if (lane < n) nxt = A4[lane];
        for (int32_t c0 = 0; c0 < n; c0 += 64) {
            const uint4 cur = nxt;
            const int32_t kend = min(64, n - c0);
            const uint32_t rr = cur.x - rbase; /* this lane's anchor: r - rbase, q, l */
            const int32_t qq = (int32_t)cur.z, ll = (int32_t)cur.w;
            if (wave == 0 && lane < kend) rq[(c0 + lane) & mask] = make_uint2(rr, cur.z);
            /* the ring (and wave 0's outputs) of the chunks so far; the
             * barrier waits for this wave's loads too, so the next chunk's
             * anchors are loaded after it */
            __syncthreads();
            if (c0 + 64 + lane < n) nxt = A4[c0 + 64 + lane];
            const int32_t lo = c0 + 64 - R; /* the oldest j in the ring */

            /* 1. st of every anchor of the chunk */
            int32_t st_lane = INT32_MAX;
            if (sorted && st >= lo) {
                /* first p in [st, i] with r'_p >= r'_i - dr (p = i always is) */
                const int64_t t = (int64_t)rr - (int64_t)dr64;
                int32_t blo = st, bhi = c0 + lane;
                while (__ballot(blo < bhi)) {
                    if (blo < bhi) {
                        const int32_t mid = (blo + bhi) >> 1;
                        if ((int64_t)rq[mid & mask].x >= t)
                            bhi = mid;
                        else
                            blo = mid + 1;
                    }
                }
                const int32_t i = c0 + lane;
                const int32_t sti = i - blo > P.max_iter ? i - P.max_iter : blo;
                st_lane = lane < kend ? sti : INT32_MAX;
                st = __builtin_amdgcn_readlane(sti, kend - 1);
            } else {
                for (int32_t k = 0; k < kend; k++) {
                    const int32_t i = c0 + k;
                    /* (readlane returns int: through uint32_t, no sign extension) */
                    const uint64_t ri64 =
                        ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(cur.y, k) << 32) |
                        (uint32_t)__builtin_amdgcn_readlane(cur.x, k);
                    /* the window must start at or before st (chunks done by
                     * the binary search leave it behind); inside the search it
                     * only moves forward, 64 at a time, so the search ends */
                    if (st < wb || st >= wb + 64) {
                        wb = st;
                        wr = wb + lane < n ? A[wb + lane].r : 0;
                    }
                    for (;;) {
                        const int32_t p = wb + lane;
                        const bool stop = p >= st && (p >= i || ri64 - wr <= dr64);
                        const uint64_t m = __ballot(stop);
                        if (m) {
                            st = wb + __builtin_ctzll(m);
                            break;
                        }
                        wb += 64;
                        wr = wb + lane < n ? A[wb + lane].r : 0;
                    }
                    if (i - st > P.max_iter) st = i - P.max_iter;
                    st_lane = lane == k ? st : st_lane;
                }
            }

            /* 2. phase A: candidates older than the chunk, one contiguous
             * range [jlo_w, jhi_w] per wave, wave 0 the highest */
            const int32_t jmin = __builtin_amdgcn_readfirstlane(st_lane); /* lane 0: the smallest */
            int32_t osc = INT32_MIN, oj = -1, ov = 0;
            const Target ta = target<AMODE>(rr, qq, ll);
            if (jmin < c0) {
                const int32_t span = c0 - jmin;
                const int32_t per = ((span + WPB - 1) / WPB + 3) / 4 * 4;
                const int32_t jhi_w = c0 - 1 - wave * per;
                const int32_t jlo_w = max(jmin, c0 - (wave + 1) * per);
                if (jhi_w >= lo)
                    phase_a<AMODE, true>(C, ta, st_lane, jhi_w, max(jlo_w, lo), mask, rq, fr, A,
                                         F, rbase, osc, oj);
                if (jlo_w < lo) /* F holds wave 0's outputs of the older chunks
                                 * (stored before the barrier) */
                    phase_a<AMODE, false>(C, ta, st_lane, min(jhi_w, lo - 1), jlo_w, mask, rq,
                                          fr, A, F, rbase, osc, oj);
            }
            if (WPB > 1) {
                if (wave > 0) part[wave * 64 + lane] = make_int2(osc, oj);
                __syncthreads();
            }
            if (wave == 0) {
                if (jmin < c0) {
                    /* the later ranges have the lower j: strict > keeps ties
                     * on the higher j */
                    for (int w = 1; w < WPB; w++) {
                        const int2 x = part[w * 64 + lane];
                        if (x.x > osc) {
                            osc = x.x;
                            oj = x.y;
                        }
                    }
                    /* anchors whose window has at most 6 candidates take the
                     * reference's scalar loop and its double gap: redo them
                     * (their old candidates are j >= c0 - 6, in the ring) */
                    const bool scal = lane < kend && c0 + lane - 1 - st_lane <= 5;
                    if (__ballot(scal)) {
                        int32_t ssc = INT32_MIN, sj = -1;
                        phase_a<GAP_SCALAR, true>(C, target<GAP_SCALAR>(rr, qq, ll), st_lane,
                                                  c0 - 1, max(jmin, c0 - 6), mask, rq, fr, A, F,
                                                  rbase, ssc, sj);
                        osc = scal ? ssc : osc;
                        oj = scal ? sj : oj;
                    }
                    if (oj >= 0) ov = V[oj];
                }
                examined += lane < kend ? (uint64_t)(c0 + lane - st_lane) : 0;
                far_anchors += lane < kend && st_lane < lo ? 1 : 0;

                /* 3. phase B: the anchors in order */
                const bool scal = lane < kend && c0 + lane - 1 - st_lane <= 5;
                int32_t fk = 0, pk = -1, vk = 0;
                if (__ballot(scal))
                    phase_b<AMODE, true>(C, c0, kend, lane, rr, qq, ll, st_lane, scal, osc, oj,
                                         ov, fk, pk, vk);
                else
                    phase_b<AMODE, false>(C, c0, kend, lane, rr, qq, ll, st_lane, scal, osc, oj,
                                          ov, fk, pk, vk);

                /* 4. the chunk's f into the ring, the outputs to HBM */
                if (lane < kend) {
                    fr[(c0 + lane) & mask] = fk;
                    B.f[a0 + c0 + lane] = (uint32_t)fk;
                    B.p[a0 + c0 + lane] = pk;
                    B.v[a0 + c0 + lane] = vk;
                }
            }
        }
    }
    examined = wave_sum_u64(examined);
    far_anchors = wave_sum_u64(far_anchors);
    if (threadIdx.x == 0) {
        atomicAdd(B.examined, (unsigned long long)examined);
        atomicAdd(B.far_anchors, (unsigned long long)far_anchors);
    }
}

hipError_t launch_chain_check(const tal_chain_anchor *a, const int64_t *off, int32_t nsets,
                              uint32_t *keys, int32_t *ids, int32_t *ctl, hipStream_t s)
{
    if (nsets <= 0) return hipSuccess;
    hipLaunchKernelGGL(chain_check, dim3(nsets), dim3(64), 0, s, a, off, nsets, keys, ids, ctl);
    return hipGetLastError();
}

#define CHAIN_KERNELS(X) X(1, true) X(2, true) X(4, true) X(1, false) X(2, false) X(4, false)

hipError_t set_chain_lds_limit(size_t bytes)
{
#define SET_LIMIT(W, T)                                                                         \
    {                                                                                              \
        const hipError_t e = hipFuncSetAttribute((const void *)chain_main<W, T>,                   \
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,       \
                                                 (int)bytes);                                      \
        if (e != hipSuccess) return e;                                                             \
    }
    CHAIN_KERNELS(SET_LIMIT)
#undef SET_LIMIT
    return hipSuccess;
}

size_t chain_lds_bytes(int ring, int table, int waves_per_block)
{
    return chain_block_lds(ring, table, waves_per_block);
}

hipError_t launch_chain_main(int waves_per_block, int grid, const ChainParams &P,
                             const ChainBuffers &B, hipStream_t s)
{
    const size_t lds = chain_lds_bytes(P.ring, P.table, waves_per_block);
#define LAUNCH(W, T)                                                                               \
    if (waves_per_block == W && (P.table > 0) == T) {                                              \
        hipLaunchKernelGGL((chain_main<W, T>), dim3(grid), dim3(64 * W), lds, s, P, B);           \
        return hipGetLastError();                                                                  \
    }
    CHAIN_KERNELS(LAUNCH)
#undef LAUNCH
    return hipErrorInvalidValue;
}

} // namespace talchain
