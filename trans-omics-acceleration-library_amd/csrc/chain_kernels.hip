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
 * >= 0.  Reduces the largest set. */
__global__ void chain_check(const tal_chain_anchor *a, const int64_t *off, int32_t nsets,
                            uint32_t *keys, int32_t *ids, int32_t *ctl,
                            unsigned long long *maxn)
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
    const int64_t n = err ? 0 : e - b;
    keys[k] = (uint32_t)min(n, (int64_t)0xFFFFFFFF);
    ids[k] = k;
    atomicMax(maxn, (unsigned long long)n);
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
    const int32_t *gt; /* the vector loop's gap cost by dd (LDS), or null */
};

/* How a candidate's gap is priced: the reference's scalar loop (windows of
 * at most 6 candidates), its vector loop from the set's table, or its vector
 * loop computed (bw too large for a table). */
enum GapMode { GAP_SCALAR, GAP_TABLE, GAP_VEC };

/*
 * Candidate j (rj = r_j - rbase, qj, f[j]) of anchor i (ri = r_i - rbase, qi,
 * li): the reference's filters and score (:262-293 for the vector loop,
 * :386-420 for the scalar one): |ddr - ddq| <= bw, ddr != 0, 0 < ddq <=
 * min(dq, dr), score f[j] + min(ddr, ddq, l_i) - gap in int32 wrap.  Returns
 * whether it passes; the score in sc.
 */
template <int MODE>
__device__ __forceinline__ bool score(const SetConst &C, uint32_t ri, int32_t qi, int32_t li,
                                      uint32_t rj, int32_t qj, int32_t fj, int32_t &sc)
{
    const int32_t ddr = (int32_t)(ri - rj);
    const int32_t ddq = (int32_t)((uint32_t)qi - (uint32_t)qj);
    const int32_t diff = (int32_t)((uint32_t)ddr - (uint32_t)ddq);
    const int32_t dd = diff < 0 ? (int32_t)(0u - (uint32_t)diff) : diff;
    const int32_t oc = min(min(ddr, ddq), li);
    int32_t gap;
    if (MODE == GAP_TABLE) /* dd in [0, bw] at its own entry; INT32_MIN (and dd > bw) at bw + 1 */
        gap = C.gt[min((uint32_t)dd, (uint32_t)C.bw + 1u)];
    else if (MODE == GAP_VEC)
        gap = gap_vec(dd, C.avg_val);
    else
        gap = gap_scalar(dd, C.avg_d);
    sc = (int32_t)((uint32_t)fj + (uint32_t)oc - (uint32_t)gap);
    /* (bitwise: one mask, no branches) */
    return !(dd > C.bw) & (ddr != 0) & ((uint32_t)ddq - 1u < C.qlim);
}

/*
 * Phase A of a chunk: lane k takes anchor c0 + k (ri, qi, li; its window
 * starts at st, INT32_MAX for lanes past the set) and scores the candidates
 * j in [max(st, jlo), jhi], all older than the chunk, j decreasing: every
 * lane reads the same j (one broadcast read), so the loop has no divergence
 * and its iterations are independent but for the running best.  Strict >
 * keeps the highest j of a lane's best score.  Candidates come from the
 * ring (RING) or from the anchor array and the stored f outputs.
 */
template <int MODE, int UNROLL, bool RING>
__device__ __forceinline__ void phase_a(const SetConst &C, uint32_t ri, int32_t qi, int32_t li,
                                        int32_t st, int32_t jhi, int32_t jlo, int mask,
                                        const uint2 *rq, const int32_t *fr,
                                        const tal_chain_anchor *A, const uint32_t *F,
                                        uint32_t rbase, int32_t &bsc, int32_t &bj)
{
    int32_t j = jhi;
    for (; j - (UNROLL - 1) >= jlo; j -= UNROLL) {
        uint2 e[UNROLL];
        int32_t g[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            if (RING) {
                e[u] = rq[(j - u) & mask];
                g[u] = fr[(j - u) & mask];
            } else {
                const tal_chain_anchor x = A[j - u];
                e[u] = make_uint2((uint32_t)x.r - rbase, (uint32_t)x.q);
                g[u] = (int32_t)F[j - u];
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            int32_t sc;
            const bool ok = score<MODE>(C, ri, qi, li, e[u].x, (int32_t)e[u].y, g[u], sc);
            if (ok & (j - u >= st) & (sc > bsc)) {
                bsc = sc;
                bj = j - u;
            }
        }
    }
    for (; j >= jlo; j--) {
        uint2 e;
        int32_t g;
        if (RING) {
            e = rq[j & mask];
            g = fr[j & mask];
        } else {
            const tal_chain_anchor x = A[j];
            e = make_uint2((uint32_t)x.r - rbase, (uint32_t)x.q);
            g = (int32_t)F[j];
        }
        int32_t sc;
        const bool ok = score<MODE>(C, ri, qi, li, e.x, (int32_t)e.y, g, sc);
        if (ok & (j >= st) & (sc > bsc)) {
            bsc = sc;
            bj = j;
        }
    }
}

/*
 * The DP: one set per wave at a time, WPB independent waves per block (each
 * with its own slice of LDS, no barriers), 64 anchors (a chunk) at a time:
 *
 *  1. st of each anchor of the chunk (:232-235).  Sets whose r are sorted
 *     and span less than 2^32 (minimap2's anchors are sorted by r): lane k
 *     binary-searches the first p >= st_prev with r_p >= r_i - dr in the
 *     ring (the first p within dr, which is where the reference's forward
 *     scan stops), then applies the max_iter clamp; for sorted r the
 *     reference's running st is max(that, i - max_iter).  Other sets (and
 *     windows past the ring): in order, from a window of 64 r values held
 *     one per lane (reloaded when st passes it), on 64-bit r.
 *  2. Phase A, lanes = anchors: the best candidate older than the chunk;
 *     then v of it, read from the v outputs.
 *  3. Phase B, in order: anchor c0 + k scores the candidates inside the
 *     chunk (lanes m < k, whose f and v are final), then takes the best of
 *     those (their j are higher, so they win ties) and of its phase-A best:
 *     f, p and v (:304-408).
 *  4. f goes to the LDS ring, the outputs to HBM.
 *
 * LDS ring of `ring` (R) slots per wave: rq[slot] = {r - rbase (low 32 bits),
 * q}, written when a chunk's anchors are loaded, and fr[slot] = f, written
 * when the chunk ends; at phase A of chunk c0 the ring holds every j >= c0 +
 * 64 - R.  Older candidates (windows longer than the ring), and v, are read
 * from the anchor array and from the outputs, which this wave stored when
 * their chunk ended.  The candidates use the low 32 bits of r, as the
 * reference's uint32 SoA does.
 */
template <int WPB, int UNROLL, bool TABLE>
__global__ void __launch_bounds__(64 * WPB) chain_main(ChainParams P, ChainBuffers B)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int R = P.ring, mask = R - 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint2 *rq = (uint2 *)(lds_raw + (size_t)wave * chain_wave_lds(R, P.table));
    int32_t *fr = (int32_t *)(rq + R);
    int32_t *gt = fr + R; /* P.table entries */
    const uint64_t dr64 = (uint64_t)(uint32_t)P.dr;
    constexpr int AMODE = TABLE ? GAP_TABLE : GAP_VEC;
    /* The task loop keeps every lane on one path: no continue, and the
     * lane-0-only updates (the counts) after the loop.  A lane-divergent
     * branch at the end of the body let the compiler's loop structure send
     * lanes 1..63 back to the loop head without lane 0, where readfirstlane
     * then read their task word, not lane 0's. */
    uint64_t examined = 0, far_anchors = 0;
    for (;;) {
        int task = 0;
        if (lane == 0) task = atomicAdd(&B.ctl[CHAIN_CTL_COUNTER], 1);
        task = __builtin_amdgcn_readfirstlane(task);
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
        /* the vector loop's gap costs for dd = 0 .. bw, and for dd =
         * INT32_MIN (|INT32_MIN| wraps to itself) at bw + 1 */
        if (TABLE)
            for (int d = lane; d <= P.bw + 1; d += 64)
                gt[d] = gap_vec(d <= P.bw ? d : INT32_MIN, C.avg_val);
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
            if (c0 + 64 + lane < n) nxt = A4[c0 + 64 + lane];
            const int32_t kend = min(64, n - c0);
            const uint32_t rr = cur.x - rbase; /* this lane's anchor: r - rbase, q, l */
            const int32_t qq = (int32_t)cur.z, ll = (int32_t)cur.w;
            if (lane < kend) rq[(c0 + lane) & mask] = make_uint2(rr, cur.z);
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
            examined += lane < kend ? (uint64_t)(c0 + lane - st_lane) : 0;

            /* 2. phase A: candidates older than the chunk */
            const int32_t jmin = __builtin_amdgcn_readfirstlane(st_lane); /* lane 0: the smallest */
            far_anchors += lane < kend && st_lane < lo ? 1 : 0;
            int32_t osc = INT32_MIN, oj = -1, ov = 0;
            if (jmin < c0) {
                phase_a<AMODE, UNROLL, true>(C, rr, qq, ll, st_lane, c0 - 1, max(jmin, lo), mask,
                                             rq, fr, A, F, rbase, osc, oj);
                /* this wave stored the outputs of the older chunks; wait
                 * until the stores are done */
                __builtin_amdgcn_s_waitcnt(0x0F70); /* vmcnt(0) */
                if (jmin < lo)
                    phase_a<AMODE, UNROLL, false>(C, rr, qq, ll, st_lane, lo - 1, jmin, mask, rq,
                                                  fr, A, F, rbase, osc, oj);
                /* anchors whose window has at most 6 candidates take the
                 * reference's scalar loop and its double gap: redo them
                 * (their old candidates are j >= c0 - 6, in the ring) */
                const bool scal = lane < kend && c0 + lane - 1 - st_lane <= 5;
                if (__ballot(scal)) {
                    int32_t ssc = INT32_MIN, sj = -1;
                    phase_a<GAP_SCALAR, 1, true>(C, rr, qq, ll, st_lane, c0 - 1,
                                                 max(jmin, c0 - 6), mask, rq, fr, A, F, rbase,
                                                 ssc, sj);
                    osc = scal ? ssc : osc;
                    oj = scal ? sj : oj;
                }
                if (oj >= 0) ov = V[oj];
            }

            /* 3. phase B: the anchors in order */
            int32_t fk = 0, pk = -1, vk = 0;
            for (int32_t k = 0; k < kend; k++) {
                const int32_t i = c0 + k;
                const int32_t sti = __builtin_amdgcn_readlane(st_lane, k);
                const uint32_t ri = __builtin_amdgcn_readlane(rr, k);
                const int32_t qi = __builtin_amdgcn_readlane(qq, k);
                const int32_t li = __builtin_amdgcn_readlane(ll, k);
                /* candidates c0 + m, m < k, within the window */
                const bool inw = lane < k && c0 + lane >= sti;
                int32_t sc = INT32_MIN;
                bool ok;
                if (i - 1 - sti <= 5)
                    ok = score<GAP_SCALAR>(C, ri, qi, li, rr, qq, fk, sc);
                else
                    ok = score<GAP_VEC>(C, ri, qi, li, rr, qq, fk, sc);
                ok &= inw;
                const int32_t s = ok ? sc : INT32_MIN;
                const int32_t msc = __ballot(ok) ? wave_max_i32(s) : INT32_MIN;
                const int32_t osc_i = __builtin_amdgcn_readlane(osc, k);
                /* the chunk's candidates have the higher j: they win ties */
                const bool intra = msc >= osc_i;
                const int32_t top = intra ? msc : osc_i;
                int32_t f = li, pj = -1, v = li;
                if (top > li) { /* top > l_i >= 0: a candidate, not a sentinel */
                    int32_t vp;
                    if (intra) {
                        const uint64_t tie = __ballot(ok && s == msc);
                        const int m = 63 - __builtin_clzll(tie); /* the highest j */
                        pj = c0 + m;
                        vp = __builtin_amdgcn_readlane(vk, m);
                    } else {
                        pj = __builtin_amdgcn_readlane(oj, k);
                        vp = __builtin_amdgcn_readlane(ov, k);
                    }
                    f = top;
                    v = vp > f ? vp : f;
                }
                fk = lane == k ? f : fk;
                pk = lane == k ? pj : pk;
                vk = lane == k ? v : vk;
            }

            /* 4. the chunk's f into the ring, the outputs to HBM */
            if (lane < kend) {
                fr[(c0 + lane) & mask] = fk;
                B.f[a0 + c0 + lane] = (uint32_t)fk;
                B.p[a0 + c0 + lane] = pk;
                B.v[a0 + c0 + lane] = vk;
            }
        }
    }
    examined = wave_sum_u64(examined);
    far_anchors = wave_sum_u64(far_anchors);
    if (lane == 0) {
        atomicAdd(B.examined, (unsigned long long)examined);
        atomicAdd(B.far_anchors, (unsigned long long)far_anchors);
    }
}

hipError_t launch_chain_check(const tal_chain_anchor *a, const int64_t *off, int32_t nsets,
                              uint32_t *keys, int32_t *ids, int32_t *ctl,
                              unsigned long long *maxn, hipStream_t s)
{
    if (nsets <= 0) return hipSuccess;
    hipLaunchKernelGGL(chain_check, dim3(nsets), dim3(64), 0, s, a, off, nsets, keys, ids, ctl,
                       maxn);
    return hipGetLastError();
}

#define CHAIN_KERNELS(X)                                                                           \
    X(1, 1, true) X(1, 2, true) X(1, 4, true) X(2, 1, true) X(2, 2, true) X(2, 4, true)            \
    X(4, 1, true) X(4, 2, true) X(4, 4, true) X(1, 1, false) X(1, 2, false) X(1, 4, false)         \
    X(2, 1, false) X(2, 2, false) X(2, 4, false) X(4, 1, false) X(4, 2, false) X(4, 4, false)

hipError_t set_chain_lds_limit(size_t bytes)
{
#define SET_LIMIT(W, U, T)                                                                         \
    {                                                                                              \
        const hipError_t e = hipFuncSetAttribute((const void *)chain_main<W, U, T>,                \
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
    return chain_wave_lds(ring, table) * (size_t)waves_per_block;
}

hipError_t launch_chain_main(int waves_per_block, int unroll, int grid, const ChainParams &P,
                             const ChainBuffers &B, hipStream_t s)
{
    const size_t lds = chain_lds_bytes(P.ring, P.table, waves_per_block);
#define LAUNCH(W, U, T)                                                                            \
    if (waves_per_block == W && unroll == U && (P.table > 0) == T) {                               \
        hipLaunchKernelGGL((chain_main<W, U, T>), dim3(grid), dim3(64 * W), lds, s, P, B);        \
        return hipGetLastError();                                                                  \
    }
    CHAIN_KERNELS(LAUNCH)
#undef LAUNCH
    return hipErrorInvalidValue;
}

} // namespace talchain
