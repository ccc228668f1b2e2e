/*
 * bsw_scalar.hip -- gfx950 kernel of the per-pair (scalar) banded extension,
 * the semantics of BandedPairWiseSW::scalarBandedSWA
 * (src/alignment/BSW/bandedSWA.cpp:116-237) behind scalarBandedSWAWrapper
 * (:242-260).
 *
 * Unlike the SIMD path, every pair is scored on its own: its own column
 * range [beg, end), narrowed from its own previous row, its own exits.  So
 * one lane is one pair with no coupling to the others, and a wave simply
 * runs 64 independent pairs.  The pairs are taken longest first (a cost
 * sort by the host runtime), so that the lanes of a wave have loops of
 * similar length.  Each lane's row state eh[j] = (H, E) lives in the wave's
 * global slab as [column][64 lanes] of int2, so that lanes working on the
 * same column touch one cache line.  Arithmetic is int32, as in the
 * reference.
 */
#include "bsw_internal.h"

namespace talbsw {

/* mat[r * 5 + q] of one reference code r, as five signed bytes of a u64. */
__device__ __forceinline__ uint64_t mat_row(const ScalarParams &P, int r)
{
    uint64_t v = 0;
#pragma unroll
    for (int c = 0; c < 5; c++) v |= (uint64_t)(uint8_t)P.mat[r * 5 + c] << (8 * c);
    return v;
}

__global__ __launch_bounds__(64) void bsw_scalar(ScalarParams P, tal_seqpair *pairs,
                                                 const uint8_t *ref, const uint8_t *qer,
                                                 const int32_t *order, int32_t ntasks,
                                                 int32_t *counter, int2 *slabs, int32_t *err)
{
    const int lane = threadIdx.x;
    int2 *eh = slabs + (size_t)blockIdx.x * (size_t)(P.max_qlen + 2) * 64 + lane;
    const int oe_del = P.o_del + P.e_del, oe_ins = P.o_ins + P.e_ins;
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(counter, 1);
        t = __shfl(t, 0, 64);
        if (t >= ntasks) break;
        const int k = order ? order[t * 64 + lane] : t * 64 + lane;
        if (k < 0 || k >= P.n) continue;
        tal_seqpair *sp = pairs + k;
        const int qlen = sp->len2, tlen = sp->len1, h0 = sp->h0;
        const uint8_t *query = qer + (int64_t)sp->idq;
        const uint8_t *target = ref + (int64_t)sp->idr;
        /* the query must be codes 0-4: the reference indexes mat with them
         * (:138-141) */
        bool bad = false;
        for (int j = 0; j < qlen; j++) bad |= query[j] > 4;
        if (bad) {
            atomicMax(err, 1);
            continue;
        }
        /* row -1 (:143-145); the rest of eh starts at zero, as calloc's */
        {
            /* the chain runs while the previous entry exceeds e_ins; past
             * its end every entry keeps calloc's zero */
            int hp = 0;
            bool chain = true;
            for (int j = 0; j <= qlen; j++) {
                int hv = 0;
                if (j == 0) {
                    hv = h0;
                } else if (j == 1) {
                    hv = h0 > oe_ins ? h0 - oe_ins : 0;
                } else {
                    chain = chain && hp > P.e_ins;
                    hv = chain ? hp - P.e_ins : 0;
                }
                hp = hv;
                eh[(size_t)j * 64] = make_int2(hv, 0);
            }
        }
        /* band clipping (:147-156): the reference divides as double */
        int w = P.w;
        {
            const int mx = P.mat_max;
            int max_ins = (int)((double)(qlen * mx + P.end_bonus - P.o_ins) / P.e_ins + 1.);
            max_ins = max_ins > 1 ? max_ins : 1;
            w = w < max_ins ? w : max_ins;
            int max_del = (int)((double)(qlen * mx + P.end_bonus - P.o_del) / P.e_del + 1.);
            max_del = max_del > 1 ? max_del : 1;
            w = w < max_del ? w : max_del;
        }
        int max = h0, max_i = -1, max_j = -1, max_ie = -1, gscore = -1, max_off = 0;
        int beg = 0, end = qlen;
        for (int i = 0; i < tlen; i++) {
            const int tc = target[i];
            if (tc > 4) {
                bad = true;
                break;
            }
            const uint64_t row = mat_row(P, tc);
            int f = 0, h1, m = 0, mj = -1, j;
            if (beg < i - w) beg = i - w;
            if (end > i + w + 1) end = i + w + 1;
            if (end > qlen) end = qlen;
            if (beg == 0) {
                h1 = h0 - (P.o_del + P.e_del * (i + 1));
                if (h1 < 0) h1 = 0;
            } else {
                h1 = 0;
            }
            /* the cell (:174-198) */
            for (j = beg; j < end; j++) {
                int2 *p = eh + (size_t)j * 64;
                const int2 c = *p;
                int M = c.x, e = c.y, h, tt;
                const int s = (int)(int8_t)(uint8_t)(row >> (8 * query[j]));
                M = M ? M + s : 0;
                h = M > e ? M : e;
                h = h > f ? h : f;
                mj = m > h ? mj : j;
                m = m > h ? m : h;
                tt = M - oe_del;
                tt = tt > 0 ? tt : 0;
                e -= P.e_del;
                e = e > tt ? e : tt;
                *p = make_int2(h1, e);
                h1 = h;
                tt = M - oe_ins;
                tt = tt > 0 ? tt : 0;
                f -= P.e_ins;
                f = f > tt ? f : tt;
            }
            eh[(size_t)end * 64] = make_int2(h1, 0);
            if (j == qlen) {
                max_ie = gscore > h1 ? max_ie : i;
                gscore = gscore > h1 ? gscore : h1;
            }
            if (m == 0) break;
            if (m > max) {
                max = m, max_i = i, max_j = mj;
                const int d = mj - i < 0 ? i - mj : mj - i;
                max_off = max_off > d ? max_off : d;
            } else if (P.zdrop > 0) {
                if (i - max_i > mj - max_j) {
                    if (max - m - ((i - max_i) - (mj - max_j)) * P.e_del > P.zdrop) break;
                } else {
                    if (max - m - ((mj - max_j) - (i - max_i)) * P.e_ins > P.zdrop) break;
                }
            }
            /* the next row's range (:227-230) */
            for (j = beg; j < end; j++) {
                const int2 c = eh[(size_t)j * 64];
                if (c.x != 0 || c.y != 0) break;
            }
            beg = j;
            for (j = end; j >= beg; j--) {
                const int2 c = eh[(size_t)j * 64];
                if (c.x != 0 || c.y != 0) break;
            }
            end = j + 2 < qlen ? j + 2 : qlen;
        }
        if (bad) {
            atomicMax(err, 1);
            continue;
        }
        sp->score = max;
        sp->qle = max_j + 1;
        sp->tle = max_i + 1;
        sp->gtle = max_ie + 1;
        sp->gscore = gscore;
        sp->max_off = max_off;
    }
}

/* Longest pairs first: key = rows x the widest range a row can sweep. */
__global__ void scalar_cost(const tal_seqpair *pairs, int32_t n, int32_t w, uint32_t *key,
                            int32_t *ids, int32_t *maxlen)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int l1 = pairs[k].len1, l2 = pairs[k].len2;
    if (l1 < 0 || l2 < 0) {
        atomicMax(maxlen, 0x7FFFFFFF); /* reported as out of range */
        key[k] = 0;
    } else {
        atomicMax(maxlen, max(l1, l2));
        const int64_t width = min((int64_t)l2, 2 * (int64_t)w + 2) + 1;
        key[k] = (uint32_t)min((int64_t)l1 * width, (int64_t)0xFFFFFFFF);
    }
    ids[k] = k;
}

hipError_t launch_scalar_cost(const tal_seqpair *pairs, int32_t n, int32_t w, uint32_t *key,
                              int32_t *ids, int32_t *maxlen, hipStream_t s)
{
    hipLaunchKernelGGL(scalar_cost, dim3((n + 255) / 256), dim3(256), 0, s, pairs, n, w, key, ids,
                       maxlen);
    return hipGetLastError();
}

hipError_t launch_scalar(int grid, const ScalarParams &P, tal_seqpair *pairs, const uint8_t *ref,
                         const uint8_t *qer, const int32_t *order, int32_t ntasks,
                         int32_t *counter, int2 *slabs, int32_t *err, hipStream_t s)
{
    hipLaunchKernelGGL(bsw_scalar, dim3(grid), dim3(64), 0, s, P, pairs, ref, qer, order, ntasks,
                       counter, slabs, err);
    return hipGetLastError();
}

} // namespace talbsw
