/*
 * chain_kernels.hip -- gfx950 kernels of the DP chaining engine
 * (include/tal_chain.h), the drop-in for dp_chain::mm_dp_vectorized
 * (src/chaining/parallel_chaining_32_bit.h:216-412).
 *
 * One wave per anchor set (sets sorted longest first, persistent grid).  The
 * anchors of a set are processed in order, as the reference does; for
 * anchor i the 64 lanes take the candidate predecessors j = i-1, i-2, ...
 * down to st in parallel, and the best (score, j) pair is reduced across the
 * wave as one 64-bit key (score in the high half, biased to compare
 * unsigned, j in the low half), so ties go to the highest j as in the
 * reference.  f and v of the last `ring` anchors live in an LDS ring (ring
 * >= max_iter + 1 or the set size), so the predecessor reads of f[j] and
 * v[p[i]] never leave the CU; the outputs are collected one per lane and
 * written 64 anchors at a time.
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

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t x)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const uint64_t y = shfl_xor_u64(x, m);
        x = y > x ? y : x;
    }
    return x;
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

/*
 * The DP: one set per wave at a time.  `ring` entries of f and of v in LDS
 * (a power of two, at least max_iter + 1 or the largest set).
 */
__global__ void __launch_bounds__(64) chain_main(ChainParams P, ChainBuffers B)
{
    extern __shared__ int32_t lds[];
    int32_t *fr = lds, *vr = lds + P.ring;
    const int lane = threadIdx.x;
    const int mask = P.ring - 1;
    /* The task loop keeps every lane on one path: no continue, and the one
     * lane-0-only update (the predecessor count) after the loop.  A
     * lane-divergent branch at the end of the body let the compiler's loop
     * structure send lanes 1..63 back to the loop head without lane 0, where
     * readfirstlane then read their task word, not lane 0's. */
    uint64_t examined = 0;
    for (;;) {
        int task = 0;
        if (lane == 0) task = atomicAdd(&B.ctl[CHAIN_CTL_COUNTER], 1);
        task = __builtin_amdgcn_readfirstlane(task);
        if (task >= P.nsets) break;
        const int set = B.order[task];
        const int64_t a0 = B.off[set];
        const int32_t n = max((int32_t)(B.off[set + 1] - a0), 0);
        const tal_chain_anchor *A = B.a + a0;
        /* avg_qspan = (float)sum(l) / n (:220-222) */
        uint64_t sum = 0;
        for (int i = lane; i < n; i += 64) sum += (uint32_t)A[i].l;
        sum = wave_sum_u64(sum);
        const float avg_qspan = (float)sum / (float)n;
        const float avg_val = (float)(0.01 * (double)avg_qspan);
        const double avg_d = (double)avg_qspan;
        int32_t st = 0;
        uint32_t of = 0;
        int32_t op = 0, ov = 0;
        for (int32_t i = 0; i < n; i++) {
            const tal_chain_anchor ai = A[i];
            const int32_t li = ai.l;
            /* st: first predecessor within dr on r (64-bit, :232), found 64
             * candidates at a time: the reference stops at the first one
             * that is not too far */
            for (;;) {
                const int32_t idx = st + lane;
                const bool far = idx < i && !(ai.r - A[idx].r <= (uint64_t)(uint32_t)P.dr);
                const uint64_t m = __ballot(far);
                const int lead = m == ~0ull ? 64 : __builtin_ctzll(~m); /* leading far lanes */
                st += lead;
                if (lead < 64) break;
            }
            if (i - st > P.max_iter) st = i - P.max_iter;
            examined += (uint64_t)(i - st);
            const bool vec = !(i - 1 - st <= 5);
            const int32_t ri = (int32_t)(uint32_t)ai.r, qi = ai.q;
            uint64_t best = 0;
            for (int32_t base = i - 1; base >= st; base -= 64) {
                const int32_t j = base - lane;
                if (j >= st) {
                    const tal_chain_anchor aj = A[j];
                    const int32_t ddr = (int32_t)((uint32_t)ri - (uint32_t)aj.r);
                    const int32_t ddq = (int32_t)((uint32_t)qi - (uint32_t)aj.q);
                    const int32_t diff = (int32_t)((uint32_t)ddr - (uint32_t)ddq);
                    const int32_t dd = diff < 0 ? (int32_t)(0u - (uint32_t)diff) : diff;
                    const bool ok = !(dd > P.bw) && ddr != 0 && ddq > 0 && !(ddq > P.dq) &&
                                    !(ddq > P.dr);
                    if (ok) {
                        int32_t oc = ddr < ddq ? ddr : ddq;
                        if (oc > li) oc = li;
                        const int32_t log_dd = dd ? ilog2_u32((uint32_t)dd) : 0;
                        int32_t gap;
                        if (vec)
                            gap = (int32_t)floorf((float)dd * avg_val) + (log_dd >> 1);
                        else
                            gap = (int32_t)((double)dd * .01 * avg_d) + (log_dd >> 1);
                        const int32_t sc =
                            (int32_t)((uint32_t)fr[j & mask] + (uint32_t)oc - (uint32_t)gap);
                        const uint64_t key =
                            ((uint64_t)((uint32_t)sc ^ 0x80000000u) << 32) | (uint32_t)j;
                        best = key > best ? key : best;
                    }
                }
            }
            best = wave_max_u64(best);
            int32_t max_f = li, max_j = -1;
            if (best) {
                const int32_t sc = (int32_t)((uint32_t)(best >> 32) ^ 0x80000000u);
                if (sc > li) {
                    max_f = sc;
                    max_j = (int32_t)(uint32_t)best;
                }
            }
            int32_t vi = max_f;
            if (max_j >= 0) {
                const int32_t vj = vr[max_j & mask];
                if (vj > max_f) vi = vj;
            }
            /* LDS operations of a wave complete in order: the next anchor's
             * reads see these */
            if (lane == 0) {
                fr[i & mask] = max_f;
                vr[i & mask] = vi;
            }
            /* the outputs of anchors i0 .. i0+63 gather one per lane */
            if (lane == (i & 63)) {
                of = (uint32_t)max_f;
                op = max_j;
                ov = vi;
            }
            if ((i & 63) == 63 || i == n - 1) {
                const int32_t i0 = i & ~63;
                if (i0 + lane <= i) {
                    B.f[a0 + i0 + lane] = of;
                    B.p[a0 + i0 + lane] = op;
                    B.v[a0 + i0 + lane] = ov;
                }
            }
        }
    }
    if (lane == 0) atomicAdd(B.examined, (unsigned long long)examined);
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

hipError_t set_chain_lds_limit(size_t bytes)
{
    return hipFuncSetAttribute((const void *)chain_main,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t launch_chain_main(int grid, const ChainParams &P, const ChainBuffers &B, hipStream_t s)
{
    const size_t lds = sizeof(int32_t) * 2 * (size_t)P.ring;
    hipLaunchKernelGGL(chain_main, dim3(grid), dim3(64), lds, s, P, B);
    return hipGetLastError();
}

} // namespace talchain
