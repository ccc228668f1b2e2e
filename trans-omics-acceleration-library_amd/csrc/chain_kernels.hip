/*
 * chain_kernels.hip -- gfx950 kernels of the DP chaining engine
 * (include/tal_chain.h), the drop-in for dp_chain::mm_dp_vectorized
 * (src/chaining/parallel_chaining_32_bit.h:216-412).
 *
 * One wave per anchor set (sets sorted longest first, persistent grid).  The
 * anchors of a set are processed in order, as the reference does; for
 * anchor i the 64 lanes take the candidate predecessors j = i-1, i-2, ...
 * down to st in parallel; the wave reduces the best score, then the highest
 * j holding it, so ties go to the highest j as in the reference.  r, q, f and v of the latest anchors live in an LDS ring (512
 * anchors by default, 20 bytes each), so the predecessor search and reads
 * never leave the CU and 16 waves fit on a CU; sets whose predecessor window
 * outgrows it are redone by an overflow pass with a ring of f and v sized for
 * max_iter.  The outputs are collected one per lane and written 64 anchors at
 * a time.
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
 * (lanes shifted in from outside the row keep their own value), and the four
 * row maxima are read as scalars. */
__device__ __forceinline__ int32_t wave_max_i32(int32_t x)
{
    x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x111, 0xF, 0xF, false)); /* row_shr:1 */
    x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x112, 0xF, 0xF, false)); /* row_shr:2 */
    x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x114, 0xF, 0xF, false)); /* row_shr:4 */
    x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x118, 0xF, 0xF, false)); /* row_shr:8 */
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

/*
 * The DP: one set per wave at a time.
 *
 * SMALL (the fast path): an LDS ring of `ring` anchors holds r (64-bit), q,
 * f and v of the latest anchors, so the predecessor search, the candidates
 * and v[p[i]] are all read from LDS; a set whose predecessor window reaches
 * ring - 1 anchors goes to the overflow list and is abandoned (its outputs
 * are rewritten by the second pass).
 * !SMALL (the overflow pass): the sets of the overflow list, with a ring of f
 * and v only, `ring` >= max_iter + 1 or the largest set, and r and q read
 * from the anchor array.
 */
template <bool SMALL>
__global__ void __launch_bounds__(64) chain_main(ChainParams P, ChainBuffers B)
{
    extern __shared__ __align__(8) unsigned char lds_raw[];
    const int R = P.ring, mask = R - 1;
    uint64_t *rr = (uint64_t *)lds_raw;              /* SMALL only */
    int32_t *qr = (int32_t *)(lds_raw + 8 * (size_t)R); /* SMALL only */
    int32_t *fr = SMALL ? qr + R : (int32_t *)lds_raw;
    int32_t *vr = fr + R;
    const int lane = threadIdx.x;
    const int ntasks = SMALL ? P.nsets : B.ctl[CHAIN_CTL_NOVER];
    int *counter = &B.ctl[SMALL ? CHAIN_CTL_COUNTER : CHAIN_CTL_COUNTER2];
    /* The task loop keeps every lane on one path: no continue, and the one
     * lane-0-only update (the predecessor count) after the loop.  A
     * lane-divergent branch at the end of the body let the compiler's loop
     * structure send lanes 1..63 back to the loop head without lane 0, where
     * readfirstlane then read their task word, not lane 0's. */
    uint64_t examined = 0;
    for (;;) {
        int task = 0;
        if (lane == 0) task = atomicAdd(counter, 1);
        task = __builtin_amdgcn_readfirstlane(task);
        if (task >= ntasks) break;
        const int set = SMALL ? B.order[task] : B.over[task];
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
        uint64_t set_examined = 0;
        bool overflow = false;
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
                bool far = false;
                if (idx < i) {
                    const uint64_t rj = SMALL ? rr[idx & mask] : A[idx].r;
                    far = !(ai.r - rj <= (uint64_t)(uint32_t)P.dr);
                }
                const uint64_t m = __ballot(far);
                const int lead = m == ~0ull ? 64 : __builtin_ctzll(~m); /* leading far lanes */
                st += lead;
                if (lead < 64) break;
            }
            if (i - st > P.max_iter) st = i - P.max_iter;
            if (SMALL && i - st >= R - 1) { /* the window outgrows the ring */
                overflow = true;
                break;
            }
            set_examined += (uint64_t)(i - st);
            const bool vec = !(i - 1 - st <= 5);
            const int32_t ri = (int32_t)(uint32_t)ai.r, qi = ai.q;
            /* per lane: the best score and, on ties, the highest j (the j of
             * a lane decrease, so the first one seen) */
            int32_t bsc = INT32_MIN, bj = -1;
            for (int32_t base = i - 1; base >= st; base -= 64) {
                const int32_t j = base - lane;
                if (j >= st) {
                    int32_t rj, qj;
                    if (SMALL) {
                        rj = (int32_t)(uint32_t)rr[j & mask];
                        qj = qr[j & mask];
                    } else {
                        const tal_chain_anchor aj = A[j];
                        rj = (int32_t)(uint32_t)aj.r;
                        qj = aj.q;
                    }
                    const int32_t ddr = (int32_t)((uint32_t)ri - (uint32_t)rj);
                    const int32_t ddq = (int32_t)((uint32_t)qi - (uint32_t)qj);
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
                        if (sc > bsc) {
                            bsc = sc;
                            bj = j;
                        }
                    }
                }
            }
            /* the best score over the wave, then the highest j holding it */
            const int32_t msc = wave_max_i32(bsc);
            int32_t max_f = li, max_j = -1;
            if (msc > li) { /* msc > li >= 0: some candidate, not the sentinel */
                max_f = msc;
                max_j = wave_max_i32(bsc == msc ? bj : -1);
            }
            int32_t vi = max_f;
            if (max_j >= 0) {
                const int32_t vj = vr[max_j & mask];
                if (vj > max_f) vi = vj;
            }
            /* this anchor's ring entries, written after its own reads (slot
             * i & mask may still hold a candidate of i); LDS operations of a
             * wave complete in order, so the next anchor reads them */
            if (lane == 0) {
                fr[i & mask] = max_f;
                vr[i & mask] = vi;
                if (SMALL) {
                    rr[i & mask] = ai.r;
                    qr[i & mask] = qi;
                }
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
        /* (overflow is wave-uniform; every lane takes part in the append, so
         * that no lane-0-only branch closes the loop body) */
        if (overflow) {
            const int slot = __shfl(atomicAdd(&B.ctl[CHAIN_CTL_NOVER], lane == 0 ? 1 : 0), 0);
            B.over[slot] = set;
        } else {
            examined += set_examined;
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
    hipError_t e = hipFuncSetAttribute((const void *)chain_main<true>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void *)chain_main<false>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

size_t chain_lds_bytes(bool small, int ring)
{
    return (small ? 20 : 8) * (size_t)ring;
}

hipError_t launch_chain_main(bool small, int grid, const ChainParams &P, const ChainBuffers &B,
                             hipStream_t s)
{
    const size_t lds = chain_lds_bytes(small, P.ring);
    if (small)
        hipLaunchKernelGGL(chain_main<true>, dim3(grid), dim3(64), lds, s, P, B);
    else
        hipLaunchKernelGGL(chain_main<false>, dim3(grid), dim3(64), lds, s, P, B);
    return hipGetLastError();
}

} // namespace talchain
