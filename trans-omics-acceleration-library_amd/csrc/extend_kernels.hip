/*
 * extend_kernels.hip -- gfx950 kernels of the extension-pair producer
 * (include/tal_extend.h): from each read's chaining result to its left and
 * right extension pairs.  Pointer-chasing work one lane per read, byte moving
 * one wave per read with lanes over its bytes.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>

#include "extend_internal.h"

namespace talext {

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

} // namespace

/* anchors of a walk's register window */
constexpr int PICK_WIN = 8;

/* The walk of read k from its chain end (end_key = f << 32 | ~i of it). */
__device__ __forceinline__ void pick_read(const ExtParams &P, const ExtBuffers &B, int k,
                                          uint64_t end_key)
{
    const int64_t a0 = B.aoff[k];
    const int64_t n64 = B.aoff[k + 1] - a0;
    const int32_t n = (int32_t)n64;
    const int64_t m64 = B.read_off[k + 1] - B.read_off[k];
    tal_extend_read o;
    o.seed_r = 0;
    o.seed_q = o.seed_len = o.chain_score = o.chain_anchors = 0;
    o.left_pair = o.right_pair = -1;
    o.left_score = o.score = 0;
    int32_t lq = 0, lt = 0, rq = 0, rt = 0;
    if (a0 < 0 || n64 < 0 || n64 > INT32_MAX || B.read_off[k] < 0 || m64 < 0) {
        atomicMax(B.err, EXT_ERR_OFFSETS); /* no anchor or base of it is read */
    } else if (m64 > 32767) {
        atomicMax(B.err, EXT_ERR_READ);
    } else if (n > 0) {
        const int32_t m = (int32_t)m64;
        const uint32_t fmax = (uint32_t)(end_key >> 32);
        const int32_t e = (int32_t)~(uint32_t)end_key;
        o.chain_score = (int32_t)fmax;
        /* walk back from e; seeds are runs of anchors on one diagonal
         * that overlap or touch in q */
        bool open = false, bad = false;
        uint64_t diag = 0, best_diag = 0;
        int32_t qlo = 0, qhi = 0, best_lo = 0, best_len = -1;
        int32_t j = e, steps = 0;
        /* the walk's next anchors [whi - PICK_WIN + 1, whi] (clipped at 0)
         * and their p, fetched together: chains mostly step to j - 1, so a
         * window serves several steps from registers */
        int32_t whi = -1;
        tal_chain_anchor wx[PICK_WIN];
        int32_t wp[PICK_WIN];
        while (j >= 0 && steps < n) {
            if (j > whi || j <= whi - PICK_WIN) {
                whi = j;
#pragma unroll
                for (int u = 0; u < PICK_WIN; u++) {
                    const int64_t at = a0 + max(j - u, 0);
                    wx[u] = B.a[at];
                    wp[u] = B.p[at];
                }
            }
            const int32_t u = whi - j;
            tal_chain_anchor x = wx[0];
            int32_t pj = wp[0];
#pragma unroll
            for (int v = 1; v < PICK_WIN; v++) {
                if (u == v) {
                    x = wx[v];
                    pj = wp[v];
                }
            }
            /* r + l <= genome_len, written without the sum (r can be near
             * 2^64; genome_len >= 0 is checked by the host) */
            if (x.q < 0 || x.l < 0 || (int64_t)x.q + x.l > m || x.r > (uint64_t)P.genome_len ||
                (uint64_t)x.l > (uint64_t)P.genome_len - x.r || x.r < (uint64_t)x.q) {
                bad = true;
                break;
            }
            const uint64_t d = x.r - (uint64_t)x.q;
            if (open && d == diag && x.q + x.l >= qlo) {
                qlo = min(qlo, x.q);
                qhi = max(qhi, x.q + x.l);
            } else {
                if (open && (qhi - qlo > best_len || (qhi - qlo == best_len && qlo < best_lo))) {
                    best_len = qhi - qlo;
                    best_lo = qlo;
                    best_diag = diag;
                }
                open = true;
                diag = d;
                qlo = x.q;
                qhi = x.q + x.l;
            }
            steps++;
            j = pj < j ? pj : -1; /* p points back; anything else ends the walk */
        }
        if (open && (qhi - qlo > best_len || (qhi - qlo == best_len && qlo < best_lo))) {
            best_len = qhi - qlo;
            best_lo = qlo;
            best_diag = diag;
        }
        if (bad) {
            atomicMax(B.err, EXT_ERR_ANCHOR);
        } else {
            o.chain_anchors = steps;
            const int32_t qb = best_lo, qe = best_lo + best_len;
            const uint64_t rb = best_diag + (uint64_t)qb;
            o.seed_r = rb;
            o.seed_q = qb;
            o.seed_len = best_len;
            lq = qb;
            lt = (int32_t)min((uint64_t)qb + (uint64_t)P.pad, rb);
            if (lq == 0 || lt == 0) lq = lt = 0;
            const uint64_t re = rb + (uint64_t)best_len;
            rq = m - qe;
            rt = (int32_t)min((uint64_t)(rq + P.pad), (uint64_t)P.genome_len - re);
            if (rq == 0 || rt == 0) rq = rt = 0;
        }
    }
    o.left_score = o.score = P.match * o.seed_len;
    B.out[k] = o;
    B.lhas[k] = lq > 0 ? 1 : 0;
    B.rhas[k] = rq > 0 ? 1 : 0;
    B.rbytes[k] = (int64_t)lt + rt;
    B.qbytes[k] = (int64_t)lq + rq;
    B.lq[k] = lq;
    B.lt[k] = lt;
    B.rq[k] = rq;
    B.rt[k] = rt;
}

/*
 * Steps 1-3 of the producer for 64 reads per wave: the chain end (the first
 * anchor with the highest f), the walk back through p, the merge of the
 * chain's anchors into seeds along diagonals, the longest seed, and the
 * lengths of the two sides.  The chain ends are found read by read with
 * lanes over the read's anchors (coalesced reads of f, one wave maximum of
 * f << 32 | ~i, so ties go to the first anchor); the walks then run one lane
 * per read: a walk is a chain of dependent loads, and one load instruction
 * serves a step of 64 walks, so their latencies overlap.
 */
__global__ void __launch_bounds__(64) extend_pick(ExtParams P, ExtBuffers B)
{
    const int lane = threadIdx.x;
    for (int base = blockIdx.x * 64; base < P.nreads; base += gridDim.x * 64) {
        const int cnt = min(64, P.nreads - base);
        /* keys of absent anchors are 0, below every real key (~i has its
         * top bit set); the first two rows of 4 reads are loaded before any
         * is reduced */
        uint64_t end_key = 0;
        for (int t = 0; t < cnt; t += 4) {
            int64_t b0[4];
            int32_t bn[4];
            uint64_t key[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                b0[u] = 0;
                bn[u] = 0;
                if (t + u < cnt) {
                    /* a set with invalid offsets reads nothing (pick_read
                     * reports it) */
                    const int64_t s0 = B.aoff[base + t + u], d = B.aoff[base + t + u + 1] - s0;
                    if (s0 >= 0 && d > 0) {
                        b0[u] = s0;
                        bn[u] = (int32_t)min(d, (int64_t)INT32_MAX - 64);
                    }
                }
                const uint64_t c0 = lane < bn[u]
                    ? ((uint64_t)B.f[b0[u] + lane] << 32) | (uint32_t)~lane : 0;
                const uint64_t c1 = lane + 64 < bn[u]
                    ? ((uint64_t)B.f[b0[u] + lane + 64] << 32) | (uint32_t)~(lane + 64) : 0;
                key[u] = c1 > c0 ? c1 : c0;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                for (int32_t i = lane + 128; i < bn[u]; i += 64) {
                    const uint64_t c = ((uint64_t)B.f[b0[u] + i] << 32) | (uint32_t)~i;
                    key[u] = c > key[u] ? c : key[u];
                }
                const uint64_t k = wave_max_u64(key[u]);
                end_key = lane == t + u ? k : end_key;
            }
        }
        const int k = base + lane;
        if (k < P.nreads) pick_read(P, B, k, end_key);
    }
}

/* The error word, the pair counts and the byte totals: the last element of
 * each exclusive scan plus its input. */
__global__ void extend_totals(ExtParams P, ExtBuffers B, ExtTotals *t)
{
    if (threadIdx.x != 0) return;
    const int64_t last = (int64_t)P.nreads - 1;
    t->err = *B.err;
    t->nleft = B.lidx[last] + B.lhas[last];
    t->nright = B.ridx[last] + B.rhas[last];
    t->ref_bytes = B.roff[last] + B.rbytes[last];
    t->qer_bytes = B.qoff[last] + B.qbytes[last];
}

/*
 * Step 4 for one read per wave: the two sides' bytes (left ones reversed)
 * into the reference and query buffers at the read's scanned offsets, and
 * the pair records.
 */
__global__ void __launch_bounds__(64) extend_build(ExtParams P, ExtBuffers B)
{
    const int lane = threadIdx.x;
    for (int k = blockIdx.x; k < P.nreads; k += gridDim.x) {
        const tal_extend_read o = B.out[k];
        const int32_t lq = B.lq[k], lt = B.lt[k], rq = B.rq[k], rt = B.rt[k];
        const int64_t roff = B.roff[k], qoff = B.qoff[k];
        const uint8_t *read = B.reads + B.read_off[k];
        const uint64_t rb = o.seed_r, re = rb + (uint64_t)o.seed_len;
        const int32_t qe = o.seed_q + o.seed_len;
        for (int32_t t = lane; t < lt; t += 64) B.ref_out[roff + t] = B.genome[rb - 1 - t];
        for (int32_t t = lane; t < lq; t += 64) B.qer_out[qoff + t] = read[o.seed_q - 1 - t];
        for (int32_t t = lane; t < rt; t += 64) B.ref_out[roff + lt + t] = B.genome[re + t];
        for (int32_t t = lane; t < rq; t += 64) B.qer_out[qoff + lq + t] = read[qe + t];
        /* (every lane writes the same records) */
        {
            tal_extend_read w = o;
            const int32_t h0 = P.match * o.seed_len;
            if (lq > 0) {
                const int32_t idx = B.lidx[k];
                tal_seqpair sp = {};
                sp.idr = (int32_t)roff;
                sp.idq = (int32_t)qoff;
                sp.id = idx;
                sp.len1 = lt;
                sp.len2 = lq;
                sp.h0 = h0;
                sp.seqid = k;
                B.left[idx] = sp;
                w.left_pair = idx;
            }
            if (rq > 0) {
                const int32_t idx = B.ridx[k];
                tal_seqpair sp = {};
                sp.idr = (int32_t)(roff + lt);
                sp.idq = (int32_t)(qoff + lq);
                sp.id = idx;
                sp.len1 = rt;
                sp.len2 = rq;
                sp.h0 = h0; /* the pipeline sets the left score here */
                sp.seqid = k;
                B.right[idx] = sp;
                w.right_pair = idx;
            }
            B.out[k] = w;
        }
    }
}

/* Between the two passes, one lane per read: the left score (or the seed
 * score) becomes the right pair's h0 and the read's left_score. */
__global__ void extend_right_h0(ExtParams P, ExtBuffers B)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.nreads) return;
    tal_extend_read o = B.out[k];
    const int32_t h0 = o.left_pair >= 0 ? B.left[o.left_pair].score : P.match * o.seed_len;
    o.left_score = h0;
    o.score = h0;
    if (o.right_pair >= 0) B.right[o.right_pair].h0 = h0;
    B.out[k] = o;
}

/* After the right pass: the read's final score. */
__global__ void extend_finish(ExtParams P, ExtBuffers B)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.nreads) return;
    const int32_t rp = B.out[k].right_pair;
    if (rp >= 0) B.out[k].score = B.right[rp].score;
}

namespace {
int read_grid(int32_t nreads) { return nreads < 65536 ? (nreads > 0 ? nreads : 1) : 65536; }
} // namespace

/* Waves of the pick kernel per CU (TAL_EXTEND_PICK_WAVES overrides).  Each
 * walk keeps a line of anchors and one of p in use; few waves per CU keep
 * them in L1/L2 between steps, many waves hide more latency. */
int pick_waves_per_cu()
{
    static const int w = [] {
        const char *e = getenv("TAL_EXTEND_PICK_WAVES");
        const int v = e ? atoi(e) : 0;
        return v > 0 ? v : 8;
    }();
    return w;
}

hipError_t launch_extend_pick(const ExtParams &P, const ExtBuffers &B, hipStream_t s)
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int grid =
        std::max(1, std::min((P.nreads + 63) / 64, std::max(1, cus) * pick_waves_per_cu()));
    hipLaunchKernelGGL(extend_pick, dim3(grid), dim3(64), 0, s, P, B);
    return hipGetLastError();
}

hipError_t launch_extend_totals(const ExtParams &P, const ExtBuffers &B, ExtTotals *t,
                                hipStream_t s)
{
    hipLaunchKernelGGL(extend_totals, dim3(1), dim3(64), 0, s, P, B, t);
    return hipGetLastError();
}

hipError_t launch_extend_build(const ExtParams &P, const ExtBuffers &B, hipStream_t s)
{
    hipLaunchKernelGGL(extend_build, dim3(read_grid(P.nreads)), dim3(64), 0, s, P, B);
    return hipGetLastError();
}

hipError_t launch_extend_right_h0(const ExtParams &P, const ExtBuffers &B, hipStream_t s)
{
    hipLaunchKernelGGL(extend_right_h0, dim3((P.nreads + 255) / 256), dim3(256), 0, s, P, B);
    return hipGetLastError();
}

hipError_t launch_extend_finish(const ExtParams &P, const ExtBuffers &B, hipStream_t s)
{
    hipLaunchKernelGGL(extend_finish, dim3((P.nreads + 255) / 256), dim3(256), 0, s, P, B);
    return hipGetLastError();
}

} // namespace talext
