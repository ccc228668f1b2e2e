/*
 * extend_kernels.hip -- gfx950 kernels of the extension-pair producer
 * (include/tal_extend.h): from each read's chaining result to its left and
 * right extension pairs.  Byte-moving and pointer-chasing work: one wave per
 * read, lanes over the read's anchors or bytes.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

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

/*
 * Steps 1-3 of the producer for one read per wave (grid-stride over reads):
 * the chain end (highest f, the first on ties, from a wave reduction of
 * keys f << 32 | ~index), the walk back through p, the merge of
 * the chain's anchors into seeds along diagonals, the longest seed, and the
 * lengths of the two sides.
 */
__global__ void __launch_bounds__(64) extend_pick(ExtParams P, ExtBuffers B)
{
    const int lane = threadIdx.x;
    for (int k = blockIdx.x; k < P.nreads; k += gridDim.x) {
        const int64_t a0 = B.aoff[k];
        const int32_t n = (int32_t)(B.aoff[k + 1] - a0);
        const int64_t m64 = B.read_off[k + 1] - B.read_off[k];
        uint64_t key = 0;
        for (int32_t i = lane; i < n; i += 64) {
            const uint64_t c = ((uint64_t)B.f[a0 + i] << 32) | (0xFFFFFFFFu - (uint32_t)i);
            key = c > key ? c : key;
        }
        key = wave_max_u64(key);
        /* every lane walks the chain on the same (broadcast) data and writes
         * the same records: no lane-divergent branch in the read loop */
        tal_extend_read o;
        o.seed_r = 0;
        o.seed_q = o.seed_len = o.chain_score = o.chain_anchors = 0;
        o.left_pair = o.right_pair = -1;
        o.left_score = o.score = 0;
        int32_t lq = 0, lt = 0, rq = 0, rt = 0;
        if (m64 < 0 || m64 > 32767) {
            atomicMax(B.err, EXT_ERR_READ);
        } else if (n > 0) {
            const int32_t m = (int32_t)m64;
            const int32_t e = (int32_t)(0xFFFFFFFFu - (uint32_t)key);
            o.chain_score = (int32_t)(key >> 32);
            /* walk back from e; seeds are runs of anchors on one diagonal
             * that overlap or touch in q */
            bool open = false, bad = false;
            uint64_t diag = 0, best_diag = 0;
            int32_t qlo = 0, qhi = 0, best_lo = 0, best_len = -1;
            int32_t j = e, steps = 0;
            while (j >= 0 && steps < n) {
                const tal_chain_anchor x = B.a[a0 + j];
                if (x.q < 0 || x.l < 0 || (int64_t)x.q + x.l > m ||
                    x.r + (uint64_t)x.l > (uint64_t)P.genome_len || x.r < (uint64_t)x.q) {
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
                const int32_t pj = B.p[a0 + j];
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

hipError_t launch_extend_pick(const ExtParams &P, const ExtBuffers &B, hipStream_t s)
{
    hipLaunchKernelGGL(extend_pick, dim3(read_grid(P.nreads)), dim3(64), 0, s, P, B);
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
