/*
 * ref_chain_bind.cpp -- C entry point over the reference's own DP chaining
 * (dp_chain::mm_dp_vectorized, src/chaining/parallel_chaining_32_bit.h),
 * compiled from the reference header by `make -C oracle ref` into
 * oracle/_ref/libchain_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY: the tests use it to check the oracle's
 * restatement against the reference itself.  Built with AVX-512 (the
 * reference's vector path); refuses to run on a CPU without it.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "parallel_chaining_32_bit.h"

extern "C" int ref_chain_dp(const uint64_t *r, const int32_t *q, const int32_t *l, int64_t n,
                            int32_t bw, int32_t max_iter, int32_t dr, int32_t dq, uint32_t *f,
                            int32_t *p, int32_t *v)
{
    if (!__builtin_cpu_supports("avx512bw") || !__builtin_cpu_supports("avx512cd")) return -2;
    if (n <= 0) return 0;
    std::vector<anchor_t> anc((size_t)n);
    for (int64_t i = 0; i < n; i++) anc[i] = anchor_t(r[i], q[i], l[i]);
    /* the vector loop reads up to 15 entries before index 0 of every array
     * (its last partial block), as bench-dp-chaining allocates them (+16) */
    uint32_t *ar, *aq, *al;
    create_SoA_Anchors_32_bit(&anc[0], (num_bits_t)n, ar, aq, al);
    uint32_t *fb = (uint32_t *)calloc((size_t)n + 16, 4);
    int32_t *pb = (int32_t *)calloc((size_t)n + 16, 4);
    int32_t *vb = (int32_t *)calloc((size_t)n + 16, 4);
    uint32_t *ff = fb + 16;
    int32_t *pp = pb + 16, *vv = vb + 16;
    /* create_SoA leaves the 16 leading entries uninitialised: clear them */
    memset(ar - 16, 0, 16 * 4);
    memset(aq - 16, 0, 16 * 4);
    memset(al - 16, 0, 16 * 4);
    dp_chain obj(5000, 5000, bw, 25, max_iter, 1.0f, 0, 1);
    obj.mm_dp_vectorized(n, &anc[0], ar, aq, al, ff, pp, vv, dr, dq, NULL, NULL);
    memcpy(f, ff, (size_t)n * 4);
    memcpy(p, pp, (size_t)n * 4);
    memcpy(v, vv, (size_t)n * 4);
    free(ar - 16);
    free(aq - 16);
    free(al - 16);
    free(fb);
    free(pb);
    free(vb);
    return 0;
}
