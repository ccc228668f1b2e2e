/*
 * bsw_simd_dispatch.c -- picks the build of the vectorised restatement
 * (bsw_simd.cpp) that the host CPU runs: AVX-512BW when the CPU has it,
 * else AVX2.  TEST INFRASTRUCTURE ONLY (the CPU baseline).
 */
#include <stdint.h>

#include "bsw_oracle.h"

int bsw_simd_run_mt_avx512(int, int, int, int, int, int, int8_t, int8_t, int, oracle_seqpair *,
                           const uint8_t *, const uint8_t *, int32_t, int32_t, int32_t, int);
int bsw_simd_run_mt_avx2(int, int, int, int, int, int, int8_t, int8_t, int, oracle_seqpair *,
                         const uint8_t *, const uint8_t *, int32_t, int32_t, int32_t, int);

/* 512 when the AVX-512BW build runs on this CPU, 256 for the AVX2 build */
int bsw_simd_width(void)
{
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512bw") ? 512 : 256;
}

int bsw_simd_run_mt(int o_del, int e_del, int o_ins, int e_ins, int zdrop, int end_bonus,
                    int8_t w_match, int8_t w_mismatch, int bits, oracle_seqpair *pairs,
                    const uint8_t *ref, const uint8_t *qer, int32_t n, int32_t w, int32_t chunk,
                    int nthreads)
{
    if (bsw_simd_width() == 512)
        return bsw_simd_run_mt_avx512(o_del, e_del, o_ins, e_ins, zdrop, end_bonus, w_match,
                                      w_mismatch, bits, pairs, ref, qer, n, w, chunk, nthreads);
    return bsw_simd_run_mt_avx2(o_del, e_del, o_ins, e_ins, zdrop, end_bonus, w_match, w_mismatch,
                                bits, pairs, ref, qer, n, w, chunk, nthreads);
}
