/*
 * bsw_oracle_mt.c -- multi-threaded driver of the oracle, used only as the
 * CPU baseline timer.  TEST INFRASTRUCTURE ONLY.
 *
 * Each thread owns its own oracle object (its own scratch), as a caller of
 * the reference must do: one BandedPairWiseSW object shares its scratch
 * through tid = 0 and is not thread-safe (bandedSWA.cpp:2756).  Threads take
 * contiguous chunks of `chunk` pairs; 20000 is bench-bsw's chunk size
 * (bandedSWA.h:51, benchmarks/bench-bsw.cpp:233-267).
 */
#include "bsw_oracle.h"

#include <omp.h>

int bsw_oracle_run_mt(int o_del, int e_del, int o_ins, int e_ins, int zdrop,
                      int end_bonus, int8_t w_match, int8_t w_mismatch,
                      int bits, oracle_seqpair *pairs, const uint8_t *ref,
                      const uint8_t *qer, int32_t n, int32_t w, int32_t chunk,
                      int nthreads)
{
    int err = 0;
    if (chunk <= 0) chunk = 20000;
#pragma omp parallel num_threads(nthreads)
    {
        bsw_oracle *o = bsw_oracle_create(o_del, e_del, o_ins, e_ins, zdrop,
                                          end_bonus, w_match, w_mismatch);
        if (!o) {
#pragma omp atomic write
            err = -1;
        }
#pragma omp for schedule(dynamic, 1)
        for (int32_t base = 0; base < n; base += chunk) {
            if (!o) continue;
            int32_t cnt = n - base < chunk ? n - base : chunk;
            if (bits == 8)
                bsw_oracle_get_scores8(o, pairs + base, ref, qer, cnt, w);
            else
                bsw_oracle_get_scores16(o, pairs + base, ref, qer, cnt, w);
        }
        bsw_oracle_destroy(o);
    }
    return err;
}
