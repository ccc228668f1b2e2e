/*
 * chain_oracle.h -- CPU restatement of TAL's DP chaining,
 * dp_chain::mm_dp_vectorized (src/chaining/parallel_chaining_32_bit.h).
 *
 * TEST INFRASTRUCTURE ONLY: built into oracle/liboracle.so for tests/ and the
 * cpu_baseline leg of bench.py; the product library never links it.
 */
#ifndef CHAIN_ORACLE_H
#define CHAIN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* anchor_t, parallel_chaining_32_bit.h:56-65 (16 bytes) */
typedef struct chain_oracle_anchor {
    uint64_t r;
    int32_t q;
    int32_t l;
} chain_oracle_anchor;

/* mm_dp_vectorized on one anchor set: f (uint32), p, v (int32) of n anchors.
 * bw and max_iter are the object's; dr and dq the call's. */
void chain_oracle_dp(const chain_oracle_anchor *a, int64_t n, int32_t bw, int32_t max_iter,
                     int32_t dr, int32_t dq, uint32_t *f, int32_t *p, int32_t *v);

/* The same over nsets concatenated sets (set k = anchors [off[k], off[k+1])),
 * sets spread over nthreads. */
void chain_oracle_dp_sets(const chain_oracle_anchor *a, const int64_t *off, int32_t nsets,
                          int32_t bw, int32_t max_iter, int32_t dr, int32_t dq, uint32_t *f,
                          int32_t *p, int32_t *v, int nthreads);

/* Plants a known mistake (tests of the pins only). */
void chain_oracle_set_mutation(int m);

#ifdef __cplusplus
}
#endif
#endif
