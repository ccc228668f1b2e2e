/*
 * sw_oracle.h -- CPU restatement of TAL's full (unbanded) inter-task
 * Smith-Waterman, PairWiseSW_{8,16,32}::getScores
 * (src/alignment/Smith-Waterman/).
 *
 * TEST INFRASTRUCTURE ONLY: built into oracle/liboracle.so for tests/,
 * smoke() and the cpu_baseline leg of bench.py; the product library never
 * links it.
 */
#ifndef SW_ORACLE_H
#define SW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SeqPair of PairWiseSW.h:44-49 (12 bytes). */
typedef struct sw_oracle_pair {
    int32_t id;
    int16_t len1, len2;
    int32_t score;
} sw_oracle_pair;

/* getScores of the bits-wide class (8, 16 or 32): groups of 64/32/16
 * consecutive pairs, sequences at seq_buf + 2*id*S (S = 128, 1024, 1024),
 * scores written for pairs [0, n).  Returns 0, or -1 for a bad width. */
int sw_oracle_get_scores(int bits, int32_t w_match, int32_t w_mismatch, int32_t w_open,
                         int32_t w_extend, sw_oracle_pair *pairs, const uint8_t *seq_buf,
                         int32_t n, int nthreads);

/* smithWatermanOnePair (PairWiseSW.cpp:63-114) on each pair's own lengths,
 * sequences located as above with slot S. */
void sw_oracle_scalar(int32_t w_match, int32_t w_mismatch, int32_t w_open, int32_t w_extend,
                      sw_oracle_pair *pairs, const uint8_t *seq_buf, int32_t n, int32_t slot);

/* Plants a known mistake in sw_oracle_get_scores (0 = none); tests of the
 * pins only. */
void sw_oracle_set_mutation(int m);

#ifdef __cplusplus
}
#endif
#endif
