/*
 * bsw_oracle.h -- CPU restatement of TAL's banded Smith-Waterman extension
 * (BandedPairWiseSW, src/alignment/BSW/bandedSWA.{h,cpp}).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; it is the checker that tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py compare the HIP path against.
 *
 * Three independent restatements live here:
 *   - bsw_oracle_get_scores16 / _8 : the AVX-512BW vector path
 *     (bandedSWA.cpp:2663-3356 and :1969-2661), emulated lane by lane with the
 *     same 32-lane (16-bit) / 64-lane (8-bit) pair groups, the same wrap-around
 *     integer widths and the same persistent per-object scratch.
 *   - bsw_oracle_scalar : the scalar ksw_extend2-style routine
 *     (bandedSWA.cpp:116-237), one pair at a time.
 *   - bsw_oracle_band_dp : a full-band textbook DP with no band narrowing,
 *     used only to pin the scalar restatement on small inputs.
 */
#ifndef TAL_BSW_ORACLE_H
#define TAL_BSW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same 56-byte layout as SeqPair (bandedSWA.h:90-99). */
typedef struct oracle_seqpair {
    int32_t idr, idq, id;
    int32_t len1, len2;
    int32_t h0;
    int32_t seqid, regid;
    int32_t score, tle, gtle, qle;
    int32_t gscore, max_off;
} oracle_seqpair;

typedef struct bsw_oracle bsw_oracle;

/* Mirrors the BandedPairWiseSW constructor (bandedSWA.cpp:50-94): w_mismatch
 * is the positive penalty; the object stores its negation (:65) and uses a
 * fixed ambiguity score of -1 (:68). */
bsw_oracle *bsw_oracle_create(int o_del, int e_del, int o_ins, int e_ins,
                              int zdrop, int end_bonus,
                              int8_t w_match, int8_t w_mismatch);
void bsw_oracle_destroy(bsw_oracle *o);

/* Restated getScores16 / getScores8.  Outputs are written into pairs[0..n).
 * Pairs past n that the reference pads a group with are supplied internally
 * (len1 = 0, len2 = 0 for 16-bit / last pair's len2 for 8-bit, h0 = 0,
 * idr = idq = 0).  Preconditions (the reference has undefined behaviour
 * otherwise): 16-bit len1,len2 <= 32767; 8-bit len1,len2 <= 127; e_ins and
 * e_del nonzero. */
void bsw_oracle_get_scores16(bsw_oracle *o, oracle_seqpair *pairs,
                             const uint8_t *ref, const uint8_t *qer,
                             int32_t n, int32_t w);
void bsw_oracle_get_scores8(bsw_oracle *o, oracle_seqpair *pairs,
                            const uint8_t *ref, const uint8_t *qer,
                            int32_t n, int32_t w);

/* Restated scalarBandedSWAWrapper (bandedSWA.cpp:242-260).  Codes must be
 * 0..4.  Match/mismatch/ambiguity come from the object's scores, laid out
 * like bwa_fill_scmat (bench-bsw.cpp:56-64). */
void bsw_oracle_scalar(bsw_oracle *o, oracle_seqpair *pairs,
                       const uint8_t *ref, const uint8_t *qer,
                       int32_t n, int32_t w);

/* The same with a caller-given 5x5 matrix, mat[r * 5 + q] (reference code r,
 * query code q), as the reference's scalar path uses the matrix its object
 * was constructed with (bandedSWA.cpp:53, :138-141). */
void bsw_oracle_scalar_mat(bsw_oracle *o, const int8_t *mat, oracle_seqpair *pairs,
                           const uint8_t *ref, const uint8_t *qer, int32_t n, int32_t w);

/* Full-band DP on one pair: every cell with |i-j| inside the per-pair band is
 * evaluated, no beg/end narrowing.  Returns score, writes the other outputs. */
int bsw_oracle_band_dp(const bsw_oracle *o, int qlen, const uint8_t *query,
                       int tlen, const uint8_t *target, int32_t w, int h0,
                       int *qle, int *tle, int *gtle, int *gscore,
                       int *max_off);

/* Experiment switches.  mode 0 (default) keeps the object's F scratch as the
 * reference does, so a group's stale F[ncol] read sees what earlier groups
 * left; mode 1 clears F[ncol] at the start of every group. */
void bsw_oracle_set_stale_mode(bsw_oracle *o, int mode);
/* Plants a known mistake in the 16-bit restatement (tests only):
 * 1 strict row-max tie rule, 2 z-drop offset sign, 3 ins/del penalties
 * swapped, 4 first-column boundary one row early, 5 ambiguity ignored,
 * 6 gscore not gated by lane exit, 7 16-bit range reset before the head/tail
 * scans dropped, 8 16-bit F of the column one past a lane's band (j == tail)
 * stored as 0.  0 restores the restatement. */
void bsw_oracle_set_mutation(bsw_oracle *o, int m);
/* Counters since creation: [0] groups, [1] stale F[ncol] reads in which the
 * stale value decided some lane's tail, [2] such reads where a lane's stale
 * value was nonzero. */
void bsw_oracle_stats(const bsw_oracle *o, uint64_t out[3]);
/* Zero mask (bit l = lane l has F != 0) of the 16-bit scratch column c. */
uint32_t bsw_oracle_f16_nonzero_mask(const bsw_oracle *o, int c);

/* Cells inside the per-pair band, the algorithmic work unit:
 * sum_i max(0, min(len2, i+b+1) - max(0, i-b)), b = per-pair band
 * (bandedSWA.cpp:2882-2904). */
uint64_t bsw_oracle_band_cells(const bsw_oracle *o, int len1, int len2,
                               int32_t w);

#ifdef __cplusplus
}
#endif
#endif
