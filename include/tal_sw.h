/*
 * tal_sw.h -- C ABI of the MI355X full (unbanded) inter-task Smith-Waterman
 * engine, a drop-in for TAL's PairWiseSW_8 / PairWiseSW_16 / PairWiseSW_32
 * (src/alignment/Smith-Waterman/PairWiseSW_{8,16,32}.{h,cpp}).
 *
 * Scores-only local alignment with affine gaps, one score per pair, written
 * into the pair record in place.  Scores are bit-identical to the reference's
 * AVX-512BW build, including the group effects of its inter-task layout: a
 * group is 64 / 32 / 16 consecutive pairs (8 / 16 / 32-bit class), every
 * pair of a group is scored over the group's padded matrix, and lane values
 * wrap at the class width.
 *
 * Status codes are those of tal_bsw.h (TAL_OK, TAL_EINVAL, ...).
 */
#ifndef TAL_SW_H
#define TAL_SW_H

#include <stddef.h>
#include <stdint.h>

#include "tal_bsw.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SeqPair of PairWiseSW.h:44-49, 12 bytes: the pair's sequences are
 * seq_buf + 2*id*S (seq1, len1 bytes) and seq_buf + (2*id+1)*S (seq2, len2
 * bytes), S = tal_sw_slot(bits); `score` is the output. */
typedef struct tal_sw_pair {
    int32_t id;
    int16_t len1, len2;
    int32_t score;
} tal_sw_pair;

typedef struct tal_sw tal_sw;

typedef struct tal_sw_stats {
    float ms_kernel;  /* the DP kernel of the last call */
    float ms_total;   /* the whole last call, device work only */
    int32_t units;    /* 32-pair (8/16-bit) or 16-pair (32-bit) units scored */
    int32_t max_rows; /* the largest padded group bounds of the last call */
    int32_t max_cols;
    int32_t regrouped; /* 1: the pairs were regrouped by length (padding could
                          not change any score, see tal_sw_get_scores) */
    int32_t split;     /* wave quarters per unit of the DP kernel (1, 2 or 4) */
} tal_sw_stats;

/* Sequence slot S of a class: MAX_SEQ_LEN_8 = 128, MAX_SEQ_LEN_16 =
 * MAX_SEQ_LEN_32 = 1024 (PairWiseSW_8.h:45, _16.h:45, _32.h:45); 0 for any
 * other width.  It is also the longest sequence the class accepts, except
 * that the 8-bit class accepts at most 127. */
int32_t tal_sw_slot(int bits);

/* PairWiseSW::PairWiseSW(w_match, w_mismatch, w_open, w_extend),
 * PairWiseSW.cpp:43-57.  The weights are used as given (the reference adds
 * w_mismatch, w_open and w_extend, so penalties are negative numbers), each
 * truncated to the class width as _mm512_set1_epi8/16/32 does. */
int tal_sw_create(int32_t w_match, int32_t w_mismatch, int32_t w_open, int32_t w_extend,
                  int device, tal_sw **out);

/* ~PairWiseSW */
void tal_sw_destroy(tal_sw *h);

const char *tal_sw_last_error(const tal_sw *h);

/* PairWiseSW_{8,16,32}::getScores(pairArray, seqBuf, numPairs, numThreads),
 * PairWiseSW_16.cpp:136-147 (and _8 :134-145, _32 :136-152), with host
 * buffers: `seq_buf_len` bytes are copied to the device, the scores of pairs
 * [0, n) are written back.  bits = 8, 16 or 32.  TAL_ERANGE: a length is
 * negative or above the class limit; TAL_EINVAL: a pair's sequences lie
 * outside seq_buf_len or its id is negative.  Blocking.
 *
 * When mismatch, gap open and gap extend are all <= 0 at the class width, no
 * lane value can wrap (match * min(len1, len2) fits the class) and no seq1
 * holds 'D' and no seq2 holds 'B' (the padding codes), a padded cell can
 * never beat a pair's own cells, so each score is the pair's own DP: the
 * engine then scores pairs regrouped by length and skips the padding work.
 * The scores are the same either way. */
int tal_sw_get_scores(tal_sw *h, int bits, tal_sw_pair *pairs, const uint8_t *seq_buf,
                      size_t seq_buf_len, int32_t n);

/* The same with the pairs and sequences already in device memory: work is
 * queued on `stream` (a hipStream_t, 0 = the handle's stream) and the call
 * waits once on it for the batch's bounds and error word. */
int tal_sw_get_scores_device(tal_sw *h, int bits, tal_sw_pair *d_pairs, const uint8_t *d_seq_buf,
                             size_t seq_buf_len, int32_t n, void *stream);

int tal_sw_get_stats(const tal_sw *h, tal_sw_stats *out);

#ifdef __cplusplus
}
#endif
#endif
