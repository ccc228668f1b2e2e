/*
 * tal_extend.h -- C ABI of the extension-pair producer and the end-to-end
 * extension pipeline on MI355X: seed anchors of reads -> DP chaining ->
 * each read's best chain -> bwa-mem2-style left and right extension pairs
 * (tal_seqpair, h0 = seed score) -> banded SW extension of both sides.
 *
 * TAL has no counterpart: its BSW benchmark reads pairs that bwa-mem2's
 * chain-to-alignment stage produced (not part of TAL).  This producer
 * builds such pairs from chains, so that the two TAL drop-ins of this
 * library (tal_chain.h for dp_chain::mm_dp_vectorized, tal_bsw.h for
 * BandedPairWiseSW::getScores8/16) run as one pipeline on realistic inputs.
 *
 * Producer rules, per read (sequence codes 0-3, 4 = N; the anchors of a
 * read are its chaining set, sorted by r):
 *  1. The chain end is the anchor with the highest f (the first one on
 *     ties); the chain is followed back through p.  A read with no anchors
 *     has no chain and no pairs.
 *  2. The chain's anchors on one diagonal (r - q) that overlap or touch in q
 *     merge into seeds; the seed is the longest (on ties, the one with the
 *     smaller q).  Its query range is [qb, qe), its reference start rb.
 *  3. Left extension: query = read[0, qb) reversed, target = genome[rb - tl,
 *     rb) reversed with tl = min(rb, qb + pad); h0 = match * (qe - qb).
 *     Right extension: query = read[qe, m), target = genome[re, re + tr) with
 *     re = rb + (qe - qb), tr = min(genome_len - re, m - qe + pad); h0 = the
 *     left extension's score, or the seed score when there is no left pair.
 *     A side with an empty query or target has no pair.
 *  4. Left pairs go in read order to one array, right pairs to another; the
 *     targets of both to one reference buffer and the queries to one query
 *     buffer, read by read (left, then right).  seqid = read index, id = the
 *     pair's index in its array, regid = 0.
 *
 * Status codes are those of tal_bsw.h.
 */
#ifndef TAL_EXTEND_H
#define TAL_EXTEND_H

#include <stddef.h>
#include <stdint.h>

#include "tal_bsw.h"
#include "tal_chain.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One read's result. */
typedef struct tal_extend_read {
    uint64_t seed_r;      /* reference start of the seed (rb) */
    int32_t seed_q, seed_len;
    int32_t chain_score;  /* f of the chain end, 0 without anchors */
    int32_t chain_anchors; /* anchors on the chain, 0 without anchors */
    int32_t left_pair, right_pair; /* index in the left / right pair array, or -1 */
    int32_t left_score;   /* the left pair's score, or h0 when there is none */
    int32_t score;        /* the right pair's score, or left_score when there is none */
} tal_extend_read;

typedef struct tal_extend_params {
    int32_t pad;  /* target length beyond the query's on each side (bwa-mem2's band); 100 */
    int32_t w;    /* band width of the extension (bench-bsw's w); 100 */
    int32_t dr, dq; /* mm_dp_vectorized's dr, dq; 5000, 5000 */
} tal_extend_params;

typedef struct tal_extend tal_extend;

typedef struct tal_extend_stats {
    float ms_chain, ms_produce, ms_left, ms_right, ms_total; /* device time of each stage */
    int32_t reads, left_pairs, right_pairs;
    int64_t ref_bytes, qer_bytes;
} tal_extend_stats;

/* The pipeline's handle: a chaining handle (chain), a BSW handle (bsw) and
 * the producer's buffers. */
int tal_extend_create(const tal_extend_params *params, const tal_chain_params *chain,
                      const tal_bsw_params *bsw, int device, tal_extend **out);
void tal_extend_destroy(tal_extend *h);
const char *tal_extend_last_error(const tal_extend *h);

/* The whole pipeline on device buffers, queued on `stream` (a hipStream_t,
 * 0 = the handle's): genome[genome_len]; read k is reads[read_off[k] ..
 * read_off[k+1]); its anchors are anchors[anchor_off[k] .. anchor_off[k+1]).
 * Fills out[nreads].  Waits once on the stream, for the pair counts.  Both
 * extension passes run on the handle's one BSW object, left then right, as
 * two getScores16 calls on one BandedPairWiseSW would (the 16-bit path for
 * every pair, as bench-bsw calls it: bench-bsw.cpp:252-253).  TAL_ERANGE: a
 * buffer passes 2^31 bytes or a read 32767 bases.  TAL_EINVAL: read or anchor
 * offsets negative or decreasing, or a chain anchor outside its read or the
 * genome (nothing outside the offsets' ranges is read). */
int tal_extend_run_device(tal_extend *h, const uint8_t *d_genome, int64_t genome_len,
                          const uint8_t *d_reads, const int64_t *d_read_off, int32_t nreads,
                          const tal_chain_anchor *d_anchors, const int64_t *d_anchor_off,
                          tal_extend_read *d_out, void *stream);

/* The same with host buffers (copied to the device and back; blocking). */
int tal_extend_run(tal_extend *h, const uint8_t *genome, int64_t genome_len, const uint8_t *reads,
                   const int64_t *read_off, int32_t nreads, const tal_chain_anchor *anchors,
                   const int64_t *anchor_off, tal_extend_read *out);

/* The producer alone (steps 1-4) from chaining results already in HBM (f
 * and p of the anchors): fills out[nreads] but the scores, and the pair
 * arrays and buffers of the handle, which tal_extend_get_pairs copies out.
 * Right pairs carry h0 = the seed score; the pipeline sets it from the left
 * scores.  Errors as for tal_extend_run_device. */
int tal_extend_produce_device(tal_extend *h, const uint8_t *d_genome, int64_t genome_len,
                              const uint8_t *d_reads, const int64_t *d_read_off, int32_t nreads,
                              const tal_chain_anchor *d_anchors, const int64_t *d_anchor_off,
                              const uint32_t *d_f, const int32_t *d_p, tal_extend_read *d_out,
                              void *stream);

/* Copies the last call's pairs and buffers to host memory (blocking); any
 * pointer may be null.  Sizes are in tal_extend_get_stats. */
int tal_extend_get_pairs(tal_extend *h, tal_seqpair *left, tal_seqpair *right, uint8_t *ref,
                         uint8_t *qer);

int tal_extend_get_stats(const tal_extend *h, tal_extend_stats *out);

#ifdef __cplusplus
}
#endif
#endif
