/*
 * bsw_synth.h -- deterministic synthetic extension-pair workloads.
 *
 * Every pair k of workload `cfg` draws from its own xoshiro256** stream,
 * seeded by splitmix64(seed + k * 0x9E3779B97F4A7C15), so any prefix or
 * sample of a workload regenerates byte-for-byte on any host running this
 * image.  Sequences are codes 0..3 (A,C,G,T) plus 4 (N).
 *
 * Workloads (pair counts are the caller's choice):
 *   1  ref U[151,250] iid; query = 151 bp from the ref start with 1% subs and
 *      0.2% 1-3 bp indels; h0 U[10,60]; w = 100.
 *   2  query U[30,250]; ref = query length + U[0,100] flank; 2% subs, 0.5%
 *      indels; h0 U[0,120]; w = 100.
 *   3  query lognormal(median 400, sigma 0.6) clipped [100,2000]; ref = query
 *      + U[0,100]; 2.5% subs + 0.5% indels; h0 U[0,120]; w in {50,100,200,500}.
 *   5  workload-1 lengths; 5% of pairs carry N at 1-5% of positions, 10% are
 *      low-complexity repeats (period 1-6), 10% have 20% mismatches;
 *      h0 log-uniform in [10,30000]; w = 100.
 */
#ifndef TAL_BSW_SYNTH_H
#define TAL_BSW_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-pair lengths, h0 and band width of pairs [first, first+n).  Any output
 * pointer may be NULL. */
int tal_synth_shape(int cfg, uint64_t seed, int64_t first, int32_t n,
                    int32_t *len1, int32_t *len2, int32_t *h0, int32_t *w);

/* Fill a 56-byte SeqPair array and the packed ref/query buffers for pairs
 * [first, first+n).  ref_off/qer_off are each pair's byte offsets (from an
 * exclusive prefix sum of len1/len2 the caller computed); the pair records get
 * idr/idq = those offsets, id = first + k, seqid = regid = -1, outputs = -1. */
int tal_synth_fill(int cfg, uint64_t seed, int64_t first, int32_t n,
                   const int64_t *ref_off, const int64_t *qer_off,
                   void *pairs, uint8_t *ref, uint8_t *qer, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
