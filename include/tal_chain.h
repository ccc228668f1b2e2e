/*
 * tal_chain.h -- C ABI of the MI355X DP chaining engine, a drop-in for TAL's
 * dp_chain::mm_dp_vectorized (src/chaining/parallel_chaining_32_bit.h), the
 * minimap2-style chaining of seed anchors.
 *
 * For each anchor i of a set (anchors sorted by r), the best predecessor j
 * within the r window and the last max_iter anchors gives
 *     f[i] = max(l_i, f[j] + min(r_i - r_j, q_i - q_j, l_i) - gap(|dr - dq|)),
 * p[i] = that j (the highest j on ties) or -1, and v[i] = the peak score on
 * the chain ending at i.  Outputs are bit-identical to the reference's
 * AVX-512BW build evaluated as written (IEEE float and double, no
 * reassociation).
 *
 * Status codes are those of tal_bsw.h.
 */
#ifndef TAL_CHAIN_H
#define TAL_CHAIN_H

#include <stddef.h>
#include <stdint.h>

#include "tal_bsw.h"

#ifdef __cplusplus
extern "C" {
#endif

/* anchor_t, parallel_chaining_32_bit.h:56-65 (16 bytes): reference position
 * r, query position q, seed length l (l >= 0). */
typedef struct tal_chain_anchor {
    uint64_t r;
    int32_t q;
    int32_t l;
} tal_chain_anchor;

/* dp_chain's constructor arguments (parallel_chaining_32_bit.h:100-118);
 * mm_dp_vectorized reads bw and max_iter, the others are kept for the
 * interface.  bench-dp-chaining uses 5000, 5000, 500, 25, 5000, 1.0, 0, 1. */
typedef struct tal_chain_params {
    int32_t max_dist_x, max_dist_y, bw, max_skip, max_iter;
    float gap_scale;
    int32_t is_cdna, n_segs;
} tal_chain_params;

typedef struct tal_chain tal_chain;

typedef struct tal_chain_stats {
    float ms_kernel;         /* the DP kernel of the last call */
    float ms_total;          /* the whole last call, device work only */
    int32_t sets;            /* anchor sets of the last call */
    int32_t lds_ring;        /* anchors in the LDS ring per wave */
    int64_t predecessors;    /* candidate predecessors examined (sum of i - st) */
    int64_t far_anchors;     /* anchors whose window reaches past the ring
                                (the older predecessors are read from HBM) */
} tal_chain_stats;

/* dp_chain::dp_chain(...) */
int tal_chain_create(const tal_chain_params *params, int device, tal_chain **out);
void tal_chain_destroy(tal_chain *h);
const char *tal_chain_last_error(const tal_chain *h);

/* mm_dp_vectorized(n, anchors, ..., f, p, v, dr, dq, NULL, NULL)
 * (parallel_chaining_32_bit.h:216-412) over nsets anchor sets at once: set k
 * is anchors[offsets[k] .. offsets[k+1]), and its outputs go to the same
 * positions of f, p and v (p indexes within the set).  Host buffers,
 * blocking.  TAL_ERANGE: a seed length is negative or a set has more than
 * INT32_MAX anchors; TAL_EINVAL: bad offsets. */
int tal_chain_dp(tal_chain *h, const tal_chain_anchor *anchors, const int64_t *offsets,
                 int32_t nsets, int32_t dr, int32_t dq, uint32_t *f, int32_t *p, int32_t *v);

/* The same with every buffer in device memory, queued on `stream` (a
 * hipStream_t, 0 = the handle's); waits once on it for the batch summary. */
int tal_chain_dp_device(tal_chain *h, const tal_chain_anchor *d_anchors,
                        const int64_t *d_offsets, int32_t nsets, int32_t dr, int32_t dq,
                        uint32_t *d_f, int32_t *d_p, int32_t *d_v, void *stream);

int tal_chain_get_stats(const tal_chain *h, tal_chain_stats *out);

#ifdef __cplusplus
}
#endif
#endif
