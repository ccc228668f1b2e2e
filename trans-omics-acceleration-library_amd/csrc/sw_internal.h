/*
 * sw_internal.h -- data shared by the full-SW kernels (sw_kernels.hip) and
 * their host runtime (tal_sw.hip).
 */
#ifndef TAL_SW_INTERNAL_H
#define TAL_SW_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/tal_sw.h"

namespace talsw {

/* control words of one call */
enum {
    SW_CTL_ERR = 0,
    SW_CTL_MAXROW = 1,
    SW_CTL_MAXCOL = 2,
    SW_CTL_COUNTER = 3,
    SW_CTL_NOT_INERT = 4, /* some pair could make the padding matter (regrouping off) */
    SW_CTL_ELIGIBLE = 5,  /* pairs outside zero-bound groups */
    SW_CTL_WORDS = 6
};
/* error word values (largest wins) */
enum { SW_ERR_BOUNDS = 1, SW_ERR_RANGE = 2 };

struct SwParams {
    int32_t n;      /* pairs */
    int32_t nunits; /* 32-pair (8/16-bit) or 16-pair (32-bit) units */
    int32_t slot;   /* sequence slot S of the class */
    int32_t rows_cap, cols_cap; /* slab capacity in rows / columns (multiples of 4) */
    int32_t aligned4;           /* the sequence buffer is 4-byte aligned */
    int32_t m;                  /* valid entries of SwBuffers::perm */
    size_t slab_words;          /* 32-bit words per wave slab */
    /* lane constants.  8/16-bit: two packed 16-bit halves (8-bit: value << 8)
     * and k_delta = mismatch - match; 32-bit: plain int32 and k_delta =
     * mismatch. */
    uint32_t k_match, k_delta, k_open, k_ext;
};

struct SwBuffers {
    tal_sw_pair *pairs;
    const uint8_t *seq;
    const int32_t *unit_dims; /* (nrow << 16) | ncol per unit */
    const int32_t *order;     /* units in sweep order */
    const int32_t *perm;      /* regrouped pairs: unit u, slot k = pair perm[u*UL + k];
                                 null = the reference's consecutive groups */
    int32_t *ctl;             /* SW_CTL_* words */
    uint32_t *slabs;          /* one slab per block of the persistent grid */
};

/* What the regrouped sweep needs of the bounds kernel: per-pair sort keys
 * and ids, and the checks that make regrouping exact. */
struct SwInert {
    int32_t on;       /* the weights allow regrouping */
    int32_t match;    /* match weight at the class width */
    int64_t lane_max; /* largest lane value of the class */
    uint32_t *pair_keys;
    int32_t *pair_ids;
};

hipError_t launch_sw_shapes(const tal_sw_pair *pairs, int32_t n, int32_t bits, size_t seq_len,
                            int32_t *unit_dims, uint32_t *keys, int32_t *ids, int32_t *ctl,
                            const SwInert &I, hipStream_t s);
hipError_t launch_sw_scan_codes(const tal_sw_pair *pairs, int32_t n, int32_t bits,
                                const uint8_t *seq, size_t seq_len, int32_t *ctl, hipStream_t s);
hipError_t launch_sw_unit_dims(const tal_sw_pair *pairs, const int32_t *perm, int32_t m,
                               int32_t bits, int32_t *unit_dims, uint32_t *keys, int32_t *ids,
                               hipStream_t s);
hipError_t launch_sw_zero_scores(tal_sw_pair *pairs, const int32_t *idx, int32_t count,
                                 hipStream_t s);
hipError_t launch_sw_main(int bits, int split, int grid, const SwParams &P, const SwBuffers &B,
                          hipStream_t s);

} // namespace talsw

#endif
