/*
 * chain_internal.h -- data shared by the DP chaining kernels
 * (chain_kernels.hip) and their host runtime (tal_chain.hip).
 */
#ifndef TAL_CHAIN_INTERNAL_H
#define TAL_CHAIN_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/tal_chain.h"

namespace talchain {

enum {
    CHAIN_CTL_ERR = 0,
    CHAIN_CTL_COUNTER = 1,  /* task counter of the ring pass */
    CHAIN_CTL_NOVER = 2,    /* sets on the overflow list */
    CHAIN_CTL_COUNTER2 = 3, /* task counter of the overflow pass */
    CHAIN_CTL_WORDS = 4
};
enum { CHAIN_ERR_OFFSETS = 1, CHAIN_ERR_RANGE = 2 };

struct ChainParams {
    int32_t nsets;
    int32_t bw, max_iter, dr, dq;
    int32_t ring; /* ring entries per wave in LDS (power of two) */
};

struct ChainBuffers {
    const tal_chain_anchor *a;
    const int64_t *off;
    uint32_t *f;
    int32_t *p, *v;
    const int32_t *order; /* sets, longest first */
    int32_t *over;        /* the overflow list (nsets entries of room) */
    int32_t *ctl;
    unsigned long long *examined; /* sum over anchors of i - st */
};

hipError_t launch_chain_check(const tal_chain_anchor *a, const int64_t *off, int32_t nsets,
                              uint32_t *keys, int32_t *ids, int32_t *ctl,
                              unsigned long long *maxn, hipStream_t s);
hipError_t set_chain_lds_limit(size_t bytes);
size_t chain_lds_bytes(bool small, int ring);
hipError_t launch_chain_main(bool small, int grid, const ChainParams &P, const ChainBuffers &B,
                             hipStream_t s);

} // namespace talchain

#endif
