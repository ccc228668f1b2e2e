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
    CHAIN_CTL_COUNTER = 1, /* task counter of the DP kernel */
    CHAIN_CTL_ANCHORS = 2, /* two words: the batch's anchors, low and high half */
    CHAIN_CTL_WORDS = 4
};
enum { CHAIN_ERR_OFFSETS = 1, CHAIN_ERR_RANGE = 2 };

struct ChainParams {
    int32_t nsets;
    int32_t bw, max_iter, dr, dq;
    int32_t ring;       /* LDS ring slots per wave (power of two, >= 128) */
    int32_t table;      /* gap-cost table entries per block (0: none; else >= bw + 2) */
};

/* LDS bytes of one block: the ring (12 bytes a slot), the gap table, the
 * waves' phase-A bests (8 bytes a lane) and the task word */
__host__ __device__ inline size_t chain_block_lds(int ring, int table, int waves_per_block)
{
    return 12 * (size_t)ring + 4 * (size_t)table + 512 * (size_t)waves_per_block + 16;
}

struct ChainBuffers {
    const tal_chain_anchor *a;
    const int64_t *off;
    uint32_t *f;
    int32_t *p, *v;
    const int32_t *order; /* sets, longest first */
    int32_t *ctl;
    unsigned long long *examined;    /* sum over anchors of i - st */
    unsigned long long *far_anchors; /* anchors with predecessors older than the ring */
};

hipError_t launch_chain_check(const tal_chain_anchor *a, const int64_t *off, int32_t nsets,
                              uint32_t *keys, int32_t *ids, int32_t *ctl, hipStream_t s);
hipError_t set_chain_lds_limit(size_t bytes);
/* LDS bytes of one block of the DP kernel */
size_t chain_lds_bytes(int ring, int table, int waves_per_block);
/* waves_per_block: 1, 2 or 4; the gap table (table mode) is used when
 * P.table > 0, which needs bw + max(min(dq, dr), 0) < 2^31 */
hipError_t launch_chain_main(int waves_per_block, int grid, const ChainParams &P,
                             const ChainBuffers &B, hipStream_t s);

} // namespace talchain

#endif
