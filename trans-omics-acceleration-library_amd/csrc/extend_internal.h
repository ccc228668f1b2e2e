/*
 * extend_internal.h -- data shared by the extension-pair producer's kernels
 * (extend_kernels.hip) and its host runtime (tal_extend.hip).
 */
#ifndef TAL_EXTEND_INTERNAL_H
#define TAL_EXTEND_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tal_extend.h"

namespace talext {

enum { EXT_ERR_ANCHOR = 1, EXT_ERR_READ = 2, EXT_ERR_OFFSETS = 3 };

struct ExtParams {
    int32_t nreads;
    int64_t genome_len;
    int32_t pad, match;
};

/* What the host reads back after the scans. */
struct ExtTotals {
    int32_t err, nleft, nright, pad_;
    int64_t ref_bytes, qer_bytes;
};

/* Per read, from the pick kernel: what the scans and the build need. */
struct ExtBuffers {
    const tal_chain_anchor *a;
    const int64_t *aoff;
    const uint32_t *f;
    const int32_t *p;
    const uint8_t *genome, *reads;
    const int64_t *read_off;
    tal_extend_read *out;
    /* per read: pair flags and byte counts (scanned into indexes and offsets) */
    int32_t *lhas, *rhas, *lidx, *ridx;
    int64_t *rbytes, *qbytes, *roff, *qoff;
    /* the seed's query range and the sides' lengths */
    int32_t *lq, *lt, *rq, *rt;
    tal_seqpair *left, *right;
    uint8_t *ref_out, *qer_out;
    int32_t *err;
};

hipError_t launch_extend_pick(const ExtParams &P, const ExtBuffers &B, hipStream_t s);
hipError_t launch_extend_totals(const ExtParams &P, const ExtBuffers &B, ExtTotals *t,
                                hipStream_t s);
hipError_t launch_extend_build(const ExtParams &P, const ExtBuffers &B, hipStream_t s);
hipError_t launch_extend_right_h0(const ExtParams &P, const ExtBuffers &B, hipStream_t s);
hipError_t launch_extend_finish(const ExtParams &P, const ExtBuffers &B, hipStream_t s);

} // namespace talext

#endif
