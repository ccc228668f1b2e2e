/*
 * bsw_internal.h -- device-side data structures shared by the kernels
 * (bsw_kernels.hip) and the host runtime (tal_bsw.hip).
 */
#ifndef TAL_BSW_INTERNAL_H
#define TAL_BSW_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tal_bsw.h"

namespace talbsw {

/* Largest query length the 16-bit path accepts (MAX_SEQ_LEN16 - 1); the
 * per-handle stale-column image covers columns 0..MAX_COL16. */
constexpr int MAX_LEN16 = 32767;
constexpr int MAX_LEN8 = 127;
constexpr int IMAGE_COLS = MAX_LEN16 + 1;

/* Scoring and band parameters, pre-scaled for the lane width: 16-bit lanes
 * hold the value itself, 8-bit lanes hold value << 8 in an int16 so that
 * 16-bit add/sub/max/min wrap and compare exactly as epi8 does. */
struct KParams {
    int32_t n;        /* pairs in this call */
    int32_t ngroups;  /* ceil(n / group lanes) */
    int32_t ntasks;   /* wave tasks in this launch */
    int32_t w;        /* band width argument */
    int32_t o_del, e_del, o_ins, e_ins, end_bonus;
    int32_t smax;      /* max(0, w_match, -w_mismatch, ambig) */
    int32_t last_len2; /* 8-bit padding lanes take the last pair's len2 */
    int32_t state_cols;  /* state columns per wave (max ncol + slack) */
    int32_t lds_cols;    /* state columns kept in LDS (multiple of 8) */
    int32_t slab_cols;   /* 64-dword rows per global slab: query words, the state
                            past lds_cols in whole 8-column blocks, and one
                            8-column scratch block */
    int32_t query_words; /* query-code words per lane, at the start of the slab */
    int16_t match, mism, ambig;
    int16_t e_del_s, oe_del_s, e_ins_s, oe_ins_s, zdrop_s;
};

struct KBuffers {
    tal_seqpair *pairs;
    const uint8_t *ref;
    const uint8_t *qer;
    const int32_t *task_groups; /* ntasks * groups-per-wave ids, -1 = none;
                                   null = identity order */
    int32_t *task_counter;      /* persistent-grid work counter */
    uint32_t *slabs;            /* per-wave slabs: query words, state past LDS */
    /* 16-bit stale-column bookkeeping (all null for the 8-bit path) */
    const uint32_t *stale_in; /* per group: lanes whose F[ncol] is nonzero */
    uint32_t *dep_out;        /* per group: lanes whose tail depended on it */
    uint8_t *wrote_out;       /* per group: F[ncol] stored during the group */
    uint32_t *fmask;          /* per group final F nonzero masks, c < ncol */
    const int64_t *fmask_off; /* per group offset into fmask */
    /* graph-replayed batches: values that live in device memory (null =
     * use the KParams field) */
    const int32_t *ntasks_dev;    /* tasks of this launch */
    const int32_t *last_len2_dev; /* 8-bit padding length */
    const int32_t *abort_flag;    /* nonzero: the batch left its plan; do nothing */
};

/* Parameters of the per-pair (scalar) kernel: the object's gaps and the
 * caller's 5x5 matrix (row = reference code, column = query code). */
struct ScalarParams {
    int32_t n, w, max_qlen;
    int32_t o_del, e_del, o_ins, e_ins, zdrop, end_bonus;
    int32_t mat_max; /* max(0, every entry of mat) */
    int8_t mat[28];
};

hipError_t launch_scalar_cost(const tal_seqpair *pairs, int32_t n, int32_t w, uint32_t *key,
                              int32_t *ids, int32_t *maxlen, hipStream_t s);
hipError_t launch_scalar(int grid, const ScalarParams &P, tal_seqpair *pairs, const uint8_t *ref,
                         const uint8_t *qer, const int32_t *order, int32_t ntasks,
                         int32_t *counter, int2 *slabs, int32_t *err, hipStream_t s);

/* Per-group shape: rows (max len1) and columns (max len2) of the group. */
struct GroupShape {
    int32_t nrow, ncol;
};

/* Host-side entry points shared by the C ABI (tal_bsw.hip) and the serving
 * loop (bsw_server.hip); not exported from the library. */
#define TAL_INTERNAL __attribute__((visibility("hidden")))

/* One width-specific batch on device buffers of the handle's device, queued on
 * stream s.  seg_first (device, one int per 16-bit group, may be null) splits
 * the batch into independent calls, each scored as on a fresh object. */
TAL_INTERNAL int run_batch(tal_bsw *h, int bits, tal_seqpair *pairs, const uint8_t *ref,
                           const uint8_t *qer, int32_t n, int32_t w, hipStream_t s,
                           const int32_t *seg_first);
TAL_INTERNAL hipStream_t handle_stream(const tal_bsw *h);
/* Largest score of the parameter set (match, -mismatch, ambiguity -1, 0): the
 * 8/16-bit dispatcher's bound. */
TAL_INTERNAL int score_bound(const tal_bsw_params &p);
/* The dispatcher rule: 1 if the pair goes to the 8-bit path. */
TAL_INTERNAL int goes_8bit(const tal_seqpair &p, int smax);

} // namespace talbsw

#endif
