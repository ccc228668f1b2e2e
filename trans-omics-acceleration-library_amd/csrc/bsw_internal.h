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

/* Per-group shape: rows (max len1) and columns (max len2) of the group. */
struct GroupShape {
    int32_t nrow, ncol;
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
    uint32_t *stale_in;       /* per group: lanes whose F[ncol] is nonzero */
    uint32_t *dep_out;        /* per group: lanes whose tail depended on it */
    uint8_t *wrote_out;       /* per group: F[ncol] stored during the group */
    uint32_t *fmask;          /* per group final F nonzero masks, c < ncol */
    const int64_t *fmask_off; /* per group offset into fmask */
    /* graph-replayed batches: values that live in device memory (null =
     * use the KParams field) */
    const int32_t *ntasks_dev;    /* tasks of this launch */
    const int32_t *last_len2_dev; /* 8-bit padding length */
    const int32_t *abort_flag;    /* nonzero: the batch left its plan; do nothing */
    /* 16-bit groups settled inside the launch (null gstate: none are) */
    const GroupShape *shapes;     /* per group rows and columns */
    const uint32_t *image;        /* F[ncol] masks before the batch (per column) */
    int32_t *gstate;              /* per group: 1 = final (zeroed before the launch) */
    int32_t *early;               /* groups run again inside the launch (counter) */
    int32_t *pending;             /* [0] groups listed, [1] taken, then one slot per
                                     group for those left pending (null: no
                                     end-of-launch pass; before the launch [0] and
                                     [1] are zeroed and the slots set to -1) */
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


/* ---- device-resident serving (bsw_server.hip, resident mode) ----
 * Everything below lives in pinned host memory, written by the host and read
 * by the persistent serving kernel, except the outputs, the stale-column
 * bookkeeping and the done flags, which the kernel writes with plain stores
 * (no atomics on host memory). */

/* One task of request slot `req`: bits 16, two 16-bit groups (one per
 * half-wave, g1 = -1 for none); bits 8, one 8-bit group g0; bits 0, the
 * stale-column resolution of the request's 16-bit groups. */
/* A task of the persistent serving kernel for request slot `req`: bits 16, a
 * pair of 16-bit groups g0, g1 (g1 = -1: one); bits 8, the 8-bit group g0;
 * bits 0, the stale-column resolution of the request's 16-bit groups; bits
 * SERVE_COPY_IN / SERVE_COPY_OUT, chunk g0 (SERVE_COPY_CHUNK bytes) of the
 * request's copy in or out. */
struct ServeTask {
    int32_t req, bits, g0, g1;
};
constexpr int32_t SERVE_COPY_IN = 1, SERVE_COPY_OUT = 2;
constexpr int64_t SERVE_COPY_CHUNK = (int64_t)256 << 10;

/* A request slot.  Group g of the 16-bit (8-bit) block holds pairs
 * [32g, 32g+32) ([64g, 64g+64)) of pairs16 (pairs8); idr/idq of every record
 * are offsets into `seq`.  Device memory: the sequences, the pair blocks
 * (outputs written in place) and the 16-bit stale-column bookkeeping; pinned
 * host memory: fmask_off, ncol16, stale and reruns, and a mirror of the
 * device layout's first cin_bytes (sequences and pair blocks), which the
 * copy tasks move in (cin_*) and, for the pair blocks, back out (cout_*).
 * Every copy address and size is a multiple of 256 bytes. */
struct ServeReq {
    const uint8_t *cin_src;
    uint8_t *cin_dst;
    const uint8_t *cout_src;
    uint8_t *cout_dst;
    int64_t cin_bytes, cout_bytes;
    tal_seqpair *pairs16, *pairs8;
    const uint8_t *seq;
    uint32_t *stale;   /* per 16-bit group: lanes whose F[ncol] starts nonzero */
    uint32_t *dep;     /* per 16-bit group */
    uint8_t *wrote;    /* per 16-bit group */
    uint32_t *fmask;   /* final F nonzero masks, ncol16[g] words per group */
    const int64_t *fmask_off;
    const int32_t *ncol16; /* columns of each 16-bit group */
    int32_t *reruns;   /* a resolve task's result: count, then the groups */
    int32_t n16, n8, w, pad;
};

/* Control words: the host publishes tasks [0, tail) of the ring and sets
 * stop when the server closes.  Each finished task t writes (int32)(t + 1)
 * into done[t % ring capacity]. */
struct ServeCtl {
    int64_t tail;
    int32_t stop;
    int32_t pad[13];
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
/* The scoring parameters of a handle as the kernels take them for one lane
 * width (n, w and the sizing fields left zero). */
TAL_INTERNAL KParams params_for(const tal_bsw_params &p, int bits);
/* State placement of a wave for groups of up to max_ncol columns with
 * waves_per_cu waves sharing a CU's LDS (tal_bsw.hip). */
struct Sizing {
    int state_cols, qwords, lds_cols, slab_cols;
    size_t lds_bytes;
};
TAL_INTERNAL Sizing state_sizing(size_t lds_per_cu, size_t lds_per_block, int waves_per_cu,
                                 int max_ncol);
/* The persistent serving kernel (bsw_kernels.hip). */
constexpr int SERVE_DEV_WORDS = 8;
constexpr int SERVE_CU_KEYS = 2048; /* per-CU rank counters past the words */
constexpr size_t SERVE_DEV_BYTES = SERVE_DEV_WORDS * sizeof(int64_t) + SERVE_CU_KEYS * sizeof(int32_t);
void serve_dev_init(int64_t words[SERVE_DEV_WORDS]);
/* dev: the kernel's SERVE_DEV_WORDS device words -- the claim head, its copy
 * of the host's tail and of the stop flag, the count of worker waves on the
 * GPU, the real-time clock of the last claim or finish, the closing flag and
 * the lowest handed-back ticket, set up by serve_dev_init before the first
 * launch; then SERVE_CU_KEYS rank counters (SERVE_DEV_BYTES in all). */
TAL_INTERNAL hipError_t launch_serve(int grid, size_t lds_bytes, const KParams &K16,
                                     const KParams &K8, const ServeTask *ring, int32_t ring_mask,
                                     const ServeReq *reqs, const ServeCtl *ctl, int32_t *done,
                                     int64_t *dev, uint32_t *slabs, int64_t idle_ticks,
                                     int32_t pollers, hipStream_t s);
TAL_INTERNAL hipError_t set_serve_lds_limit(size_t bytes);

} // namespace talbsw

#endif
