/*
 * tal_bsw.h -- C ABI of the MI355X banded Smith-Waterman extension engine.
 *
 * Drop-in for TAL's BandedPairWiseSW (src/alignment/BSW/bandedSWA.h:113-336):
 * the same 56-byte SeqPair records, the same packed reference/query byte
 * buffers addressed by per-pair offsets, the same scoring parameters, and the
 * same six outputs written in place, bit-identical to a line-by-line CPU
 * restatement of the AVX-512BW build of getScores16 / getScores8 (the
 * reference's BSW source needs bwa-mem2's macro.h, which it does not ship, so
 * it is not compiled for comparison).  Plain C types only; every call returns
 * a status code and never exits the process.
 *
 * Group semantics.  Like the reference, getScores16 processes pairs in
 * consecutive groups of 32 and getScores8 in groups of 64; pairs of one group
 * share the row loop, the column range and the band narrowing, so a pair's
 * outputs may depend on the pairs next to it.  The engine reproduces those
 * groups exactly.  The last group of a call is padded with pairs of
 * len1 = 0, h0 = 0, idr = idq = 0 and len2 = 0 (16-bit) or the last pair's
 * len2 (8-bit); the reference pads the same lengths but reads h0/idq from
 * whatever lies past the caller's array (bandedSWA.cpp:2710-2717,
 * :2013-2018), and unlike the reference nothing is written past pairs[n-1].
 *
 * Persistent scratch.  The reference object keeps its 16-bit F scratch
 * across groups and calls, and its tail scan can read column F[ncol] before
 * the current group stores it (bandedSWA.cpp:3034-3035, :3267, :3293-3302).
 * A handle keeps the same state on the device, starting from zero like a
 * freshly constructed object, so a sequence of calls on one handle matches
 * the same sequence on one reference object.
 */
#ifndef TAL_BSW_H
#define TAL_BSW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define TAL_OK 0
#define TAL_EINVAL 1  /* bad argument (null pointer, n < 0, bad parameter) */
#define TAL_ENOMEM 2  /* device or pinned-host allocation failed */
#define TAL_EDEVICE 3 /* HIP runtime error; see tal_bsw_last_error */
#define TAL_ERANGE 4  /* a pair exceeds the path's limits (see below) */
#define TAL_EAGAIN 5  /* a request is not finished yet (serving loop) */

/* SeqPair, bandedSWA.h:90-99.  idr/idq: byte offsets of the reference
 * (rows, len1) and query (columns, len2) in their buffers; h0: initial score;
 * seqid/regid: passed through.  Outputs: score, tle, gtle, qle, gscore,
 * max_off. */
typedef struct tal_seqpair {
    int32_t idr, idq, id;
    int32_t len1, len2;
    int32_t h0;
    int32_t seqid, regid;
    int32_t score, tle, gtle, qle;
    int32_t gscore, max_off;
} tal_seqpair;

/* Constructor arguments of BandedPairWiseSW (bandedSWA.cpp:50-94).
 * w_mismatch is the positive penalty, as the constructor takes it; the
 * ambiguity score is fixed at -1 (DEFAULT_AMBIG, bandedSWA.h:54).  The
 * reference's mat/numThreads arguments only feed its scalar path and its
 * per-thread scratch sizing, so they have no counterpart here. */
typedef struct tal_bsw_params {
    int32_t o_del, e_del, o_ins, e_ins;
    int32_t zdrop, end_bonus;
    int32_t w_match, w_mismatch;
} tal_bsw_params;

typedef struct tal_bsw tal_bsw;

/* Replaces BandedPairWiseSW::BandedPairWiseSW (bandedSWA.cpp:50-94).
 * Binds the handle to HIP device `device`.  e_del and e_ins must be nonzero
 * (the reference divides by them, :2887-2903); w_match and w_mismatch must
 * fit int8_t as in the reference signature. */
int tal_bsw_create(const tal_bsw_params *params, int device, tal_bsw **out);

/* Replaces BandedPairWiseSW::~BandedPairWiseSW (bandedSWA.cpp:97-100). */
void tal_bsw_destroy(tal_bsw *h);

/* Human-readable description of the last failing call on this handle. */
const char *tal_bsw_last_error(const tal_bsw *h);

/* Replaces BandedPairWiseSW::getScores16 (bandedSWA.cpp:2664-2688).
 * pairs/ref/qer are host memory; ref_len/qer_len are the buffer extents
 * (every idr+len1 <= ref_len, idq+len2 <= qer_len).  Limits: 0 <= len1,len2
 * <= 32767 (MAX_SEQ_LEN16 - 1), w >= 0.  Blocking. */
int tal_bsw_get_scores16(tal_bsw *h, tal_seqpair *pairs, const uint8_t *ref,
                         size_t ref_len, const uint8_t *qer, size_t qer_len,
                         int32_t n, int32_t w);

/* Replaces BandedPairWiseSW::getScores8 (bandedSWA.cpp:1970-1995).  Limits:
 * 0 <= len1,len2 <= 127 (MAX_SEQ_LEN8 - 1), w >= 0.  Blocking. */
int tal_bsw_get_scores8(tal_bsw *h, tal_seqpair *pairs, const uint8_t *ref,
                        size_t ref_len, const uint8_t *qer, size_t qer_len,
                        int32_t n, int32_t w);

/* 8/16-bit dispatcher.  The reference leaves this split to its caller
 * (bench-bsw always calls getScores16, benchmarks/bench-bsw.cpp:253).  A pair
 * goes to the 8-bit path iff len1 <= 127, len2 <= 127 and
 * h0 + max(w_match,0) * min(len1,len2) <= 127, so no 8-bit score can wrap;
 * the rest go to the 16-bit path.  Results equal getScores8 on the 8-bit pairs
 * (in input order) followed by getScores16 on the others, on this handle. */
int tal_bsw_get_scores_auto(tal_bsw *h, tal_seqpair *pairs, const uint8_t *ref,
                            size_t ref_len, const uint8_t *qer, size_t qer_len,
                            int32_t n, int32_t w);

/* Device-resident variants: pairs/ref/qer are device pointers on the handle's
 * device; outputs are written in place into the device pairs.  `stream` is a
 * hipStream_t (NULL = the handle's own stream).  The call waits on the stream
 * for a few status words: the batch's largest group (the kernels' buffers are
 * sized from it), the 8-bit path's padding length, the dispatcher's 8-bit
 * count, and, in the 16-bit path, the groups whose stale-column read needs a
 * rerun; the DP itself and the result writes stay queued.  No host copies of
 * the inputs are made.  As in the reference, which reads seqBufRef + idr and
 * seqBufQer + idq unchecked (bandedSWA.cpp:2806, 2848), keeping every pair's
 * sequences inside ref and qer is the caller's part; a negative length, or
 * one over the path's limit, gives TAL_ERANGE. */
int tal_bsw_get_scores16_device(tal_bsw *h, tal_seqpair *pairs,
                                const uint8_t *ref, const uint8_t *qer,
                                int32_t n, int32_t w, void *stream);
int tal_bsw_get_scores8_device(tal_bsw *h, tal_seqpair *pairs,
                               const uint8_t *ref, const uint8_t *qer,
                               int32_t n, int32_t w, void *stream);
int tal_bsw_get_scores_auto_device(tal_bsw *h, tal_seqpair *pairs,
                                   const uint8_t *ref, const uint8_t *qer,
                                   int32_t n, int32_t w, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code that must not call hipMalloc, tight serving loops): the same
 * call as tal_bsw_get_scores_auto_device, with the same results, but every
 * piece of call scratch comes from a static device array of
 * TAL_BSW_STATIC_SCRATCH_BYTES that the library holds on each device until
 * the process ends; the call makes no hipMalloc or hipFree.  The handle's own
 * buffers from tal_bsw_create (the 16-bit F-scratch image) are used as in the
 * other calls.  A call needs about 70 bytes per pair, plus per 16-bit group of
 * 32 pairs 4 bytes per column of its longest query, plus one wave-state slab
 * (about 256 bytes per column of the longest query); the kernels then run as
 * many blocks as the rest of the array holds slabs for.  A call whose
 * scratch does not fit returns TAL_ERANGE: before any work when the pairs
 * alone do not fit, else possibly after the 8-bit pairs' outputs were written.
 * All handles on one device share the one array, which keeps the last call's
 * contents: calls that overlap on one device race and their results are
 * undefined. */
#define TAL_BSW_STATIC_SCRATCH_BYTES ((size_t)64 << 20)
int tal_bsw_get_scores_auto_device_noalloc(tal_bsw *h, tal_seqpair *pairs,
                                           const uint8_t *ref,
                                           const uint8_t *qer, int32_t n,
                                           int32_t w, void *stream);

/* Replaces BandedPairWiseSW::scalarBandedSWAWrapper (bandedSWA.cpp:242-260),
 * and with n = 1 scalarBandedSWA (:116-237): the textbook per-pair extension,
 * each pair scored on its own (no group coupling; its own column range,
 * narrowing and exits), in int32 arithmetic, with the 5x5 scoring matrix
 * `mat` (host memory; mat[r * 5 + q] scores reference code r against query
 * code q, as the reference indexes the matrix it was constructed with, which
 * bench-bsw fills with bwa_fill_scmat, benchmarks/bench-bsw.cpp:56-64).  The
 * gap and end-bonus parameters are the handle's.  Every sequence byte must be
 * a code 0-4, since the reference indexes mat with it; another byte, or a
 * length over 32767, gives TAL_ERANGE.  The host form is blocking; the device
 * form takes pairs/ref/qer in device memory and waits on the stream for the
 * batch's longest length and the error word. */
int tal_bsw_scalar_scores(tal_bsw *h, const int8_t *mat, tal_seqpair *pairs, const uint8_t *ref,
                          size_t ref_len, const uint8_t *qer, size_t qer_len, int32_t n,
                          int32_t w);
int tal_bsw_scalar_scores_device(tal_bsw *h, const int8_t *mat, tal_seqpair *pairs,
                                 const uint8_t *ref, const uint8_t *qer, int32_t n, int32_t w,
                                 void *stream);

/* Number of pairs of pairs[0..n) that the dispatcher sends to the 8-bit path
 * (host memory). */
int32_t tal_bsw_count_8bit(const tal_bsw *h, const tal_seqpair *pairs,
                           int32_t n);

/* Timing and shape of the last call on the handle (no reference
 * counterpart; the reference's RDT tick counters, bandedSWA.cpp:2947-2953,
 * play this role on the CPU).  Kernel times come from HIP events recorded on
 * the stream the kernels ran on; reading them waits for that stream. */
typedef struct tal_bsw_stats {
    float ms_main16;    /* 16-bit DP kernel, including stale-column reruns */
    float ms_main8;     /* 8-bit DP kernel */
    float ms_rerun16;   /* the rerun part of ms_main16 */
    int32_t groups16;   /* 16-bit groups (32 pairs) in the last call */
    int32_t groups8;    /* 8-bit groups (64 pairs) */
    int32_t reruns16;   /* groups rerun for a stale F[ncol] value */
    int32_t lds_state16; /* 1 if the 16-bit state lived in LDS */
    int32_t pairs8;     /* pairs sent to the 8-bit path */
    int32_t pairs16;    /* pairs sent to the 16-bit path */
    float ms_replay;    /* whole graph replay of the last plan launch (the
                           per-kernel fields stay 0 for plan launches) */
    int32_t settled16;  /* of reruns16: groups run again inside the 16-bit
                           main launch (their predecessors were final when
                           they ended), the rest ran after it */
} tal_bsw_stats;

int tal_bsw_get_stats(tal_bsw *h, tal_bsw_stats *out);

/* ---------------------------------------------------------------------------
 * Replayable batch plans (no reference counterpart: the reference has no
 * device or launch layer).  A plan binds a device batch (pairs/ref/qer/n/w,
 * as for tal_bsw_get_scores_auto_device) to its handle, sizes every buffer
 * from the batch as it is at creation (one synchronous pass), and captures the
 * whole dispatcher pipeline -- the 8/16-bit split, both DP launches, the
 * stale-column resolution with two rerun rounds, the update of the handle's
 * stale-column state and the result scatters -- in one hipGraph.
 *
 * tal_bsw_plan_launch replays it on `stream` (NULL = the handle's stream) and
 * scores the batch's current contents exactly as tal_bsw_get_scores_auto_device
 * would; it waits for the replay and reads one status word, and finishes any
 * rerun the graph's rounds left.  The pair records may change between launches
 * as long as the 8/16-bit split keeps its size and no 16-bit group gets wider
 * than at creation; otherwise the launch returns TAL_ERANGE (the outputs are
 * then unspecified, the handle's stale-column state is untouched) and a new
 * plan is needed.  While a plan exists its handle may be used for other calls;
 * the plan owns its own buffers.
 */
typedef struct tal_bsw_plan tal_bsw_plan;

int tal_bsw_plan_create(tal_bsw *h, tal_seqpair *pairs, const uint8_t *ref,
                        const uint8_t *qer, int32_t n, int32_t w, tal_bsw_plan **out);
int tal_bsw_plan_launch(tal_bsw_plan *plan, void *stream);
void tal_bsw_plan_destroy(tal_bsw_plan *plan);

/* Stale-column rerun rounds captured in the graph of plans created after this
 * call (0..8, default 2).  Groups still pending after the replay's rounds are
 * rerun eagerly by tal_bsw_plan_launch; results are the same for every value.
 * After a launch, tal_bsw_get_stats reports in reruns16 the groups rerun in the
 * graph and after it. */
int tal_bsw_set_plan_rounds(tal_bsw *h, int32_t rounds);

/* Autotuner (no reference counterpart).  Times candidate launch
 * configurations -- waves per CU, which sets how many DP state columns each
 * wave keeps in LDS (the LDS tile width), and the task order (input order, or
 * longest groups first with similar groups paired in a wave) -- by replaying
 * a plan of the batch's first `sample` pairs (0 = all) on a scratch handle
 * and a scratch copy of the pairs, so neither the batch nor h's stale-column
 * state changes; then applies the fastest to h.  Results are the same under
 * every configuration; only speed differs. */
#define TAL_TUNING_MAX_TRIALS 16
typedef struct tal_bsw_tuning {
    int32_t waves_per_cu; /* chosen */
    int32_t order;        /* chosen: 0 input order, 1 longest first */
    float ms;             /* replay time of the chosen configuration */
    int32_t trials;
    int32_t trial_waves[TAL_TUNING_MAX_TRIALS];
    int32_t trial_order[TAL_TUNING_MAX_TRIALS];
    float trial_ms[TAL_TUNING_MAX_TRIALS];
} tal_bsw_tuning;

int tal_bsw_autotune(tal_bsw *h, const tal_seqpair *pairs, const uint8_t *ref,
                     const uint8_t *qer, int32_t n, int32_t w, int32_t sample,
                     tal_bsw_tuning *out);

/* Applies a configuration (e.g. a cached autotuner result). */
int tal_bsw_set_tuning(tal_bsw *h, int32_t waves_per_cu, int32_t order);

/* ---------------------------------------------------------------------------
 * Streamed serving.  No reference counterpart: TAL has no serving loop (its
 * callers batch pairs themselves, benchmarks/bench-bsw.cpp:216-253); this is
 * the continuous-batching front end of the same engine.
 *
 * Requests (a SeqPair array with its packed sequence buffers, as for
 * tal_bsw_get_scores_auto) are queued by tal_bsw_server_submit and admitted in
 * arrival order while their device footprint fits in `hbm_budget`.  Admitted
 * requests are copied into device memory; every free worker then takes all
 * admitted requests with the head's band width (up to max_batch_pairs pairs)
 * and scores them as one batch on its own stream, so requests arriving while
 * others are on the GPU join the next batch, and up to `workers` batches are
 * on the GPU at once.  Each request is scored as one tal_bsw_get_scores_auto
 * call on a freshly constructed handle: its groups never mix with other
 * requests' pairs and its stale-column state starts from zero, so its outputs
 * do not depend on what it was batched with.
 *
 * With resident = 1 no batch is launched: one persistent kernel stays on the
 * GPU and takes tasks (a pair of 16-bit groups, an 8-bit group, or a
 * request's stale-column resolution) from a ring in pinned host memory.
 * Admission lays a request's sequences and padded pair blocks out in pinned
 * memory and publishes its tasks at once: the kernel copies them to device
 * memory under the same budget (its device footprint is those plus its
 * 16-bit F masks), scores them and copies the results back, so a request
 * admitted while others run is taken up by waves as they free up.  The
 * server queues no work on other streams; a stream of the caller's that
 * shares a hardware queue with the kernel waits while it runs.  Queries longer than max_query_len are
 * TAL_ERANGE in this mode.  The kernel leaves after 50 ms without work and
 * is launched again when work arrives; the same independence holds.
 */
typedef struct tal_bsw_server tal_bsw_server;

typedef struct tal_bsw_server_config {
    size_t hbm_budget;       /* device bytes for admitted requests, < 2 GiB
                                (sequence offsets are int32); 0 = 1 GiB */
    int32_t workers;         /* batches on the GPU at once; 0 = 2 (host-batched mode) */
    int32_t max_batch_pairs; /* pairs per batch at most; 0 = 4194304 (host-batched mode) */
    int32_t resident;        /* 1: the device-resident loop; 0: host-batched */
    int32_t max_query_len;   /* resident mode: longest query; 0 = 1024 */
} tal_bsw_server_config;

/* Per-request record filled by tal_bsw_server_wait. */
typedef struct tal_bsw_request_info {
    int32_t status;     /* TAL_OK or the request's error code */
    int32_t batch;      /* batch the request ran in (-1 if it did not run;
                           resident mode: 0 if it ran) */
    int32_t batch_requests; /* requests in that batch (resident mode: requests
                               on the device when it was admitted, itself
                               included) */
    int32_t pad;
    double queue_ms;    /* submit -> admitted (copied to the device) */
    double latency_ms;  /* submit -> outputs written into the caller's pairs */
} tal_bsw_request_info;

typedef struct tal_bsw_server_stats {
    int64_t requests;   /* completed requests */
    int64_t pairs;      /* pairs of the completed requests */
    int64_t batches;    /* batches run (resident mode: kernel launches) */
    int64_t peak_bytes; /* peak device bytes held by admitted requests */
    double busy_ms;     /* summed wall time of the batches (host-batched mode) */
    int64_t reruns;     /* resident mode: 16-bit groups run again after a
                           stale-column resolution */
} tal_bsw_server_stats;

int tal_bsw_server_create(const tal_bsw_params *params, int device,
                          const tal_bsw_server_config *config, tal_bsw_server **out);

/* Queues a request and returns at once with its ticket.  pairs/ref/qer are
 * host memory that must stay valid until the request is waited for; the six
 * outputs are written into pairs.  Thread-safe. */
int tal_bsw_server_submit(tal_bsw_server *s, tal_seqpair *pairs, const uint8_t *ref,
                          size_t ref_len, const uint8_t *qer, size_t qer_len,
                          int32_t n, int32_t w, uint64_t *ticket);

/* Waits up to timeout_ms (< 0: no limit) for a request.  Returns TAL_EAGAIN if
 * it is not finished, else the request's status; info (may be null) receives
 * its record.  A finished request is forgotten once waited for. */
int tal_bsw_server_wait(tal_bsw_server *s, uint64_t ticket, int32_t timeout_ms,
                        tal_bsw_request_info *info);

int tal_bsw_server_get_stats(tal_bsw_server *s, tal_bsw_server_stats *out);

/* Finishes every queued request, then releases the server. */
void tal_bsw_server_destroy(tal_bsw_server *s);

/* ---------------------------------------------------------------------------
 * Pair files.  The text format is bench-bsw's (benchmarks/bench-bsw.cpp:
 * 117-169, loadPairs): three lines per pair -- h0 in decimal, the reference,
 * the query -- each base an ASCII character whose code is the character
 * minus '0' ('0'-'3' = ACGT, '4' = N).
 *
 * tal_bsw_pairfile_read reads the next pairs (up to max_pairs) the way
 * loadPairs fills its buffers: pair k's reference at ref + k * ref_slot and
 * its query at qer + k * qer_slot (bench-bsw uses slots of MAX_SEQ_LEN_REF =
 * 256 and MAX_SEQ_LEN_QER = 128 bytes, bandedSWA.h:47-48), idr/idq those
 * offsets, id = k, the pass-through and output fields -1.  *n_read is 0 at the
 * end of the file; an h0 line with no reference after it also ends the file,
 * as in bench-bsw.  A sequence longer than its slot minus two is TAL_ERANGE
 * (bench-bsw's fgets stores the newline and a NUL in the slot, and would split
 * a longer line); a reference with no query line after it is TAL_EINVAL.
 * tal_bsw_pairfile_write writes pairs in the same format. */
typedef struct tal_bsw_pairfile tal_bsw_pairfile;

int tal_bsw_pairfile_open(const char *path, tal_bsw_pairfile **out);
int tal_bsw_pairfile_read(tal_bsw_pairfile *f, tal_seqpair *pairs, uint8_t *ref,
                          size_t ref_slot, uint8_t *qer, size_t qer_slot,
                          int32_t max_pairs, int32_t *n_read);
void tal_bsw_pairfile_close(tal_bsw_pairfile *f);
int tal_bsw_pairfile_write(const char *path, const tal_seqpair *pairs,
                           const uint8_t *ref, const uint8_t *qer, int32_t n);

/* Packed batch files (no reference counterpart): a 64-byte header, the
 * SeqPair records, the reference bytes, the query bytes.  tal_bsw_pack_map
 * maps a file privately (copy-on-write), so outputs may be written into the
 * records and the sequences passed straight to tal_bsw_get_scores_*. */
typedef struct tal_bsw_pack {
    tal_seqpair *pairs;
    int64_t n;
    uint8_t *ref;
    size_t ref_len;
    uint8_t *qer;
    size_t qer_len;
    void *mapping;
    size_t mapping_len;
} tal_bsw_pack;

int tal_bsw_pack_write(const char *path, const tal_seqpair *pairs, int64_t n,
                       const uint8_t *ref, size_t ref_len, const uint8_t *qer,
                       size_t qer_len);
int tal_bsw_pack_map(const char *path, tal_bsw_pack *out);
void tal_bsw_pack_unmap(tal_bsw_pack *p);

#ifdef __cplusplus
}
#endif
#endif
