/*
 * tal_sw.hip -- host runtime and C ABI of the full inter-task Smith-Waterman
 * engine (include/tal_sw.h): the drop-in for
 * PairWiseSW_{8,16,32}::getScores (src/alignment/Smith-Waterman/).
 *
 * A call: one kernel computes every group's padded bounds, validates the
 * pairs and reduces the batch maxima (one wait on the stream to read them and
 * the error word), the units are sorted by their bounds, and the persistent
 * DP kernel scores four units per wave (sw_kernels.hip).  The reference's
 * getScores allocates its scratch per call; here the handle keeps its device
 * buffers and grows them as needed.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <new>
#include <string>

#include "sw_internal.h"

using namespace talsw;

namespace {

struct Buf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (p) {
            const hipError_t e = hipFree(p);
            p = nullptr;
            cap = 0;
            if (e != hipSuccess) return e;
        }
        size_t want = std::max(bytes, cap * 3 / 2);
        want = (want + 4095) & ~(size_t)4095;
        const hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
    template <class T> T *as() const { return (T *)p; }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

} // namespace

struct tal_sw {
    int32_t w_match = 0, w_mismatch = 0, w_open = 0, w_extend = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    /* persistent waves per CU (TAL_SW_WAVES_PER_CU): the DP kernel's 168
     * VGPRs leave room for three waves per SIMD */
    int waves_per_cu = 12;
    Buf dims, keys, keys2, ids, order, ctl, slabs, cubtmp, d_pairs, d_seq;
    Buf pkeys, pkeys2, pids, perm; /* regrouping: per-pair keys and order */
    bool regroup = true; /* TAL_SW_REGROUP=0 keeps the reference's groups */
    int split_force = 0; /* TAL_SW_SPLIT=1/2/4 fixes the unit split */
    int32_t *ctl_host = nullptr; /* pinned copy of the control words */
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    bool timed = false;
    tal_sw_stats st{};
    std::string err;
};

namespace {

int fail(tal_sw *h, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

#define SWCHK(h, expr, what)                                                                       \
    do {                                                                                           \
        const hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                      \
            return fail((h), e_ == hipErrorOutOfMemory ? TAL_ENOMEM : TAL_EDEVICE, "%s: %s",       \
                        (what), hipGetErrorString(e_));                                            \
    } while (0)

int group_lanes(int bits) { return bits == 8 ? 64 : bits == 16 ? 32 : 16; }
int unit_lanes(int bits) { return bits == 32 ? 16 : 32; }

/* lane constants of the class: each weight truncated to the class width as
 * _mm512_set1_epi8/16/32 does; the 8-bit class holds value << 8 */
SwParams lane_constants(const tal_sw *h, int bits)
{
    SwParams P{};
    if (bits == 32) {
        P.k_match = (uint32_t)h->w_match;
        P.k_delta = (uint32_t)h->w_mismatch;
        P.k_open = (uint32_t)h->w_open;
        P.k_ext = (uint32_t)h->w_extend;
        return P;
    }
    auto half = [&](int32_t v) -> uint32_t {
        return bits == 16 ? (uint32_t)(uint16_t)v : ((uint32_t)(uint8_t)v << 8) & 0xFFFF;
    };
    auto pack = [](uint32_t x) { return x | (x << 16); };
    const uint32_t match = half(h->w_match), mism = half(h->w_mismatch);
    P.k_match = pack(match);
    P.k_delta = pack((mism - match) & 0xFFFF);
    P.k_open = pack(half(h->w_open));
    P.k_ext = pack(half(h->w_extend));
    return P;
}

/* Lane value of a weight in the class (the reference's set1 truncation). */
int64_t lane_value(int bits, int32_t w)
{
    return bits == 8 ? (int8_t)w : bits == 16 ? (int16_t)w : w;
}

/*
 * Can padding change a score?  Every pair of a reference group is scored
 * over the group's padded matrix; a padded cell holds a padding code on at
 * least one side.  With mismatch, gap open and gap extend all <= 0 at the
 * class width, a padded cell whose codes differ scores at most the best of
 * the cells above and to its left, so by induction no padded cell exceeds
 * the pair's own best, and cells inside the pair's own lengths never depend
 * on padded ones.  The codes differ unless seq1 holds 'D' or seq2 holds 'B'
 * (scanned on the device), and no lane value wraps when match * min(len1,
 * len2) fits the lane and open + extend does not underflow (checked per
 * pair).  When all of this holds a pair's score is its own DP's, whatever
 * its group, and pairs may be regrouped by length: this removes the work on
 * padded cells.  Zero-bound groups of the 8-bit class still score 0.
 */
bool weights_allow_regroup(const tal_sw *h, int bits)
{
    const int64_t mis = lane_value(bits, h->w_mismatch), open = lane_value(bits, h->w_open);
    const int64_t ext = lane_value(bits, h->w_extend);
    const int64_t lane_min = bits == 8 ? -128 : bits == 16 ? -32768 : INT32_MIN;
    return mis <= 0 && open <= 0 && ext <= 0 && open + ext >= lane_min;
}

int run(tal_sw *h, int bits, tal_sw_pair *d_pairs, const uint8_t *d_seq, size_t seq_len,
        int32_t n, hipStream_t s)
{
    const int GL = group_lanes(bits), UL = unit_lanes(bits);
    const int32_t G = (n + GL - 1) / GL;
    int32_t U = G * (GL / UL);
    SwInert I{};
    I.on = h->regroup && weights_allow_regroup(h, bits);
    I.match = (int32_t)lane_value(bits, h->w_match);
    I.lane_max = bits == 8 ? 127 : bits == 16 ? 32767 : INT32_MAX;
    if (I.on) {
        SWCHK(h, h->pkeys.ensure(sizeof(uint32_t) * n), "alloc pair keys");
        SWCHK(h, h->pkeys2.ensure(sizeof(uint32_t) * n), "alloc pair keys");
        SWCHK(h, h->pids.ensure(sizeof(int32_t) * n), "alloc pair ids");
        SWCHK(h, h->perm.ensure(sizeof(int32_t) * n), "alloc pair order");
        I.pair_keys = h->pkeys.as<uint32_t>();
        I.pair_ids = h->pids.as<int32_t>();
    }
    SWCHK(h, h->dims.ensure(sizeof(int32_t) * U), "alloc dims");
    SWCHK(h, h->keys.ensure(sizeof(uint32_t) * U), "alloc keys");
    SWCHK(h, h->keys2.ensure(sizeof(uint32_t) * U), "alloc keys");
    SWCHK(h, h->ids.ensure(sizeof(int32_t) * U), "alloc ids");
    SWCHK(h, h->order.ensure(sizeof(int32_t) * U), "alloc order");
    SWCHK(h, h->ctl.ensure(sizeof(int32_t) * SW_CTL_WORDS), "alloc ctl");
    int32_t *ctl = h->ctl.as<int32_t>();
    SWCHK(h, hipEventRecord(h->ev[0], s), "event");
    SWCHK(h, hipMemsetAsync(ctl, 0, sizeof(int32_t) * SW_CTL_WORDS, s), "clear ctl");
    SWCHK(h, launch_sw_shapes(d_pairs, n, bits, seq_len, h->dims.as<int32_t>(),
                              h->keys.as<uint32_t>(), h->ids.as<int32_t>(), ctl, I, s),
          "shapes");
    if (I.on) SWCHK(h, launch_sw_scan_codes(d_pairs, n, bits, d_seq, seq_len, ctl, s), "scan");
    SWCHK(h, hipMemcpyAsync(h->ctl_host, ctl, sizeof(int32_t) * SW_CTL_WORDS,
                            hipMemcpyDeviceToHost, s),
          "read ctl");
    SWCHK(h, hipStreamSynchronize(s), "sync shapes");
    const int32_t err = h->ctl_host[SW_CTL_ERR];
    if (err == SW_ERR_RANGE)
        return fail(h, TAL_ERANGE, "a pair length is negative or exceeds %d",
                    bits == 8 ? 127 : 1024);
    if (err == SW_ERR_BOUNDS)
        return fail(h, TAL_EINVAL, "a pair's id is negative or its sequences lie outside the "
                                   "%zu-byte sequence buffer", seq_len);
    const int max_rows = h->ctl_host[SW_CTL_MAXROW], max_cols = h->ctl_host[SW_CTL_MAXCOL];
    const bool regrouped = I.on && h->ctl_host[SW_CTL_NOT_INERT] == 0;
    int32_t m = 0;
    if (regrouped) {
        /* pairs by Morton key (zero-bound groups last, score 0), then units of
         * UL consecutive sorted pairs */
        m = h->ctl_host[SW_CTL_ELIGIBLE];
        size_t ptmp = 0;
        SWCHK(h, hipcub::DeviceRadixSort::SortPairsDescending(
                     nullptr, ptmp, h->pkeys.as<uint32_t>(), h->pkeys2.as<uint32_t>(),
                     h->pids.as<int32_t>(), h->perm.as<int32_t>(), n, 0, 24, s),
              "sort size");
        SWCHK(h, h->cubtmp.ensure(ptmp + 16), "alloc sort tmp");
        SWCHK(h, hipcub::DeviceRadixSort::SortPairsDescending(
                     h->cubtmp.p, ptmp, h->pkeys.as<uint32_t>(), h->pkeys2.as<uint32_t>(),
                     h->pids.as<int32_t>(), h->perm.as<int32_t>(), n, 0, 24, s),
              "sort pairs");
        SWCHK(h, launch_sw_zero_scores(d_pairs, h->perm.as<int32_t>() + m, n - m, s), "zero");
        SWCHK(h, launch_sw_unit_dims(d_pairs, h->perm.as<int32_t>(), m, bits,
                                     h->dims.as<int32_t>(), h->keys.as<uint32_t>(),
                                     h->ids.as<int32_t>(), s),
              "unit bounds");
        U = (m + UL - 1) / UL;
    }

    size_t tmp = 0;
    SWCHK(h, hipcub::DeviceRadixSort::SortPairsDescending(
                 nullptr, tmp, h->keys.as<uint32_t>(), h->keys2.as<uint32_t>(),
                 h->ids.as<int32_t>(), h->order.as<int32_t>(), U, 0, 20, s),
          "sort size");
    SWCHK(h, h->cubtmp.ensure(tmp + 16), "alloc sort tmp");
    if (U > 0)
        SWCHK(h, hipcub::DeviceRadixSort::SortPairsDescending(
                     h->cubtmp.p, tmp, h->keys.as<uint32_t>(), h->keys2.as<uint32_t>(),
                     h->ids.as<int32_t>(), h->order.as<int32_t>(), U, 0, 20, s),
              "sort units");

    SwParams P = lane_constants(h, bits);
    P.n = n;
    P.nunits = U;
    P.slot = bits == 8 ? 128 : 1024;
    /* split units sweep 16-row strips, staged up to a multiple of 16 */
    P.rows_cap = std::max(16, (max_rows + 15) & ~15);
    P.cols_cap = std::max(4, max_cols);
    P.aligned4 = ((uintptr_t)d_seq & 3) == 0;
    P.m = m;
    P.slab_words = (size_t)64 * (P.rows_cap + 3 * (size_t)P.cols_cap);
    /* Split each unit over K quarters of a wave.  A split pays K - 1 pipeline
     * steps per round of K strips (a few percent once a row has more than 60
     * tiles) and saves slab traffic and task length; measured on MI355X, K = 4
     * wins from about 250 columns, K = 2 from about 64.  Batches with too few
     * units to give every wave slot two tasks split further. */
    const int slots = h->cu_count * h->waves_per_cu;
    int split = max_cols >= 256 ? 4 : max_cols >= 64 ? 2 : 1;
    while (split < 4 && (int64_t)U * split < 8LL * slots) split *= 2;
    if (h->split_force) split = h->split_force;
    const int ntasks = (U + 4 / split - 1) / (4 / split);
    h->st.regrouped = regrouped ? 1 : 0;
    h->st.units = U;
    h->st.split = split;
    if (ntasks == 0) return TAL_OK; /* every pair in a zero-bound group */
    const int grid = std::max(1, std::min(ntasks, slots));
    SWCHK(h, h->slabs.ensure(sizeof(uint32_t) * P.slab_words * grid), "alloc slabs");
    SwBuffers B{};
    B.pairs = d_pairs;
    B.seq = d_seq;
    B.unit_dims = h->dims.as<int32_t>();
    B.order = h->order.as<int32_t>();
    B.perm = regrouped ? h->perm.as<int32_t>() : nullptr;
    B.ctl = ctl;
    B.slabs = h->slabs.as<uint32_t>();
    SWCHK(h, hipEventRecord(h->ev[1], s), "event");
    SWCHK(h, launch_sw_main(bits, split, grid, P, B, s), "launch");
    SWCHK(h, hipEventRecord(h->ev[2], s), "event");
    h->timed = true;
    h->st.max_rows = max_rows;
    h->st.max_cols = max_cols;
    return TAL_OK;
}

int check_args(tal_sw *h, int bits, int32_t n)
{
    if (!h) return TAL_EINVAL;
    if (bits != 8 && bits != 16 && bits != 32)
        return fail(h, TAL_EINVAL, "bits must be 8, 16 or 32 (got %d)", bits);
    if (n < 0) return fail(h, TAL_EINVAL, "n < 0");
    return TAL_OK;
}

} // namespace

extern "C" {

int32_t tal_sw_slot(int bits)
{
    return bits == 8 ? 128 : (bits == 16 || bits == 32) ? 1024 : 0;
}

int tal_sw_create(int32_t w_match, int32_t w_mismatch, int32_t w_open, int32_t w_extend,
                  int device, tal_sw **out)
{
    if (!out) return TAL_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TAL_EDEVICE;
    if (device < 0 || device >= ndev) return TAL_EINVAL;
    tal_sw *h = new (std::nothrow) tal_sw();
    if (!h) return TAL_ENOMEM;
    h->w_match = w_match;
    h->w_mismatch = w_mismatch;
    h->w_open = w_open;
    h->w_extend = w_extend;
    h->device = device;
    if (const char *v = getenv("TAL_SW_REGROUP")) h->regroup = atoi(v) != 0;
    if (const char *v = getenv("TAL_SW_SPLIT")) {
        const int x = atoi(v);
        if (x == 1 || x == 2 || x == 4) h->split_force = x;
    }
    if (const char *v = getenv("TAL_SW_WAVES_PER_CU")) {
        const int x = atoi(v);
        if (x >= 1 && x <= 32) h->waves_per_cu = x;
    }
    int rc = TAL_OK;
    do {
        if (hipSetDevice(device) != hipSuccess) { rc = TAL_EDEVICE; break; }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess) { rc = TAL_EDEVICE; break; }
        h->cu_count = prop.multiProcessorCount;
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
            rc = TAL_EDEVICE;
            break;
        }
        for (auto &e : h->ev)
            if (hipEventCreate(&e) != hipSuccess) rc = TAL_EDEVICE;
        if (rc) break;
        if (hipHostMalloc((void **)&h->ctl_host, sizeof(int32_t) * SW_CTL_WORDS) != hipSuccess) {
            rc = TAL_ENOMEM;
            break;
        }
    } while (0);
    if (rc != TAL_OK) {
        tal_sw_destroy(h);
        return rc;
    }
    *out = h;
    return TAL_OK;
}

void tal_sw_destroy(tal_sw *h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (Buf *b : {&h->dims, &h->keys, &h->keys2, &h->ids, &h->order, &h->ctl, &h->slabs,
                   &h->cubtmp, &h->d_pairs, &h->d_seq, &h->pkeys, &h->pkeys2, &h->pids, &h->perm})
        b->release();
    if (h->ctl_host) (void)hipHostFree(h->ctl_host);
    for (auto e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char *tal_sw_last_error(const tal_sw *h) { return h ? h->err.c_str() : "null handle"; }

int tal_sw_get_scores_device(tal_sw *h, int bits, tal_sw_pair *d_pairs, const uint8_t *d_seq_buf,
                             size_t seq_buf_len, int32_t n, void *stream)
{
    const int a = check_args(h, bits, n);
    if (a != TAL_OK) return a;
    h->err.clear();
    if (n == 0) return TAL_OK;
    if (!d_pairs || !d_seq_buf) return fail(h, TAL_EINVAL, "null buffer");
    SWCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    return run(h, bits, d_pairs, d_seq_buf, seq_buf_len, n, s);
}

int tal_sw_get_scores(tal_sw *h, int bits, tal_sw_pair *pairs, const uint8_t *seq_buf,
                      size_t seq_buf_len, int32_t n)
{
    const int a = check_args(h, bits, n);
    if (a != TAL_OK) return a;
    h->err.clear();
    if (n == 0) return TAL_OK;
    if (!pairs || !seq_buf) return fail(h, TAL_EINVAL, "null buffer");
    SWCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = h->stream;
    SWCHK(h, h->d_pairs.ensure(sizeof(tal_sw_pair) * (size_t)n), "alloc pairs");
    SWCHK(h, h->d_seq.ensure(std::max<size_t>(seq_buf_len, 4)), "alloc sequences");
    SWCHK(h, hipMemcpyAsync(h->d_pairs.p, pairs, sizeof(tal_sw_pair) * (size_t)n,
                            hipMemcpyHostToDevice, s),
          "copy pairs");
    SWCHK(h, hipMemcpyAsync(h->d_seq.p, seq_buf, seq_buf_len, hipMemcpyHostToDevice, s),
          "copy sequences");
    const int rc = run(h, bits, h->d_pairs.as<tal_sw_pair>(), h->d_seq.as<uint8_t>(),
                       seq_buf_len, n, s);
    if (rc != TAL_OK) return rc;
    SWCHK(h, hipMemcpyAsync(pairs, h->d_pairs.p, sizeof(tal_sw_pair) * (size_t)n,
                            hipMemcpyDeviceToHost, s),
          "copy scores");
    SWCHK(h, hipStreamSynchronize(s), "sync");
    return TAL_OK;
}

int tal_sw_get_stats(const tal_sw *hc, tal_sw_stats *out)
{
    tal_sw *h = const_cast<tal_sw *>(hc);
    if (!h || !out) return TAL_EINVAL;
    if (h->timed) {
        SWCHK(h, hipSetDevice(h->device), "set device");
        SWCHK(h, hipEventSynchronize(h->ev[2]), "event sync");
        SWCHK(h, hipEventElapsedTime(&h->st.ms_kernel, h->ev[1], h->ev[2]), "event time");
        SWCHK(h, hipEventElapsedTime(&h->st.ms_total, h->ev[0], h->ev[2]), "event time");
    }
    *out = h->st;
    return TAL_OK;
}

} /* extern "C" */
