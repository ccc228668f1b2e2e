/*
 * tal_chain.hip -- host runtime and C ABI of the DP chaining engine
 * (include/tal_chain.h): the drop-in for dp_chain::mm_dp_vectorized
 * (src/chaining/parallel_chaining_32_bit.h:216-412), over many anchor sets
 * per call.
 *
 * A call: one kernel validates the sets (one wait on the stream to read the
 * error word), the sets are sorted longest first, and the persistent DP
 * kernel runs one set per wave with an LDS ring of the latest anchors
 * (chain_kernels.hip).
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <new>
#include <string>

#include "chain_internal.h"

using namespace talchain;

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

/* control words the host reads back after the check kernel */
struct Summary {
    int32_t ctl[CHAIN_CTL_WORDS];
    unsigned long long examined;
    unsigned long long far_anchors;
};

/* mean anchors a set at or below which a batch runs one wave a set */
constexpr int SHORT_SET = 256;

} // namespace

struct tal_chain {
    tal_chain_params prm{};
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    size_t lds_per_block = 0, lds_per_cu = 0;
    Buf keys, keys2, ids, order, sum, cubtmp, d_a, d_off, d_f, d_p, d_v;
    /* LDS ring slots per block (TAL_CHAIN_RING, a power of two in
     * 128..4096), 12 bytes each; waves per block, which share a set
     * (TAL_CHAIN_WPB: 1, 2 or 4), and waves per CU (TAL_CHAIN_WAVES_PER_CU,
     * at most what the LDS holds).  Measured on MI355X with 4000 sets of
     * 200..4000 anchors: a 1024-slot ring holds the windows (512 sends 3% of
     * the anchors' older candidates to HBM, 7% slower), 2 waves a set at 20
     * a CU run 16% faster than 1 at 8, 4 a set are slower (phase B, one
     * wave's, waits for the shared phase A).  Batches of short sets (at most
     * SHORT_SET anchors a set on average) run one wave a set with a ring of
     * at most 512 slots instead: a set of one or two chunks has little
     * phase A to share, and the second wave would wait.  Measured with
     * 200,000 sets of about 111 anchors: 3.36 ms against 4.12 ms.  The
     * short shape is not taken when TAL_CHAIN_RING or TAL_CHAIN_WPB is set. */
    int ring = 1024, wpb = 2, waves_per_cu = 20;
    bool auto_shape = true;
    /* gap-cost table entries per wave: bw + 2, rounded up to 64, when at
     * most 2048 (8 KB), else (or TAL_CHAIN_TABLE=0) none: costs computed per
     * candidate */
    int table = 0;

    Summary *host = nullptr; /* pinned */
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    bool timed = false;
    tal_chain_stats st{};
    std::string err;
};

namespace {

int fail(tal_chain *h, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

#define CHCHK(h, expr, what)                                                                       \
    do {                                                                                           \
        const hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                      \
            return fail((h), e_ == hipErrorOutOfMemory ? TAL_ENOMEM : TAL_EDEVICE, "%s: %s",       \
                        (what), hipGetErrorString(e_));                                            \
    } while (0)

int run(tal_chain *h, const tal_chain_anchor *a, const int64_t *off, int32_t nsets, int32_t dr,
        int32_t dq, uint32_t *f, int32_t *p, int32_t *v, hipStream_t s)
{
    CHCHK(h, h->keys.ensure(sizeof(uint32_t) * nsets), "alloc keys");
    CHCHK(h, h->keys2.ensure(sizeof(uint32_t) * nsets), "alloc keys");
    CHCHK(h, h->ids.ensure(sizeof(int32_t) * nsets), "alloc ids");
    CHCHK(h, h->order.ensure(sizeof(int32_t) * nsets), "alloc order");
    CHCHK(h, h->sum.ensure(sizeof(Summary)), "alloc summary");
    Summary *ds = h->sum.as<Summary>();
    CHCHK(h, hipEventRecord(h->ev[0], s), "event");
    CHCHK(h, hipMemsetAsync(ds, 0, sizeof(Summary), s), "clear summary");
    CHCHK(h, launch_chain_check(a, off, nsets, h->keys.as<uint32_t>(), h->ids.as<int32_t>(),
                                ds->ctl, s),
          "check");
    CHCHK(h, hipMemcpyAsync(h->host, ds, sizeof(Summary), hipMemcpyDeviceToHost, s), "read");
    CHCHK(h, hipStreamSynchronize(s), "sync check");
    if (h->host->ctl[CHAIN_CTL_ERR] == CHAIN_ERR_RANGE)
        return fail(h, TAL_ERANGE, "a seed length is negative or a set has more than INT32_MAX "
                                   "anchors");
    if (h->host->ctl[CHAIN_CTL_ERR] == CHAIN_ERR_OFFSETS)
        return fail(h, TAL_EINVAL, "set offsets must be non-negative and non-decreasing");
    const int64_t anchors = (int64_t)(uint32_t)h->host->ctl[CHAIN_CTL_ANCHORS] |
                            ((int64_t)h->host->ctl[CHAIN_CTL_ANCHORS + 1] << 32);
    int ring = h->ring, wpb = h->wpb;
    if (h->auto_shape && anchors <= (int64_t)SHORT_SET * nsets) {
        wpb = 1;
        ring = std::min(ring, 512);
    }

    size_t tmp = 0;
    CHCHK(h, hipcub::DeviceRadixSort::SortPairsDescending(
                 nullptr, tmp, h->keys.as<uint32_t>(), h->keys2.as<uint32_t>(),
                 h->ids.as<int32_t>(), h->order.as<int32_t>(), nsets, 0, 32, s),
          "sort size");
    CHCHK(h, h->cubtmp.ensure(tmp + 16), "alloc sort tmp");
    CHCHK(h, hipcub::DeviceRadixSort::SortPairsDescending(
                 h->cubtmp.p, tmp, h->keys.as<uint32_t>(), h->keys2.as<uint32_t>(),
                 h->ids.as<int32_t>(), h->order.as<int32_t>(), nsets, 0, 32, s),
          "sort sets");
    ChainParams P{};
    P.nsets = nsets;
    P.bw = h->prm.bw;
    P.max_iter = h->prm.max_iter;
    P.dr = dr;
    P.dq = dq;
    P.ring = ring;
    /* table mode needs bw + min(dq, dr) < 2^31 (see the kernels) */
    const int64_t qlim = std::max<int64_t>(std::min(dq, dr), 0);
    P.table = (int64_t)h->prm.bw + qlim < ((int64_t)1 << 31) ? h->table : 0;
    ChainBuffers B{};
    B.a = a;
    B.off = off;
    B.f = f;
    B.p = p;
    B.v = v;
    B.order = h->order.as<int32_t>();
    B.ctl = ds->ctl;
    B.examined = &ds->examined;
    B.far_anchors = &ds->far_anchors;
    /* a persistent grid of blocks of wpb waves, one set per block at a
     * time: as many blocks as the LDS and waves_per_cu allow on every CU,
     * fewer when there are fewer sets */
    const int per_cu = std::max(
        1, std::min<int>(h->waves_per_cu / wpb,
                         (int)(h->lds_per_cu / chain_lds_bytes(P.ring, P.table, wpb))));
    const int grid = std::max(1, std::min(nsets, h->cu_count * per_cu));
    CHCHK(h, hipEventRecord(h->ev[1], s), "event");
    CHCHK(h, launch_chain_main(wpb, grid, P, B, s), "launch");
    CHCHK(h, hipEventRecord(h->ev[2], s), "event");
    h->timed = true;
    h->st.sets = nsets;
    h->st.lds_ring = ring;
    return TAL_OK;
}

int check_args(tal_chain *h, int32_t nsets)
{
    if (!h) return TAL_EINVAL;
    if (nsets < 0) return fail(h, TAL_EINVAL, "nsets < 0");
    return TAL_OK;
}

} // namespace

extern "C" {

int tal_chain_create(const tal_chain_params *params, int device, tal_chain **out)
{
    if (!out || !params) return TAL_EINVAL;
    *out = nullptr;
    if (params->max_iter < 0 || params->bw < 0) return TAL_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TAL_EDEVICE;
    if (device < 0 || device >= ndev) return TAL_EINVAL;
    tal_chain *h = new (std::nothrow) tal_chain();
    if (!h) return TAL_ENOMEM;
    h->prm = *params;
    h->device = device;
    if (getenv("TAL_CHAIN_RING") || getenv("TAL_CHAIN_WPB")) h->auto_shape = false;
    if (const char *e = getenv("TAL_CHAIN_RING")) {
        const int x = atoi(e);
        if (x >= 128 && x <= 4096 && (x & (x - 1)) == 0) h->ring = x;
    }
    if (const char *e = getenv("TAL_CHAIN_WPB")) {
        const int x = atoi(e);
        if (x == 1 || x == 2 || x == 4) h->wpb = x;
    }
    if (const char *e = getenv("TAL_CHAIN_WAVES_PER_CU")) {
        const int x = atoi(e);
        if (x >= 1 && x <= 32) h->waves_per_cu = x;
    }
    if ((int64_t)params->bw + 2 <= 2048) h->table = (params->bw + 2 + 63) & ~63;
    if (const char *e = getenv("TAL_CHAIN_TABLE"))
        if (atoi(e) == 0) h->table = 0;
    int rc = TAL_OK;
    do {
        if (hipSetDevice(device) != hipSuccess) { rc = TAL_EDEVICE; break; }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess) { rc = TAL_EDEVICE; break; }
        h->cu_count = prop.multiProcessorCount;
        h->lds_per_block = prop.sharedMemPerBlock;
        h->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                              : prop.sharedMemPerBlock;
        /* the ring's LDS bounds the waves per CU and the ring per block */
        while (chain_lds_bytes(h->ring, h->table, h->wpb) > h->lds_per_block && h->wpb > 1)
            h->wpb >>= 1;
        while (chain_lds_bytes(h->ring, h->table, h->wpb) > h->lds_per_block && h->ring > 128)
            h->ring >>= 1;
        if (set_chain_lds_limit(h->lds_per_block) != hipSuccess) { rc = TAL_EDEVICE; break; }
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
            rc = TAL_EDEVICE;
            break;
        }
        for (auto &e : h->ev)
            if (hipEventCreate(&e) != hipSuccess) rc = TAL_EDEVICE;
        if (rc) break;
        if (hipHostMalloc((void **)&h->host, sizeof(Summary)) != hipSuccess) {
            rc = TAL_ENOMEM;
            break;
        }
    } while (0);
    if (rc != TAL_OK) {
        tal_chain_destroy(h);
        return rc;
    }
    *out = h;
    return TAL_OK;
}

void tal_chain_destroy(tal_chain *h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (Buf *b : {&h->keys, &h->keys2, &h->ids, &h->order, &h->sum, &h->cubtmp, &h->d_a,
                   &h->d_off, &h->d_f, &h->d_p, &h->d_v})
        b->release();
    if (h->host) (void)hipHostFree(h->host);
    for (auto e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char *tal_chain_last_error(const tal_chain *h) { return h ? h->err.c_str() : "null handle"; }

int tal_chain_dp_device(tal_chain *h, const tal_chain_anchor *d_anchors, const int64_t *d_offsets,
                        int32_t nsets, int32_t dr, int32_t dq, uint32_t *d_f, int32_t *d_p,
                        int32_t *d_v, void *stream)
{
    const int a = check_args(h, nsets);
    if (a != TAL_OK) return a;
    h->err.clear();
    if (nsets == 0) return TAL_OK;
    if (!d_anchors || !d_offsets || !d_f || !d_p || !d_v)
        return fail(h, TAL_EINVAL, "null buffer");
    CHCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    return run(h, d_anchors, d_offsets, nsets, dr, dq, d_f, d_p, d_v, s);
}

int tal_chain_dp(tal_chain *h, const tal_chain_anchor *anchors, const int64_t *offsets,
                 int32_t nsets, int32_t dr, int32_t dq, uint32_t *f, int32_t *p, int32_t *v)
{
    const int a = check_args(h, nsets);
    if (a != TAL_OK) return a;
    h->err.clear();
    if (nsets == 0) return TAL_OK;
    if (!anchors || !offsets || !f || !p || !v) return fail(h, TAL_EINVAL, "null buffer");
    /* the anchor count is the last offset; the host checks the offsets it
     * needs to size the copies */
    const int64_t n = offsets[nsets];
    if (offsets[0] != 0 || n < 0) return fail(h, TAL_EINVAL, "offsets must start at 0");
    for (int32_t k = 0; k < nsets; k++)
        if (offsets[k + 1] < offsets[k])
            return fail(h, TAL_EINVAL, "set offsets must be non-decreasing");
    CHCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = h->stream;
    const size_t na = (size_t)std::max<int64_t>(n, 1);
    CHCHK(h, h->d_a.ensure(sizeof(tal_chain_anchor) * na), "alloc anchors");
    CHCHK(h, h->d_off.ensure(sizeof(int64_t) * ((size_t)nsets + 1)), "alloc offsets");
    CHCHK(h, h->d_f.ensure(sizeof(uint32_t) * na), "alloc f");
    CHCHK(h, h->d_p.ensure(sizeof(int32_t) * na), "alloc p");
    CHCHK(h, h->d_v.ensure(sizeof(int32_t) * na), "alloc v");
    CHCHK(h, hipMemcpyAsync(h->d_a.p, anchors, sizeof(tal_chain_anchor) * (size_t)n,
                            hipMemcpyHostToDevice, s),
          "copy anchors");
    CHCHK(h, hipMemcpyAsync(h->d_off.p, offsets, sizeof(int64_t) * ((size_t)nsets + 1),
                            hipMemcpyHostToDevice, s),
          "copy offsets");
    const int rc = run(h, h->d_a.as<tal_chain_anchor>(), h->d_off.as<int64_t>(), nsets, dr, dq,
                       h->d_f.as<uint32_t>(), h->d_p.as<int32_t>(), h->d_v.as<int32_t>(), s);
    if (rc != TAL_OK) return rc;
    CHCHK(h, hipMemcpyAsync(f, h->d_f.p, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost, s),
          "copy f");
    CHCHK(h, hipMemcpyAsync(p, h->d_p.p, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, s),
          "copy p");
    CHCHK(h, hipMemcpyAsync(v, h->d_v.p, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, s),
          "copy v");
    CHCHK(h, hipStreamSynchronize(s), "sync");
    return TAL_OK;
}

int tal_chain_get_stats(const tal_chain *hc, tal_chain_stats *out)
{
    tal_chain *h = const_cast<tal_chain *>(hc);
    if (!h || !out) return TAL_EINVAL;
    if (h->timed) {
        CHCHK(h, hipSetDevice(h->device), "set device");
        CHCHK(h, hipEventSynchronize(h->ev[2]), "event sync");
        CHCHK(h, hipEventElapsedTime(&h->st.ms_kernel, h->ev[1], h->ev[2]), "event time");
        CHCHK(h, hipEventElapsedTime(&h->st.ms_total, h->ev[0], h->ev[2]), "event time");
        Summary sm;
        CHCHK(h, hipMemcpy(&sm, h->sum.p, sizeof sm, hipMemcpyDeviceToHost), "read counts");
        h->st.predecessors = (int64_t)sm.examined;
        h->st.far_anchors = (int64_t)sm.far_anchors;
    }
    *out = h->st;
    return TAL_OK;
}

} /* extern "C" */
