/*
 * tal_extend.hip -- host runtime and C ABI of the extension-pair producer
 * and the end-to-end extension pipeline (include/tal_extend.h).
 *
 * A run: DP chaining of every read's anchors (tal_chain_dp_device), the
 * producer's pick kernel, four hipCUB scans (pair indexes, byte offsets),
 * one totals kernel and one wait for it, the build kernel, the left extension pass
 * (tal_bsw_get_scores16_device, bench-bsw's getScores16), the left scores into the right pairs'
 * h0, the right pass, and the final scores.  Everything stays in HBM.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <new>
#include <string>

#include "extend_internal.h"

using namespace talext;

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

struct tal_extend {
    tal_extend_params prm{};
    tal_bsw_params bsw_prm{};
    int device = 0;
    hipStream_t stream = nullptr;
    tal_chain *chain = nullptr;
    tal_bsw *bsw = nullptr;
    Buf f, p, v, lhas, rhas, lidx, ridx, rbytes, qbytes, roff, qoff, lq, lt, rq, rt, left, right,
        ref, qer, err, cubtmp, totals;
    /* the host entry point's device copies */
    Buf h_genome, h_reads, h_read_off, h_anchors, h_anchor_off, h_out;
    ExtTotals *host = nullptr; /* pinned */
    hipEvent_t ev[6] = {};
    int nev = 0;
    tal_extend_stats st{};
    std::string errmsg;
};

namespace {

int fail(tal_extend *h, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->errmsg = buf;
    return code;
}

#define EXCHK(h, expr, what)                                                                       \
    do {                                                                                           \
        const hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                      \
            return fail((h), e_ == hipErrorOutOfMemory ? TAL_ENOMEM : TAL_EDEVICE, "%s: %s",       \
                        (what), hipGetErrorString(e_));                                            \
    } while (0)

/* hipCUB temporary storage of one call (sized by a dry run) */
hipError_t cub_tmp(tal_extend *h, size_t bytes) { return h->cubtmp.ensure(bytes + 16); }

template <class T> hipError_t excl_sum(tal_extend *h, const T *in, T *out, int32_t n, hipStream_t s)
{
    size_t tmp = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n, s);
    if (e == hipSuccess) e = cub_tmp(h, tmp);
    if (e != hipSuccess) return e;
    return hipcub::DeviceScan::ExclusiveSum(h->cubtmp.p, tmp, in, out, n, s);
}

/* Sizes the hipCUB storage for every scan of a run up front, so that no
 * scan reallocates it while an earlier one is queued. */
hipError_t cub_reserve(tal_extend *h, int32_t n)
{
    size_t a = 0;
    const hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, a, (const int64_t *)nullptr,
                                                          (int64_t *)nullptr, n);
    if (e != hipSuccess) return e;
    return cub_tmp(h, a);
}

ExtParams ext_params(const tal_extend *h, int32_t nreads, int64_t genome_len)
{
    return ExtParams{nreads, genome_len, h->prm.pad, h->bsw_prm.w_match};
}

/* Steps 1-4 of the producer; fills the stats' pair counts and byte sizes. */
int produce(tal_extend *h, const uint8_t *genome, int64_t genome_len, const uint8_t *reads,
            const int64_t *read_off, int32_t nreads, const tal_chain_anchor *anchors,
            const int64_t *anchor_off, const uint32_t *f, const int32_t *p, tal_extend_read *out,
            hipStream_t s)
{
    const size_t n = (size_t)std::max(nreads, 1);
    EXCHK(h, h->lhas.ensure(4 * n), "alloc");
    EXCHK(h, h->rhas.ensure(4 * n), "alloc");
    EXCHK(h, h->lidx.ensure(4 * n), "alloc");
    EXCHK(h, h->ridx.ensure(4 * n), "alloc");
    EXCHK(h, h->rbytes.ensure(8 * n), "alloc");
    EXCHK(h, h->qbytes.ensure(8 * n), "alloc");
    EXCHK(h, h->roff.ensure(8 * n), "alloc");
    EXCHK(h, h->qoff.ensure(8 * n), "alloc");
    EXCHK(h, h->lq.ensure(4 * n), "alloc");
    EXCHK(h, h->lt.ensure(4 * n), "alloc");
    EXCHK(h, h->rq.ensure(4 * n), "alloc");
    EXCHK(h, h->rt.ensure(4 * n), "alloc");
    EXCHK(h, h->err.ensure(4), "alloc");
    const ExtParams P = ext_params(h, nreads, genome_len);
    ExtBuffers B{};
    B.a = anchors;
    B.aoff = anchor_off;
    B.f = f;
    B.p = p;
    B.genome = genome;
    B.reads = reads;
    B.read_off = read_off;
    B.out = out;
    B.lhas = h->lhas.as<int32_t>();
    B.rhas = h->rhas.as<int32_t>();
    B.lidx = h->lidx.as<int32_t>();
    B.ridx = h->ridx.as<int32_t>();
    B.rbytes = h->rbytes.as<int64_t>();
    B.qbytes = h->qbytes.as<int64_t>();
    B.roff = h->roff.as<int64_t>();
    B.qoff = h->qoff.as<int64_t>();
    B.lq = h->lq.as<int32_t>();
    B.lt = h->lt.as<int32_t>();
    B.rq = h->rq.as<int32_t>();
    B.rt = h->rt.as<int32_t>();
    B.err = h->err.as<int32_t>();
    ExtTotals *dt = h->totals.as<ExtTotals>();
    EXCHK(h, cub_reserve(h, nreads), "alloc");
    EXCHK(h, hipMemsetAsync(B.err, 0, 4, s), "clear");
    EXCHK(h, launch_extend_pick(P, B, s), "pick");
    EXCHK(h, excl_sum(h, B.lhas, B.lidx, nreads, s), "scan");
    EXCHK(h, excl_sum(h, B.rhas, B.ridx, nreads, s), "scan");
    EXCHK(h, excl_sum(h, B.rbytes, B.roff, nreads, s), "scan");
    EXCHK(h, excl_sum(h, B.qbytes, B.qoff, nreads, s), "scan");
    EXCHK(h, launch_extend_totals(P, B, dt, s), "totals");
    EXCHK(h, hipMemcpyAsync(h->host, dt, sizeof(ExtTotals), hipMemcpyDeviceToHost, s), "read");
    EXCHK(h, hipStreamSynchronize(s), "sync");
    const ExtTotals t = *h->host;
    if (t.err == EXT_ERR_OFFSETS)
        return fail(h, TAL_EINVAL, "read or anchor offsets must be non-negative and non-decreasing");
    if (t.err == EXT_ERR_READ) return fail(h, TAL_ERANGE, "a read is longer than 32767 bases");
    if (t.err == EXT_ERR_ANCHOR)
        return fail(h, TAL_EINVAL, "an anchor of a chain lies outside its read or the genome");
    const int32_t nleft = t.nleft, nright = t.nright;
    const int64_t rb = t.ref_bytes, qb = t.qer_bytes;
    if (rb >= ((int64_t)1 << 31) || qb >= ((int64_t)1 << 31))
        return fail(h, TAL_ERANGE, "the extension buffers pass 2^31 bytes");
    EXCHK(h, h->left.ensure(sizeof(tal_seqpair) * std::max(nleft, 1)), "alloc pairs");
    EXCHK(h, h->right.ensure(sizeof(tal_seqpair) * std::max(nright, 1)), "alloc pairs");
    EXCHK(h, h->ref.ensure((size_t)std::max<int64_t>(rb, 1)), "alloc ref");
    EXCHK(h, h->qer.ensure((size_t)std::max<int64_t>(qb, 1)), "alloc qer");
    B.left = h->left.as<tal_seqpair>();
    B.right = h->right.as<tal_seqpair>();
    B.ref_out = h->ref.as<uint8_t>();
    B.qer_out = h->qer.as<uint8_t>();
    EXCHK(h, launch_extend_build(P, B, s), "build");
    h->st.reads = nreads;
    h->st.left_pairs = nleft;
    h->st.right_pairs = nright;
    h->st.ref_bytes = rb;
    h->st.qer_bytes = qb;
    return TAL_OK;
}

ExtBuffers pass_buffers(tal_extend *h, tal_extend_read *out)
{
    ExtBuffers B{};
    B.out = out;
    B.left = h->left.as<tal_seqpair>();
    B.right = h->right.as<tal_seqpair>();
    return B;
}

int check_common(tal_extend *h, const uint8_t *genome, int64_t genome_len, const uint8_t *reads,
                 const int64_t *read_off, int32_t nreads, const tal_chain_anchor *anchors,
                 const int64_t *anchor_off, tal_extend_read *out)
{
    if (!h) return TAL_EINVAL;
    h->errmsg.clear();
    if (nreads < 0) return fail(h, TAL_EINVAL, "nreads < 0");
    if (genome_len < 0) return fail(h, TAL_EINVAL, "genome_len < 0");
    if (nreads > 0 && (!genome || !reads || !read_off || !anchors || !anchor_off || !out))
        return fail(h, TAL_EINVAL, "null buffer");
    return TAL_OK;
}

} // namespace

extern "C" {

int tal_extend_create(const tal_extend_params *params, const tal_chain_params *chain,
                      const tal_bsw_params *bsw, int device, tal_extend **out)
{
    if (!out || !params || !chain || !bsw) return TAL_EINVAL;
    *out = nullptr;
    if (params->pad < 0 || params->w < 0) return TAL_EINVAL;
    tal_extend *h = new (std::nothrow) tal_extend();
    if (!h) return TAL_ENOMEM;
    h->prm = *params;
    h->bsw_prm = *bsw;
    h->device = device;
    int rc = tal_chain_create(chain, device, &h->chain);
    if (rc == TAL_OK) rc = tal_bsw_create(bsw, device, &h->bsw);
    if (rc == TAL_OK) {
        if (hipSetDevice(device) != hipSuccess ||
            hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
            rc = TAL_EDEVICE;
    }
    for (auto &e : h->ev)
        if (rc == TAL_OK && hipEventCreate(&e) != hipSuccess) rc = TAL_EDEVICE;
    if (rc == TAL_OK && hipHostMalloc((void **)&h->host, sizeof(ExtTotals)) != hipSuccess)
        rc = TAL_ENOMEM;
    if (rc == TAL_OK && h->totals.ensure(sizeof(ExtTotals)) != hipSuccess) rc = TAL_ENOMEM;
    if (rc != TAL_OK) {
        tal_extend_destroy(h);
        return rc;
    }
    *out = h;
    return TAL_OK;
}

void tal_extend_destroy(tal_extend *h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (Buf *b : {&h->f, &h->p, &h->v, &h->lhas, &h->rhas, &h->lidx, &h->ridx, &h->rbytes,
                   &h->qbytes, &h->roff, &h->qoff, &h->lq, &h->lt, &h->rq, &h->rt, &h->left,
                   &h->right, &h->ref, &h->qer, &h->err, &h->cubtmp, &h->totals, &h->h_genome,
                   &h->h_reads, &h->h_read_off, &h->h_anchors, &h->h_anchor_off, &h->h_out})
        b->release();
    if (h->host) (void)hipHostFree(h->host);
    for (auto e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    tal_bsw_destroy(h->bsw);
    tal_chain_destroy(h->chain);
    delete h;
}

const char *tal_extend_last_error(const tal_extend *h) { return h ? h->errmsg.c_str() : "null handle"; }

int tal_extend_produce_device(tal_extend *h, const uint8_t *d_genome, int64_t genome_len,
                              const uint8_t *d_reads, const int64_t *d_read_off, int32_t nreads,
                              const tal_chain_anchor *d_anchors, const int64_t *d_anchor_off,
                              const uint32_t *d_f, const int32_t *d_p, tal_extend_read *d_out,
                              void *stream)
{
    const int c = check_common(h, d_genome, genome_len, d_reads, d_read_off, nreads, d_anchors, d_anchor_off,
                               d_out);
    if (c != TAL_OK) return c;
    h->st = tal_extend_stats{};
    if (nreads == 0) return TAL_OK;
    if (!d_f || !d_p) return fail(h, TAL_EINVAL, "null buffer");
    EXCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    return produce(h, d_genome, genome_len, d_reads, d_read_off, nreads, d_anchors, d_anchor_off,
                   d_f, d_p, d_out, s);
}

int tal_extend_run_device(tal_extend *h, const uint8_t *d_genome, int64_t genome_len,
                          const uint8_t *d_reads, const int64_t *d_read_off, int32_t nreads,
                          const tal_chain_anchor *d_anchors, const int64_t *d_anchor_off,
                          tal_extend_read *d_out, void *stream)
{
    const int c = check_common(h, d_genome, genome_len, d_reads, d_read_off, nreads, d_anchors, d_anchor_off,
                               d_out);
    if (c != TAL_OK) return c;
    h->st = tal_extend_stats{};
    if (nreads == 0) return TAL_OK;
    EXCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    /* anchors per read -> f, p, v */
    int64_t na = 0;
    EXCHK(h, hipMemcpyAsync(&h->host->ref_bytes, d_anchor_off + nreads, 8, hipMemcpyDeviceToHost,
                            s),
          "read anchor count");
    EXCHK(h, hipStreamSynchronize(s), "sync");
    na = h->host->ref_bytes;
    if (na < 0) return fail(h, TAL_EINVAL, "anchor offsets");
    const size_t nn = (size_t)std::max<int64_t>(na, 1);
    EXCHK(h, h->f.ensure(4 * nn), "alloc f");
    EXCHK(h, h->p.ensure(4 * nn), "alloc p");
    EXCHK(h, h->v.ensure(4 * nn), "alloc v");
    EXCHK(h, hipEventRecord(h->ev[0], s), "event");
    int rc = tal_chain_dp_device(h->chain, d_anchors, d_anchor_off, nreads, h->prm.dr, h->prm.dq,
                                 h->f.as<uint32_t>(), h->p.as<int32_t>(), h->v.as<int32_t>(), s);
    if (rc != TAL_OK) return fail(h, rc, "chaining: %s", tal_chain_last_error(h->chain));
    EXCHK(h, hipEventRecord(h->ev[1], s), "event");
    rc = produce(h, d_genome, genome_len, d_reads, d_read_off, nreads, d_anchors, d_anchor_off,
                 h->f.as<uint32_t>(), h->p.as<int32_t>(), d_out, s);
    if (rc != TAL_OK) return rc;
    const tal_extend_stats counts = h->st;
    EXCHK(h, hipEventRecord(h->ev[2], s), "event");
    const ExtParams P = ext_params(h, nreads, genome_len);
    const ExtBuffers B = pass_buffers(h, d_out);
    if (counts.left_pairs > 0) {
        rc = tal_bsw_get_scores16_device(h->bsw, h->left.as<tal_seqpair>(), h->ref.as<uint8_t>(),
                                            h->qer.as<uint8_t>(), counts.left_pairs, h->prm.w, s);
        if (rc != TAL_OK) return fail(h, rc, "left extension: %s", tal_bsw_last_error(h->bsw));
    }
    EXCHK(h, launch_extend_right_h0(P, B, s), "right h0");
    EXCHK(h, hipEventRecord(h->ev[3], s), "event");
    if (counts.right_pairs > 0) {
        rc = tal_bsw_get_scores16_device(h->bsw, h->right.as<tal_seqpair>(),
                                            h->ref.as<uint8_t>(), h->qer.as<uint8_t>(),
                                            counts.right_pairs, h->prm.w, s);
        if (rc != TAL_OK) return fail(h, rc, "right extension: %s", tal_bsw_last_error(h->bsw));
    }
    EXCHK(h, launch_extend_finish(P, B, s), "finish");
    EXCHK(h, hipEventRecord(h->ev[4], s), "event");
    EXCHK(h, hipEventSynchronize(h->ev[4]), "sync");
    float ms[4] = {};
    for (int i = 0; i < 4; i++) EXCHK(h, hipEventElapsedTime(&ms[i], h->ev[i], h->ev[i + 1]), "time");
    h->st = counts;
    h->st.ms_chain = ms[0];
    h->st.ms_produce = ms[1];
    h->st.ms_left = ms[2];
    h->st.ms_right = ms[3];
    h->st.ms_total = ms[0] + ms[1] + ms[2] + ms[3];
    return TAL_OK;
}

int tal_extend_run(tal_extend *h, const uint8_t *genome, int64_t genome_len, const uint8_t *reads,
                   const int64_t *read_off, int32_t nreads, const tal_chain_anchor *anchors,
                   const int64_t *anchor_off, tal_extend_read *out)
{
    const int c = check_common(h, genome, genome_len, reads, read_off, nreads, anchors, anchor_off, out);
    if (c != TAL_OK) return c;
    h->st = tal_extend_stats{};
    if (nreads == 0) return TAL_OK;
    if (genome_len < 0 || read_off[0] != 0 || anchor_off[0] != 0)
        return fail(h, TAL_EINVAL, "offsets must start at 0");
    for (int32_t k = 0; k < nreads; k++)
        if (read_off[k + 1] < read_off[k] || anchor_off[k + 1] < anchor_off[k])
            return fail(h, TAL_EINVAL, "offsets must be non-decreasing");
    const int64_t nb = read_off[nreads], na = anchor_off[nreads];
    EXCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = h->stream;
    EXCHK(h, h->h_genome.ensure((size_t)std::max<int64_t>(genome_len, 1)), "alloc");
    EXCHK(h, h->h_reads.ensure((size_t)std::max<int64_t>(nb, 1)), "alloc");
    EXCHK(h, h->h_read_off.ensure(8 * ((size_t)nreads + 1)), "alloc");
    EXCHK(h, h->h_anchors.ensure(sizeof(tal_chain_anchor) * (size_t)std::max<int64_t>(na, 1)),
          "alloc");
    EXCHK(h, h->h_anchor_off.ensure(8 * ((size_t)nreads + 1)), "alloc");
    EXCHK(h, h->h_out.ensure(sizeof(tal_extend_read) * (size_t)nreads), "alloc");
    EXCHK(h, hipMemcpyAsync(h->h_genome.p, genome, (size_t)genome_len, hipMemcpyHostToDevice, s),
          "copy");
    EXCHK(h, hipMemcpyAsync(h->h_reads.p, reads, (size_t)nb, hipMemcpyHostToDevice, s), "copy");
    EXCHK(h, hipMemcpyAsync(h->h_read_off.p, read_off, 8 * ((size_t)nreads + 1),
                            hipMemcpyHostToDevice, s), "copy");
    EXCHK(h, hipMemcpyAsync(h->h_anchors.p, anchors, sizeof(tal_chain_anchor) * (size_t)na,
                            hipMemcpyHostToDevice, s), "copy");
    EXCHK(h, hipMemcpyAsync(h->h_anchor_off.p, anchor_off, 8 * ((size_t)nreads + 1),
                            hipMemcpyHostToDevice, s), "copy");
    const int rc = tal_extend_run_device(h, h->h_genome.as<uint8_t>(), genome_len,
                                         h->h_reads.as<uint8_t>(), h->h_read_off.as<int64_t>(),
                                         nreads, h->h_anchors.as<tal_chain_anchor>(),
                                         h->h_anchor_off.as<int64_t>(),
                                         h->h_out.as<tal_extend_read>(), s);
    if (rc != TAL_OK) return rc;
    EXCHK(h, hipMemcpyAsync(out, h->h_out.p, sizeof(tal_extend_read) * (size_t)nreads,
                            hipMemcpyDeviceToHost, s), "copy");
    EXCHK(h, hipStreamSynchronize(s), "sync");
    return TAL_OK;
}

int tal_extend_get_pairs(tal_extend *h, tal_seqpair *left, tal_seqpair *right, uint8_t *ref,
                         uint8_t *qer)
{
    if (!h) return TAL_EINVAL;
    EXCHK(h, hipSetDevice(h->device), "set device");
    EXCHK(h, hipStreamSynchronize(h->stream), "sync");
    if (left && h->st.left_pairs > 0)
        EXCHK(h, hipMemcpy(left, h->left.p, sizeof(tal_seqpair) * h->st.left_pairs,
                           hipMemcpyDeviceToHost), "copy");
    if (right && h->st.right_pairs > 0)
        EXCHK(h, hipMemcpy(right, h->right.p, sizeof(tal_seqpair) * h->st.right_pairs,
                           hipMemcpyDeviceToHost), "copy");
    if (ref && h->st.ref_bytes > 0)
        EXCHK(h, hipMemcpy(ref, h->ref.p, (size_t)h->st.ref_bytes, hipMemcpyDeviceToHost), "copy");
    if (qer && h->st.qer_bytes > 0)
        EXCHK(h, hipMemcpy(qer, h->qer.p, (size_t)h->st.qer_bytes, hipMemcpyDeviceToHost), "copy");
    return TAL_OK;
}

int tal_extend_get_stats(const tal_extend *h, tal_extend_stats *out)
{
    if (!h || !out) return TAL_EINVAL;
    *out = h->st;
    return TAL_OK;
}

} /* extern "C" */
