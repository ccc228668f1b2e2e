/*
 * bsw_server.hip -- the streamed serving loop behind tal_bsw_server_* in
 * include/tal_bsw.h: arrival-order admission under a device-memory budget,
 * continuous batching over several worker streams, and per-request
 * fresh-object semantics.
 *
 * Threads.  One admission thread moves requests from `pending` to `ready`:
 * it waits for room in the arena, prepares the request's padded 8-bit and
 * 16-bit SeqPair blocks on the host and copies them with the sequence bytes
 * into the arena.  Each worker owns an engine handle (its own stream and
 * scratch) and repeatedly takes every ready request with the head's band
 * width, runs them as one batch, writes the outputs into the callers' arrays
 * and frees their arena slots.
 *
 * Independence.  A request's 16-bit pairs are padded to a multiple of 32 and
 * its 8-bit pairs to a multiple of 64 with explicit padding records (len1 0,
 * h0 0, offsets 0 in the request's own buffers, len2 0 or the request's last
 * 8-bit len2 -- the padding of one getScores call), so a batch's groups never
 * span requests; the 16-bit stale-column search runs per request from the
 * zero image (run_batch with seg_first).
 *
 * Resident mode (config.resident = 1).  One persistent kernel (bsw_serve)
 * stays on the GPU and pulls tasks from a ring in pinned host memory; the
 * host never launches a batch.  The admission thread copies a request's
 * sequences into the device arena (by DMA, under the same budget), writes
 * its padded pair blocks into a pinned arena and publishes one task per
 * 16-bit group pair and per 8-bit group, longest first, by raising the
 * ring's tail; waves that free up take them at once, so a request admitted
 * while others run joins the running kernel as slots free up.  A monitor
 * thread watches the tasks' done flags; when a request's groups are done it
 * publishes the request's stale-column resolution as one more task, then any
 * reruns that resolution asks for, and finally copies the outputs into the
 * caller's array.  The kernel leaves after a stretch without work and the
 * monitor launches it again when work is pending.
 */
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bsw_internal.h"

using namespace talbsw;

namespace {

constexpr size_t ALIGN = 256;
constexpr size_t MAX_BUDGET = ((size_t)1 << 31) - ALIGN;

size_t align_up(size_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

/* First-fit allocator over [0, size) with coalescing frees. */
struct Arena {
    std::map<size_t, size_t> free_; /* offset -> length */
    size_t used = 0, peak = 0;
    void reset(size_t size)
    {
        free_.clear();
        free_[0] = size;
        used = peak = 0;
    }
    bool take(size_t bytes, size_t *off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used += bytes;
            peak = std::max(peak, used);
            return true;
        }
        return false;
    }
    void give(size_t off, size_t bytes)
    {
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Request {
    uint64_t ticket = 0;
    tal_seqpair *pairs = nullptr;
    const uint8_t *ref = nullptr, *qer = nullptr;
    size_t ref_len = 0, qer_len = 0;
    int32_t n = 0, w = 0;
    /* prepared */
    std::vector<int32_t> idx8, idx16; /* caller positions of the 8/16-bit pairs */
    std::vector<tal_seqpair> hb8, hb16; /* host-batched mode: the padded blocks, built at admission */
    int32_t n8p = 0, n16p = 0;        /* padded counts */
    size_t off = 0, bytes = 0;        /* arena slot */
    size_t ref_off = 0, qer_off = 0, p8_off = 0, p16_off = 0; /* absolute arena offsets */
    bool admitted = false;
    /* resident mode */
    int slot = -1;                  /* ServeReq slot */
    size_t pin_off = 0, pin_bytes = 0; /* pinned-arena slot */
    tal_seqpair *hp16 = nullptr, *hp8 = nullptr; /* pinned mirrors of the pair blocks */
    std::vector<ServeTask> gtasks;  /* group tasks, published once the copy-in is done */
    int32_t G16 = 0, G8 = 0;        /* groups of each width */
    int64_t t_first = 0, t_end = 0; /* the tasks being waited for: [t_first, t_end) */
    int64_t cursor = 0;             /* first of them not seen done */
    int phase = 0;                  /* 0 groups, 1 resolution, 2 reruns */
    int in_flight = 0;              /* requests on the device when admitted (itself included) */
    bool keep = false;              /* its tasks may still be in the ring: never reuse its memory */
    /* result */
    int status = TAL_OK;
    int32_t batch = -1, batch_requests = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    bool done = false;
};

} // namespace

namespace {

/* Host copies split over a few helper threads (the admission thread takes
 * the first slice): a request's sequences go into its pinned mirror at the
 * rate of several cores. */
struct CopyPool {
    struct Slice {
        uint8_t *dst = nullptr;
        const uint8_t *src = nullptr;
        size_t n = 0;
    };
    std::mutex mu;
    std::condition_variable go, done;
    std::vector<Slice> slices; /* one per helper */
    std::vector<std::thread> helpers;
    uint64_t round = 0;
    int pending = 0;
    bool stop = false;

    void start(int k)
    {
        slices.resize((size_t)k);
        for (int i = 0; i < k; i++) helpers.emplace_back([this, i] { run(i); });
    }
    void run(int i)
    {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            go.wait(lk, [&] { return stop || round != seen; });
            if (stop) return;
            seen = round;
            const Slice sl = slices[(size_t)i];
            lk.unlock();
            if (sl.n) std::memcpy(sl.dst, sl.src, sl.n);
            lk.lock();
            if (--pending == 0) done.notify_all();
        }
    }
    void copy(uint8_t *dst, const uint8_t *src, size_t n)
    {
        const size_t parts = std::min(helpers.size() + 1, std::max<size_t>(1, n >> 20));
        if (parts <= 1) {
            std::memcpy(dst, src, n);
            return;
        }
        const size_t per = (n / parts + 255) & ~(size_t)255;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (size_t i = 0; i < helpers.size(); i++) {
                const size_t lo = std::min(n, per * (i + 1)), hi = std::min(n, per * (i + 2));
                slices[i] = Slice{dst + lo, src + lo, hi - lo};
            }
            pending = (int)helpers.size();
            round++;
        }
        go.notify_all();
        std::memcpy(dst, src, std::min(per, n));
        std::unique_lock<std::mutex> lk(mu);
        done.wait(lk, [&] { return pending == 0; });
    }
    void shutdown()
    {
        {
            std::unique_lock<std::mutex> lk(mu);
            stop = true;
        }
        go.notify_all();
        for (std::thread &t : helpers)
            if (t.joinable()) t.join();
        helpers.clear();
    }
};

/* Resident-mode state: the task ring and request slots in pinned host
 * memory, the claim counters and state slabs of the persistent kernel in
 * device memory. */
struct Resident {
    int64_t cap = 0;           /* ring capacity (a power of two) */
    ServeTask *ring = nullptr; /* pinned */
    int32_t *done = nullptr;   /* pinned, one flag per ring entry */
    ServeReq *reqs = nullptr;  /* pinned */
    ServeCtl *ctl = nullptr;   /* pinned */
    uint8_t *pin = nullptr;    /* pinned arena: pair blocks, group records, staging */
    size_t pin_size = 0;
    Arena pin_alloc;
    int64_t *dev = nullptr;    /* device: the kernel's words (launch_serve) */
    uint32_t *slabs = nullptr;                     /* device */
    int grid = 0;
    size_t lds_bytes = 0;
    KParams K16{}, K8{};
    int max_qlen = 0;
    int64_t tail = 0; /* tasks published */
    std::vector<int> free_slots;
    std::vector<Request *> live; /* requests with tasks on the device */
    hipStream_t kstream = nullptr;
    hipEvent_t kdone = nullptr;
    bool launched = false;
    int64_t idle_ticks = 0;
    std::thread monitor;
    CopyPool copies;
};

constexpr int RES_SLOTS = 4096;
constexpr int64_t RES_RING = (int64_t)1 << 20;
constexpr size_t RES_PINNED = (size_t)1 << 30;
constexpr int RES_WAVES_PER_CU = 8; /* two per SIMD: room for other kernels' waves */

} // namespace

struct tal_bsw_server {
    tal_bsw_params prm{};
    bool resident = false;
    Resident *res = nullptr;
    int device = 0;
    int smax = 0;
    size_t budget = 0;
    int32_t max_batch = 0;
    uint8_t *arena = nullptr;
    Arena alloc;
    std::vector<tal_bsw *> engines;
    std::mutex mu;
    std::condition_variable cv; /* every state change */
    std::deque<Request *> pending, ready;
    std::unordered_map<uint64_t, Request *> all;
    uint64_t next_ticket = 1;
    int32_t next_batch = 0;
    int active_batches = 0;
    bool admitting = false; /* the admission thread holds a request between queues */
    bool stop = false;
    std::thread admitter;
    std::vector<std::thread> workers;
    tal_bsw_server_stats st{};
};

namespace {

/* Host-side checks of one request (the limits of the 16-bit path and the
 * buffer extents); the dispatcher split and padded blocks. */
int prepare(tal_bsw_server *s, Request *r)
{
    for (int32_t k = 0; k < r->n; k++) {
        const tal_seqpair &p = r->pairs[k];
        if (p.len1 < 0 || p.len2 < 0 || p.idr < 0 || p.idq < 0) return TAL_EINVAL;
        if (p.len1 > MAX_LEN16 || p.len2 > MAX_LEN16) return TAL_ERANGE;
        if ((size_t)p.idr + (size_t)p.len1 > r->ref_len ||
            (size_t)p.idq + (size_t)p.len2 > r->qer_len)
            return TAL_EINVAL;
        (goes_8bit(p, s->smax) ? r->idx8 : r->idx16).push_back(k);
    }
    const int32_t n8 = (int32_t)r->idx8.size(), n16 = (int32_t)r->idx16.size();
    r->n8p = (n8 + 63) / 64 * 64;
    r->n16p = (n16 + 31) / 32 * 32;
    /* host-batched footprint: the sequences (a batch's blocks live in its
     * worker's buffers; resident mode sizes its own layout) */
    r->bytes = align_up(std::max<size_t>(r->ref_len, 1)) + align_up(std::max<size_t>(r->qer_len, 1));
    return TAL_OK;
}

/* The padded blocks with offsets moved into the arena, into out[0, npad). */
void build_block_at(const Request *r, const std::vector<int32_t> &idx, int32_t npad, bool eight,
                    tal_seqpair *out)
{
    const int32_t m = (int32_t)idx.size();
    for (int32_t t = 0; t < m; t++) {
        tal_seqpair p = r->pairs[idx[t]];
        p.idr = (int32_t)(r->ref_off + (size_t)p.idr);
        p.idq = (int32_t)(r->qer_off + (size_t)p.idq);
        out[t] = p;
    }
    for (int32_t t = m; t < npad; t++) { /* the padding of one getScores call */
        tal_seqpair p{};
        p.idr = (int32_t)r->ref_off;
        p.idq = (int32_t)r->qer_off;
        p.len2 = eight ? r->pairs[idx[m - 1]].len2 : 0;
        out[t] = p;
    }
}
void build_block(const Request *r, const std::vector<int32_t> &idx, int32_t npad, bool eight,
                 std::vector<tal_seqpair> &out)
{
    out.assign(npad, tal_seqpair{});
    build_block_at(r, idx, npad, eight, out.data());
}

void finish(tal_bsw_server *s, Request *r, int status)
{
    /* called with s->mu held */
    r->status = status;
    r->t_done = now_ms();
    r->done = true;
    if (r->admitted && !r->keep) s->alloc.give(r->off, r->bytes);
    if (status == TAL_OK) {
        s->st.requests++;
        s->st.pairs += r->n;
    }
}

/* Admission of the host-batched mode: arrival order, room in the arena, the
 * padded blocks built on the host.  The copies to the device are left to the
 * worker that runs the request, on its own stream (a copy stream of its own
 * here could share a hardware queue with a worker's stream, and each
 * admission would then wait for a whole batch). */
void admit_loop(tal_bsw_server *s)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->pending.empty(); });
            if (s->pending.empty()) return; /* stop with nothing queued */
            r = s->pending.front();
            s->admitting = true;
        }
        int rc = prepare(s, r);
        if (rc == TAL_OK && r->bytes > s->budget) rc = TAL_ERANGE;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (rc == TAL_OK) /* arrival order: wait for room for the head */
                s->cv.wait(lk, [&] { return s->alloc.take(r->bytes, &r->off); });
            s->pending.pop_front();
            if (rc != TAL_OK) {
                finish(s, r, rc);
                s->admitting = false;
                s->cv.notify_all();
                continue;
            }
            r->admitted = true;
        }
        r->ref_off = r->off;
        r->qer_off = r->ref_off + align_up(std::max<size_t>(r->ref_len, 1));
        if (r->n8p) build_block(r, r->idx8, r->n8p, true, r->hb8);
        if (r->n16p) build_block(r, r->idx16, r->n16p, false, r->hb16);
        std::unique_lock<std::mutex> lk(s->mu);
        r->t_admit = now_ms();
        s->ready.push_back(r);
        s->admitting = false;
        s->cv.notify_all();
    }
}

struct WorkerBufs {
    tal_seqpair *p8 = nullptr, *p16 = nullptr;
    int32_t *seg = nullptr;
    size_t cap8 = 0, cap16 = 0, capseg = 0;
    std::vector<tal_seqpair> host;
    std::vector<int32_t> hseg;
    ~WorkerBufs()
    {
        if (p8) (void)hipFree(p8);
        if (p16) (void)hipFree(p16);
        if (seg) (void)hipFree(seg);
    }
    template <class T> static hipError_t grow(T **p, size_t *cap, size_t n)
    {
        if (n <= *cap) return hipSuccess;
        /* 1.5x growth from the old capacity, so that a slowly growing batch
         * size does not reallocate every time */
        const size_t want = std::max(n, *cap * 3 / 2);
        if (*p) {
            const hipError_t f = hipFree(*p);
            *p = nullptr;
            *cap = 0;
            if (f != hipSuccess) return f;
        }
        hipError_t e = hipMalloc((void **)p, sizeof(T) * want);
        if (e == hipSuccess) *cap = want;
        return e;
    }
};

/* One batch: gather the requests' blocks, run both widths, write outputs. */
int run_one(tal_bsw_server *s, tal_bsw *eng, WorkerBufs &wb, std::vector<Request *> &batch)
{
    hipStream_t st = handle_stream(eng);
    const int32_t w = batch[0]->w;
    int64_t n8 = 0, n16 = 0;
    for (Request *r : batch) {
        n8 += r->n8p;
        n16 += r->n16p;
    }
    if (wb.grow(&wb.p8, &wb.cap8, (size_t)std::max<int64_t>(n8, 1)) != hipSuccess ||
        wb.grow(&wb.p16, &wb.cap16, (size_t)std::max<int64_t>(n16, 1)) != hipSuccess ||
        wb.grow(&wb.seg, &wb.capseg, (size_t)std::max<int64_t>(n16 / 32, 1)) != hipSuccess)
        return TAL_ENOMEM;
    wb.hseg.clear();
    int64_t o8 = 0, o16 = 0;
    /* each request's sequences into its arena slot and its blocks into the
     * batch, on this worker's stream, ahead of its kernels */
    for (Request *r : batch) {
        if (r->ref_len && hipMemcpyAsync(s->arena + r->ref_off, r->ref, r->ref_len,
                                         hipMemcpyHostToDevice, st) != hipSuccess)
            return TAL_EDEVICE;
        if (r->qer_len && hipMemcpyAsync(s->arena + r->qer_off, r->qer, r->qer_len,
                                         hipMemcpyHostToDevice, st) != hipSuccess)
            return TAL_EDEVICE;
        if (r->n8p &&
            hipMemcpyAsync(wb.p8 + o8, r->hb8.data(), sizeof(tal_seqpair) * r->n8p,
                           hipMemcpyHostToDevice, st) != hipSuccess)
            return TAL_EDEVICE;
        if (r->n16p &&
            hipMemcpyAsync(wb.p16 + o16, r->hb16.data(), sizeof(tal_seqpair) * r->n16p,
                           hipMemcpyHostToDevice, st) != hipSuccess)
            return TAL_EDEVICE;
        const int32_t first = (int32_t)(o16 / 32);
        for (int32_t g = 0; g < r->n16p / 32; g++) wb.hseg.push_back(first);
        o8 += r->n8p;
        o16 += r->n16p;
    }
    if (n16 && hipMemcpyAsync(wb.seg, wb.hseg.data(), sizeof(int32_t) * wb.hseg.size(),
                              hipMemcpyHostToDevice, st) != hipSuccess)
        return TAL_EDEVICE;
    const uint8_t *base = s->arena;
    if (n8 > INT32_MAX || n16 > INT32_MAX) return TAL_ERANGE;
    int rc = TAL_OK;
    if (n8) rc = run_batch(eng, 8, wb.p8, base, base, (int32_t)n8, w, st, nullptr);
    if (rc == TAL_OK && n16) rc = run_batch(eng, 16, wb.p16, base, base, (int32_t)n16, w, st, wb.seg);
    if (rc != TAL_OK) return rc;
    /* outputs back into the callers' arrays */
    wb.host.resize((size_t)(n8 + n16));
    if (n8 && hipMemcpyAsync(wb.host.data(), wb.p8, sizeof(tal_seqpair) * n8,
                             hipMemcpyDeviceToHost, st) != hipSuccess)
        return TAL_EDEVICE;
    if (n16 && hipMemcpyAsync(wb.host.data() + n8, wb.p16, sizeof(tal_seqpair) * n16,
                              hipMemcpyDeviceToHost, st) != hipSuccess)
        return TAL_EDEVICE;
    if (hipStreamSynchronize(st) != hipSuccess) return TAL_EDEVICE;
    o8 = 0;
    o16 = n8;
    for (Request *r : batch) {
        auto put = [&](const std::vector<int32_t> &idx, int64_t o) {
            for (size_t t = 0; t < idx.size(); t++) {
                const tal_seqpair &q = wb.host[(size_t)o + t];
                tal_seqpair &d = r->pairs[idx[t]];
                d.score = q.score;
                d.tle = q.tle;
                d.gtle = q.gtle;
                d.qle = q.qle;
                d.gscore = q.gscore;
                d.max_off = q.max_off;
            }
        };
        put(r->idx8, o8);
        put(r->idx16, o16);
        o8 += r->n8p;
        o16 += r->n16p;
    }
    return TAL_OK;
}

void work_loop(tal_bsw_server *s, int k)
{
    (void)hipSetDevice(s->device);
    tal_bsw *eng = s->engines[k];
    WorkerBufs wb;
    for (;;) {
        std::vector<Request *> batch;
        int32_t id = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] {
                return !s->ready.empty() || (s->stop && s->pending.empty() && !s->admitting);
            });
            if (s->ready.empty()) return;
            /* every ready request with the head's band width, in order */
            const int32_t w = s->ready.front()->w;
            int64_t total = 0;
            for (auto it = s->ready.begin(); it != s->ready.end();) {
                Request *r = *it;
                if (r->w == w && (batch.empty() || total + r->n <= s->max_batch)) {
                    total += r->n;
                    batch.push_back(r);
                    it = s->ready.erase(it);
                } else {
                    ++it;
                }
            }
            id = s->next_batch++;
            s->active_batches++;
        }
        const double t0 = now_ms();
        const int rc = run_one(s, eng, wb, batch);
        const double t1 = now_ms();
        for (Request *r : batch) { /* (the blocks are on the device or no longer needed) */
            std::vector<tal_seqpair>().swap(r->hb8);
            std::vector<tal_seqpair>().swap(r->hb16);
        }
        std::unique_lock<std::mutex> lk(s->mu);
        s->active_batches--;
        s->st.batches++;
        s->st.busy_ms += t1 - t0;
        for (Request *r : batch) {
            r->batch = id;
            r->batch_requests = (int32_t)batch.size();
            finish(s, r, rc);
        }
        s->st.peak_bytes = (int64_t)s->alloc.peak;
        s->cv.notify_all();
    }
}

/* ---------------------------------------------------------- resident mode */

bool res_running(const Resident *R)
{
    return R->launched && hipEventQuery(R->kdone) == hipErrorNotReady;
}

/* Launches the persistent kernel unless it is running (s->mu held). */
int res_ensure_kernel(tal_bsw_server *s)
{
    Resident *R = s->res;
    if (res_running(R)) return TAL_OK;
    if (launch_serve(R->grid, R->lds_bytes, R->K16, R->K8, R->ring, (int32_t)(R->cap - 1), R->reqs,
                     R->ctl, R->done, R->dev, R->slabs, R->idle_ticks, 4, R->kstream) != hipSuccess ||
        hipEventRecord(R->kdone, R->kstream) != hipSuccess)
        return TAL_EDEVICE;
    R->launched = true;
    s->st.batches++;
    return TAL_OK;
}

/* The first task index still on the device (every earlier ring entry may be
 * written again). */
int64_t res_oldest(const Resident *R)
{
    int64_t m = R->tail;
    for (const Request *r : R->live)
        if (r->cursor < r->t_end) m = std::min(m, r->cursor);
    return m;
}

/* Publishes tasks for request r (s->mu held): they become the range the
 * monitor waits for.  Returns false, publishing nothing, if the ring has no
 * room for them yet. */
bool res_publish(tal_bsw_server *s, Request *r, const std::vector<ServeTask> &tasks)
{
    Resident *R = s->res;
    const int64_t n = (int64_t)tasks.size();
    if (R->tail + n - res_oldest(R) > R->cap) return false;
    for (int64_t k = 0; k < n; k++) R->ring[(R->tail + k) & (R->cap - 1)] = tasks[(size_t)k];
    r->t_first = r->cursor = R->tail;
    r->t_end = R->tail + n;
    R->tail += n;
    __atomic_store_n(&R->ctl->tail, R->tail, __ATOMIC_RELEASE);
    return true;
}

/* Group shapes and longest-first tasks of a prepared request whose blocks
 * are in place. */
std::vector<ServeTask> res_group_tasks(const Request *r, const tal_seqpair *p16,
                                       const tal_seqpair *p8, int32_t *ncol16)
{
    struct G {
        int64_t cost;
        int bits, g;
    };
    std::vector<G> gs;
    auto cost = [&](const tal_seqpair *p, int lanes, int w, int *ncol_out) {
        int nr = 0, nc = 0;
        for (int l = 0; l < lanes; l++) {
            nr = std::max(nr, p[l].len1);
            nc = std::max(nc, p[l].len2);
        }
        if (ncol_out) *ncol_out = nc;
        return (int64_t)nr * std::min(nc, 2 * w + 2);
    };
    for (int g = 0; g < r->G16; g++) gs.push_back({cost(p16 + 32 * g, 32, r->w, &ncol16[g]), 16, g});
    for (int g = 0; g < r->G8; g++) gs.push_back({cost(p8 + 64 * g, 64, r->w, nullptr), 8, g});
    std::stable_sort(gs.begin(), gs.end(), [](const G &a, const G &b) { return a.cost > b.cost; });
    std::vector<ServeTask> t;
    int pend16 = -1; /* a 16-bit group waiting for its partner */
    for (const G &x : gs) {
        if (x.bits == 8) {
            t.push_back({r->slot, 8, x.g, -1});
        } else if (pend16 < 0) {
            pend16 = (int)t.size();
            t.push_back({r->slot, 16, x.g, -1});
        } else {
            t[(size_t)pend16].g1 = x.g;
            pend16 = -1;
        }
    }
    return t;
}

/* Outputs into the caller's array, slots and memory released (s->mu held). */
void res_finish(tal_bsw_server *s, Request *r, int status)
{
    Resident *R = s->res;
    if (status == TAL_OK) {
        /* (the copy-out tasks brought the pair blocks back to the mirror) */
        const tal_seqpair *p16 = r->hp16, *p8 = r->hp8;
        auto put = [&](const std::vector<int32_t> &idx, const tal_seqpair *blk) {
            for (size_t t = 0; t < idx.size(); t++) {
                const tal_seqpair &q = blk[t];
                tal_seqpair &d = r->pairs[idx[t]];
                d.score = q.score;
                d.tle = q.tle;
                d.gtle = q.gtle;
                d.qle = q.qle;
                d.gscore = q.gscore;
                d.max_off = q.max_off;
            }
        };
        put(r->idx16, p16);
        put(r->idx8, p8);
    }
    if (!r->keep) {
        if (r->slot >= 0) R->free_slots.push_back(r->slot);
        if (r->pin_bytes) R->pin_alloc.give(r->pin_off, r->pin_bytes);
    }
    R->live.erase(std::remove(R->live.begin(), R->live.end(), r), R->live.end());
    r->batch = status == TAL_OK ? 0 : -1;
    r->batch_requests = r->in_flight;
    finish(s, r, status);
    s->cv.notify_all();
}

/* Device-arena layout of a request (offsets from its slot): the sequences,
 * the pair blocks, then the 16-bit bookkeeping; `mirror` bytes from the
 * start (sequences and pair blocks) are also laid out in pinned memory. */
struct DevLayout {
    size_t ref, qer, p16, p8, dep, wrote, fmask, bytes, mirror;
};
DevLayout dev_layout(const Request *r, size_t fcols)
{
    DevLayout D{};
    size_t o = 0;
    auto take = [&](size_t b) {
        const size_t at = o;
        o += align_up(std::max<size_t>(b, 1));
        return at;
    };
    D.ref = take(r->ref_len);
    D.qer = take(r->qer_len);
    D.p16 = take(sizeof(tal_seqpair) * (size_t)r->n16p);
    D.p8 = take(sizeof(tal_seqpair) * (size_t)r->n8p);
    D.mirror = o;
    D.dep = take(sizeof(uint32_t) * (size_t)r->G16);
    D.wrote = take((size_t)r->G16);
    D.fmask = take(sizeof(uint32_t) * fcols);
    D.bytes = o;
    return D;
}

/* Pinned-arena layout of a request: the mirror of the device layout's
 * sequences and pair blocks, then the per-group records. */
struct PinLayout {
    size_t ncol, foff, stale, reruns, bytes;
};
PinLayout pin_layout(const Request *r, const DevLayout &D)
{
    PinLayout L{};
    size_t o = D.mirror;
    auto take = [&](size_t b) {
        const size_t at = o;
        o += align_up(std::max<size_t>(b, 1));
        return at;
    };
    L.ncol = take(sizeof(int32_t) * (size_t)r->G16);
    L.foff = take(sizeof(int64_t) * (size_t)r->G16);
    L.stale = take(sizeof(uint32_t) * (size_t)r->G16);
    L.reruns = take(sizeof(int32_t) * ((size_t)r->G16 + 1));
    L.bytes = o;
    return L;
}

void res_admit_loop(tal_bsw_server *s)
{
    (void)hipSetDevice(s->device);
    Resident *R = s->res;
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->pending.empty(); });
            if (s->pending.empty()) return; /* stop with nothing queued */
            r = s->pending.front();
            s->admitting = true;
        }
        int rc = prepare(s, r);
        /* the kernel's state slabs hold max_qlen columns */
        for (int32_t k = 0; k < r->n && rc == TAL_OK; k++)
            if (r->pairs[k].len2 > R->max_qlen) rc = TAL_ERANGE;
        r->G16 = r->n16p / 32;
        r->G8 = r->n8p / 64;
        PinLayout L{};
        DevLayout D{};
        int64_t nin = 0; /* copy-in chunks */
        if (rc == TAL_OK) {
            size_t fcols = 0; /* F-mask words of the 16-bit groups */
            for (int32_t g = 0; g < r->G16; g++) {
                int nc = 0;
                for (int l = 0; l < 32 && 32 * g + l < (int32_t)r->idx16.size(); l++)
                    nc = std::max(nc, r->pairs[r->idx16[32 * g + l]].len2);
                fcols += (size_t)nc;
            }
            D = dev_layout(r, fcols);
            L = pin_layout(r, D);
            r->bytes = D.bytes; /* device footprint */
            nin = ((int64_t)D.mirror + SERVE_COPY_CHUNK - 1) / SERVE_COPY_CHUNK;
            if (r->bytes > s->budget || L.bytes > R->pin_size) rc = TAL_ERANGE;
        }
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (rc == TAL_OK) /* arrival order: wait for room for the head */
                s->cv.wait(lk, [&] {
                    if (R->free_slots.empty()) return false;
                    if (R->tail - res_oldest(R) + r->G16 + r->G8 + 2 * nin + 2 > R->cap / 2)
                        return false;
                    size_t po = 0, dof = 0;
                    if (!R->pin_alloc.take(L.bytes, &po)) return false;
                    if (!s->alloc.take(r->bytes, &dof)) {
                        R->pin_alloc.give(po, L.bytes);
                        return false;
                    }
                    r->pin_off = po;
                    r->pin_bytes = L.bytes;
                    r->off = dof;
                    return true;
                });
            s->pending.pop_front();
            if (rc != TAL_OK) {
                finish(s, r, rc);
                s->admitting = false;
                s->cv.notify_all();
                continue;
            }
            r->admitted = true;
            r->slot = R->free_slots.back();
            R->free_slots.pop_back();
        }
        /* device arena: ref, qer, pair blocks, dep, wrote, fmask */
        r->ref_off = r->off + D.ref;
        r->qer_off = r->off + D.qer;
        r->p16_off = r->off + D.p16;
        r->p8_off = r->off + D.p8;
        /* The pinned mirror: the host fills it, the kernel's copy-in tasks
         * move it to the device and its copy-out tasks bring the pair blocks
         * back.  (No DMA and no other stream: work queued on another stream
         * can sit behind the persistent kernel when the two share a
         * hardware queue.) */
        uint8_t *pin = R->pin + r->pin_off;
        if (r->ref_len) R->copies.copy(pin + D.ref, r->ref, r->ref_len);
        if (r->qer_len) R->copies.copy(pin + D.qer, r->qer, r->qer_len);
        tal_seqpair *p16 = (tal_seqpair *)(pin + D.p16), *p8 = (tal_seqpair *)(pin + D.p8);
        r->hp16 = p16;
        r->hp8 = p8;
        int32_t *ncol = (int32_t *)(pin + L.ncol);
        int64_t *foff = (int64_t *)(pin + L.foff);
        uint32_t *stale = (uint32_t *)(pin + L.stale);
        int32_t *reruns = (int32_t *)(pin + L.reruns);
        if (r->n16p) build_block_at(r, r->idx16, r->n16p, false, p16);
        if (r->n8p) build_block_at(r, r->idx8, r->n8p, true, p8);
        r->gtasks = res_group_tasks(r, p16, p8, ncol);
        int64_t fo = 0;
        for (int32_t g = 0; g < r->G16; g++) {
            foff[g] = fo;
            fo += ncol[g];
        }
        std::memset(stale, 0, sizeof(uint32_t) * (size_t)r->G16);
        reruns[0] = 0;
        ServeReq q{};
        q.cin_src = pin;
        q.cin_dst = s->arena + r->off;
        q.cin_bytes = (int64_t)D.mirror;
        q.cout_src = s->arena + r->p16_off;
        q.cout_dst = pin + D.p16;
        q.cout_bytes = (int64_t)(D.mirror - D.p16);
        q.pairs16 = (tal_seqpair *)(s->arena + r->p16_off);
        q.pairs8 = (tal_seqpair *)(s->arena + r->p8_off);
        q.seq = s->arena;
        q.stale = stale;
        q.dep = (uint32_t *)(s->arena + r->off + D.dep);
        q.wrote = s->arena + r->off + D.wrote;
        q.fmask = (uint32_t *)(s->arena + r->off + D.fmask);
        q.fmask_off = foff;
        q.ncol16 = ncol;
        q.reruns = reruns;
        q.n16 = r->n16p;
        q.n8 = r->n8p;
        q.w = r->w;
        std::vector<ServeTask> cin;
        for (int64_t k = 0; k < nin; k++) cin.push_back({r->slot, SERVE_COPY_IN, (int32_t)k, -1});
        std::unique_lock<std::mutex> lk(s->mu);
        R->reqs[r->slot] = q;
        r->t_admit = now_ms();
        R->live.push_back(r);
        r->in_flight = (int)R->live.size();
        r->phase = 3;
        /* room was checked at admission; should the monitor's tasks have
         * taken it meanwhile, wait for the device to drain some */
        while (!res_publish(s, r, cin)) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(50));
            lk.lock();
        }
        const int krc = res_ensure_kernel(s);
        if (krc != TAL_OK) {
            r->keep = true;
            res_finish(s, r, krc);
        }
        s->admitting = false;
        s->cv.notify_all();
    }
}

/* Advances each live request through its phases as its tasks finish. */
void res_monitor_loop(tal_bsw_server *s)
{
    (void)hipSetDevice(s->device);
    Resident *R = s->res;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        if (s->stop && s->pending.empty() && !s->admitting && R->live.empty()) return;
        bool moved = false; /* a request changed phase or finished */
        std::vector<Request *> live = R->live;
        for (Request *r : live) {
            while (r->cursor < r->t_end &&
                   __atomic_load_n(&R->done[r->cursor & (R->cap - 1)], __ATOMIC_ACQUIRE) ==
                       (int32_t)(r->cursor + 1))
                r->cursor++;
            if (r->cursor < r->t_end) continue;
            /* the copy-out of the pair blocks: the request's last phase */
            auto copy_out = [&]() {
                const int64_t nb = R->reqs[r->slot].cout_bytes;
                std::vector<ServeTask> t;
                for (int64_t k = 0; k * SERVE_COPY_CHUNK < nb; k++)
                    t.push_back({r->slot, SERVE_COPY_OUT, (int32_t)k, -1});
                /* (no room in the ring yet: the next round tries again) */
                if (res_publish(s, r, t)) {
                    r->phase = 4;
                    moved = true;
                }
            };
            if (r->phase == 3) { /* copied in: the groups */
                if (res_publish(s, r, r->gtasks)) {
                    r->phase = 0;
                    moved = true;
                }
                continue;
            }
            if (r->phase == 4) { /* copied out: done */
                res_finish(s, r, TAL_OK);
                moved = true;
                continue;
            }
            if (r->phase == 1) { /* a resolution finished: reruns, or done */
                const int32_t *rr = R->reqs[r->slot].reruns;
                const int32_t nrr = rr[0];
                if (nrr == 0) {
                    copy_out();
                    continue;
                }
                std::vector<ServeTask> t;
                for (int32_t k = 0; k < nrr; k += 2)
                    t.push_back({r->slot, 16, rr[1 + k], k + 1 < nrr ? rr[2 + k] : -1});
                /* (no room in the ring yet: the next round tries again) */
                if (res_publish(s, r, t)) {
                    r->phase = 2;
                    s->st.reruns += nrr;
                    moved = true;
                }
            } else if (r->G16 > 0) { /* groups or reruns finished: resolve (again) */
                if (res_publish(s, r, {ServeTask{r->slot, 0, 0, -1}})) {
                    r->phase = 1;
                    moved = true;
                }
            } else { /* no 16-bit groups */
                copy_out();
            }
        }
        if (!R->live.empty() && !res_running(R)) {
            const int krc = res_ensure_kernel(s);
            if (krc != TAL_OK)
                for (Request *r : std::vector<Request *>(R->live)) {
                    r->keep = true;
                    res_finish(s, r, krc);
                }
        }
        if (!moved) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lk.lock();
        }
    }
}

void res_release(tal_bsw_server *s)
{
    Resident *R = s->res;
    if (!R) return;
    R->copies.shutdown();
    if (R->ctl) __atomic_store_n(&R->ctl->stop, 1, __ATOMIC_RELEASE);
    if (R->launched) (void)hipEventSynchronize(R->kdone);
    if (R->kdone) (void)hipEventDestroy(R->kdone);
    if (R->kstream) (void)hipStreamDestroy(R->kstream);
    for (void *p : {(void *)R->dev, (void *)R->slabs})
        if (p) (void)hipFree(p);
    for (void *p : {(void *)R->ring, (void *)R->done, (void *)R->reqs, (void *)R->ctl, (void *)R->pin})
        if (p) (void)hipHostFree(p);
    delete R;
    s->res = nullptr;
}

int res_create(tal_bsw_server *s, int max_qlen)
{
    Resident *R = new (std::nothrow) Resident();
    if (!R) return TAL_ENOMEM;
    s->res = R;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, s->device) != hipSuccess) return TAL_EDEVICE;
    const size_t lds_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                                : prop.sharedMemPerBlock;
    int wpc = RES_WAVES_PER_CU;
    if (const char *e = getenv("TAL_BSW_SERVE_WAVES_PER_CU")) wpc = std::max(1, std::min(atoi(e), 12));
    const Sizing z = state_sizing(lds_cu, prop.sharedMemPerBlock, wpc, max_qlen);
    R->max_qlen = max_qlen;
    R->grid = prop.multiProcessorCount * wpc;
    R->lds_bytes = z.lds_bytes;
    for (int bits : {16, 8}) {
        KParams &K = bits == 16 ? R->K16 : R->K8;
        K = params_for(s->prm, bits);
        K.state_cols = z.state_cols;
        K.lds_cols = z.lds_cols;
        K.slab_cols = z.slab_cols;
        K.query_words = z.qwords;
    }
    R->cap = RES_RING;
    R->pin_size = RES_PINNED;
    R->idle_ticks = 5000000; /* 50 ms of the 100 MHz clock */
    const unsigned fl = hipHostMallocCoherent | hipHostMallocMapped;
    if (hipHostMalloc((void **)&R->ring, sizeof(ServeTask) * (size_t)R->cap, fl) != hipSuccess ||
        hipHostMalloc((void **)&R->done, sizeof(int32_t) * (size_t)R->cap, fl) != hipSuccess ||
        hipHostMalloc((void **)&R->reqs, sizeof(ServeReq) * RES_SLOTS, fl) != hipSuccess ||
        hipHostMalloc((void **)&R->ctl, sizeof(ServeCtl), fl) != hipSuccess ||
        hipHostMalloc((void **)&R->pin, R->pin_size, fl) != hipSuccess)
        return TAL_ENOMEM;
    std::memset(R->done, 0, sizeof(int32_t) * (size_t)R->cap);
    std::memset(R->reqs, 0, sizeof(ServeReq) * RES_SLOTS);
    std::memset(R->ctl, 0, sizeof(ServeCtl));
    R->pin_alloc.reset(R->pin_size);
    R->copies.start(3);
    for (int k = RES_SLOTS - 1; k >= 0; k--) R->free_slots.push_back(k);
    const size_t slab_bytes = (size_t)z.slab_cols * 64 * sizeof(uint32_t);
    int64_t words[SERVE_DEV_WORDS];
    serve_dev_init(words);
    if (hipMalloc((void **)&R->dev, SERVE_DEV_BYTES) != hipSuccess) return TAL_ENOMEM;
    if (hipMemcpy(R->dev, words, sizeof(words), hipMemcpyHostToDevice) != hipSuccess)
        return TAL_EDEVICE;
    if (hipMalloc((void **)&R->slabs, slab_bytes * (size_t)R->grid) != hipSuccess) return TAL_ENOMEM;
    if (set_serve_lds_limit(prop.sharedMemPerBlock) != hipSuccess ||
        hipStreamCreateWithFlags(&R->kstream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&R->kdone, hipEventDisableTiming) != hipSuccess)
        return TAL_EDEVICE;
    return TAL_OK;
}

} // namespace

extern "C" {

int tal_bsw_server_create(const tal_bsw_params *params, int device,
                          const tal_bsw_server_config *config, tal_bsw_server **out)
{
    if (!params || !out) return TAL_EINVAL;
    *out = nullptr;
    tal_bsw_server_config cfg{};
    if (config) cfg = *config;
    if (cfg.hbm_budget == 0) cfg.hbm_budget = (size_t)1 << 30;
    if (cfg.workers <= 0) cfg.workers = 2;
    if (cfg.max_batch_pairs <= 0) cfg.max_batch_pairs = 1 << 22;
    if (cfg.max_query_len <= 0) cfg.max_query_len = 1024;
    if (cfg.hbm_budget > MAX_BUDGET || cfg.workers > 8 || cfg.max_query_len > MAX_LEN16 ||
        (cfg.resident != 0 && cfg.resident != 1))
        return TAL_EINVAL;
    tal_bsw_server *s = new (std::nothrow) tal_bsw_server();
    if (!s) return TAL_ENOMEM;
    s->prm = *params;
    s->device = device;
    s->smax = score_bound(*params);
    s->budget = cfg.hbm_budget;
    s->max_batch = cfg.max_batch_pairs;
    s->resident = cfg.resident == 1;
    int rc = TAL_OK;
    for (int k = 0; k < (s->resident ? 0 : cfg.workers) && rc == TAL_OK; k++) {
        tal_bsw *e = nullptr;
        rc = tal_bsw_create(params, device, &e);
        if (rc == TAL_OK) s->engines.push_back(e);
    }
    if (rc == TAL_OK && hipSetDevice(device) != hipSuccess) rc = TAL_EDEVICE;
    if (rc == TAL_OK && s->resident) rc = res_create(s, cfg.max_query_len);
    if (rc == TAL_OK && hipMalloc((void **)&s->arena, s->budget) != hipSuccess) {
        s->arena = nullptr;
        rc = TAL_ENOMEM;
    }
    if (rc != TAL_OK) {
        for (tal_bsw *e : s->engines) tal_bsw_destroy(e);
        res_release(s);
        if (s->arena) (void)hipFree(s->arena);
        delete s;
        return rc;
    }
    s->alloc.reset(s->budget);
    if (s->resident) {
        s->admitter = std::thread(res_admit_loop, s);
        s->res->monitor = std::thread(res_monitor_loop, s);
    } else {
        s->admitter = std::thread(admit_loop, s);
        for (int k = 0; k < cfg.workers; k++) s->workers.emplace_back(work_loop, s, k);
    }
    *out = s;
    return TAL_OK;
}

int tal_bsw_server_submit(tal_bsw_server *s, tal_seqpair *pairs, const uint8_t *ref,
                          size_t ref_len, const uint8_t *qer, size_t qer_len, int32_t n, int32_t w,
                          uint64_t *ticket)
{
    if (!s || !ticket || n <= 0 || w < 0 || !pairs) return TAL_EINVAL;
    if ((!ref && ref_len) || (!qer && qer_len)) return TAL_EINVAL;
    Request *r = new (std::nothrow) Request();
    if (!r) return TAL_ENOMEM;
    r->pairs = pairs;
    r->ref = ref;
    r->ref_len = ref_len;
    r->qer = qer;
    r->qer_len = qer_len;
    r->n = n;
    r->w = w;
    r->t_submit = now_ms();
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->stop) {
        delete r;
        return TAL_EINVAL;
    }
    r->ticket = s->next_ticket++;
    s->all[r->ticket] = r;
    s->pending.push_back(r);
    *ticket = r->ticket;
    s->cv.notify_all();
    return TAL_OK;
}

int tal_bsw_server_wait(tal_bsw_server *s, uint64_t ticket, int32_t timeout_ms,
                        tal_bsw_request_info *info)
{
    if (!s) return TAL_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return TAL_EINVAL;
    Request *r = it->second;
    auto done = [&] { return r->done; };
    if (timeout_ms < 0)
        s->cv.wait(lk, done);
    else if (!s->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
        return TAL_EAGAIN;
    if (info) {
        info->status = r->status;
        info->batch = r->batch;
        info->batch_requests = r->batch_requests;
        info->pad = 0;
        info->queue_ms = r->admitted ? r->t_admit - r->t_submit : 0.0;
        info->latency_ms = r->t_done - r->t_submit;
    }
    const int status = r->status;
    s->all.erase(it);
    delete r;
    return status;
}

int tal_bsw_server_get_stats(tal_bsw_server *s, tal_bsw_server_stats *out)
{
    if (!s || !out) return TAL_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    s->st.peak_bytes = (int64_t)s->alloc.peak;
    *out = s->st;
    return TAL_OK;
}

void tal_bsw_server_destroy(tal_bsw_server *s)
{
    if (!s) return;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->admitter.joinable()) s->admitter.join();
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv.notify_all();
    }
    for (std::thread &t : s->workers)
        if (t.joinable()) t.join();
    if (s->res && s->res->monitor.joinable()) s->res->monitor.join();
    res_release(s);
    for (auto &kv : s->all) delete kv.second;
    for (tal_bsw *e : s->engines) tal_bsw_destroy(e);
    (void)hipSetDevice(s->device);
    if (s->arena) (void)hipFree(s->arena);
    delete s;
}

} /* extern "C" */
