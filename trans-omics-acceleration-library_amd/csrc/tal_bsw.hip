/*
 * tal_bsw.hip -- host runtime behind include/tal_bsw.h.
 *
 * A handle owns one HIP stream, the device buffers a call needs (grown on
 * demand, never shrunk), and the device image of the reference object's
 * persistent 16-bit F scratch (one zero-mask per column).  Every entry point
 * returns a status code; errors are recorded in the handle and never abort.
 */
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "bsw_internal.h"

/* Scratch of tal_bsw_get_scores_auto_device_noalloc.  Every device the
 * library runs on holds its own copy, at its own address, until the process
 * ends.  In host code `scratch` names a host-side stand-in, never the device
 * memory: host code uses the per-device address that hipGetSymbolAddress
 * gives (static_scratch below). */
static __device__ __align__(256) uint8_t scratch[TAL_BSW_STATIC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::vector<uint8_t *> g_scratch_addr; /* per device; null = not looked up yet */

/* This device's address of `scratch`, looked up once per device with that
 * device current. */
hipError_t static_scratch(int device, uint8_t **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_addr.size() <= device) g_scratch_addr.resize(device + 1, nullptr);
    if (!g_scratch_addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[device] = (uint8_t *)p;
    }
    *out = g_scratch_addr[device];
    return hipSuccess;
}

} // namespace

namespace talbsw {
hipError_t launch_main(int bits, bool lds, int grid, size_t lds_bytes, const KParams &K,
                       const KBuffers &B, hipStream_t s);
hipError_t set_main_lds_limit(size_t bytes);
hipError_t launch_shapes(int bits, const tal_seqpair *pairs, int32_t n, int32_t ngroups,
                         int32_t pad_len2, GroupShape *out, int64_t *ncol_plus1, hipStream_t s,
                         const int32_t *pad_len2_dev = nullptr);
hipError_t launch_resolve(int32_t ngroups, const GroupShape *shape, const uint32_t *dep,
                          const uint8_t *wrote, const uint32_t *fmask, const int64_t *fmask_off,
                          const uint32_t *image, const int32_t *seg_first, uint32_t *stale,
                          int32_t *rerun_list, int32_t *rerun_count, hipStream_t s,
                          const int32_t *abort_flag = nullptr);
hipError_t launch_image_update(int32_t ngroups, int32_t ncols, const GroupShape *shape,
                               const uint8_t *wrote, int32_t *last, const uint32_t *fmask,
                               const int64_t *fmask_off, uint32_t *image, hipStream_t s,
                               const int32_t *skip = nullptr);
hipError_t launch_partition(const tal_seqpair *pairs, int32_t n, int32_t match_pos,
                            int32_t planned8, int32_t *blk, int32_t *sel8, int32_t *sel16,
                            int32_t *err, hipStream_t s);
hipError_t launch_plan_check16(const GroupShape *shape, int32_t ngroups, int32_t max_ncol,
                               int32_t *err, hipStream_t s);
hipError_t launch_plan_pad8(const tal_seqpair *pairs8, int32_t n8, int32_t *out, int32_t *err,
                            hipStream_t s);
hipError_t launch_copy_i32(const int32_t *src, int32_t *dst, hipStream_t s);
hipError_t launch_flag8(const tal_seqpair *pairs, int32_t n, int32_t match_pos, uint8_t *flag,
                        hipStream_t s);
hipError_t launch_gather(const tal_seqpair *src, const int32_t *idx, int32_t m,
                         tal_seqpair *dst, hipStream_t s);
hipError_t launch_scatter_out(const tal_seqpair *src, const int32_t *idx, int32_t m,
                              tal_seqpair *dst, hipStream_t s);
} // namespace talbsw

using namespace talbsw;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            cap = 0;
            if (e != hipSuccess) return e;
        }
        size_t want = std::max(bytes, cap * 3 / 2);
        want = (want + 4095) & ~(size_t)4095;
        hipError_t e = hipMalloc(&p, want);
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

/* The scratch of a no-allocation call: consecutive 256-byte-aligned slices
 * of the static device array, handed out in call order. */
struct Arena {
    uint8_t *base = nullptr;
    size_t cap = 0, used = 0;
};

/* Per-width scratch of one call. */
struct PathBufs {
    DevBuf shapes, ncp1, foff, dep, wrote, stale, fmask, rerun, tasks, counter, slabs, last,
        cubtmp, summary, ckey, ckey2, gids, order, gstate, pending;
};

} // namespace

struct tal_bsw {
    tal_bsw_params prm{};
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    size_t lds_per_block = 0;
    size_t lds_per_cu = 0;
    int waves_per_cu = 8;  /* target waves per CU for the LDS/global split */
    int order = 1;         /* task order: 0 = input order, 1 = longest groups first */
    int plan_rounds = 2;   /* stale-column rerun rounds captured in a plan's graph */
    DevBuf image; /* IMAGE_COLS u32: F[c] nonzero lanes of the 16-bit scratch */
    DevBuf zero_image; /* the image of a freshly constructed object */
    PathBufs b16, b8;
    /* dispatcher scratch */
    DevBuf flag, idx_all, idx_sel, nsel, pairs8, cubtmp;
    /* scalar path scratch */
    DevBuf sc_key, sc_key2, sc_ids, sc_order, sc_ctl, sc_slabs;
    /* host-API staging */
    DevBuf d_pairs, d_ref, d_qer;
    /* this device's copy of the static scratch array, and the arena of the
     * no-allocation call in progress (null in every other call) */
    uint8_t *scratch = nullptr;
    Arena *arena = nullptr;
    /* fault injection (TAL_BSW_FAIL_ALLOC=k at creation): the k-th scratch
     * request of the handle fails as if hipMalloc had run out of memory */
    long fail_alloc = 0, allocs = 0;
    /* per-call timing */
    /* 16-bit start/end, 8-bit start/end, rerun start/end */
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    bool ev16 = false, ev8 = false;
    bool in_call = false; /* inside a dispatcher/host call */
    tal_bsw_stats st{};
    std::string err;
};

namespace {

int fail(tal_bsw *h, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

int hip_fail(tal_bsw *h, hipError_t e, const char *what)
{
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? TAL_ENOMEM
                                                                                 : TAL_EDEVICE;
    return fail(h, code, "%s: %s", what, hipGetErrorString(e));
}

#define HIPCHK(h, expr, what)                                                                      \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return hip_fail((h), e_, (what));                                    \
    } while (0)

/* `bytes` of call scratch: the handle's buffer b, grown with hipMalloc, or,
 * in a no-allocation call, the next slice of the static array (TAL_ERANGE
 * when the array is exhausted; b is left alone). */
template <class T> int take(tal_bsw *h, DevBuf &b, size_t bytes, const char *what, T **out)
{
    if (!h->arena) {
        if (h->fail_alloc > 0 && ++h->allocs == h->fail_alloc)
            HIPCHK(h, hipErrorOutOfMemory, what);
        HIPCHK(h, b.ensure(bytes), what);
        *out = (T *)b.p;
        return TAL_OK;
    }
    Arena &a = *h->arena;
    const size_t off = (a.used + 255) & ~(size_t)255;
    if (off > a.cap || bytes > a.cap - off)
        return fail(h, TAL_ERANGE, "%s: %zu bytes do not fit the %zu-byte static scratch (%zu in use)",
                    what, bytes, a.cap, off);
    a.used = off + bytes;
    *out = (T *)(a.base + off);
    return TAL_OK;
}

#define TAKE(h, buf, bytes, what, out)                                                             \
    do {                                                                                           \
        const int rc_ = take((h), (buf), (bytes), (what), (out));                                  \
        if (rc_) return rc_;                                                                       \
    } while (0)

int smax_of(const tal_bsw_params &p)
{
    int s = 0;
    const int mm = -(int)(int8_t)p.w_mismatch; /* stored negated, :65 */
    s = std::max(s, (int)(int8_t)p.w_match);
    s = std::max(s, (int)(int8_t)mm);
    s = std::max(s, -1);
    return s;
}

KParams base_params(const tal_bsw *h, int bits, int32_t n, int32_t w)
{
    KParams K = params_for(h->prm, bits);
    K.n = n;
    K.w = w;
    return K;
}

/* Summary of the group shapes, computed on the device and read back. */
struct ShapeSummary {
    int32_t max_ncol;
    int32_t max_nrow;
    int64_t total_cols; /* sum over groups of ncol + 1 (16-bit only) */
};

__global__ void summarize(const GroupShape *shapes, const int64_t *foff_incl, int32_t ngroups,
                          ShapeSummary *out)
{
    __shared__ int32_t mc[256], mr[256];
    int32_t c = 0, r = 0;
    for (int g = threadIdx.x; g < ngroups; g += blockDim.x) {
        c = max(c, shapes[g].ncol);
        r = max(r, shapes[g].nrow);
    }
    mc[threadIdx.x] = c;
    mr[threadIdx.x] = r;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            mc[threadIdx.x] = max(mc[threadIdx.x], mc[threadIdx.x + o]);
            mr[threadIdx.x] = max(mr[threadIdx.x], mr[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out->max_ncol = mc[0];
        out->max_nrow = mr[0];
        out->total_cols = (foff_incl && ngroups > 0) ? foff_incl[ngroups - 1] : 0;
    }
}

/* Inclusive prefix sums -> exclusive offsets in place (shift by one). */
__global__ void to_exclusive(const int64_t *incl, int64_t *excl, int32_t n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) excl[k] = k == 0 ? 0 : incl[k - 1];
}

__global__ void iota32(int32_t *x, int32_t n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = k;
}

/* Cost key of each group for the task order: rows times the widest column
 * range a row can sweep (the group's columns, at most 2w+2). */
__global__ void group_cost(const GroupShape *shapes, int32_t ngroups, int32_t w, uint32_t *key,
                           int32_t *ids)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    const int64_t width = min((int64_t)shapes[g].ncol, 2 * (int64_t)w + 2) + 1;
    const int64_t c = (int64_t)shapes[g].nrow * width;
    key[g] = (uint32_t)min(c, (int64_t)0xFFFFFFFF);
    ids[g] = g;
}

/* Marks the unused task slots past the last group. */
__global__ void fill_i32(int32_t *x, int32_t n, int32_t v)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = v;
}

/* Fixed-stride F-mask offsets of a planned batch (stride = planned
 * columns + 1), so that they stay valid for any batch within the plan. */
__global__ void stride_offsets(int64_t *off, int32_t n, int64_t stride)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) off[k] = (int64_t)k * stride;
}

/* Tasks of a rerun launch from the device rerun count. */
__global__ void tasks_from_count(const int32_t *count, int32_t gpw, int32_t *ntasks)
{
    *ntasks = (*count + gpw - 1) / gpw;
}

/* Adds a rerun round's group count to the replay's total. */
__global__ void add_count(const int32_t *count, int32_t *total) { *total += *count; }

/* Control word of a replayed batch: skip the in-graph image update when
 * reruns are still pending or the batch left the plan's bounds. */
__global__ void make_skip(const int32_t *count, const int32_t *err, int32_t *skip,
                          int32_t *count_out)
{
    *skip = (*count != 0 || *err != 0) ? 1 : 0;
    *count_out = *count;
}

/* Pairs up a rerun list into 16-bit wave tasks (two groups per task). */
__global__ void make_tasks(const int32_t *list, const int32_t *count, int32_t *tasks,
                           int32_t cap_tasks, int gpw)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cap_tasks) return;
    const int c = *count;
    for (int k = 0; k < gpw; k++) {
        const int i = t * gpw + k;
        tasks[i] = i < c ? list[i] : -1;
    }
}

int grid_for(const tal_bsw *h, bool lds, size_t lds_bytes, int ntasks)
{
    int per_cu = 8; /* waves per CU with state in global memory */
    if (lds) {
        per_cu = (int)(h->lds_per_cu / std::max<size_t>(lds_bytes, 1));
        per_cu = std::max(1, std::min(per_cu, 16));
    }
    long long g = (long long)h->cu_count * per_cu;
    return (int)std::max<long long>(1, std::min<long long>(g, ntasks));
}

/* The state placement for this handle's waves per CU (state_sizing). */
Sizing sizing(const tal_bsw *h, int max_ncol)
{
    return state_sizing(h->lds_per_cu, h->lds_per_block, h->waves_per_cu, max_ncol);
}

/* Task order: longest groups first, and the two groups of a 16-bit wave of
 * similar size, so that the persistent grid ends evenly and the two
 * half-waves of a wave run loops of similar length.  Any order gives the
 * same results; this one only changes the schedule.  Leaves the task list
 * (ntasks * GPW ids, -1 padded) in *order_out. */
int make_order(tal_bsw *h, PathBufs &pb, const GroupShape *shapes, int32_t G, int32_t w,
               int ntasks, int GPW, hipStream_t s, int32_t **order_out)
{
    uint32_t *ckey, *ckey2;
    int32_t *gids, *order;
    void *cubtmp;
    TAKE(h, pb.ckey, sizeof(uint32_t) * G, "alloc keys", &ckey);
    TAKE(h, pb.ckey2, sizeof(uint32_t) * G, "alloc keys", &ckey2);
    TAKE(h, pb.gids, sizeof(int32_t) * G, "alloc ids", &gids);
    TAKE(h, pb.order, sizeof(int32_t) * ((size_t)ntasks * GPW), "alloc order", &order);
    hipLaunchKernelGGL(group_cost, dim3((G + 255) / 256), dim3(256), 0, s, shapes, G, w, ckey, gids);
    HIPCHK(h, hipGetLastError(), "group cost");
    size_t tmp = 0;
    HIPCHK(h, hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp, ckey, ckey2, gids, order,
                                                           G, 0, 32, s),
           "sort size");
    TAKE(h, pb.cubtmp, tmp + 16, "alloc sort tmp", &cubtmp);
    HIPCHK(h, hipcub::DeviceRadixSort::SortPairsDescending(cubtmp, tmp, ckey, ckey2, gids, order,
                                                           G, 0, 32, s),
           "sort groups");
    const int pad = ntasks * GPW - G;
    if (pad > 0) {
        hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, s, order + G, pad, -1);
        HIPCHK(h, hipGetLastError(), "pad tasks");
    }
    *order_out = order;
    return TAL_OK;
}

/*
 * One width-specific batch on device buffers.  16-bit runs the stale-column
 * resolution loop and updates the handle's image afterwards.
 */
int run_path(tal_bsw *h, int bits, tal_seqpair *pairs, const uint8_t *ref, const uint8_t *qer,
             int32_t n, int32_t w, int32_t last_len2, hipStream_t s,
             const int32_t *seg_first = nullptr)
{
    /* seg_first (device, per 16-bit group: the first group of its segment)
     * marks a batch of independent calls, each scored as on a freshly
     * constructed object: the stale-column search stays inside a segment,
     * starts from the zero image, and the handle's image is left alone. */
    const bool fresh = seg_first != nullptr;
    if (n == 0) return TAL_OK;
    PathBufs &pb = bits == 16 ? h->b16 : h->b8;
    const int GL = bits == 16 ? 32 : 64, GPW = 64 / GL;
    const int32_t G = (n + GL - 1) / GL;
    KParams K = base_params(h, bits, n, w);
    K.ngroups = G;
    K.last_len2 = last_len2;

    GroupShape *shapes;
    ShapeSummary *summary;
    int64_t *ncp1 = nullptr, *foff = nullptr;
    TAKE(h, pb.shapes, sizeof(GroupShape) * G, "alloc shapes", &shapes);
    TAKE(h, pb.summary, sizeof(ShapeSummary), "alloc summary", &summary);
    if (bits == 16) {
        TAKE(h, pb.ncp1, sizeof(int64_t) * G, "alloc ncol", &ncp1);
        TAKE(h, pb.foff, sizeof(int64_t) * G, "alloc offsets", &foff);
        HIPCHK(h, launch_shapes(16, pairs, n, G, 0, shapes, ncp1, s), "shapes");
        size_t tmp = 0;
        void *cubtmp;
        HIPCHK(h, hipcub::DeviceScan::InclusiveSum(nullptr, tmp, ncp1, ncp1, G, s), "scan size");
        TAKE(h, pb.cubtmp, tmp + 16, "alloc scan tmp", &cubtmp);
        HIPCHK(h, hipcub::DeviceScan::InclusiveSum(cubtmp, tmp, ncp1, ncp1, G, s), "scan");
        hipLaunchKernelGGL(summarize, dim3(1), dim3(256), 0, s, shapes, ncp1, G, summary);
        hipLaunchKernelGGL(to_exclusive, dim3((G + 255) / 256), dim3(256), 0, s, ncp1, foff, G);
    } else {
        HIPCHK(h, launch_shapes(8, pairs, n, G, last_len2, shapes, nullptr, s), "shapes");
        hipLaunchKernelGGL(summarize, dim3(1), dim3(256), 0, s, shapes, (const int64_t *)nullptr,
                           G, summary);
    }
    HIPCHK(h, hipGetLastError(), "summary");
    ShapeSummary sum{};
    HIPCHK(h, hipMemcpyAsync(&sum, summary, sizeof sum, hipMemcpyDeviceToHost, s),
           "read summary");
    HIPCHK(h, hipStreamSynchronize(s), "sync summary");
    const int limit = bits == 16 ? MAX_LEN16 : MAX_LEN8;
    if (sum.max_ncol > limit || sum.max_nrow > limit || last_len2 < 0 || last_len2 > limit)
        return fail(h, TAL_ERANGE, "a pair length is negative or exceeds %d", limit);
    const int32_t max_ncol = bits == 16 ? sum.max_ncol : MAX_LEN8;
    const int64_t total_cols = sum.total_cols;

    const Sizing z = sizing(h, max_ncol);
    const int state_cols = z.state_cols, qwords = z.qwords, lds_cols = z.lds_cols;
    const int slab_cols = z.slab_cols;
    const size_t lds_bytes = z.lds_bytes;
    const bool lds = true;
    const int ntasks = (G + GPW - 1) / GPW;
    int grid = grid_for(h, lds, lds_bytes, ntasks);
    K.ntasks = ntasks;
    K.state_cols = state_cols;
    K.lds_cols = lds_cols;
    K.slab_cols = slab_cols;
    K.query_words = qwords;

    KBuffers B{};
    B.pairs = pairs;
    B.ref = ref;
    B.qer = qer;
    if (h->order == 1 && G > 1) {
        int32_t *order;
        const int rc = make_order(h, pb, shapes, G, w, ntasks, GPW, s, &order);
        if (rc) return rc;
        B.task_groups = order;
    }
    TAKE(h, pb.counter, sizeof(int32_t) * 4, "alloc counter", &B.task_counter);
    int32_t *list = nullptr, *tasks = nullptr, *last = nullptr;
    uint32_t *stale = nullptr;
    if (bits == 16) {
        TAKE(h, pb.dep, sizeof(uint32_t) * G, "alloc dep", &B.dep_out);
        TAKE(h, pb.wrote, G, "alloc wrote", &B.wrote_out);
        TAKE(h, pb.stale, sizeof(uint32_t) * G, "alloc stale", &stale);
        TAKE(h, pb.fmask, sizeof(uint32_t) * (size_t)std::max<int64_t>(total_cols, 1),
             "alloc fmask", &B.fmask);
        TAKE(h, pb.rerun, sizeof(int32_t) * (G + 1), "alloc rerun", &list);
        TAKE(h, pb.tasks, sizeof(int32_t) * (size_t)(G + 2) * 2, "alloc tasks", &tasks);
        if (!fresh) TAKE(h, pb.last, sizeof(int32_t) * IMAGE_COLS, "alloc last", &last);
        HIPCHK(h, hipMemsetAsync(stale, 0, sizeof(uint32_t) * G, s), "clear stale");
        B.stale_in = stale;
        B.fmask_off = foff;
        if (!seg_first) { /* one call: groups are settled inside the launch too */
            TAKE(h, pb.gstate, sizeof(int32_t) * G, "alloc group states", &B.gstate);
            HIPCHK(h, hipMemsetAsync(B.gstate, 0, sizeof(int32_t) * G, s), "clear group states");
            TAKE(h, pb.pending, sizeof(int32_t) * (G + 2), "alloc pending list", &B.pending);
            HIPCHK(h, hipMemsetAsync(B.pending, 0, sizeof(int32_t) * 2, s), "clear pending");
            HIPCHK(h, hipMemsetAsync(B.pending + 2, 0xFF, sizeof(int32_t) * G, s),
                   "clear pending slots");
            B.shapes = shapes;
            B.image = fresh ? h->zero_image.as<uint32_t>() : h->image.as<uint32_t>();
        }
    }
    /* The state slabs come last: a no-allocation call runs as many blocks as
     * the rest of the static scratch holds slabs for (the grid is persistent,
     * so any grid of at least one block scores every task). */
    const size_t slab_bytes = (size_t)slab_cols * 64 * sizeof(uint32_t);
    if (h->arena) {
        const Arena &a = *h->arena;
        const size_t off = (a.used + 255) & ~(size_t)255;
        const size_t fit = off < a.cap ? (a.cap - off) / slab_bytes : 0;
        grid = (int)std::min<size_t>((size_t)grid, std::max<size_t>(fit, 1));
    }
    TAKE(h, pb.slabs, (size_t)grid * slab_bytes, "alloc state slabs", &B.slabs);
    if (B.gstate) B.early = B.task_counter + 2;
    HIPCHK(h, hipMemsetAsync(B.task_counter, 0, sizeof(int32_t) * 3, s), "clear counters");
    hipEvent_t ev_beg = h->ev[bits == 16 ? 0 : 2], ev_end = h->ev[bits == 16 ? 1 : 3];
    HIPCHK(h, hipEventRecord(ev_beg, s), "event");
    HIPCHK(h, launch_main(bits, lds, grid, lds ? lds_bytes : 0, K, B, s), "main kernel");
    HIPCHK(h, hipEventRecord(ev_end, s), "event");
    int32_t reruns = 0, early = 0;
    float ms_rerun = 0.f;

    if (bits == 16) {
        int32_t *cnt = B.task_counter + 1;
        for (int iter = 0;; iter++) {
            HIPCHK(h, hipMemsetAsync(cnt, 0, sizeof(int32_t), s), "clear rerun count");
            HIPCHK(h, launch_resolve(G, shapes, B.dep_out, B.wrote_out, B.fmask, B.fmask_off,
                                     fresh ? h->zero_image.as<uint32_t>() : h->image.as<uint32_t>(),
                                     seg_first, stale, list, cnt, s),
                   "resolve");
            int32_t cc[2] = {0, 0}; /* reruns listed; groups run again inside launches */
            HIPCHK(h, hipMemcpyAsync(cc, cnt, sizeof cc, hipMemcpyDeviceToHost, s), "read reruns");
            HIPCHK(h, hipStreamSynchronize(s), "sync reruns");
            const int32_t c = cc[0];
            early = cc[1];
            if (c == 0) break;
            reruns += c;
            if (iter > G + 1) return fail(h, TAL_EDEVICE, "stale-column resolution did not converge");
            /* rerun the listed groups with their corrected F[ncol] masks */
            const int rt = (c + GPW - 1) / GPW;
            hipLaunchKernelGGL(make_tasks, dim3((rt + 255) / 256), dim3(256), 0, s, list, cnt,
                               tasks, rt, GPW);
            HIPCHK(h, hipGetLastError(), "make tasks");
            KParams K2 = K;
            K2.ntasks = rt;
            KBuffers B2 = B;
            B2.task_groups = tasks;
            B2.pending = nullptr;
            HIPCHK(h, hipMemsetAsync(B.task_counter, 0, sizeof(int32_t), s), "clear counter");
            const int g2 = grid_for(h, lds, lds_bytes, rt);
            HIPCHK(h, hipEventRecord(h->ev[4], s), "event");
            HIPCHK(h, launch_main(16, lds, std::min(g2, grid), lds ? lds_bytes : 0, K2, B2, s),
                   "rerun kernel");
            HIPCHK(h, hipEventRecord(h->ev[5], s), "event");
            HIPCHK(h, hipEventSynchronize(h->ev[5]), "event sync");
            float ms = 0.f;
            HIPCHK(h, hipEventElapsedTime(&ms, h->ev[4], h->ev[5]), "event time");
            ms_rerun += ms;
        }
        h->ev16 = true;
        h->st.ms_rerun16 = ms_rerun;
        h->st.groups16 = G;
        h->st.reruns16 = reruns + (B.gstate ? early : 0);
        h->st.settled16 = B.gstate ? early : 0;
        h->st.lds_state16 = lds_cols >= ((state_cols + 7) & ~7) ? 1 : 0;
        if (!fresh) {
            HIPCHK(h, hipMemsetAsync(last, 0xFF, sizeof(int32_t) * IMAGE_COLS, s), "clear last");
            HIPCHK(h, launch_image_update(G, max_ncol + 1, shapes, B.wrote_out, last, B.fmask,
                                          B.fmask_off, h->image.as<uint32_t>(), s),
                   "image update");
        }
    } else {
        h->ev8 = true;
        h->st.groups8 = G;
    }
    return TAL_OK;
}


/* ------------------------------------------------------------ batch plans */

/* One width of a plan: its sizing, launch shapes and buffers. */
struct WidthPlan {
    int32_t n = 0, G = 0, ntasks = 0, grid = 0, rgrid = 0, max_ncol = 0;
    Sizing z{};
    KParams K{};
    KBuffers B{};
    PathBufs pb;
};

} // namespace

struct tal_bsw_plan {
    tal_bsw *h = nullptr;
    tal_seqpair *pairs = nullptr;
    const uint8_t *ref = nullptr, *qer = nullptr;
    int32_t n = 0, w = 0, n8 = 0, n16 = 0;
    int rounds = 2; /* rerun rounds inside the graph */
    WidthPlan p8, p16;
    /* ctl: [0] error, [1] 8-bit padding len2, [2] rerun tasks, [3] skip,
     * [4] pending reruns after the graph, [5] groups rerun inside the graph */
    DevBuf blk, sel8, sel16, pairs8, pairs16, ctl;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    /* the 8-bit branch of the graph: captured on its own stream, forked
     * after the 16-bit main launch and joined at the end, so that it runs
     * beside the 16-bit rerun rounds */
    hipStream_t s8 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

namespace {

void plan_release(tal_bsw_plan *P)
{
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->fork) (void)hipEventDestroy(P->fork);
    if (P->join) (void)hipEventDestroy(P->join);
    if (P->s8) (void)hipStreamDestroy(P->s8);
    DevBuf *all[] = {&P->blk, &P->sel8, &P->sel16, &P->pairs8, &P->pairs16, &P->ctl};
    for (DevBuf *b : all) b->release();
    for (PathBufs *pb : {&P->p8.pb, &P->p16.pb}) {
        DevBuf *bs[] = {&pb->shapes, &pb->ncp1, &pb->foff,    &pb->dep,   &pb->wrote,
                        &pb->stale,  &pb->fmask, &pb->rerun,  &pb->tasks, &pb->counter,
                        &pb->slabs,  &pb->last,  &pb->cubtmp, &pb->summary, &pb->ckey,
                        &pb->ckey2,  &pb->gids,  &pb->order};
        for (DevBuf *b : bs) b->release();
    }
}

/* Sizes one width of a plan from the batch as it is now (eager, with host
 * reads) and allocates everything its replay needs. */
int plan_width(tal_bsw *h, tal_bsw_plan *P, int bits, hipStream_t s)
{
    WidthPlan &W = bits == 16 ? P->p16 : P->p8;
    W.n = bits == 16 ? P->n16 : P->n8;
    if (W.n == 0) return TAL_OK;
    PathBufs &pb = W.pb;
    const int GL = bits == 16 ? 32 : 64, GPW = 64 / GL;
    tal_seqpair *pairs = bits == 16 ? P->pairs16.as<tal_seqpair>() : P->pairs8.as<tal_seqpair>();
    W.G = (W.n + GL - 1) / GL;
    const int32_t G = W.G;
    int32_t pad_len2 = 0;
    if (bits == 8) {
        HIPCHK(h, hipMemcpyAsync(&pad_len2, &pairs[W.n - 1].len2, sizeof pad_len2,
                                 hipMemcpyDeviceToHost, s),
               "read last len2");
        HIPCHK(h, hipStreamSynchronize(s), "sync last len2");
    }
    HIPCHK(h, pb.shapes.ensure(sizeof(GroupShape) * G), "alloc shapes");
    HIPCHK(h, pb.summary.ensure(sizeof(ShapeSummary)), "alloc summary");
    GroupShape *shapes = pb.shapes.as<GroupShape>();
    HIPCHK(h, launch_shapes(bits, pairs, W.n, G, pad_len2, shapes, nullptr, s), "shapes");
    hipLaunchKernelGGL(summarize, dim3(1), dim3(256), 0, s, shapes, (const int64_t *)nullptr, G,
                       pb.summary.as<ShapeSummary>());
    HIPCHK(h, hipGetLastError(), "summary");
    ShapeSummary sum{};
    HIPCHK(h, hipMemcpyAsync(&sum, pb.summary.p, sizeof sum, hipMemcpyDeviceToHost, s),
           "read summary");
    HIPCHK(h, hipStreamSynchronize(s), "sync summary");
    const int limit = bits == 16 ? MAX_LEN16 : MAX_LEN8;
    if (sum.max_ncol > limit || sum.max_nrow > limit)
        return fail(h, TAL_ERANGE, "a pair length is negative or exceeds %d", limit);
    W.max_ncol = bits == 16 ? sum.max_ncol : MAX_LEN8;
    W.z = sizing(h, W.max_ncol);
    W.ntasks = (G + GPW - 1) / GPW;
    W.grid = grid_for(h, true, W.z.lds_bytes, W.ntasks);
    W.rgrid = std::min(W.grid, h->cu_count * 2);
    KParams &K = W.K;
    K = base_params(h, bits, W.n, P->w);
    K.ngroups = G;
    K.ntasks = W.ntasks;
    K.state_cols = W.z.state_cols;
    K.lds_cols = W.z.lds_cols;
    K.slab_cols = W.z.slab_cols;
    K.query_words = W.z.qwords;
    KBuffers &B = W.B;
    B = KBuffers{};
    B.pairs = pairs;
    B.ref = P->ref;
    B.qer = P->qer;
    if (h->order == 1 && G > 1) {
        int32_t *order;
        const int rc = make_order(h, pb, shapes, G, P->w, W.ntasks, GPW, s, &order);
        if (rc) return rc;
        B.task_groups = order;
    }
    HIPCHK(h, pb.counter.ensure(sizeof(int32_t) * 4), "alloc counter");
    B.task_counter = pb.counter.as<int32_t>();
    B.abort_flag = P->ctl.as<int32_t>();
    HIPCHK(h, pb.slabs.ensure((size_t)W.grid * W.z.slab_cols * 64 * sizeof(uint32_t)),
           "alloc state slabs");
    B.slabs = pb.slabs.as<uint32_t>();
    if (bits == 8) {
        B.last_len2_dev = P->ctl.as<int32_t>() + 1;
    } else {
        const int64_t stride = (int64_t)W.max_ncol + 1;
        HIPCHK(h, pb.foff.ensure(sizeof(int64_t) * G), "alloc offsets");
        HIPCHK(h, pb.dep.ensure(sizeof(uint32_t) * G), "alloc dep");
        HIPCHK(h, pb.wrote.ensure(G), "alloc wrote");
        HIPCHK(h, pb.stale.ensure(sizeof(uint32_t) * G), "alloc stale");
        HIPCHK(h, pb.fmask.ensure(sizeof(uint32_t) * (size_t)G * stride), "alloc fmask");
        HIPCHK(h, pb.rerun.ensure(sizeof(int32_t) * (G + 1)), "alloc rerun");
        HIPCHK(h, pb.tasks.ensure(sizeof(int32_t) * (size_t)(G + 2) * 2), "alloc tasks");
        HIPCHK(h, pb.last.ensure(sizeof(int32_t) * IMAGE_COLS), "alloc last");
        HIPCHK(h, pb.gstate.ensure(sizeof(int32_t) * G), "alloc group states");
        HIPCHK(h, pb.pending.ensure(sizeof(int32_t) * (G + 2)), "alloc pending list");
        hipLaunchKernelGGL(stride_offsets, dim3((G + 255) / 256), dim3(256), 0, s,
                           pb.foff.as<int64_t>(), G, stride);
        HIPCHK(h, hipGetLastError(), "offsets");
        B.stale_in = pb.stale.as<uint32_t>();
        B.dep_out = pb.dep.as<uint32_t>();
        B.wrote_out = pb.wrote.as<uint8_t>();
        B.fmask = pb.fmask.as<uint32_t>();
        B.fmask_off = pb.foff.as<int64_t>();
        B.shapes = pb.shapes.as<GroupShape>();
        B.image = h->image.as<uint32_t>();
        B.gstate = pb.gstate.as<int32_t>();
        B.early = P->ctl.as<int32_t>() + 6;
        B.pending = pb.pending.as<int32_t>();
    }
    return TAL_OK;
}

/* The replayed pipeline: dispatcher split, gathers, both DP launches, the
 * stale-column resolution with `rounds` rerun rounds, the image update when
 * nothing is pending, and the result scatters.  No host reads, no memsets. */
int plan_enqueue(tal_bsw *h, tal_bsw_plan *P, hipStream_t s)
{
    int32_t *ctl = P->ctl.as<int32_t>();
    hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, s, ctl, 8, 0);
    HIPCHK(h, launch_partition(P->pairs, P->n, smax_of(h->prm), P->n8, P->blk.as<int32_t>(),
                               P->sel8.as<int32_t>(), P->sel16.as<int32_t>(), ctl, s),
           "partition");
    if (P->n8)
        HIPCHK(h, launch_gather(P->pairs, P->sel8.as<int32_t>(), P->n8,
                                P->pairs8.as<tal_seqpair>(), s),
               "gather8");
    if (P->n16)
        HIPCHK(h, launch_gather(P->pairs, P->sel16.as<int32_t>(), P->n16,
                                P->pairs16.as<tal_seqpair>(), s),
               "gather16");
    /* the 16-bit main launch, with the 8-bit branch beside it and the
     * 16-bit resolution and rerun rounds (they share no data) */
    WidthPlan &W = P->p16;
    const int32_t G = W.G;
    GroupShape *shapes = W.pb.shapes.as<GroupShape>();
    int32_t *cnt = W.B.task_counter + 1;
    int32_t *list = W.pb.rerun.as<int32_t>();
    if (P->n16) {
        HIPCHK(h, launch_shapes(16, P->pairs16.as<tal_seqpair>(), W.n, G, 0, shapes, nullptr, s),
               "shapes16");
        HIPCHK(h, launch_plan_check16(shapes, G, W.max_ncol, ctl, s), "check16");
        hipLaunchKernelGGL(fill_i32, dim3((G + 255) / 256), dim3(256), 0, s,
                           (int32_t *)W.pb.stale.p, G, 0);
        hipLaunchKernelGGL(fill_i32, dim3((G + 255) / 256), dim3(256), 0, s, W.B.gstate, G, 0);
        hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, s, W.B.pending, 2, 0);
        hipLaunchKernelGGL(fill_i32, dim3((G + 255) / 256), dim3(256), 0, s, W.B.pending + 2, G,
                           -1);
        hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, s, W.B.task_counter, 2, 0);
    }
    /* The 8-bit branch forks just before the 16-bit launch: its blocks take
     * the slots the 16-bit launch's waves free as its last tasks end (most
     * stale-column groups are settled inside that launch, so the rounds
     * after it are short and would leave the branch exposed). */
    hipStream_t s8 = s;
    if (P->n8 && P->n16 && P->s8) {
        HIPCHK(h, hipEventRecord(P->fork, s), "fork");
        HIPCHK(h, hipStreamWaitEvent(P->s8, P->fork, 0), "fork wait");
        s8 = P->s8;
    }
    if (P->n16) HIPCHK(h, launch_main(16, true, W.grid, W.z.lds_bytes, W.K, W.B, s), "main16");
    if (P->n8) {
        WidthPlan &W8 = P->p8;
        HIPCHK(h, launch_plan_pad8(P->pairs8.as<tal_seqpair>(), P->n8, ctl + 1, ctl, s8), "pad8");
        hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, s8, W8.B.task_counter, 1, 0);
        HIPCHK(h, launch_main(8, true, W8.grid, W8.z.lds_bytes, W8.K, W8.B, s8), "main8");
        HIPCHK(h, launch_scatter_out(P->pairs8.as<tal_seqpair>(), P->sel8.as<int32_t>(), P->n8,
                                     P->pairs, s8),
               "scatter8");
    }
    if (s8 != s) HIPCHK(h, hipEventRecord(P->join, s8), "join");
    if (P->n16) {
        HIPCHK(h, launch_resolve(G, shapes, W.B.dep_out, W.B.wrote_out, W.B.fmask,
                                 W.B.fmask_off, h->image.as<uint32_t>(), nullptr,
                                 W.pb.stale.as<uint32_t>(), list, cnt, s, ctl),
               "resolve");
        const int cap = (G + 1) / 2;
        for (int r = 0; r < P->rounds; r++) {
            hipLaunchKernelGGL(add_count, dim3(1), dim3(1), 0, s, cnt, ctl + 5);
            hipLaunchKernelGGL(make_tasks, dim3((cap + 255) / 256), dim3(256), 0, s, list, cnt,
                               W.pb.tasks.as<int32_t>(), cap, 2);
            hipLaunchKernelGGL(tasks_from_count, dim3(1), dim3(1), 0, s, cnt, 2, ctl + 2);
            hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, s, W.B.task_counter, 2, 0);
            KBuffers B2 = W.B;
            B2.task_groups = W.pb.tasks.as<int32_t>();
            B2.ntasks_dev = ctl + 2;
            B2.pending = nullptr;
            HIPCHK(h, launch_main(16, true, W.rgrid, W.z.lds_bytes, W.K, B2, s), "rerun16");
            HIPCHK(h, launch_resolve(G, shapes, W.B.dep_out, W.B.wrote_out, W.B.fmask,
                                     W.B.fmask_off, h->image.as<uint32_t>(), nullptr,
                                     W.pb.stale.as<uint32_t>(), list, cnt, s, ctl),
                   "resolve");
        }
        hipLaunchKernelGGL(make_skip, dim3(1), dim3(1), 0, s, cnt, ctl, ctl + 3, ctl + 4);
        hipLaunchKernelGGL(fill_i32, dim3(IMAGE_COLS / 256), dim3(256), 0, s, W.pb.last.as<int32_t>(),
                           IMAGE_COLS, -1);
        HIPCHK(h, launch_image_update(G, W.max_ncol + 1, shapes, W.B.wrote_out,
                                      W.pb.last.as<int32_t>(), W.B.fmask, W.B.fmask_off,
                                      h->image.as<uint32_t>(), s, ctl + 3),
               "image update");
        HIPCHK(h, launch_scatter_out(P->pairs16.as<tal_seqpair>(), P->sel16.as<int32_t>(),
                                     P->n16, P->pairs, s),
               "scatter16");
    }
    if (s8 != s) HIPCHK(h, hipStreamWaitEvent(s, P->join, 0), "join wait");
    return TAL_OK;
}

/* After a replay: the reruns the graph's rounds left, eagerly, then the
 * image update. */
int plan_finish(tal_bsw *h, tal_bsw_plan *P, hipStream_t s, int32_t pending)
{
    WidthPlan &W = P->p16;
    const int32_t G = W.G;
    GroupShape *shapes = W.pb.shapes.as<GroupShape>();
    int32_t *cnt = W.B.task_counter + 1;
    int32_t *list = W.pb.rerun.as<int32_t>();
    int32_t c = pending;
    for (int iter = 0; c > 0; iter++) {
        if (iter > G + 1) return fail(h, TAL_EDEVICE, "stale-column resolution did not converge");
        h->st.reruns16 += c;
        const int rt = (c + 1) / 2;
        hipLaunchKernelGGL(make_tasks, dim3((rt + 255) / 256), dim3(256), 0, s, list, cnt,
                           W.pb.tasks.as<int32_t>(), rt, 2);
        HIPCHK(h, hipGetLastError(), "make tasks");
        KParams K2 = W.K;
        K2.ntasks = rt;
        KBuffers B2 = W.B;
        B2.task_groups = W.pb.tasks.as<int32_t>();
        B2.pending = nullptr;
        HIPCHK(h, hipMemsetAsync(W.B.task_counter, 0, sizeof(int32_t), s), "clear counter");
        HIPCHK(h, launch_main(16, true, std::min(W.grid, grid_for(h, true, W.z.lds_bytes, rt)),
                              W.z.lds_bytes, K2, B2, s),
               "rerun kernel");
        HIPCHK(h, hipMemsetAsync(cnt, 0, sizeof(int32_t), s), "clear rerun count");
        HIPCHK(h, launch_resolve(G, shapes, W.B.dep_out, W.B.wrote_out, W.B.fmask,
                                 W.B.fmask_off, h->image.as<uint32_t>(), nullptr,
                                 W.pb.stale.as<uint32_t>(), list, cnt, s),
               "resolve");
        HIPCHK(h, hipMemcpyAsync(&c, cnt, sizeof c, hipMemcpyDeviceToHost, s), "read reruns");
        HIPCHK(h, hipStreamSynchronize(s), "sync reruns");
    }
    HIPCHK(h, hipMemsetAsync(W.pb.last.p, 0xFF, sizeof(int32_t) * IMAGE_COLS, s), "clear last");
    HIPCHK(h, launch_image_update(G, W.max_ncol + 1, shapes, W.B.wrote_out,
                                  W.pb.last.as<int32_t>(), W.B.fmask, W.B.fmask_off,
                                  h->image.as<uint32_t>(), s),
           "image update");
    /* the graph scattered the 16-bit results before these reruns */
    HIPCHK(h, launch_scatter_out(P->pairs16.as<tal_seqpair>(), P->sel16.as<int32_t>(), P->n16,
                                 P->pairs, s),
           "scatter16");
    return TAL_OK;
}

int check_common(tal_bsw *h, tal_seqpair *pairs, int32_t n, int32_t w)
{
    if (!h) return TAL_EINVAL;
    if (!h->in_call) { /* a new top-level call: forget the previous timing */
        h->st = tal_bsw_stats{};
        h->ev16 = h->ev8 = false;
    }
    if (n < 0) return fail(h, TAL_EINVAL, "n < 0");
    if (w < 0) return fail(h, TAL_EINVAL, "w < 0");
    if (n > 0 && !pairs) return fail(h, TAL_EINVAL, "null pairs");
    return TAL_OK;
}

/* Host-side validation of the pairs of a host-buffer call. */
int check_pairs(tal_bsw *h, const tal_seqpair *pairs, int32_t n, size_t ref_len, size_t qer_len,
                int maxlen)
{
    for (int32_t k = 0; k < n; k++) {
        const tal_seqpair &p = pairs[k];
        if (p.len1 < 0 || p.len2 < 0 || p.idr < 0 || p.idq < 0)
            return fail(h, TAL_EINVAL, "pair %d: negative length or offset", k);
        if (p.len1 > maxlen || p.len2 > maxlen)
            return fail(h, TAL_ERANGE, "pair %d: length %d/%d exceeds %d", k, p.len1, p.len2,
                        maxlen);
        if ((size_t)p.idr + (size_t)p.len1 > ref_len || (size_t)p.idq + (size_t)p.len2 > qer_len)
            return fail(h, TAL_EINVAL, "pair %d: sequence outside its buffer", k);
    }
    return TAL_OK;
}

int host_call(tal_bsw *h, int mode, tal_seqpair *pairs, const uint8_t *ref, size_t ref_len,
              const uint8_t *qer, size_t qer_len, int32_t n, int32_t w,
              const int8_t *mat = nullptr)
{
    int rc = check_common(h, pairs, n, w);
    if (rc) return rc;
    if (n > 0 && ((!ref && ref_len) || (!qer && qer_len)))
        return fail(h, TAL_EINVAL, "null sequence buffer");
    rc = check_pairs(h, pairs, n, ref_len, qer_len, mode == 8 ? MAX_LEN8 : MAX_LEN16);
    if (rc) return rc;
    if (n == 0) return TAL_OK;
    HIPCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = h->stream;
    void *unused;
    TAKE(h, h->d_pairs, sizeof(tal_seqpair) * n, "alloc pairs", &unused);
    TAKE(h, h->d_ref, std::max<size_t>(ref_len, 1), "alloc ref", &unused);
    TAKE(h, h->d_qer, std::max<size_t>(qer_len, 1), "alloc qer", &unused);
    HIPCHK(h, hipMemcpyAsync(h->d_pairs.p, pairs, sizeof(tal_seqpair) * n, hipMemcpyHostToDevice, s),
           "copy pairs");
    if (ref_len)
        HIPCHK(h, hipMemcpyAsync(h->d_ref.p, ref, ref_len, hipMemcpyHostToDevice, s), "copy ref");
    if (qer_len)
        HIPCHK(h, hipMemcpyAsync(h->d_qer.p, qer, qer_len, hipMemcpyHostToDevice, s), "copy qer");
    tal_seqpair *dp = h->d_pairs.as<tal_seqpair>();
    struct InCall {
        tal_bsw *h;
        ~InCall() { h->in_call = false; }
    } guard{h};
    h->in_call = true;
    if (mode == 1)
        rc = tal_bsw_scalar_scores_device(h, mat, dp, h->d_ref.as<uint8_t>(),
                                          h->d_qer.as<uint8_t>(), n, w, s);
    else if (mode == 16)
        rc = tal_bsw_get_scores16_device(h, dp, h->d_ref.as<uint8_t>(), h->d_qer.as<uint8_t>(), n,
                                         w, s);
    else if (mode == 8)
        rc = tal_bsw_get_scores8_device(h, dp, h->d_ref.as<uint8_t>(), h->d_qer.as<uint8_t>(), n,
                                        w, s);
    else
        rc = tal_bsw_get_scores_auto_device(h, dp, h->d_ref.as<uint8_t>(),
                                            h->d_qer.as<uint8_t>(), n, w, s);
    if (rc) return rc;
    HIPCHK(h, hipMemcpyAsync(pairs, dp, sizeof(tal_seqpair) * n, hipMemcpyDeviceToHost, s),
           "copy results");
    HIPCHK(h, hipStreamSynchronize(s), "sync results");
    return TAL_OK;
}

} // namespace

namespace talbsw {

int run_batch(tal_bsw *h, int bits, tal_seqpair *pairs, const uint8_t *ref, const uint8_t *qer,
              int32_t n, int32_t w, hipStream_t s, const int32_t *seg_first)
{
    if (!h) return TAL_EINVAL;
    if (n < 0 || w < 0 || (n > 0 && !pairs)) return fail(h, TAL_EINVAL, "bad batch arguments");
    h->st = tal_bsw_stats{};
    h->ev16 = h->ev8 = false;
    HIPCHK(h, hipSetDevice(h->device), "set device");
    return run_path(h, bits, pairs, ref, qer, n, w, 0, s, seg_first);
}

hipStream_t handle_stream(const tal_bsw *h) { return h->stream; }

int score_bound(const tal_bsw_params &p) { return smax_of(p); }

KParams params_for(const tal_bsw_params &p, int bits)
{
    KParams K{};
    const int sh = bits == 16 ? 0 : 8;
    auto sc = [&](int v) -> int16_t {
        return (int16_t)(uint16_t)((bits == 16 ? (uint32_t)(uint16_t)v : (uint32_t)(uint8_t)v) << sh);
    };
    K.o_del = p.o_del;
    K.e_del = p.e_del;
    K.o_ins = p.o_ins;
    K.e_ins = p.e_ins;
    K.end_bonus = p.end_bonus;
    K.smax = smax_of(p);
    K.match = sc((int8_t)p.w_match);
    K.mism = sc((int8_t)(-(int)(int8_t)p.w_mismatch));
    K.ambig = sc(-1);
    K.e_del_s = sc(p.e_del);
    K.oe_del_s = sc(p.o_del + p.e_del);
    K.e_ins_s = sc(p.e_ins);
    K.oe_ins_s = sc(p.o_ins + p.e_ins);
    K.zdrop_s = sc(p.zdrop);
    return K;
}

/* Placement of a wave's state: columns 0..max_ncol plus 17 columns of
 * slack for the 8-column block loads that run past a group's last column,
 * and the fast path's query-code words (8 columns per word, one word of
 * prefetch slack).  A window of lds_cols state columns lives in LDS, sized so
 * that `waves_per_cu` waves fit on a CU; the query words and the home of
 * every state column go to a per-block slab in global memory. */
Sizing state_sizing(size_t lds_per_cu, size_t lds_per_block, int waves_per_cu, int max_ncol)
{
    Sizing z{};
    z.state_cols = max_ncol + 18;
    z.qwords = (max_ncol >> 3) + 2;
    const int wpc = std::max(1, waves_per_cu);
    const long budget_cols = (long)(lds_per_cu / wpc) / 256;
    const int fit_all = (z.state_cols + 7) & ~7;
    int lds_cols = (int)std::max<long>(8, budget_cols) & ~7;
    if (fit_all <= lds_cols) lds_cols = fit_all;
    const size_t max_cols = lds_per_block / 256;
    if ((size_t)lds_cols > max_cols) lds_cols = (int)(max_cols & ~(size_t)7);
    z.lds_cols = lds_cols;
    /* every state column in whole 8-column blocks */
    z.slab_cols = z.qwords + ((z.state_cols + 7) & ~7);
    z.lds_bytes = (size_t)lds_cols * 64 * sizeof(uint32_t);
    return z;
}

int goes_8bit(const tal_seqpair &p, int smax)
{
    const int mn = std::min(p.len1, p.len2);
    return p.len1 >= 0 && p.len2 >= 0 && p.len1 <= MAX_LEN8 && p.len2 <= MAX_LEN8 && p.h0 >= 0 &&
           (long long)p.h0 + (long long)smax * mn <= 127;
}

} // namespace talbsw

extern "C" {

int tal_bsw_create(const tal_bsw_params *params, int device, tal_bsw **out)
{
    if (!params || !out) return TAL_EINVAL;
    *out = nullptr;
    if (params->e_del == 0 || params->e_ins == 0) return TAL_EINVAL;
    if (params->w_match < -128 || params->w_match > 127 || params->w_mismatch < -128 ||
        params->w_mismatch > 127)
        return TAL_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TAL_EDEVICE;
    if (device < 0 || device >= ndev) return TAL_EINVAL;
    tal_bsw *h = new (std::nothrow) tal_bsw();
    if (!h) return TAL_ENOMEM;
    h->prm = *params;
    h->device = device;
    int rc = TAL_OK;
    do {
        if (hipSetDevice(device) != hipSuccess) { rc = TAL_EDEVICE; break; }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess) { rc = TAL_EDEVICE; break; }
        h->cu_count = prop.multiProcessorCount;
        h->lds_per_block = prop.sharedMemPerBlock;
        h->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                              : prop.sharedMemPerBlock;
        if (set_main_lds_limit(h->lds_per_block) != hipSuccess) { rc = TAL_EDEVICE; break; }
        if (const char *e = getenv("TAL_BSW_WAVES_PER_CU")) h->waves_per_cu = std::max(1, atoi(e));
        if (const char *e = getenv("TAL_BSW_ORDER")) h->order = atoi(e) ? 1 : 0;
        if (const char *e = getenv("TAL_BSW_FAIL_ALLOC")) h->fail_alloc = atol(e);
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) { rc = TAL_EDEVICE; break; }
        bool evok = true;
        for (hipEvent_t &e : h->ev) evok = evok && hipEventCreate(&e) == hipSuccess;
        if (!evok) { rc = TAL_EDEVICE; break; }
        if (static_scratch(device, &h->scratch) != hipSuccess) { rc = TAL_EDEVICE; break; }
        hipError_t e = h->image.ensure(sizeof(uint32_t) * IMAGE_COLS);
        if (e != hipSuccess) { rc = TAL_ENOMEM; break; }
        e = h->zero_image.ensure(sizeof(uint32_t) * IMAGE_COLS);
        if (e != hipSuccess) { rc = TAL_ENOMEM; break; }
        if (hipMemset(h->image.p, 0, sizeof(uint32_t) * IMAGE_COLS) != hipSuccess ||
            hipMemset(h->zero_image.p, 0, sizeof(uint32_t) * IMAGE_COLS) != hipSuccess) {
            rc = TAL_EDEVICE;
            break;
        }
    } while (0);
    if (rc) {
        tal_bsw_destroy(h);
        return rc;
    }
    *out = h;
    return TAL_OK;
}

void tal_bsw_destroy(tal_bsw *h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    DevBuf *all[] = {&h->image,  &h->zero_image, &h->flag,    &h->idx_all, &h->idx_sel,
                     &h->nsel,   &h->pairs8,     &h->cubtmp,  &h->d_pairs, &h->d_ref,
                     &h->d_qer};
    for (DevBuf *b : all) b->release();
    for (PathBufs *pb : {&h->b16, &h->b8}) {
        DevBuf *bs[] = {&pb->shapes, &pb->ncp1, &pb->foff,    &pb->dep,   &pb->wrote,
                        &pb->stale,  &pb->fmask, &pb->rerun,  &pb->tasks, &pb->counter,
                        &pb->slabs,  &pb->last,  &pb->cubtmp, &pb->summary, &pb->ckey,
                        &pb->ckey2,  &pb->gids,  &pb->order};
        for (DevBuf *b : bs) b->release();
    }
    for (hipEvent_t e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char *tal_bsw_last_error(const tal_bsw *h) { return h ? h->err.c_str() : "null handle"; }

int tal_bsw_get_scores16(tal_bsw *h, tal_seqpair *pairs, const uint8_t *ref, size_t ref_len,
                         const uint8_t *qer, size_t qer_len, int32_t n, int32_t w)
{
    return host_call(h, 16, pairs, ref, ref_len, qer, qer_len, n, w);
}

int tal_bsw_get_scores8(tal_bsw *h, tal_seqpair *pairs, const uint8_t *ref, size_t ref_len,
                        const uint8_t *qer, size_t qer_len, int32_t n, int32_t w)
{
    return host_call(h, 8, pairs, ref, ref_len, qer, qer_len, n, w);
}

int tal_bsw_get_scores_auto(tal_bsw *h, tal_seqpair *pairs, const uint8_t *ref, size_t ref_len,
                            const uint8_t *qer, size_t qer_len, int32_t n, int32_t w)
{
    return host_call(h, 0, pairs, ref, ref_len, qer, qer_len, n, w);
}

int tal_bsw_get_scores16_device(tal_bsw *h, tal_seqpair *pairs, const uint8_t *ref,
                                const uint8_t *qer, int32_t n, int32_t w, void *stream)
{
    int rc = check_common(h, pairs, n, w);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    return run_path(h, 16, pairs, ref, qer, n, w, 0, s);
}

int tal_bsw_get_scores8_device(tal_bsw *h, tal_seqpair *pairs, const uint8_t *ref,
                               const uint8_t *qer, int32_t n, int32_t w, void *stream)
{
    int rc = check_common(h, pairs, n, w);
    if (rc) return rc;
    if (n == 0) return TAL_OK;
    HIPCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    int32_t last_len2 = 0;
    HIPCHK(h, hipMemcpyAsync(&last_len2, &pairs[n - 1].len2, sizeof last_len2,
                             hipMemcpyDeviceToHost, s),
           "read last len2");
    HIPCHK(h, hipStreamSynchronize(s), "sync last len2");
    return run_path(h, 8, pairs, ref, qer, n, w, last_len2, s);
}

int tal_bsw_get_scores_auto_device(tal_bsw *h, tal_seqpair *pairs, const uint8_t *ref,
                                   const uint8_t *qer, int32_t n, int32_t w, void *stream)
{
    int rc = check_common(h, pairs, n, w);
    if (rc) return rc;
    if (n == 0) return TAL_OK;
    HIPCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    uint8_t *flag;
    int32_t *all, *sel, *cnt;
    void *cubtmp;
    tal_seqpair *pairs8, *pairs16;
    TAKE(h, h->flag, 2 * (size_t)n, "alloc flags", &flag);
    TAKE(h, h->idx_all, sizeof(int32_t) * n, "alloc index", &all);
    TAKE(h, h->idx_sel, sizeof(int32_t) * n, "alloc index", &sel);
    TAKE(h, h->nsel, sizeof(int32_t) * 2, "alloc count", &cnt);
    HIPCHK(h, launch_flag8(pairs, n, smax_of(h->prm), flag, s), "flag8");
    hipLaunchKernelGGL(iota32, dim3((n + 255) / 256), dim3(256), 0, s, all, n);
    HIPCHK(h, hipGetLastError(), "iota");
    /* two stable selections: 8-bit pair indices, then 16-bit ones, each in
     * input order (one temporary, sized for the larger of the two) */
    size_t tmp = 0, tmp2 = 0;
    HIPCHK(h, hipcub::DeviceSelect::Flagged(nullptr, tmp, all, flag, sel, cnt, n, s),
           "select size");
    HIPCHK(h, hipcub::DeviceSelect::Flagged(nullptr, tmp2, all, flag + n, sel, cnt + 1, n, s),
           "select size");
    tmp = std::max(tmp, tmp2);
    TAKE(h, h->cubtmp, tmp + 16, "alloc select tmp", &cubtmp);
    TAKE(h, h->pairs8, sizeof(tal_seqpair) * n, "alloc pairs8", &pairs8);
    HIPCHK(h, hipcub::DeviceSelect::Flagged(cubtmp, tmp, all, flag, sel, cnt, n, s),
           "select 8-bit");
    int32_t n8 = 0;
    HIPCHK(h, hipMemcpyAsync(&n8, cnt, sizeof n8, hipMemcpyDeviceToHost, s), "read n8");
    HIPCHK(h, hipStreamSynchronize(s), "sync n8");
    const int32_t n16 = n - n8;
    if (n16)
        HIPCHK(h, hipcub::DeviceSelect::Flagged(cubtmp, tmp, all, flag + n, sel + n8, cnt + 1, n,
                                                s),
               "select 16-bit");
    /* one buffer holds both halves: the 8-bit pairs, then the 16-bit ones */
    pairs16 = pairs8 + n8;
    h->st.pairs8 = n8;
    h->st.pairs16 = n16;
    const bool outer = !h->in_call;
    h->in_call = true;
    struct Restore {
        tal_bsw *h;
        bool outer;
        ~Restore()
        {
            if (outer) h->in_call = false;
        }
    } restore{h, outer};
    /* In a no-allocation call the 8-bit path's scratch is free again once its
     * work is queued: the 16-bit path's kernels run after it on the stream. */
    const size_t mark = h->arena ? h->arena->used : 0;
    if (n8) {
        HIPCHK(h, launch_gather(pairs, sel, n8, pairs8, s), "gather8");
        rc = tal_bsw_get_scores8_device(h, pairs8, ref, qer, n8, w, s);
        if (rc) return rc;
        HIPCHK(h, launch_scatter_out(pairs8, sel, n8, pairs, s), "scatter8");
    }
    if (h->arena) h->arena->used = mark;
    if (n16) {
        HIPCHK(h, launch_gather(pairs, sel + n8, n16, pairs16, s), "gather16");
        rc = tal_bsw_get_scores16_device(h, pairs16, ref, qer, n16, w, s);
        if (rc) return rc;
        HIPCHK(h, launch_scatter_out(pairs16, sel + n8, n16, pairs, s), "scatter16");
    }
    return TAL_OK;
}

int tal_bsw_get_scores_auto_device_noalloc(tal_bsw *h, tal_seqpair *pairs, const uint8_t *ref,
                                           const uint8_t *qer, int32_t n, int32_t w, void *stream)
{
    if (!h) return TAL_EINVAL;
    if (n < 0) return fail(h, TAL_EINVAL, "n < 0");
    /* the dispatcher's own part (flags, indices, one copy of the pairs) must
     * fit before anything is queued */
    const size_t own = (size_t)n * (2 + 4 + 4 + sizeof(tal_seqpair)) + 4 * 256;
    if (own > TAL_BSW_STATIC_SCRATCH_BYTES)
        return fail(h, TAL_ERANGE, "%d pairs need more than the %zu-byte static scratch", n,
                    (size_t)TAL_BSW_STATIC_SCRATCH_BYTES);
    Arena a;
    a.base = h->scratch;
    a.cap = TAL_BSW_STATIC_SCRATCH_BYTES;
    struct Bind {
        tal_bsw *h;
        ~Bind() { h->arena = nullptr; }
    } bind{h};
    h->arena = &a;
    return tal_bsw_get_scores_auto_device(h, pairs, ref, qer, n, w, stream);
}

int tal_bsw_plan_create(tal_bsw *h, tal_seqpair *pairs, const uint8_t *ref, const uint8_t *qer,
                        int32_t n, int32_t w, tal_bsw_plan **out)
{
    if (!out) return TAL_EINVAL;
    *out = nullptr;
    int rc = check_common(h, pairs, n, w);
    if (rc) return rc;
    if (n <= 0) return fail(h, TAL_EINVAL, "a plan needs n > 0");
    HIPCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = h->stream;
    tal_bsw_plan *P = new (std::nothrow) tal_bsw_plan();
    if (!P) return TAL_ENOMEM;
    P->h = h;
    P->pairs = pairs;
    P->ref = ref;
    P->qer = qer;
    P->n = n;
    P->w = w;
    P->rounds = h->plan_rounds;
    struct Guard {
        tal_bsw_plan *P;
        bool keep = false;
        ~Guard()
        {
            if (!keep) {
                plan_release(P);
                delete P;
            }
        }
    } guard{P};
    const int nblk = (n + 255) / 256;
    HIPCHK(h, P->blk.ensure(sizeof(int32_t) * nblk), "alloc partition");
    HIPCHK(h, P->sel8.ensure(sizeof(int32_t) * n), "alloc index");
    HIPCHK(h, P->sel16.ensure(sizeof(int32_t) * n), "alloc index");
    HIPCHK(h, P->ctl.ensure(sizeof(int32_t) * 8), "alloc control");
    /* the planned split, from the batch as it is now */
    HIPCHK(h, hipMemsetAsync(P->ctl.p, 0, sizeof(int32_t) * 8, s), "clear control");
    std::vector<int32_t> hb(nblk);
    HIPCHK(h, launch_partition(pairs, n, smax_of(h->prm), -1, P->blk.as<int32_t>(),
                               P->sel8.as<int32_t>(), P->sel16.as<int32_t>(),
                               P->ctl.as<int32_t>(), s),
           "partition");
    /* part_offsets left exclusive offsets; the 8-bit total is the last
     * offset plus the last block's count, recounted on the host */
    std::vector<tal_seqpair> tail(std::min(n, 256));
    HIPCHK(h, hipMemcpyAsync(hb.data(), P->blk.p, sizeof(int32_t) * nblk, hipMemcpyDeviceToHost, s),
           "read partition");
    HIPCHK(h, hipMemcpyAsync(tail.data(), pairs + (size_t)(nblk - 1) * 256,
                             sizeof(tal_seqpair) * (size_t)(n - (nblk - 1) * 256),
                             hipMemcpyDeviceToHost, s),
           "read last block");
    HIPCHK(h, hipStreamSynchronize(s), "sync partition");
    int32_t last8 = 0;
    for (int32_t k = 0; k < n - (nblk - 1) * 256; k++) last8 += goes_8bit(tail[k], smax_of(h->prm));
    P->n8 = hb[nblk - 1] + last8;
    P->n16 = n - P->n8;
    HIPCHK(h, launch_partition(pairs, n, smax_of(h->prm), P->n8, P->blk.as<int32_t>(),
                               P->sel8.as<int32_t>(), P->sel16.as<int32_t>(),
                               P->ctl.as<int32_t>(), s),
           "partition");
    HIPCHK(h, P->pairs8.ensure(sizeof(tal_seqpair) * std::max(P->n8, 1)), "alloc pairs8");
    HIPCHK(h, P->pairs16.ensure(sizeof(tal_seqpair) * std::max(P->n16, 1)), "alloc pairs16");
    if (P->n8)
        HIPCHK(h, launch_gather(pairs, P->sel8.as<int32_t>(), P->n8, P->pairs8.as<tal_seqpair>(), s),
               "gather8");
    if (P->n16)
        HIPCHK(h, launch_gather(pairs, P->sel16.as<int32_t>(), P->n16,
                                P->pairs16.as<tal_seqpair>(), s),
               "gather16");
    rc = plan_width(h, P, 8, s);
    if (rc) return rc;
    rc = plan_width(h, P, 16, s);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(s), "sync plan");
    if (P->n8 && P->n16) {
        HIPCHK(h, hipStreamCreateWithFlags(&P->s8, hipStreamNonBlocking), "branch stream");
        HIPCHK(h, hipEventCreateWithFlags(&P->fork, hipEventDisableTiming), "fork event");
        HIPCHK(h, hipEventCreateWithFlags(&P->join, hipEventDisableTiming), "join event");
    }
    /* capture (kernel events are not recorded inside the graph: the replay
     * is timed as a whole by tal_bsw_plan_launch) */
    HIPCHK(h, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    rc = plan_enqueue(h, P, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != TAL_OK || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return rc ? rc : fail(h, TAL_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
    }
    P->graph = g;
    HIPCHK(h, hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0), "graph instantiate");
    guard.keep = true;
    *out = P;
    return TAL_OK;
}

int tal_bsw_plan_launch(tal_bsw_plan *P, void *stream)
{
    if (!P) return TAL_EINVAL;
    tal_bsw *h = P->h;
    h->st = tal_bsw_stats{};
    h->ev16 = h->ev8 = false;
    HIPCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    HIPCHK(h, hipEventRecord(h->ev[4], s), "event");
    HIPCHK(h, hipGraphLaunch(P->exec, s), "graph launch");
    HIPCHK(h, hipEventRecord(h->ev[5], s), "event");
    int32_t ctl[8];
    HIPCHK(h, hipMemcpyAsync(ctl, P->ctl.p, sizeof ctl, hipMemcpyDeviceToHost, s), "read control");
    HIPCHK(h, hipStreamSynchronize(s), "sync graph");
    HIPCHK(h, hipEventElapsedTime(&h->st.ms_replay, h->ev[4], h->ev[5]), "event time");
    h->st.pairs8 = P->n8;
    h->st.pairs16 = P->n16;
    h->st.groups8 = P->p8.G;
    h->st.groups16 = P->p16.G;
    h->st.reruns16 = ctl[5] + ctl[6];
    h->st.settled16 = ctl[6];
    if (ctl[0] != 0)
        return fail(h, TAL_ERANGE, "the batch no longer fits its plan (dispatcher split or "
                                   "group columns changed); build a new plan");
    if (P->n16 && ctl[4] != 0) return plan_finish(h, P, s, ctl[4]);
    return TAL_OK;
}

void tal_bsw_plan_destroy(tal_bsw_plan *P)
{
    if (!P) return;
    (void)hipSetDevice(P->h->device);
    (void)hipStreamSynchronize(P->h->stream);
    plan_release(P);
    delete P;
}

int tal_bsw_set_plan_rounds(tal_bsw *h, int32_t rounds)
{
    if (!h) return TAL_EINVAL;
    if (rounds < 0 || rounds > 8) return fail(h, TAL_EINVAL, "plan rounds must be 0..8");
    h->plan_rounds = rounds;
    return TAL_OK;
}

int tal_bsw_set_tuning(tal_bsw *h, int32_t waves_per_cu, int32_t order)
{
    if (!h) return TAL_EINVAL;
    if (waves_per_cu < 1 || waves_per_cu > 32 || (order != 0 && order != 1))
        return fail(h, TAL_EINVAL, "bad tuning");
    h->waves_per_cu = waves_per_cu;
    h->order = order;
    return TAL_OK;
}

int tal_bsw_autotune(tal_bsw *h, const tal_seqpair *pairs, const uint8_t *ref,
                     const uint8_t *qer, int32_t n, int32_t w, int32_t sample,
                     tal_bsw_tuning *out)
{
    if (!h || !out) return TAL_EINVAL;
    if (n <= 0 || !pairs || w < 0 || sample < 0) return fail(h, TAL_EINVAL, "bad autotune arguments");
    const int32_t m = sample > 0 ? std::min(sample, n) : n;
    HIPCHK(h, hipSetDevice(h->device), "set device");
    tal_bsw *t = nullptr;
    int rc = tal_bsw_create(&h->prm, h->device, &t);
    if (rc) return fail(h, rc, "autotune: scratch handle");
    DevBuf copy;
    struct Cleanup {
        tal_bsw *t;
        DevBuf *copy;
        ~Cleanup()
        {
            copy->release();
            tal_bsw_destroy(t);
        }
    } cleanup{t, &copy};
    HIPCHK(h, copy.ensure(sizeof(tal_seqpair) * (size_t)m), "alloc autotune copy");
    HIPCHK(h, hipMemcpyAsync(copy.p, pairs, sizeof(tal_seqpair) * (size_t)m,
                             hipMemcpyDeviceToDevice, t->stream),
           "copy batch");
    HIPCHK(h, hipStreamSynchronize(t->stream), "sync copy");
    /* every distinct LDS window from 160 to 40 columns but 128 (5 waves per
     * CU, never the fastest): 9 and 10 waves give the same window, as do 12
     * and 13, 14 and 16 */
    static const int waves[] = {4, 6, 7, 8, 10, 11, 12, 16};
    tal_bsw_tuning res{};
    res.ms = -1.f;
    for (int order = 0; order < 2; order++)
        for (int wv : waves) {
            t->waves_per_cu = wv;
            t->order = order;
            tal_bsw_plan *P = nullptr;
            rc = tal_bsw_plan_create(t, copy.as<tal_seqpair>(), ref, qer, m, w, &P);
            if (rc) return fail(h, rc, "autotune: %s", t->err.c_str());
            float best = -1.f;
            for (int rep = 0; rep < 3 && rc == TAL_OK; rep++) {
                const auto t0 = std::chrono::steady_clock::now();
                rc = tal_bsw_plan_launch(P, nullptr);
                const auto t1 = std::chrono::steady_clock::now();
                const float ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
                if (rep > 0 && (best < 0 || ms < best)) best = ms; /* first replay warms */
            }
            tal_bsw_plan_destroy(P);
            if (rc) return fail(h, rc, "autotune: %s", t->err.c_str());
            if (res.trials < TAL_TUNING_MAX_TRIALS) {
                res.trial_waves[res.trials] = wv;
                res.trial_order[res.trials] = order;
                res.trial_ms[res.trials] = best;
                res.trials++;
            }
            if (res.ms < 0 || best < res.ms) {
                res.ms = best;
                res.waves_per_cu = wv;
                res.order = order;
            }
        }
    h->waves_per_cu = res.waves_per_cu;
    h->order = res.order;
    *out = res;
    return TAL_OK;
}

int tal_bsw_scalar_scores_device(tal_bsw *h, const int8_t *mat, tal_seqpair *pairs,
                                 const uint8_t *ref, const uint8_t *qer, int32_t n, int32_t w,
                                 void *stream)
{
    int rc = check_common(h, pairs, n, w);
    if (rc) return rc;
    if (!mat) return fail(h, TAL_EINVAL, "null scoring matrix");
    if (n == 0) return TAL_OK;
    HIPCHK(h, hipSetDevice(h->device), "set device");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    ScalarParams P{};
    P.n = n;
    P.w = w;
    P.o_del = h->prm.o_del;
    P.e_del = h->prm.e_del;
    P.o_ins = h->prm.o_ins;
    P.e_ins = h->prm.e_ins;
    P.zdrop = h->prm.zdrop;
    P.end_bonus = h->prm.end_bonus;
    P.mat_max = 0; /* :148-150 */
    for (int k = 0; k < 25; k++) {
        P.mat[k] = mat[k];
        P.mat_max = std::max(P.mat_max, (int)mat[k]);
    }
    const int ntasks = (n + 63) / 64;
    HIPCHK(h, h->sc_key.ensure(sizeof(uint32_t) * n), "alloc keys");
    HIPCHK(h, h->sc_key2.ensure(sizeof(uint32_t) * n), "alloc keys");
    HIPCHK(h, h->sc_ids.ensure(sizeof(int32_t) * n), "alloc ids");
    HIPCHK(h, h->sc_order.ensure(sizeof(int32_t) * (size_t)ntasks * 64), "alloc order");
    HIPCHK(h, h->sc_ctl.ensure(sizeof(int32_t) * 4), "alloc control");
    int32_t *ctl = h->sc_ctl.as<int32_t>(); /* [0] task counter, [1] max length, [2] error */
    HIPCHK(h, hipMemsetAsync(ctl, 0, sizeof(int32_t) * 4, s), "clear control");
    HIPCHK(h, launch_scalar_cost(pairs, n, w, h->sc_key.as<uint32_t>(), h->sc_ids.as<int32_t>(),
                                 ctl + 1, s),
           "scalar cost");
    size_t tmp = 0;
    HIPCHK(h, hipcub::DeviceRadixSort::SortPairsDescending(
                  nullptr, tmp, h->sc_key.as<uint32_t>(), h->sc_key2.as<uint32_t>(),
                  h->sc_ids.as<int32_t>(), h->sc_order.as<int32_t>(), n, 0, 32, s),
           "sort size");
    HIPCHK(h, h->cubtmp.ensure(tmp + 16), "alloc sort tmp");
    HIPCHK(h, hipcub::DeviceRadixSort::SortPairsDescending(
                  h->cubtmp.p, tmp, h->sc_key.as<uint32_t>(), h->sc_key2.as<uint32_t>(),
                  h->sc_ids.as<int32_t>(), h->sc_order.as<int32_t>(), n, 0, 32, s),
           "sort pairs");
    if (ntasks * 64 > n) {
        hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, s, h->sc_order.as<int32_t>() + n,
                           ntasks * 64 - n, -1);
        HIPCHK(h, hipGetLastError(), "pad order");
    }
    int32_t maxlen = 0;
    HIPCHK(h, hipMemcpyAsync(&maxlen, ctl + 1, sizeof maxlen, hipMemcpyDeviceToHost, s),
           "read lengths");
    HIPCHK(h, hipStreamSynchronize(s), "sync lengths");
    if (maxlen > MAX_LEN16)
        return fail(h, TAL_ERANGE, "a pair length is negative or exceeds %d", MAX_LEN16);
    P.max_qlen = maxlen;
    /* one slab of (longest length + 2) columns x 64 lanes of (H, E) per wave;
     * as many waves as fit in 1 GiB, at most 8 per CU */
    const size_t slab = (size_t)(maxlen + 2) * 64 * sizeof(int2);
    const long long fit = std::max<long long>(1, (1ll << 30) / (long long)slab);
    const int grid = (int)std::max<long long>(
        1, std::min<long long>({(long long)h->cu_count * 8, (long long)ntasks, fit}));
    HIPCHK(h, h->sc_slabs.ensure(slab * grid), "alloc scalar slabs");
    HIPCHK(h, launch_scalar(grid, P, pairs, ref, qer, h->sc_order.as<int32_t>(), ntasks, ctl,
                            h->sc_slabs.as<int2>(), ctl + 2, s),
           "scalar kernel");
    int32_t err = 0;
    HIPCHK(h, hipMemcpyAsync(&err, ctl + 2, sizeof err, hipMemcpyDeviceToHost, s), "read error");
    HIPCHK(h, hipStreamSynchronize(s), "sync scalar");
    if (err) return fail(h, TAL_ERANGE, "a sequence byte is not a base code 0-4");
    return TAL_OK;
}

int tal_bsw_scalar_scores(tal_bsw *h, const int8_t *mat, tal_seqpair *pairs, const uint8_t *ref,
                          size_t ref_len, const uint8_t *qer, size_t qer_len, int32_t n, int32_t w)
{
    if (h && !mat) return fail(h, TAL_EINVAL, "null scoring matrix");
    return host_call(h, 1, pairs, ref, ref_len, qer, qer_len, n, w, mat);
}

int tal_bsw_get_stats(tal_bsw *h, tal_bsw_stats *out)
{
    if (!h || !out) return TAL_EINVAL;
    HIPCHK(h, hipSetDevice(h->device), "set device");
    const float rerun = h->st.ms_rerun16;
    h->st.ms_main16 = h->st.ms_main8 = 0.f;
    if (h->ev16) {
        HIPCHK(h, hipEventSynchronize(h->ev[1]), "event sync");
        HIPCHK(h, hipEventElapsedTime(&h->st.ms_main16, h->ev[0], h->ev[1]), "event time");
        h->st.ms_main16 += rerun;
    }
    if (h->ev8) {
        HIPCHK(h, hipEventSynchronize(h->ev[3]), "event sync");
        HIPCHK(h, hipEventElapsedTime(&h->st.ms_main8, h->ev[2], h->ev[3]), "event time");
    }
    *out = h->st;
    return TAL_OK;
}

int32_t tal_bsw_count_8bit(const tal_bsw *h, const tal_seqpair *pairs, int32_t n)
{
    if (!h || (!pairs && n > 0)) return -1;
    const int sm = smax_of(h->prm);
    int32_t c = 0;
    for (int32_t k = 0; k < n; k++) c += goes_8bit(pairs[k], sm);
    return c;
}

} /* extern "C" */
