// kifi_api.hip -- the extern "C" boundary of libkifi_hip.so (declared in include/kifi.h).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>

#include "kifi_internal.hpp"
#include "kifi_prims.hpp"

namespace kifi {

kifi_status integrate_dispatch(kifi_volume *vol, DepthIn d_depth, const kifi_intrinsics *cam,
                               const float *w2e, const Mat4 *w2e_dev, const SplatParams *splat);
kifi_status mesh_dispatch(kifi_volume *vol, bool tris, float *verts, size_t vcap, uint32_t *idx, size_t icap,
                          size_t *nverts, size_t *nidx, int *stalled, bool async_copy);
kifi_status bin_host(kifi_ctx *ctx, const float *points, size_t n, const float w2c[16], int W, int H,
                     float *out);

kifi_status set_error(kifi_ctx *ctx, kifi_status st, const std::string &msg)
{
    if (ctx)
        ctx->err = msg;
    return st;
}

kifi_status hip_error(kifi_ctx *ctx, hipError_t e, const char *what)
{
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s failed: %s", what, hipGetErrorString(e));
    return set_error(ctx, e == hipErrorOutOfMemory ? KIFI_ERR_NOMEM : KIFI_ERR_HIP, buf);
}

static kifi_status ensure_buf(kifi_ctx *ctx, void **buf, size_t *have, size_t bytes)
{
    if (bytes <= *have)
        return KIFI_OK;
    size_t want = bytes + bytes / 4 + (1 << 20);
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (*buf)
        KIFI_HIP(ctx, hipFree(*buf));
    *buf = nullptr;
    *have = 0;
    KIFI_HIP(ctx, hipMalloc(buf, want));
    *have = want;
    return KIFI_OK;
}

kifi_status ensure_scratch(kifi_ctx *ctx, size_t bytes)
{
    return ensure_buf(ctx, &ctx->scratch, &ctx->scratch_bytes, bytes);
}

kifi_status out_settle(kifi_ctx *ctx)
{
    if (!ctx->out_pending)
        return KIFI_OK;
    ctx->out_pending = false;
    KIFI_HIP(ctx, hipEventSynchronize(ctx->out_copied));
    return KIFI_OK;
}

kifi_status ensure_out(kifi_ctx *ctx, size_t bytes)
{
    if (bytes > ctx->out_bytes) {      // the buffer is freed: no copy may still read it
        kifi_status s = out_settle(ctx);
        if (s != KIFI_OK)
            return s;
    }
    return ensure_buf(ctx, &ctx->out_scratch, &ctx->out_bytes, bytes);
}

// Grow both buffer pairs of a volume to hold `needed` voxels, keeping content.
kifi_status grow_volume(kifi_volume *vol, size_t needed)
{
    kifi_ctx *ctx = vol->ctx;
    size_t cap = vol->cap ? vol->cap : (size_t)1 << 20;
    while (cap < needed)
        cap = cap + cap / 2;
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint64_t n = 0;
    KIFI_HIP(ctx, hipMemcpy(&n, vol->d_size, 8, hipMemcpyDeviceToHost));
    void *nk[2] = { nullptr, nullptr }, *nv[2] = { nullptr, nullptr };
    void *ns1[2] = { nullptr, nullptr }, *ns2[2] = { nullptr, nullptr };
    // key samples: every 16th and every 4096th key; the key buffers carry 32
    // keys of padding (the integration reads aligned windows of up to 32 keys
    // that start below the size)
    const size_t s1b = ((cap + 15) / 16 + 1) * vol->key_bytes, s2b = ((cap + 4095) / 4096 + 1) * vol->key_bytes;
    for (int b = 0; b < 2; ++b) {
        hipError_t e = hipMalloc(&nk[b], (cap + 32) * vol->key_bytes);
        if (e == hipSuccess) e = hipMalloc(&nv[b], cap * vol->val_bytes);
        if (e == hipSuccess) e = hipMalloc(&ns1[b], s1b);
        if (e == hipSuccess) e = hipMalloc(&ns2[b], s2b);
        if (e != hipSuccess) {
            for (int q = 0; q <= b; ++q) {
                if (nk[q]) (void)hipFree(nk[q]);
                if (nv[q]) (void)hipFree(nv[q]);
                if (ns1[q]) (void)hipFree(ns1[q]);
                if (ns2[q]) (void)hipFree(ns2[q]);
            }
            return hip_error(ctx, e, "hipMalloc(volume)");
        }
    }
    if (n) {
        KIFI_HIP(ctx, hipMemcpy(nk[0], vol->keys[vol->cur], n * vol->key_bytes, hipMemcpyDeviceToDevice));
        KIFI_HIP(ctx, hipMemcpy(nv[0], vol->vals[vol->cur], n * vol->val_bytes, hipMemcpyDeviceToDevice));
    }
    for (int b = 0; b < 2; ++b) {
        if (vol->keys[b]) KIFI_HIP(ctx, hipFree(vol->keys[b]));
        if (vol->vals[b]) KIFI_HIP(ctx, hipFree(vol->vals[b]));
        if (vol->samp1[b]) KIFI_HIP(ctx, hipFree(vol->samp1[b]));
        if (vol->samp2[b]) KIFI_HIP(ctx, hipFree(vol->samp2[b]));
        vol->keys[b] = nk[b];
        vol->vals[b] = nv[b];
        vol->samp1[b] = ns1[b];
        vol->samp2[b] = ns2[b];
    }
    vol->cur = 0;
    vol->samples_valid = false;
    vol->cap = cap;
    ring_settle(vol, n);
    return KIFI_OK;
}

void ring_refresh(kifi_volume *v)
{
    for (;;) {
        int s = -1;
        for (int i = 0; i < kifi_volume::kSizeRing; ++i)
            if (v->ring_frame[i] && (s < 0 || v->ring_frame[i] < v->ring_frame[s]))
                s = i;
        if (s < 0 || hipEventQuery(v->ring_ev[s]) != hipSuccess)
            return;
        const FrameCounters &c = v->h_ring[s];
        if (c.overflow) {
            v->ring_overflow = true;     // the volume kept its size; known_size stays a bound
        } else if (v->ring_frame[s] > v->known_frame) {
            v->known_size = (size_t)c.v_new;
            v->known_frame = v->ring_frame[s];
        }
        v->ring_frame[s] = 0;
    }
}

void ring_settle(kifi_volume *v, size_t n)
{
    ring_refresh(v);                      // keeps the overflow flags of the frames just finished
    for (int i = 0; i < kifi_volume::kSizeRing; ++i)
        v->ring_frame[i] = 0;
    v->known_size = n;
    v->known_frame = v->frames_enq;
}

// Segmented scans and reduction with the workgroup primitives every kernel
// uses (kifi_prims.hpp): one segment per workgroup of NT = segment threads.
// OP 0: inclusive scan, 1: exclusive scan, 2: sum per segment.
template <int NT, int OP>
__global__ __launch_bounds__(NT) void k_segmented(const uint32_t *__restrict__ in, uint32_t n,
                                                  uint32_t *__restrict__ out)
{
    __shared__ uint32_t tmp[NT / kWave + 1];
    const size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
    const uint32_t v = i < n ? in[i] : 0u;
    if (OP == 2) {
        const uint32_t s = block_reduce_sum<NT>(v, tmp);
        if (threadIdx.x == 0)
            out[blockIdx.x] = s;
    } else {
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<NT>(v, tmp, tot);
        if (i < n)
            out[i] = OP == 0 ? ex + v : ex;
    }
}

template <int OP>
static kifi_status segmented(kifi_ctx *ctx, const uint32_t *data, uint32_t size, uint32_t segment, uint32_t *out)
{
    if (!ctx || (size && (!data || !out)) || (segment != 64 && segment != 256 && segment != 1024))
        return ctx ? set_error(ctx, KIFI_ERR_ARG, "segmented scan/reduce: segment must be 64, 256 or 1024")
                   : KIFI_ERR_ARG;
    if (!size)
        return KIFI_OK;
    const uint32_t segs = (size + segment - 1) / segment;
    const size_t outn = OP == 2 ? segs : size;
    const size_t inb = ((size_t)size * 4 + 255) & ~size_t(255);
    kifi_status st = ensure_scratch(ctx, inb + outn * 4);
    if (st != KIFI_OK)
        return st;
    uint32_t *din = (uint32_t *)ctx->scratch, *dout = (uint32_t *)((char *)ctx->scratch + inb);
    KIFI_HIP(ctx, hipMemcpyAsync(din, data, (size_t)size * 4, hipMemcpyHostToDevice, ctx->stream));
    if (segment == 64)
        hipLaunchKernelGGL((k_segmented<64, OP>), dim3(segs), dim3(64), 0, ctx->stream, din, size, dout);
    else if (segment == 256)
        hipLaunchKernelGGL((k_segmented<256, OP>), dim3(segs), dim3(256), 0, ctx->stream, din, size, dout);
    else
        hipLaunchKernelGGL((k_segmented<1024, OP>), dim3(segs), dim3(1024), 0, ctx->stream, din, size, dout);
    KIFI_HIP(ctx, hipGetLastError());
    KIFI_HIP(ctx, hipMemcpyAsync(out, dout, outn * 4, hipMemcpyDeviceToHost, ctx->stream));
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KIFI_OK;
}

__global__ void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

static hipEvent_t pool_event(kifi_ctx *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

int timer_begin(kifi_ctx *ctx, int stage)
{
    if (stage < 0 || stage >= KIFI_STAGE_COUNT || !(ctx->timing_mask >> stage & 1))
        return -1;
    hipEvent_t a = pool_event(ctx), b = pool_event(ctx);
    if (!a || !b)
        return -1;
    (void)hipEventRecord(a, ctx->stream);
    ctx->timers[stage].push_back({ a, b });
    return (int)ctx->timers[stage].size() - 1;
}

void timer_end(kifi_ctx *ctx, int stage, int slot)
{
    if (slot < 0)
        return;
    (void)hipEventRecord(ctx->timers[stage][slot].second, ctx->stream);
}

} // namespace kifi

using namespace kifi;

extern "C" {

kifi_status kifi_timing_enable(kifi_ctx *ctx, int stage_mask)
{
    if (!ctx)
        return KIFI_ERR_ARG;
    ctx->timing_mask = (unsigned)stage_mask & ((1u << KIFI_STAGE_COUNT) - 1u);
    return KIFI_OK;
}

kifi_status kifi_timing_read(kifi_ctx *ctx, int stage, double *total_ms, long long *count)
{
    if (!ctx || stage < 0 || stage >= KIFI_STAGE_COUNT || !total_ms || !count)
        return KIFI_ERR_ARG;
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double tot = 0.0;
    for (auto &pr : ctx->timers[stage]) {
        float ms = 0.0f;
        KIFI_HIP(ctx, hipEventElapsedTime(&ms, pr.first, pr.second));
        tot += ms;
        ctx->event_pool.push_back(pr.first);
        ctx->event_pool.push_back(pr.second);
    }
    *count = (long long)ctx->timers[stage].size();
    *total_ms = tot;
    ctx->timers[stage].clear();
    return KIFI_OK;
}

kifi_status kifi_device_alloc(kifi_ctx *ctx, size_t bytes, void **out)
{
    if (!ctx || !out)
        return KIFI_ERR_ARG;
    *out = nullptr;
    KIFI_HIP(ctx, hipSetDevice(ctx->device));
    KIFI_HIP(ctx, hipMalloc(out, bytes ? bytes : 1));
    return KIFI_OK;
}

kifi_status kifi_device_free(kifi_ctx *ctx, void *ptr)
{
    if (!ctx)
        return KIFI_ERR_ARG;
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ptr)
        KIFI_HIP(ctx, hipFree(ptr));
    return KIFI_OK;
}

kifi_status kifi_host_alloc(kifi_ctx *ctx, size_t bytes, void **out)
{
    if (!ctx || !out)
        return KIFI_ERR_ARG;
    *out = nullptr;
    KIFI_HIP(ctx, hipSetDevice(ctx->device));
    KIFI_HIP(ctx, hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault));
    return KIFI_OK;
}

kifi_status kifi_host_free(kifi_ctx *ctx, void *ptr)
{
    if (!ctx)
        return KIFI_ERR_ARG;
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    kifi_status s = out_settle(ctx);       // a mesh copy may still write into ptr
    if (s != KIFI_OK)
        return s;
    if (ptr)
        KIFI_HIP(ctx, hipHostFree(ptr));
    return KIFI_OK;
}

kifi_status kifi_device_copy(kifi_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src)) || bytes % 16 != 0)
        return KIFI_ERR_ARG;
    const size_t n = bytes / 16;
    if (n) {
        const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)ctx->num_cus * 32);
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, ctx->stream, (const uint4 *)src, (uint4 *)dst, n);
        KIFI_HIP(ctx, hipGetLastError());
    }
    return KIFI_OK;
}

kifi_status kifi_segmented_inclusive_scan(kifi_ctx *ctx, const uint32_t *data, uint32_t size, uint32_t segment,
                                          uint32_t *out)
{
    return segmented<0>(ctx, data, size, segment, out);
}

kifi_status kifi_segmented_exclusive_scan(kifi_ctx *ctx, const uint32_t *data, uint32_t size, uint32_t segment,
                                          uint32_t *out)
{
    return segmented<1>(ctx, data, size, segment, out);
}

kifi_status kifi_segmented_reduce(kifi_ctx *ctx, const uint32_t *data, uint32_t size, uint32_t segment,
                                  uint32_t *sums)
{
    return segmented<2>(ctx, data, size, segment, sums);
}

kifi_status kifi_integrate_totals(kifi_volume *v, uint64_t totals[6])
{
    if (!v || !totals)
        return KIFI_ERR_ARG;
    kifi_ctx *ctx = v->ctx;
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    FrameCounters c;
    KIFI_HIP(ctx, hipMemcpy(&c, v->d_counters, sizeof c, hipMemcpyDeviceToHost));
    uint64_t upd[kUpdSlots * kUpdStride];
    KIFI_HIP(ctx, hipMemcpy(upd, v->d_upd, sizeof upd, hipMemcpyDeviceToHost));
    totals[0] = c.tot_frames;
    totals[1] = c.tot_v_old;
    totals[2] = c.tot_new;
    totals[3] = c.tot_v_new;
    totals[4] = c.tot_keys;
    totals[5] = 0;
    for (int i = 0; i < kUpdSlots; ++i)
        totals[5] += upd[i * kUpdStride];
    return KIFI_OK;
}

kifi_status kifi_memcpy_to_device(kifi_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src)))
        return KIFI_ERR_ARG;
    KIFI_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KIFI_OK;
}

kifi_status kifi_ctx_create(int device, kifi_ctx **out)
{
    if (!out)
        return KIFI_ERR_ARG;
    *out = nullptr;
    kifi_ctx *ctx = new (std::nothrow) kifi_ctx;
    if (!ctx)
        return KIFI_ERR_NOMEM;
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
        delete ctx;
        return e == hipErrorOutOfMemory ? KIFI_ERR_NOMEM : KIFI_ERR_HIP;
    }
    *out = ctx;
    return KIFI_OK;
}

void kifi_ctx_destroy(kifi_ctx *ctx)
{
    if (!ctx)
        return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    if (ctx->out_stream)
        (void)hipStreamSynchronize(ctx->out_stream);
    if (ctx->out_ready) (void)hipEventDestroy(ctx->out_ready);
    if (ctx->out_copied) (void)hipEventDestroy(ctx->out_copied);
    if (ctx->out_stream) (void)hipStreamDestroy(ctx->out_stream);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->out_scratch) (void)hipFree(ctx->out_scratch);
    if (ctx->copy_stream)
        (void)hipStreamSynchronize(ctx->copy_stream);
    for (int i = 0; i < 2; ++i) {
        if (ctx->stage[i]) (void)hipFree(ctx->stage[i]);
        if (ctx->stage_copied[i]) (void)hipEventDestroy(ctx->stage_copied[i]);
        if (ctx->stage_used[i]) (void)hipEventDestroy(ctx->stage_used[i]);
    }
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    for (auto &v : ctx->timers)
        for (auto &pr : v) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (hipEvent_t e : ctx->event_pool)
        (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *kifi_last_error(const kifi_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : "no context";
}

kifi_status kifi_ctx_synchronize(kifi_ctx *ctx)
{
    if (!ctx)
        return KIFI_ERR_ARG;
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return out_settle(ctx);
}

void *kifi_ctx_stream(kifi_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

static bool pow2(int x) { return x > 0 && !(x & (x - 1)); }

kifi_status kifi_volume_create(kifi_ctx *ctx, const kifi_volume_desc *d, kifi_volume **out)
{
    if (!ctx || !d || !out)
        return KIFI_ERR_ARG;
    *out = nullptr;
    // Volume.cpp:22-26 asserts, extended to 64-bit keys
    const int maxres = d->key_bits == 32 ? 1024 : (1 << 20);
    if ((d->key_bits != 32 && d->key_bits != 64) || !pow2(d->res) || d->res > maxres || d->res < 2 ||
        !(d->side > 0.0f) || !(d->trunc > 0.0f) ||
        (d->storage != KIFI_STORAGE_F32 && d->storage != KIFI_STORAGE_F16U8))
        return set_error(ctx, KIFI_ERR_ARG, "kifi_volume_create: invalid volume description");
    KIFI_HIP(ctx, hipSetDevice(ctx->device));
    kifi_volume *v = new (std::nothrow) kifi_volume;
    if (!v)
        return KIFI_ERR_NOMEM;
    v->ctx = ctx;
    v->desc = *d;
    v->key_bytes = d->key_bits / 8;
    v->val_bytes = d->storage == KIFI_STORAGE_F32 ? 8 : 4;
    // Volume.cpp:28-36
    const float minx = -0.5f * d->side;
    v->consts.res = d->res;
    v->consts.vl = d->side / (float)d->res;
    v->consts.vlHalfPlusMin = v->consts.vl * 0.5f + minx;
    v->consts.vlInv = 1.0f / v->consts.vl;
    v->consts.minOverNegVl = v->consts.vlInv * -minx;
    v->consts.trunc = d->trunc;
    const uint64_t zero = 0;
    hipError_t e = hipMalloc(&v->d_counters, sizeof(FrameCounters));
    if (e == hipSuccess) e = hipMalloc(&v->d_upd, kUpdSlots * kUpdStride * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemset(v->d_upd, 0, kUpdSlots * kUpdStride * sizeof(uint64_t));
    if (e == hipSuccess) e = hipHostMalloc(&v->h_ring, kifi_volume::kSizeRing * sizeof(FrameCounters),
                                           hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&v->h_ring_dev, v->h_ring, 0);
    if (e == hipSuccess) e = hipMalloc(&v->d_size, 8);
    if (e == hipSuccess) e = hipMalloc(&v->d_mats, sizeof(FrameMats));
    for (int i = 0; i < kifi_volume::kSizeRing && e == hipSuccess; ++i)
        e = hipEventCreateWithFlags(&v->ring_ev[i], hipEventDisableTiming);
    if (e == hipSuccess) e = hipMemcpy(v->d_size, &zero, 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(v->d_counters, 0, sizeof(FrameCounters));
    if (e != hipSuccess) {
        kifi_volume_destroy(v);
        return hip_error(ctx, e, "kifi_volume_create");
    }
    std::memset(v->h_ring, 0, kifi_volume::kSizeRing * sizeof(FrameCounters));
    kifi_status st = grow_volume(v, (size_t)1 << 20);
    if (st != KIFI_OK) {
        kifi_volume_destroy(v);
        return st;
    }
    *out = v;
    return KIFI_OK;
}

void kifi_volume_destroy(kifi_volume *v)
{
    if (!v)
        return;
    (void)hipSetDevice(v->ctx->device);
    (void)hipStreamSynchronize(v->ctx->stream);
    for (int b = 0; b < 2; ++b) {
        if (v->keys[b]) (void)hipFree(v->keys[b]);
        if (v->vals[b]) (void)hipFree(v->vals[b]);
        if (v->samp1[b]) (void)hipFree(v->samp1[b]);
        if (v->samp2[b]) (void)hipFree(v->samp2[b]);
    }
    if (v->d_counters) (void)hipFree(v->d_counters);
    if (v->d_upd) (void)hipFree(v->d_upd);
    if (v->h_ring) (void)hipHostFree(v->h_ring);
    if (v->d_size) (void)hipFree(v->d_size);
    if (v->d_mats) (void)hipFree(v->d_mats);
    for (hipEvent_t ev : v->ring_ev)
        if (ev) (void)hipEventDestroy(ev);
    delete v;
}

static kifi_status sync_size(kifi_volume *v, uint64_t *n)
{
    kifi_ctx *ctx = v->ctx;
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ring_refresh(v);
    const bool overflow = v->ring_overflow;
    v->ring_overflow = false;
    KIFI_HIP(ctx, hipMemcpy(n, v->d_size, 8, hipMemcpyDeviceToHost));
    ring_settle(v, (size_t)*n);
    if (overflow)
        return set_error(ctx, KIFI_ERR_STATE, "device-side capacity overflow in an integration since the last check");
    return KIFI_OK;
}

kifi_status kifi_volume_clear(kifi_volume *v)
{
    if (!v)
        return KIFI_ERR_ARG;
    kifi_ctx *ctx = v->ctx;
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint64_t zero = 0;
    KIFI_HIP(ctx, hipMemcpy(v->d_size, &zero, 8, hipMemcpyHostToDevice));
    ring_settle(v, 0);
    return KIFI_OK;
}

kifi_status kifi_volume_size(kifi_volume *v, size_t *n)
{
    if (!v || !n)
        return KIFI_ERR_ARG;
    uint64_t s = 0;
    kifi_status st = sync_size(v, &s);
    *n = (size_t)s;
    return st;
}

kifi_status kifi_volume_download(kifi_volume *v, void *keys, void *values, size_t cap, size_t *n)
{
    if (!v || !n)
        return KIFI_ERR_ARG;
    uint64_t s = 0;
    kifi_status st = sync_size(v, &s);
    if (st != KIFI_OK)
        return st;
    *n = (size_t)s;
    if (cap < s)
        return set_error(v->ctx, KIFI_ERR_CAPACITY, "kifi_volume_download: buffer too small");
    if (s && keys)
        KIFI_HIP(v->ctx, hipMemcpy(keys, v->keys[v->cur], s * v->key_bytes, hipMemcpyDeviceToHost));
    if (s && values)
        KIFI_HIP(v->ctx, hipMemcpy(values, v->vals[v->cur], s * v->val_bytes, hipMemcpyDeviceToHost));
    return KIFI_OK;
}

kifi_status kifi_volume_upload(kifi_volume *v, const void *keys, const void *values, size_t n)
{
    if (!v || (n && (!keys || !values)))
        return KIFI_ERR_ARG;
    kifi_ctx *ctx = v->ctx;
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (n > v->cap) {
        uint64_t zero = 0;
        KIFI_HIP(ctx, hipMemcpy(v->d_size, &zero, 8, hipMemcpyHostToDevice));
        kifi_status st = grow_volume(v, n);
        if (st != KIFI_OK)
            return st;
    }
    if (n) {
        KIFI_HIP(ctx, hipMemcpy(v->keys[v->cur], keys, n * v->key_bytes, hipMemcpyHostToDevice));
        KIFI_HIP(ctx, hipMemcpy(v->vals[v->cur], values, n * v->val_bytes, hipMemcpyHostToDevice));
    }
    uint64_t s = n;
    KIFI_HIP(ctx, hipMemcpy(v->d_size, &s, 8, hipMemcpyHostToDevice));
    v->samples_valid = false;
    ring_settle(v, n);
    return KIFI_OK;
}

// ---- volume file: checkpoint / resume
// "KIFIVOL\n", uint32 version 1, int32 res, float side, float trunc, int32
// key_bits, int32 storage, uint64 n, n keys, n values (the device layouts),
// little endian.
static const char kVolMagic[8] = { 'K', 'I', 'F', 'I', 'V', 'O', 'L', '\n' };

kifi_status kifi_volume_save(kifi_volume *v, const char *path)
{
    if (!v || !path)
        return KIFI_ERR_ARG;
    kifi_ctx *ctx = v->ctx;
    size_t n = 0;
    kifi_status st = kifi_volume_download(v, nullptr, nullptr, 0, &n);
    if (st != KIFI_OK && st != KIFI_ERR_CAPACITY)
        return st;
    std::vector<char> keys(n * v->key_bytes), vals(n * v->val_bytes);
    size_t got = 0;
    st = kifi_volume_download(v, keys.data(), vals.data(), n, &got);
    if (st != KIFI_OK)
        return st;
    FILE *f = std::fopen(path, "wb");
    if (!f)
        return set_error(ctx, KIFI_ERR_ARG, "kifi_volume_save: cannot open the file");
    const uint32_t version = 1;
    const int32_t res = v->desc.res, kb = v->desc.key_bits, storage = v->desc.storage;
    const float side = v->desc.side, trunc = v->desc.trunc;
    const uint64_t n64 = n;
    bool ok = std::fwrite(kVolMagic, 1, 8, f) == 8 && std::fwrite(&version, 4, 1, f) == 1 &&
              std::fwrite(&res, 4, 1, f) == 1 && std::fwrite(&side, 4, 1, f) == 1 &&
              std::fwrite(&trunc, 4, 1, f) == 1 && std::fwrite(&kb, 4, 1, f) == 1 &&
              std::fwrite(&storage, 4, 1, f) == 1 && std::fwrite(&n64, 8, 1, f) == 1 &&
              std::fwrite(keys.data(), 1, keys.size(), f) == keys.size() &&
              std::fwrite(vals.data(), 1, vals.size(), f) == vals.size();
    ok = std::fclose(f) == 0 && ok;
    return ok ? KIFI_OK : set_error(ctx, KIFI_ERR_ARG, "kifi_volume_save: write failed");
}

} // extern "C"

// Keys of a loaded file: strictly increasing (the merge relies on it), no bits
// above the three axes, every axis below res.
template <typename K>
static bool valid_keys(const K *k, uint64_t n, int res)
{
    constexpr int kb = KeyTraits<K>::kb;
    for (uint64_t i = 0; i < n; ++i) {
        if (i > 0 && !(k[i - 1] < k[i]))
            return false;
        if (k[i] >> (3 * kb))
            return false;
        uint32_t x, y, z;
        unpack_key<K>(k[i], x, y, z);
        if (x >= (uint32_t)res || y >= (uint32_t)res || z >= (uint32_t)res)
            return false;
    }
    return true;
}

extern "C" {

kifi_status kifi_volume_load(kifi_ctx *ctx, const char *path, kifi_volume **out)
{
    if (!ctx || !path || !out)
        return KIFI_ERR_ARG;
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f)
        return set_error(ctx, KIFI_ERR_ARG, "kifi_volume_load: cannot open the file");
    char magic[8];
    uint32_t version = 0;
    kifi_volume_desc d{};
    int32_t res = 0, kb = 0, storage = 0;
    uint64_t n = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kVolMagic, 8) == 0 &&
              std::fread(&version, 4, 1, f) == 1 && version == 1 && std::fread(&res, 4, 1, f) == 1 &&
              std::fread(&d.side, 4, 1, f) == 1 && std::fread(&d.trunc, 4, 1, f) == 1 &&
              std::fread(&kb, 4, 1, f) == 1 && std::fread(&storage, 4, 1, f) == 1 && std::fread(&n, 8, 1, f) == 1;
    d.res = res;
    d.key_bits = kb;
    d.storage = storage;
    // The header is checked before anything is allocated: a valid description,
    // n at most res^3 voxels, and exactly n keys and n values left in the file.
    const char *why = ok ? nullptr : "not a volume file (header)";
    const int maxres = kb == 32 ? 1024 : (1 << 20);
    if (!why && ((kb != 32 && kb != 64) || res < 2 || res > maxres || (res & (res - 1)) ||
                 (storage != KIFI_STORAGE_F32 && storage != KIFI_STORAGE_F16U8) || !(d.side > 0.0f) ||
                 !(d.trunc > 0.0f)))
        why = "invalid volume description";
    const uint64_t rec = (uint64_t)(kb / 8) + (storage == KIFI_STORAGE_F32 ? 8u : 4u);
    if (!why && n > (uint64_t)res * (uint64_t)res * (uint64_t)res)   // res <= 2^20: res^3 <= 2^60
        why = "voxel count above res^3";
    if (!why) {
        const long here = std::ftell(f);
        if (here < 0 || std::fseek(f, 0, SEEK_END) != 0)
            why = "cannot size the file";
        const long end = why ? -1 : std::ftell(f);
        if (!why && (end < here || std::fseek(f, here, SEEK_SET) != 0))
            why = "cannot size the file";
        if (!why) {
            const uint64_t left = (uint64_t)(end - here);
            if (left % rec != 0 || left / rec != n)
                why = "file size does not match its voxel count";
        }
    }
    if (why) {
        std::fclose(f);
        return set_error(ctx, KIFI_ERR_ARG, std::string("kifi_volume_load: ") + why);
    }
    kifi_volume *v = nullptr;
    kifi_status st = kifi_volume_create(ctx, &d, &v);
    if (st == KIFI_OK) {
        std::vector<char> keys, vals;
        try {
            keys.resize(n * v->key_bytes);
            vals.resize(n * v->val_bytes);
        } catch (const std::exception &) {
            st = set_error(ctx, KIFI_ERR_NOMEM, "kifi_volume_load: host memory for the voxels");
        }
        if (st == KIFI_OK) {
            ok = std::fread(keys.data(), 1, keys.size(), f) == keys.size() &&
                 std::fread(vals.data(), 1, vals.size(), f) == vals.size();
            ok = ok && (v->key_bytes == 4 ? valid_keys((const uint32_t *)keys.data(), n, res)
                                          : valid_keys((const uint64_t *)keys.data(), n, res));
            st = ok ? kifi_volume_upload(v, keys.data(), vals.data(), n)
                    : set_error(ctx, KIFI_ERR_ARG, "kifi_volume_load: keys not strictly increasing or "
                                                   "outside the volume");
        }
    }
    std::fclose(f);
    if (st != KIFI_OK) {
        if (v)
            kifi_volume_destroy(v);
        return st;
    }
    *out = v;
    return KIFI_OK;
}

static kifi_status check_cam(kifi_ctx *ctx, const kifi_intrinsics *cam)
{
    // Integrator.cpp:35-36, 89-90 asserts
    if (!cam || cam->width <= 0 || cam->height <= 0 || cam->width % 4 != 0 || !(cam->fx > 0.0f) ||
        !(cam->fy > 0.0f))
        return set_error(ctx, KIFI_ERR_ARG, "invalid intrinsics (width must be a positive multiple of 4)");
    return KIFI_OK;
}

static kifi_status integrate_device(kifi_volume *v, const void *depth_dev, bool metres, const kifi_intrinsics *cam,
                                   const float worldToEye[16])
{
    if (!v || !depth_dev || !worldToEye)
        return KIFI_ERR_ARG;
    kifi_status st = check_cam(v->ctx, cam);
    if (st != KIFI_OK)
        return st;
    DepthIn d;
    d.p = depth_dev;
    d.metres = metres;
    return integrate_dispatch(v, d, cam, worldToEye, nullptr, nullptr);
}

} // extern "C"

namespace kifi {

kifi_status stage_frame(kifi_ctx *ctx, const void *host, size_t bytes, const void **dev, int *slot)
{
    if (!ctx->copy_stream) {
        KIFI_HIP(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
        for (int i = 0; i < 2; ++i) {
            KIFI_HIP(ctx, hipEventCreateWithFlags(&ctx->stage_copied[i], hipEventDisableTiming));
            KIFI_HIP(ctx, hipEventCreateWithFlags(&ctx->stage_used[i], hipEventDisableTiming));
        }
    }
    if (bytes > ctx->stage_bytes) {
        KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
        KIFI_HIP(ctx, hipStreamSynchronize(ctx->copy_stream));
        for (int i = 0; i < 2; ++i) {
            if (ctx->stage[i])
                KIFI_HIP(ctx, hipFree(ctx->stage[i]));
            ctx->stage[i] = nullptr;
            ctx->stage_busy[i] = false;
        }
        ctx->stage_bytes = 0;
        for (int i = 0; i < 2; ++i)
            KIFI_HIP(ctx, hipMalloc(&ctx->stage[i], bytes));
        ctx->stage_bytes = bytes;
    }
    const int s = ctx->stage_slot;
    ctx->stage_slot ^= 1;
    if (ctx->stage_busy[s])       // the frame before last may still read this buffer
        KIFI_HIP(ctx, hipStreamWaitEvent(ctx->copy_stream, ctx->stage_used[s], 0));
    KIFI_HIP(ctx, hipMemcpyAsync(ctx->stage[s], host, bytes, hipMemcpyHostToDevice, ctx->copy_stream));
    KIFI_HIP(ctx, hipEventRecord(ctx->stage_copied[s], ctx->copy_stream));
    KIFI_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->stage_copied[s], 0));
    // the caller may reuse its host buffer when the call returns: wait until the
    // copy has read it (the kernels of this frame still run asynchronously)
    KIFI_HIP(ctx, hipEventSynchronize(ctx->stage_copied[s]));
    *dev = ctx->stage[s];
    *slot = s;
    return KIFI_OK;
}

void stage_done(kifi_ctx *ctx, int slot)
{
    if (hipEventRecord(ctx->stage_used[slot], ctx->stream) == hipSuccess)
        ctx->stage_busy[slot] = true;
}

} // namespace kifi

extern "C" {

static kifi_status integrate_host(kifi_volume *v, const void *depth, size_t texel, const kifi_intrinsics *cam,
                                  const float worldToEye[16])
{
    if (!v || !depth || !worldToEye)
        return KIFI_ERR_ARG;
    kifi_ctx *ctx = v->ctx;
    kifi_status st = check_cam(ctx, cam);
    if (st != KIFI_OK)
        return st;
    const size_t P = (size_t)cam->width * cam->height;
    DepthIn d;
    int slot = 0;
    st = stage_frame(ctx, depth, P * texel, &d.p, &slot);
    if (st != KIFI_OK)
        return st;
    d.metres = texel == 4;
    st = integrate_dispatch(v, d, cam, worldToEye, nullptr, nullptr);
    stage_done(ctx, slot);
    return st;
}

kifi_status kifi_integrate_device(kifi_volume *v, const uint16_t *depth_dev, const kifi_intrinsics *cam,
                                  const float worldToEye[16])
{
    return integrate_device(v, depth_dev, false, cam, worldToEye);
}

kifi_status kifi_integrate(kifi_volume *v, const uint16_t *depth_mm, const kifi_intrinsics *cam,
                           const float worldToEye[16])
{
    return integrate_host(v, depth_mm, 2, cam, worldToEye);
}

kifi_status kifi_integrate_device_f32(kifi_volume *v, const float *depth_dev, const kifi_intrinsics *cam,
                                      const float worldToEye[16])
{
    return integrate_device(v, depth_dev, true, cam, worldToEye);
}

kifi_status kifi_integrate_f32(kifi_volume *v, const float *depth_m, const kifi_intrinsics *cam,
                               const float worldToEye[16])
{
    return integrate_host(v, depth_m, 4, cam, worldToEye);
}

kifi_status kifi_integrate_stats(kifi_volume *v, size_t stats[4])
{
    if (!v || !stats)
        return KIFI_ERR_ARG;
    kifi_ctx *ctx = v->ctx;
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    FrameCounters c;
    KIFI_HIP(ctx, hipMemcpy(&c, v->d_counters, sizeof c, hipMemcpyDeviceToHost));
    if (c.overflow)
        return set_error(ctx, KIFI_ERR_STATE, "device-side capacity overflow in the last integration");
    stats[0] = c.n_keys;
    stats[1] = c.n_cand;
    stats[2] = c.n_new;
    stats[3] = (size_t)c.v_new;
    return KIFI_OK;
}

kifi_status kifi_mesh(kifi_volume *v, float *verts, size_t vcap, size_t *nverts)
{
    if (!v || !nverts || (vcap && !verts))
        return KIFI_ERR_ARG;
    kifi_status st = mesh_dispatch(v, false, verts, vcap, nullptr, 0, nverts, nullptr, nullptr, false);
    if (st == KIFI_ERR_CAPACITY && *nverts <= 0xffffffffull)   // else the 32-bit limit's message stands
        set_error(v->ctx, st, "kifi_mesh: vertex buffer too small");
    return st;
}

kifi_status kifi_mesh_async(kifi_volume *v, float *verts, size_t vcap, size_t *nverts)
{
    if (!v || !nverts || (vcap && !verts))
        return KIFI_ERR_ARG;
    kifi_status st = mesh_dispatch(v, false, verts, vcap, nullptr, 0, nverts, nullptr, nullptr, true);
    if (st == KIFI_ERR_CAPACITY && *nverts <= 0xffffffffull)
        set_error(v->ctx, st, "kifi_mesh_async: vertex buffer too small");
    return st;
}

kifi_status kifi_mesh_wait(kifi_ctx *ctx)
{
    if (!ctx)
        return KIFI_ERR_ARG;
    return out_settle(ctx);
}

kifi_status kifi_mesh_indexed(kifi_volume *v, float *verts, size_t vcap, uint32_t *idx, size_t icap,
                              size_t *nverts, size_t *nidx, int *stalled)
{
    if (!v || !nverts || !nidx || (vcap && !verts) || (icap && !idx))
        return KIFI_ERR_ARG;
    kifi_status st = mesh_dispatch(v, true, verts, vcap, idx, icap, nverts, nidx, stalled, false);
    if (st == KIFI_ERR_CAPACITY && *nverts <= 0xffffffffull && *nidx <= 0xffffffffull)
        set_error(v->ctx, st, "kifi_mesh_indexed: buffer too small");
    return st;
}

kifi_status kifi_bin(kifi_ctx *ctx, const float *points, size_t n, const float worldToClip[16], int width,
                     int height, float *out_points)
{
    if (!ctx || (n && !points) || !worldToClip || width <= 0 || height <= 0 || !out_points)
        return KIFI_ERR_ARG;
    if (n >= 0xffffffffull)
        return set_error(ctx, KIFI_ERR_ARG, "kifi_bin: too many points");
    return bin_host(ctx, points, n, worldToClip, width, height, out_points);
}

} // extern "C"
