// kifi_pipeline.hip -- C ABI of model prediction, ICP, Horn and the frame loop.
#include <cstddef>
#include <cstring>
#include <mutex>
#include <new>

#include "kifi_icp.hpp"
#include "kifi_internal.hpp"

namespace kifi {
kifi_status integrate_dispatch(kifi_volume *vol, DepthIn d_depth, const kifi_intrinsics *cam,
                               const float *w2e, const Mat4 *w2e_dev, const SplatParams *splat);
void icp_set_pose(IcpWork *w, const double *e2w, const float *w2e, hipStream_t st);
void icp_set_prev(IcpWork *w, const float *w2e, hipStream_t st);
bool twistnet_guess(kifi_twistnet *n, const uint16_t *d_mm, int W, int H, const double *prev_pose_dev,
                    double *guess_dev, hipStream_t st);
void twistnet_observe(kifi_twistnet *n, const uint16_t *d_mm, int W, int H, hipStream_t st);
void twistnet_reset_frames(kifi_twistnet *n);
} // namespace kifi

using namespace kifi;

// ---- allocation-free ICP: its work memory is this file-scope array ----
// One copy per GPU, held until the process ends.  In host code `scratch` is a
// host-side stand-in for the symbol, never a device address: its device
// address comes from hipGetSymbolAddress, once per device with that device
// current, and only that address is used afterwards.
namespace {
constexpr size_t kNoAllocBytes = size_t(2) << 20;              // 2 MiB
constexpr size_t kNoAllocMaxPixels = size_t(640) * 480;        // frames up to 640 x 480
constexpr int kNoAllocDevices = 64;
struct NoAllocSlot {
    std::once_flag once;
    void          *ptr = nullptr;
    hipError_t     err = hipSuccess;
};
NoAllocSlot g_noalloc[kNoAllocDevices];
} // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// The device address of this device's `scratch` (looked up on the first call).
static kifi_status noalloc_scratch(kifi_ctx *ctx, char **out)
{
    if (ctx->device < 0 || ctx->device >= kNoAllocDevices)
        return set_error(ctx, KIFI_ERR_ARG, "allocation-free entry point: device id above 63");
    NoAllocSlot &slot = g_noalloc[ctx->device];
    std::call_once(slot.once, [&] {
        int prev = -1;
        hipError_t e = hipGetDevice(&prev);
        if (e == hipSuccess)
            e = hipSetDevice(ctx->device);
        void *p = nullptr;
        if (e == hipSuccess)
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev >= 0)
            (void)hipSetDevice(prev);
        slot.err = e;
        slot.ptr = e == hipSuccess ? p : nullptr;
    });
    if (slot.err != hipSuccess || !slot.ptr)
        return hip_error(ctx, slot.err != hipSuccess ? slot.err : hipErrorInvalidSymbol,
                         "hipGetSymbolAddress(scratch)");
    *out = (char *)slot.ptr;
    return KIFI_OK;
}

// Fill kernel (no memset node): the tracking state starts at zero on every call.
__global__ void k_zero_words(uint32_t *__restrict__ p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

struct kifi_pipeline {
    kifi_ctx        *ctx = nullptr;
    kifi_volume     *vol = nullptr;
    kifi_model      *model = nullptr;
    kifi_intrinsics  cam{};
    kifi_icp_params  params{};
    IcpWork          work;
    Mat4            *d_w2e = nullptr;   // world->eye of the current frame (binary32)
    double          *d_guess = nullptr;
    bool             have_guess = false;
    kifi_twistnet   *net = nullptr;     // learned initial guess (optional, not owned)
    double           guess[16];
    long long        frame = 0;
};

static const double *prev_pose_dev(const IcpWork *w)
{
    return (const double *)((const char *)w->d_state + offsetof(IcpState, prev));
}

static kifi_status bad_cam(kifi_ctx *ctx, const kifi_intrinsics *cam)
{
    if (!cam || cam->width <= 0 || cam->height <= 0 || cam->width % 4 != 0 || cam->height % 4 != 0 ||
        !(cam->fx > 0.0f) || !(cam->fy > 0.0f))
        return set_error(ctx, KIFI_ERR_ARG, "invalid intrinsics (width and height must be multiples of 4)");
    return KIFI_OK;
}

extern "C" {

kifi_status kifi_icp_params_default(kifi_icp_params *p)
{
    if (!p)
        return KIFI_ERR_ARG;
    p->mode = KIFI_ICP_POINT_TO_PLANE;
    p->levels = 3;
    p->iters[0] = 10;   // 512 x 424
    p->iters[1] = 5;    // 256 x 212
    p->iters[2] = 4;    // 128 x 106
    p->dist_thresh = 0.1f;
    p->eps_rot = 2e-4f;
    p->eps_trans = 2e-4f;
    p->min_inliers = 64;
    p->fallback_rms = 0.0f;
    p->level_gate_growth = 1.0f;
    p->fallback_iters = 0;
    p->naive_sums = 0;
    p->eps_level_growth = 4.0f;   // half the resolution and a quarter of the samples per coarser level
    p->spin_limit = 0;
    for (int l = 0; l < 3; ++l)
        p->first_iters[l] = 0;    // the first pass runs iters
    p->motion_prior = 0.0f;
    return KIFI_OK;
}

kifi_status kifi_model_create(kifi_ctx *ctx, const kifi_intrinsics *cam, kifi_model **out)
{
    if (!ctx || !out)
        return KIFI_ERR_ARG;
    *out = nullptr;
    kifi_status s = bad_cam(ctx, cam);
    if (s != KIFI_OK)
        return s;
    kifi_model *m = new (std::nothrow) kifi_model;
    if (!m)
        return KIFI_ERR_NOMEM;
    m->ctx = ctx;
    m->cam = *cam;
    m->mp.W = cam->width;
    m->mp.H = cam->height;
    m->mp.fx = cam->fx;
    m->mp.fy = cam->fy;
    m->mp.cx = cam->cx;
    m->mp.cy = cam->cy;
    m->mp.znear = 0.1f;
    s = model_alloc(m);
    if (s != KIFI_OK) {
        model_free(m);
        delete m;
        return s;
    }
    *out = m;
    return KIFI_OK;
}

void kifi_model_destroy(kifi_model *m)
{
    if (!m)
        return;
    (void)hipStreamSynchronize(m->ctx->stream);
    model_free(m);
    delete m;
}

kifi_status kifi_model_predict(kifi_model *m, kifi_volume *vol, const float worldToEye[16])
{
    if (!m || !vol || !worldToEye || vol->ctx != m->ctx)
        return KIFI_ERR_ARG;
    m->mp.dmax = vol->consts.vl;   // near-surface band: |D| < one voxel
    return model_predict(m, vol, worldToEye, nullptr);
}

kifi_status kifi_model_download(kifi_model *m, float *points, float *normals)
{
    if (!m)
        return KIFI_ERR_ARG;
    kifi_ctx *ctx = m->ctx;
    const size_t P = (size_t)m->cam.width * m->cam.height;
    if (points)
        KIFI_HIP(ctx, hipMemcpyAsync(points, m->d_pts, P * 16, hipMemcpyDeviceToHost, ctx->stream));
    if (normals)
        KIFI_HIP(ctx, hipMemcpyAsync(normals, m->d_nrm, P * 16, hipMemcpyDeviceToHost, ctx->stream));
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KIFI_OK;
}

// A multi-iteration ICP launch whose workgroups stopped waiting for each other
// left the tracking state incomplete.
static kifi_status state_ok(kifi_ctx *ctx, const IcpState &s)
{
    return s.spin_timeouts ? set_error(ctx, KIFI_ERR_STATE, "ICP: a multi-iteration launch timed out waiting for "
                                                           "its workgroups")
                           : KIFI_OK;
}

static void fill_stats(const IcpState &s, kifi_icp_stats *st)
{
    st->iterations = s.total_iters;
    for (int l = 0; l < 3; ++l)
        st->iters[l] = s.iters[l];
    st->rms = s.rms;
    st->rms_1m = s.rms_1m;
    st->inliers = s.inliers;
    st->failed = s.failed;
    st->fallback = s.fallback;
    st->sum_iters = s.sum_iters;
    st->frames = s.frames_tracked;
    st->fallbacks = s.fallbacks;
    for (int l = 0; l < 3; ++l)
        st->sum_level_iters[l] = s.sum_level_iters[l];
    for (int i = 0; i < 16; ++i)
        st->start_pose[i] = s.guess[i];
}

// Host frame -> one ICP run from `guess`; the final device state in *hs.
static kifi_status align_host(kifi_model *m, const uint16_t *depth_mm, const kifi_intrinsics *cam,
                              const double guess[16], const kifi_icp_params *params, IcpState &hs)
{
    kifi_ctx *ctx = m->ctx;
    kifi_status s = bad_cam(ctx, cam);
    if (s != KIFI_OK)
        return s;
    if (cam->width != m->cam.width || cam->height != m->cam.height)
        return set_error(ctx, KIFI_ERR_ARG, "kifi_icp_align: frame size differs from the model's");
    IcpWork w;
    s = icp_alloc(&w, ctx, cam->width, cam->height);
    if (s != KIFI_OK) {
        icp_free(&w);
        return s;
    }
    const size_t P = (size_t)cam->width * cam->height;
    uint16_t *d_depth = nullptr;
    hipError_t e = hipMalloc(&d_depth, P * 2);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_depth, depth_mm, P * 2, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        icp_free(&w);
        if (d_depth) (void)hipFree(d_depth);
        return hip_error(ctx, e, "kifi_icp_align upload");
    }
    icp_set_pose(&w, guess, nullptr, ctx->stream);   // previous pose for a fallback = the guess
    s = icp_run(m, &w, d_depth, cam, guess, nullptr, params);
    if (s == KIFI_OK) {
        e = hipMemcpyAsync(&hs, w.d_state, sizeof hs, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess)
            s = hip_error(ctx, e, "kifi_icp_align");
        else
            s = state_ok(ctx, hs);
    }
    (void)hipFree(d_depth);
    icp_free(&w);
    return s;
}

kifi_status kifi_icp_align(kifi_model *m, const uint16_t *depth_mm, const kifi_intrinsics *cam,
                           const double guess[16], const kifi_icp_params *params, double out[16],
                           kifi_icp_stats *stats)
{
    if (!m || !depth_mm || !guess || !params || !out)
        return KIFI_ERR_ARG;
    IcpState hs;
    const kifi_status s = align_host(m, depth_mm, cam, guess, params, hs);
    if (s == KIFI_OK) {
        for (int i = 0; i < 16; ++i)
            out[i] = hs.pose[i];
        if (stats)
            fill_stats(hs, stats);
    }
    return s;
}

kifi_status kifi_icp_align_device_noalloc(kifi_model *m, const uint16_t *depth_mm_dev, const kifi_intrinsics *cam,
                                         const double guess[16], const kifi_icp_params *params, double out[16],
                                         kifi_icp_stats *stats)
{
    if (!m || !depth_mm_dev || !guess || !params || !out)
        return KIFI_ERR_ARG;
    kifi_ctx *ctx = m->ctx;
    kifi_status s = bad_cam(ctx, cam);
    if (s != KIFI_OK)
        return s;
    if (cam->width != m->cam.width || cam->height != m->cam.height)
        return set_error(ctx, KIFI_ERR_ARG, "kifi_icp_align_device_noalloc: frame size differs from the model's");
    const size_t P = (size_t)cam->width * cam->height;
    const size_t work = icp_work_bytes(cam->width, cam->height);
    const size_t state_off = (work + 255) & ~size_t(255);
    if (P > kNoAllocMaxPixels || state_off + sizeof(IcpState) > kNoAllocBytes)
        return set_error(ctx, KIFI_ERR_CAPACITY, "kifi_icp_align_device_noalloc: frames above 640 x 480 pixels "
                                                 "do not fit its 2 MiB work array");
    char *base = nullptr;
    s = noalloc_scratch(ctx, &base);
    if (s != KIFI_OK)
        return s;
    hipStream_t st = ctx->stream;
    IcpWork w;
    icp_work_carve(&w, base, cam->width, cam->height);
    w.d_state = (IcpState *)(base + state_off);
    // stage timers create their events on demand: off for this call
    const unsigned mask = ctx->timing_mask;
    ctx->timing_mask = 0;
    static_assert(sizeof(IcpState) % 4 == 0, "whole words");
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(256), 0, st, (uint32_t *)w.d_state, sizeof(IcpState) / 4);
    icp_set_pose(&w, guess, nullptr, st);   // previous pose for a fallback = the guess
    s = icp_run(m, &w, depth_mm_dev, cam, guess, nullptr, params);
    ctx->timing_mask = mask;
    if (s != KIFI_OK)
        return s;
    IcpState hs;
    KIFI_HIP(ctx, hipMemcpyAsync(&hs, w.d_state, sizeof hs, hipMemcpyDeviceToHost, st));
    KIFI_HIP(ctx, hipStreamSynchronize(st));
    s = state_ok(ctx, hs);
    if (s != KIFI_OK)
        return s;
    for (int i = 0; i < 16; ++i)
        out[i] = hs.pose[i];
    if (stats)
        fill_stats(hs, stats);
    return KIFI_OK;
}

kifi_status kifi_icp_normal_equations(kifi_model *m, const uint16_t *depth_mm, const kifi_intrinsics *cam,
                                      const double eyeToWorld[16], int level, float dist_thresh, int naive_sums,
                                      double out[30])
{
    if (!m || !depth_mm || !eyeToWorld || !out)
        return KIFI_ERR_ARG;
    kifi_ctx *ctx = m->ctx;
    kifi_status s = bad_cam(ctx, cam);
    if (s != KIFI_OK)
        return s;
    if (cam->width != m->cam.width || cam->height != m->cam.height)
        return set_error(ctx, KIFI_ERR_ARG, "kifi_icp_normal_equations: frame size differs from the model's");
    if (level < 0 || level > 2 || !(dist_thresh > 0.0f))
        return set_error(ctx, KIFI_ERR_ARG, "kifi_icp_normal_equations: level must be 0..2 and the gate positive");
    // one association on `level`: the schedule runs that level once and no other
    kifi_icp_params ip;
    kifi_icp_params_default(&ip);
    ip.levels = level + 1;
    ip.iters[0] = ip.iters[1] = ip.iters[2] = 0;
    ip.iters[level] = 1;
    ip.dist_thresh = dist_thresh;
    ip.min_inliers = 0;
    ip.naive_sums = naive_sums ? 1 : 0;
    IcpState hs;
    s = align_host(m, depth_mm, cam, eyeToWorld, &ip, hs);
    if (s == KIFI_OK)
        for (int j = 0; j < 30; ++j)
            out[j] = hs.sums[j];
    return s;
}

static kifi_status align_points(kifi_ctx *ctx, const void *depth, bool metres, const kifi_intrinsics *cam,
                                const float guess[16], const float *synth_xyz, int synth_width, int synth_height,
                                const float synthEyeToWorld[16], int iterations, float out[16],
                                kifi_icp_stats *stats)
{
    if (!ctx || !depth || !guess || !synth_xyz || !synthEyeToWorld || !out || iterations < 0)
        return KIFI_ERR_ARG;
    kifi_status s = bad_cam(ctx, cam);
    if (s != KIFI_OK)
        return s;
    if (synth_width <= 0 || synth_height <= 0)
        return set_error(ctx, KIFI_ERR_ARG, "kifi_icp_align_points: empty point buffer");
    // the model is the caller's buffer: its size, projected with the raw camera
    kifi_intrinsics mcam = *cam;
    mcam.width = synth_width;
    mcam.height = synth_height;
    kifi_model m;
    m.ctx = ctx;
    m.cam = mcam;
    m.mp.W = synth_width;
    m.mp.H = synth_height;
    m.mp.fx = cam->fx;
    m.mp.fy = cam->fy;
    m.mp.cx = cam->cx;
    m.mp.cy = cam->cy;
    IcpWork w;
    void *d_depth = nullptr;
    const size_t P = (size_t)cam->width * cam->height;
    const size_t texel = metres ? 4 : 2;
    s = model_alloc(&m);
    if (s == KIFI_OK) {
        float w2e[16];
        for (int i = 0; i < 16; ++i)
            w2e[i] = synthEyeToWorld[i];
        invert_transform(w2e);                       // dstWorldToEye (ICP.cpp:48)
        s = model_set_points(&m, synth_xyz, w2e, cam);
    }
    if (s == KIFI_OK)
        s = icp_alloc(&w, ctx, cam->width, cam->height);
    if (s == KIFI_OK) {
        hipError_t e = hipMalloc(&d_depth, P * texel);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_depth, depth, P * texel, hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess)
            s = hip_error(ctx, e, "kifi_icp_align_points upload");
    }
    IcpState hs;
    if (s == KIFI_OK) {
        // ICP::Align (ICP.cpp:9-31): `iterations` Horn steps, no convergence test
        kifi_icp_params ip;
        kifi_icp_params_default(&ip);
        ip.mode = KIFI_ICP_HORN;
        ip.levels = 1;
        ip.iters[0] = iterations;
        ip.eps_rot = 0.0f;
        ip.eps_trans = 0.0f;
        ip.min_inliers = 1;
        double g[16];
        for (int i = 0; i < 16; ++i)
            g[i] = guess[i];
        icp_set_pose(&w, g, nullptr, ctx->stream);
        DepthIn din;
        din.p = d_depth;
        din.metres = metres;
        s = icp_run(&m, &w, din, cam, g, nullptr, &ip);
        if (s == KIFI_OK) {
            hipError_t e = hipMemcpyAsync(&hs, w.d_state, sizeof hs, hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess)
                e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess)
                s = hip_error(ctx, e, "kifi_icp_align_points");
        }
    }
    if (s == KIFI_OK) {
        for (int i = 0; i < 16; ++i)
            out[i] = (float)hs.pose[i];
        if (stats)
            fill_stats(hs, stats);
    }
    (void)hipStreamSynchronize(ctx->stream);
    if (d_depth)
        (void)hipFree(d_depth);
    icp_free(&w);
    model_free(&m);
    return s;
}

kifi_status kifi_icp_align_points(kifi_ctx *ctx, const uint16_t *depth_mm, const kifi_intrinsics *cam,
                                  const float guess[16], const float *synth_xyz, int synth_width, int synth_height,
                                  const float synthEyeToWorld[16], int iterations, float out[16],
                                  kifi_icp_stats *stats)
{
    return align_points(ctx, depth_mm, false, cam, guess, synth_xyz, synth_width, synth_height, synthEyeToWorld,
                        iterations, out, stats);
}

kifi_status kifi_icp_align_points_f32(kifi_ctx *ctx, const float *depth_m, const kifi_intrinsics *cam,
                                      const float guess[16], const float *synth_xyz, int synth_width,
                                      int synth_height, const float synthEyeToWorld[16], int iterations,
                                      float out[16], kifi_icp_stats *stats)
{
    return align_points(ctx, depth_m, true, cam, guess, synth_xyz, synth_width, synth_height, synthEyeToWorld,
                        iterations, out, stats);
}

kifi_status kifi_horn_align(kifi_ctx *ctx, const float *src, const float *dst, size_t n, float out[16])
{
    if (!ctx || !src || !dst || !out || n < 3 || n >= 0xffffffffull)
        return KIFI_ERR_ARG;
    return horn_align(ctx, src, dst, n, out);
}

kifi_status kifi_pipeline_create(kifi_ctx *ctx, const kifi_volume_desc *vd, const kifi_intrinsics *cam,
                                 const kifi_icp_params *params, kifi_pipeline **out)
{
    if (!ctx || !vd || !out)
        return KIFI_ERR_ARG;
    *out = nullptr;
    kifi_status s = bad_cam(ctx, cam);
    if (s != KIFI_OK)
        return s;
    kifi_pipeline *p = new (std::nothrow) kifi_pipeline;
    if (!p)
        return KIFI_ERR_NOMEM;
    p->ctx = ctx;
    p->cam = *cam;
    if (params)
        p->params = *params;
    else
        kifi_icp_params_default(&p->params);
    s = kifi_volume_create(ctx, vd, &p->vol);
    if (s == KIFI_OK)
        s = kifi_model_create(ctx, cam, &p->model);
    if (s == KIFI_OK)
        s = icp_alloc(&p->work, ctx, cam->width, cam->height);
    if (s == KIFI_OK) {
        hipError_t e = hipMalloc(&p->d_w2e, sizeof(Mat4));
        if (e == hipSuccess)
            e = hipMalloc(&p->d_guess, 16 * sizeof(double));
        if (e != hipSuccess)
            s = hip_error(ctx, e, "kifi_pipeline_create");
    }
    if (s != KIFI_OK) {
        kifi_pipeline_destroy(p);
        return s;
    }
    // the first tracked frame is integrated at the start pose; the reference
    // starts at identity moved 2 m along z (Pipeline.cpp:19-21), kifi_pipeline_set_pose
    // sets any other
    const double I[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
    icp_set_pose(&p->work, I, nullptr, ctx->stream);
    *out = p;
    return KIFI_OK;
}

void kifi_pipeline_destroy(kifi_pipeline *p)
{
    if (!p)
        return;
    (void)hipStreamSynchronize(p->ctx->stream);
    if (p->vol) kifi_volume_destroy(p->vol);
    if (p->model) kifi_model_destroy(p->model);
    icp_free(&p->work);
    if (p->d_w2e) (void)hipFree(p->d_w2e);
    if (p->d_guess) (void)hipFree(p->d_guess);
    delete p;
}

kifi_volume *kifi_pipeline_volume(kifi_pipeline *p) { return p ? p->vol : nullptr; }
kifi_model *kifi_pipeline_model(kifi_pipeline *p) { return p ? p->model : nullptr; }

kifi_status kifi_pipeline_set_guess(kifi_pipeline *p, const double e2w[16])
{
    if (!p || !e2w)
        return KIFI_ERR_ARG;
    std::memcpy(p->guess, e2w, sizeof p->guess);
    p->have_guess = true;
    return KIFI_OK;
}

kifi_status kifi_pipeline_reset(kifi_pipeline *p)
{
    if (!p)
        return KIFI_ERR_ARG;
    kifi_status s = kifi_volume_clear(p->vol);   // synchronises the stream first
    if (s != KIFI_OK)
        return s;
    const double I[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
    icp_set_pose(&p->work, I, nullptr, p->ctx->stream);
    KIFI_HIP(p->ctx, hipGetLastError());
    p->frame = 0;
    p->have_guess = false;
    if (p->net)
        twistnet_reset_frames(p->net);
    return KIFI_OK;
}

kifi_status kifi_pipeline_set_twistnet(kifi_pipeline *p, kifi_twistnet *net)
{
    if (!p)
        return KIFI_ERR_ARG;
    if (net)
        twistnet_reset_frames(net);
    p->net = net;
    return KIFI_OK;
}

kifi_status kifi_pipeline_set_pose(kifi_pipeline *p, const double e2w[16])
{
    if (!p || !e2w)
        return KIFI_ERR_ARG;
    icp_set_pose(&p->work, e2w, nullptr, p->ctx->stream);
    KIFI_HIP(p->ctx, hipGetLastError());
    return KIFI_OK;
}

// One frame of the loop; depth is uint16 mm or binary32 metres (a metre frame
// does not feed the learned initial pose, which reads mm).
static kifi_status pipeline_integrate(kifi_pipeline *p, const void *depth, bool metres, int on_device,
                                      const float w2e[16], const float *fuse_w2e = nullptr)
{
    if (!p || !depth)
        return KIFI_ERR_ARG;
    kifi_ctx *ctx = p->ctx;
    hipStream_t st = ctx->stream;
    const size_t P = (size_t)p->cam.width * p->cam.height;
    DepthIn din;
    din.p = depth;
    din.metres = metres;
    int slot = -1;            // a host frame goes through the context's copy stream
    if (!on_device) {
        kifi_status s0 = stage_frame(ctx, depth, P * (metres ? 4 : 2), &din.p, &slot);
        if (s0 != KIFI_OK)
            return s0;
    }
    struct Release {          // the staging buffer is free once this frame's kernels are
        kifi_ctx *ctx;
        int slot;
        ~Release()
        {
            if (slot >= 0)
                stage_done(ctx, slot);
        }
    } release{ ctx, slot };
    const uint16_t *d_depth = (const uint16_t *)din.p;
    kifi_twistnet *net = metres ? nullptr : p->net;
    // the network predicts the motion between consecutive mm frames: a metre
    // frame breaks that sequence, so the next mm frame starts a new one
    if (metres && p->net)
        twistnet_reset_frames(p->net);
    kifi_status s = KIFI_OK;
    bool tracked = false;     // icp_run wrote this frame's worldToEye
    bool defer_end = false;   // ... or the integration's keygen writes it, ending the tracking
    if (w2e) {
        // Pipeline.cpp:32-35: integrate at the given pose
        icp_set_pose(&p->work, nullptr, w2e, st);
        if (net)
            twistnet_observe(net, d_depth, p->cam.width, p->cam.height, st);
    } else if (p->frame > 0) {
        // Pipeline.cpp:37-39: track against the model of the previous frame,
        // from a caller's guess, the network's prediction, or the previous pose
        kifi_icp_params ip = p->params;
        const double *guess_dev = prev_pose_dev(&p->work);
        bool guessed = p->have_guess;
        if (!guessed && net &&
            twistnet_guess(net, d_depth, p->cam.width, p->cam.height, prev_pose_dev(&p->work), p->d_guess, st)) {
            guess_dev = p->d_guess;
            guessed = true;
        }
        if (!guessed) {
            ip.fallback_rms = 0.0f;   // the guess is the previous pose: nothing to fall back to
            ip.fallback_iters = 0;
        }
        // a frame started from the caller's guess still becomes the network's
        // previous frame: the next prediction is then the motion from this one
        if (p->have_guess && net)
            twistnet_observe(net, d_depth, p->cam.width, p->cam.height, st);
        // a frame fused at its tracked pose ends its tracking in the
        // integration's first kernel (k_keygen's frame start)
        defer_end = !fuse_w2e;
        s = icp_run(p->model, &p->work, din, &p->cam, p->have_guess ? p->guess : nullptr,
                    p->have_guess ? nullptr : guess_dev, &ip, p->d_w2e, defer_end);
        tracked = true;
        p->have_guess = false;
        if (s != KIFI_OK) {
            if (defer_end)
                icp_frame_end(&p->work, p->d_w2e, st);
            return s;
        }
    } else if (net) {
        twistnet_observe(net, d_depth, p->cam.width, p->cam.height, st);
    }
    // Pipeline.cpp:41-46: worldToEye = invert(eyeToWorld), integrate; the
    // model for the next frame (Pipeline.cpp:48-52, seen from this pose) is
    // splatted inside the integration's voxel pass, then resolved
    if (w2e) {
        store_w2e(w2e, p->d_w2e, st);         // the given worldToEye as is (Pipeline.cpp:35)
    } else if (fuse_w2e) {
        // fused at the caller's pose, which the next frame is tracked from; the
        // tracked pose stays this frame's result
        store_w2e(fuse_w2e, p->d_w2e, st);
        if (tracked)
            icp_set_prev(&p->work, fuse_w2e, st);
        else
            icp_set_pose(&p->work, nullptr, fuse_w2e, st);
    } else if (!tracked) {
        icp_w2e(&p->work, p->d_w2e, st);
    }
    // the model pose and the z-buffer clear ride on the integration's first
    // kernels (its camera is the model's)
    model_prepare(p->model, p->vol);
    SplatParams sp = model_splat_params(p->model, p->d_w2e);
    sp.pose_out = p->model->d_pose;
    sp.icp_end = defer_end ? p->work.d_state : nullptr;
    s = integrate_dispatch(p->vol, din, &p->cam, nullptr, p->d_w2e, &sp);
    if (s != KIFI_OK) {
        if (defer_end)   // the integration stopped before its first kernel: end the tracking here
            icp_frame_end(&p->work, p->d_w2e, st);
        return s;
    }
    const int t_m = timer_begin(ctx, KIFI_STAGE_MODEL);
    model_end(p->model, p->vol, false);
    timer_end(ctx, KIFI_STAGE_MODEL, t_m);
    KIFI_HIP(ctx, hipGetLastError());
    ++p->frame;
    return KIFI_OK;
}

kifi_status kifi_pipeline_integrate(kifi_pipeline *p, const uint16_t *depth, int on_device, const float w2e[16])
{
    return pipeline_integrate(p, depth, false, on_device, w2e);
}

kifi_status kifi_pipeline_integrate_f32(kifi_pipeline *p, const float *depth_m, int on_device, const float w2e[16])
{
    return pipeline_integrate(p, depth_m, true, on_device, w2e);
}

kifi_status kifi_pipeline_integrate_at(kifi_pipeline *p, const uint16_t *depth, int on_device,
                                       const float fuse_w2e[16])
{
    if (!fuse_w2e)
        return KIFI_ERR_ARG;
    return pipeline_integrate(p, depth, false, on_device, nullptr, fuse_w2e);
}

kifi_status kifi_pipeline_pose(kifi_pipeline *p, double e2w[16])
{
    if (!p || !e2w)
        return KIFI_ERR_ARG;
    IcpState hs;
    KIFI_HIP(p->ctx, hipMemcpyAsync(&hs, p->work.d_state, sizeof hs, hipMemcpyDeviceToHost, p->ctx->stream));
    KIFI_HIP(p->ctx, hipStreamSynchronize(p->ctx->stream));
    for (int i = 0; i < 16; ++i)
        e2w[i] = hs.pose[i];
    return state_ok(p->ctx, hs);
}

kifi_status kifi_pipeline_stats(kifi_pipeline *p, kifi_icp_stats *stats)
{
    if (!p || !stats)
        return KIFI_ERR_ARG;
    IcpState hs;
    KIFI_HIP(p->ctx, hipMemcpyAsync(&hs, p->work.d_state, sizeof hs, hipMemcpyDeviceToHost, p->ctx->stream));
    KIFI_HIP(p->ctx, hipStreamSynchronize(p->ctx->stream));
    fill_stats(hs, stats);
    return state_ok(p->ctx, hs);
}

} // extern "C"
