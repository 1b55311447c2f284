// kifi_internal.hpp -- shared device helpers and host-side state of libkifi_hip.so.
//
// CDNA4 (gfx950) only: 64-lane wavefronts, 64-bit ballots, LDS per workgroup.
// Everything here is compiled with -ffp-contract=off so that binary32 results
// match the oracle's source-order arithmetic bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "kifi.h"

namespace kifi {

constexpr int kWave = 64;

// ---------------------------------------------------------------- wave64 ----
__device__ __forceinline__ uint64_t lanemask_lt()
{
    const uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ uint32_t popc64(uint64_t m) { return (uint32_t)__popcll(m); }

// Lanes of the wave (among `active`) whose `bits` low bits of `d` equal this lane's.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, int bits, uint64_t active)
{
    uint64_t m = active;
    for (int b = 0; b < bits; ++b) {
        const uint64_t on = __ballot((d >> b) & 1u);
        m &= ((d >> b) & 1u) ? on : ~on;
    }
    return m;
}

// ------------------------------------------------------------------ keys ----
template <typename K> struct KeyTraits;
template <> struct KeyTraits<uint32_t> {
    static constexpr int kb = 10;       // bits per axis (util/math.h:442-457)
    static constexpr int sort_bits = 32;
    static constexpr int digit_bits = 8;    // radix digit: 30 used bits in 4 passes (10-bit digits, 3
                                            // passes, measured slower: the sort is latency-bound)
};
template <> struct KeyTraits<uint64_t> {
    static constexpr int kb = 21;
    static constexpr int sort_bits = 64;
    static constexpr int digit_bits = 9;    // 2 kb + log2(res) bits: 6 passes up to res 2^12
};

template <typename K>
__host__ __device__ __forceinline__ K pack_key(K x, K y, K z)
{
    constexpr int kb = KeyTraits<K>::kb;
    return z << (2 * kb) | y << kb | x;
}

template <typename K>
__host__ __device__ __forceinline__ void unpack_key(K k, uint32_t &x, uint32_t &y, uint32_t &z)
{
    constexpr int kb = KeyTraits<K>::kb;
    constexpr K m = (K(1) << kb) - 1;
    x = (uint32_t)(k & m);
    y = (uint32_t)(k >> kb & m);
    z = (uint32_t)(k >> (2 * kb));
}

// ------------------------------------------------------------ parameters ----
// Host-precomputed binary32 constants, produced in the reference's order
// (Integrator.cpp:41-42, 94-105; Volume.cpp:28-36) and passed by value.
struct VolumeConsts {
    int   res;
    float vl;             // voxel length
    float vlHalfPlusMin;  // vl * 0.5 + min
    float vlInv;          // 1 / vl
    float minOverNegVl;   // vlInv * -min
    float trunc;
};

struct Mat4 {
    float m[16];          // column storage
};

// (col0*x + col1*y) + (col2*z + col3*w), util/math.h:362-365, 381-397
__host__ __device__ __forceinline__ void mat_vec(const float *m, const float v[4], float out[4])
{
    float r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        r[i] = (m[i] * v[0] + m[i + 4] * v[1]) + (m[i + 8] * v[2] + m[i + 12] * v[3]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        out[i] = r[i];
}

__host__ __device__ __forceinline__ void mat_vec(const Mat4 &M, const float v[4], float out[4])
{
    mat_vec(M.m, v, out);
}

// float4x4 * float4x4, util/math.h:367-377
__host__ __device__ __forceinline__ void mat_mul(const float *m, const float *n, float *out)
{
    float r[16];
#pragma unroll
    for (int c = 0; c < 4; ++c)
        mat_vec(m, n + 4 * c, r + 4 * c);
#pragma unroll
    for (int i = 0; i < 16; ++i)
        out[i] = r[i];
}

// util/math.h:615-625
__host__ __device__ __forceinline__ void invert_transform(float *m)
{
    float t[4] = { -m[12], -m[13], -m[14], 1.0f };
    m[12] = 0.0f; m[13] = 0.0f; m[14] = 0.0f; m[15] = 1.0f;
    float s;
    s = m[1]; m[1] = m[4]; m[4] = s;
    s = m[2]; m[2] = m[8]; m[8] = s;
    s = m[6]; m[6] = m[9]; m[9] = s;
    float c3[4];
    mat_vec(m, t, c3);
    m[12] = c3[0]; m[13] = c3[1]; m[14] = c3[2]; m[15] = c3[3];
}

// DepthSensorParams::EyeToClipRH, DepthSensorParams.cpp:47-63
__host__ __device__ __forceinline__ void eye_to_clip_rh(int Wi, int Hi, float fx, float fy, float cx, float cy,
                                                         float zn, float zf, float *out)
{
    for (int i = 0; i < 16; ++i)
        out[i] = 0.0f;
    const float W = (float)Wi, H = (float)Hi;
    out[0] = 2.0f * fx / W;
    out[8] = W - 2.0f * cx;
    out[5] = 2.0f * fy / H;
    out[9] = H - 2.0f * cy;
    out[10] = -(zf + zn) / (zf - zn);
    out[14] = -2.0f * zf * zn / (zf - zn);
    out[11] = -1.0f;
}

// Per-frame matrices of an integration, computed on the device from worldToEye
// in the reference's order (Integrator.cpp:41-42).
// Clip rows x, y and w of a voxel centre as affine functions of its integer
// coordinates (w2c * (vl * (x, y, z) + c0)), evaluated with fused multiply-adds,
// and a bound `eps` that exceeds twice their distance from the reference-order
// binary32 values: a cheap conservative test of the update's cone.
struct ClipPre {
    float x[4], y[4], w[4];
    float eps;
};

struct FrameMats {
    Mat4 w2e;   // worldToEye as given
    Mat4 w2c;   // EyeToClipRH() * worldToEye
    Mat4 e2w;   // invert_transform(worldToEye)
    ClipPre pre;
};

// A depth frame on the device: uint16 millimetres (converted as DepthStream.cpp:89-90,
// mm * 0.001f) or binary32 metres (DepthStream's FLOAT texels, used as they are).
struct DepthIn {
    const void *p = nullptr;
    bool metres = false;
};

__host__ __device__ __forceinline__ float depth_m(const uint16_t *d, size_t i) { return (float)d[i] * 0.001f; }
__host__ __device__ __forceinline__ float depth_m(const float *d, size_t i) { return d[i]; }

// world->eye of the model camera and its centre in world coordinates
struct ModelPose {
    Mat4  w2e;
    Mat4  proj;     // EyeToClipRH() * w2e (the reference's association projection)
    float cam[3];
};

// Optional model splat folded into the merge + update voxel pass.
struct IcpState;   // the tracker's device state (kifi_icp.hpp)

struct SplatParams {
    const Mat4 *w2e;                 // device: world->eye of this frame
    unsigned long long *zbuf;        // W*H, cleared to ~0
    int   W, H;
    float fx, fy, cx, cy, vl, c0, dmax, znear;
    // When set, the integration also starts the prediction: k_keygen writes
    // the model pose of this frame (the camera is the integration's) and
    // clears zbuf.  Null when the caller has done both.
    ModelPose *pose_out;
    // When set, the integration's k_keygen ends the frame's tracking (the
    // tracker's state, its worldToEye into w2e) and takes the frame's
    // worldToEye from the tracker's pose.
    IcpState *icp_end;
};

// Reset list for per-frame device counters (zeroed by a fill kernel, never by a
// memset node, so that the integration sequence can be captured in a graph).
struct FrameCounters {
    uint32_t n_cand;      // candidates written by keygen (up to 8 per valid pixel)
    uint32_t n_keys;      // valid pixels (K)
    uint32_t n_unique;    // unique candidates
    uint32_t n_new;       // E = n_unique - 1 (the reference's x-pass drops the max)
    uint64_t v_old;       // V0
    uint64_t v_new;       // V1
    uint32_t n_tiles;     // merge tiles this frame
    uint32_t overflow;    // set when a device-side capacity check fails
    uint32_t n_kept;      // new voxels: V1 - V0
    // running totals over all integrations (never reset by a frame)
    uint64_t tot_frames, tot_v_old, tot_new, tot_v_new, tot_keys;
};

// Running count of the voxels whose TSDF an update changed (A), spread over
// kUpdSlots counters kUpdStride words apart (one 128-byte line each): merge
// tile t adds to counter t % kUpdSlots, so the tiles' atomics do not queue on
// one address.  The total is the counters' sum.
constexpr int kUpdSlots = 64;
constexpr int kUpdStride = 16;

} // namespace kifi

// ------------------------------------------------------------ host state ----
struct kifi_ctx {
    int          device = 0;
    hipStream_t  stream = nullptr;
    std::string  err;
    // reusable device scratch, grown on demand
    void        *scratch = nullptr;
    size_t       scratch_bytes = 0;
    void        *out_scratch = nullptr;  // output staging (mesh, model maps)
    size_t       out_bytes = 0;
    // asynchronous mesh delivery: the vertex copy out of out_scratch runs on
    // out_stream after out_ready (recorded on stream); out_copied marks its end
    hipStream_t  out_stream = nullptr;
    hipEvent_t   out_ready = nullptr;
    hipEvent_t   out_copied = nullptr;
    bool         out_pending = false;   // a copy out of out_scratch may still run
    // host frames: uploaded on a copy stream into two staging buffers in turn,
    // so the upload of frame t+1 overlaps the kernels of frame t
    hipStream_t  copy_stream = nullptr;
    void        *stage[2] = { nullptr, nullptr };
    size_t       stage_bytes = 0;
    hipEvent_t   stage_copied[2] = { nullptr, nullptr };   // on copy_stream: the upload is done
    hipEvent_t   stage_used[2] = { nullptr, nullptr };     // on stream: the frame's kernels are done
    bool         stage_busy[2] = { false, false };
    int          stage_slot = 0;
    int          num_cus = 0;
    int          multi_ok = -1;    // multi-iteration ICP launch fits one workgroup per CU (-1: not yet asked)
    int          tail_ok = -1;     // the same for the level-tail launch
    // stage timers (HIP event pairs recorded on `stream` while enabled)
    unsigned     timing_mask = 0;   // bit s: time stage s
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timers[KIFI_STAGE_COUNT];
    std::vector<hipEvent_t> event_pool;
};

namespace kifi {
// Uploads a host frame of `bytes` into the next staging buffer on the copy stream
// and makes the context's stream wait for it; returns once the copy has read the
// host buffer.  *dev receives the device copy.
// stage_done(ctx, slot) after the frame's last kernel releases the buffer.
kifi_status stage_frame(kifi_ctx *ctx, const void *host, size_t bytes, const void **dev, int *slot);
void        stage_done(kifi_ctx *ctx, int slot);
// Records a start event for `stage` when timing is enabled; returns the slot.
int  timer_begin(kifi_ctx *ctx, int stage);
void timer_end(kifi_ctx *ctx, int stage, int slot);
} // namespace kifi

struct kifi_volume {
    kifi_ctx        *ctx = nullptr;
    kifi_volume_desc desc{};
    kifi::VolumeConsts consts{};
    size_t           key_bytes = 4, val_bytes = 8;
    void            *keys[2] = { nullptr, nullptr };
    void            *vals[2] = { nullptr, nullptr };
    // every 64th and every 4096th key of each key buffer (the integration's
    // searches); valid for the current buffer unless the host wrote the volume
    void            *samp1[2] = { nullptr, nullptr };
    void            *samp2[2] = { nullptr, nullptr };
    bool             samples_valid = false;
    int              cur = 0;         // which buffer pair holds the volume
    size_t           cap = 0;         // capacity of each buffer (voxels)
    kifi::FrameCounters *d_counters = nullptr;   // device
    uint64_t        *d_upd = nullptr;            // device: kUpdSlots x kUpdStride, A's running counters
    uint64_t        *d_size = nullptr;           // device-resident volume size
    kifi::FrameMats *d_mats = nullptr;           // matrices of the last integration
    // Size read-back without stalling the stream: each integration's last
    // bookkeeping kernel copies the frame's counters into ring slot
    // (frame % kSizeRing) of pinned, device-mapped host memory, and an event
    // marks when the slot may be read.  The host bounds the device-side size by
    // the newest size it has read plus 8P per frame enqueued after that one.  An
    // integration reusing a slot first waits for the frame that last used it,
    // so at most kSizeRing frames are in flight.
    static constexpr int kSizeRing = 4;
    kifi::FrameCounters *h_ring = nullptr;       // kSizeRing slots
    kifi::FrameCounters *h_ring_dev = nullptr;   // their device address
    hipEvent_t       ring_ev[kSizeRing] = {};
    uint64_t         ring_frame[kSizeRing] = {}; // frame number written into the slot; 0: nothing pending
    uint64_t         frames_enq = 0;             // integrations enqueued (frames numbered from 1)
    uint64_t         known_frame = 0;            // the newest frame whose resulting size the host knows
    size_t           known_size = 0;             // the volume size after known_frame
    bool             ring_overflow = false;      // a completed frame reported a capacity overflow
};

namespace kifi {

// error helpers
kifi_status set_error(kifi_ctx *ctx, kifi_status st, const std::string &msg);
kifi_status hip_error(kifi_ctx *ctx, hipError_t e, const char *what);

#define KIFI_HIP(ctx, call)                                          \
    do {                                                             \
        hipError_t e_ = (call);                                      \
        if (e_ != hipSuccess)                                        \
            return ::kifi::hip_error((ctx), e_, #call);              \
    } while (0)

// Takes the counters of every completed frame in the size ring, oldest first.
void ring_refresh(kifi_volume *v);
// After a stream synchronisation: every enqueued frame is done and the size is n.
void ring_settle(kifi_volume *v, size_t n);

// scratch allocators (grow; contents not preserved across a growth)
kifi_status ensure_scratch(kifi_ctx *ctx, size_t bytes);
kifi_status ensure_out(kifi_ctx *ctx, size_t bytes);
// Waits for a pending asynchronous copy out of out_scratch (no-op when none).
kifi_status out_settle(kifi_ctx *ctx);

} // namespace kifi
