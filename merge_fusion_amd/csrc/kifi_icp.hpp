// kifi_icp.hpp -- device state shared by the model-prediction / ICP kernels and
// the pipeline host code.
#pragma once

#include "kifi_internal.hpp"

namespace kifi {

struct Pose64 {
    double m[16];   // eye->world, column storage
};

struct ModelParams {
    int   W, H;
    float fx, fy, cx, cy;   // model camera intrinsics (full resolution)
    float vl, c0;           // voxel length, VoxelCenter offset
    float dmax;             // near-surface band |D| < dmax
    float znear;
};

struct IcpLevel {
    int   W, H;
    float fx, fy, cx, cy;
    const float *depth;     // metres
};

struct IcpSolveParams {
    double eps_rot, eps_trans;
    int    min_inliers;
    double prior;           // motion prior weight per inlier (0: off)
};

// Device-resident tracking state of one frame (plus running totals).
struct IcpState {
    double pose[16];        // current eye->world estimate (FP64)
    double guess[16];       // the initial guess of this frame
    double prev[16];        // the previous frame's pose (fallback start)
    Mat4   pose_f;          // binary32 copy used by the association kernels
    int    iters[4];
    int    total_iters;
    int    level_done[4];   // per pyramid level: converged, failed or not run
    int    failed;
    int    fallback;        // this frame re-ran plain ICP
    float  rms;
    float  rms_1m;          // RMS of r / d^2 x 1 m^2 (noise-normalised residual)
    int    inliers;
    // first pass of a frame that fell back (kept if it beats the second)
    double alt_pose[16];
    float  alt_rms;
    int    alt_inliers;
    unsigned ticket;        // workgroups done in the current association launch
    unsigned gen;           // iterations solved by multi-iteration launches (monotone)
    unsigned spin_timeouts; // waits of a multi-iteration launch that gave up (0 unless broken)
    unsigned tail_gen[6];   // hand-offs of each level-tail launch of the frame (pass x 3 + level)
    double sums[30];        // reduced sums of the last association (icp_finish layout)
    double prior_acc[6];    // motion since the pass's start: rotation, camera-centre displacement
    double prior_c[3];      // camera centre of the pass's start pose
    // running totals over frames
    long long sum_iters;
    long long frames_tracked;
    long long fallbacks;
    long long sum_level_iters[3];   // solves per pyramid level
};

// End of a tracked frame: the pose becomes the previous pose, the running
// totals advance, and (w2e set) the frame's worldToEye is written.  One thread.
__device__ __forceinline__ void icp_frame_end_dev(IcpState *st, Mat4 *w2e)
{
    for (int i = 0; i < 16; ++i)
        st->prev[i] = st->pose[i];
    st->sum_iters += st->total_iters;
    st->frames_tracked += 1;
    st->fallbacks += st->fallback;
    for (int l = 0; l < 3; ++l)
        st->sum_level_iters[l] += st->iters[l];
    if (w2e) {
        Mat4 m = st->pose_f;
        invert_transform(m.m);
        *w2e = m;
    }
}

struct IcpWork {
    void    *buf = nullptr;
    float   *lv[3] = { nullptr, nullptr, nullptr };
    void    *partials = nullptr;
    IcpState *d_state = nullptr;
    int      W = 0, H = 0;
};

} // namespace kifi

struct kifi_model {
    kifi_ctx           *ctx = nullptr;
    kifi_intrinsics     cam{};
    kifi::ModelParams   mp{};
    float               disc = 0.15f;    // depth-discontinuity gate for normals (m)
    float4             *d_pts = nullptr; // world points, w = valid
    float4             *d_nrm = nullptr; // unit normals, w = valid
    unsigned long long *d_zbuf = nullptr;
    kifi::ModelPose    *d_pose = nullptr;
};

namespace kifi {
kifi_status model_alloc(kifi_model *m);
void        model_free(kifi_model *m);
kifi_status model_predict(kifi_model *m, kifi_volume *vol, const float *w2e_host, const Mat4 *w2e_dev);
void        model_begin(kifi_model *m, kifi_volume *vol, const float *w2e_host, const Mat4 *w2e_dev);
// The host side of model_begin only: an integration given model_splat_params
// with pose_out set then writes the pose and clears the z-buffer itself.
void        model_prepare(kifi_model *m, kifi_volume *vol);
SplatParams model_splat_params(kifi_model *m, const Mat4 *w2e_dev);
void        model_end(kifi_model *m, kifi_volume *vol, bool splat);
kifi_status model_set_points(kifi_model *m, const float *xyz, const float *w2e, const kifi_intrinsics *proj_cam);
kifi_status model_from_depth(kifi_model *m, const uint16_t *d_mm, const float e2w[16], const float w2e[16]);
kifi_status icp_alloc(IcpWork *w, kifi_ctx *ctx, int W, int H);
size_t      icp_work_bytes(int W, int H);                        // work buffer bytes (256-aligned parts)
void        icp_work_carve(IcpWork *w, void *buf, int W, int H); // the levels and partials in buf
void        icp_free(IcpWork *w);
// defer_end: the frame's end (icp_frame_end_dev) is left to the caller's next
// kernel (the integration's k_keygen, workgroup 0), one launch fewer per tracked frame.
kifi_status icp_run(kifi_model *m, IcpWork *w, DepthIn depth, const kifi_intrinsics *cam,
                    const double *guess_host, const double *guess_dev, const kifi_icp_params *ip,
                    Mat4 *w2e_out = nullptr, bool defer_end = false);
void        icp_frame_end(IcpWork *w, Mat4 *w2e_out, hipStream_t st);
inline kifi_status icp_run(kifi_model *m, IcpWork *w, const uint16_t *d_depth_mm, const kifi_intrinsics *cam,
                           const double *guess_host, const double *guess_dev, const kifi_icp_params *ip,
                           Mat4 *w2e_out = nullptr)
{
    DepthIn d;
    d.p = d_depth_mm;
    return icp_run(m, w, d, cam, guess_host, guess_dev, ip, w2e_out);
}
void        icp_w2e(IcpWork *w, Mat4 *d_out, hipStream_t st);
void        store_w2e(const float *w2e, Mat4 *d_out, hipStream_t st);
kifi_status horn_align(kifi_ctx *ctx, const float *src, const float *dst, size_t n, float out[16]);
} // namespace kifi
