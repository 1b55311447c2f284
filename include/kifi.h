/*
 * kifi.h -- C ABI of the MI355X fusion library (libkifi_hip.so).
 *
 * The reference (KinectFission/kifi) exposes a C++ class API with no FFI; this
 * ABI carries exactly what those entry points take and return, as plain
 * pointers, sizes and PODs, so that the C++ drop-in classes in
 * merge_fusion_amd/cpp/kifi/ (and ctypes/cgo-style bindings) sit on top of it.
 *
 * Conventions shared with the reference:
 *   - matrices are float[16] in column storage, element (r, c) at m[r + 4c]
 *     (util/math.h:59-80, 259-265);
 *   - depth frames are row-major, pixel (x, y) at y*width + x (util/vector2d.h:88-99);
 *     they are passed as uint16 millimetres and converted on the device exactly as
 *     DepthStream.cpp:89-90 does ((float)mm * 0.001f);
 *   - 32-bit voxel keys are z<<20 | y<<10 | x (util/math.h:442-457); 64-bit keys
 *     use 21 bits per axis (z<<42 | y<<21 | x) for resolutions above 1024.
 *
 * Errors: every call returns a kifi_status; kifi_last_error() gives the text.
 * No exception crosses the ABI.  One context per host thread; calls on one
 * context are ordered on its HIP stream, and a call that returns host-visible
 * data is synchronous.  The library owns device memory; callers own host buffers.
 */
#ifndef KIFI_H
#define KIFI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    KIFI_OK = 0,
    KIFI_ERR_ARG = 1,      /* invalid argument (the reference asserts)           */
    KIFI_ERR_HIP = 2,      /* HIP runtime error                                  */
    KIFI_ERR_NOMEM = 3,    /* device allocation failed                           */
    KIFI_ERR_CAPACITY = 4, /* caller buffer too small; *count gives the size     */
    KIFI_ERR_STATE = 5     /* invalid state (e.g. device-side overflow detected) */
} kifi_status;

typedef struct kifi_ctx kifi_ctx;
typedef struct kifi_volume kifi_volume;

/* DepthSensorParams (DepthSensorParams.h:23-45): resolution, focal length and
 * principal point in pixels, sensible range in metres. */
typedef struct {
    int   width, height;
    float fx, fy;
    float cx, cy;
    float znear, zfar;
} kifi_intrinsics;

enum { KIFI_STORAGE_F32 = 0, KIFI_STORAGE_F16U8 = 1 };

/* Volume(resolution, sideLength, truncationMargin) (Volume.h:15-31) plus the
 * MI355X layout choices: key width and voxel storage. */
typedef struct {
    int   res;        /* power of two; <= 1024 for 32-bit keys, <= 2^20 for 64-bit */
    float side;       /* metres */
    float trunc;      /* metres */
    int   key_bits;   /* 32 or 64 */
    int   storage;    /* KIFI_STORAGE_F32: {float sum, float weight} (Voxel.h:18-20);
                         KIFI_STORAGE_F16U8: uint32 = binary16 distance | weight << 16 */
} kifi_volume_desc;

/* ---- context ---- */
kifi_status kifi_ctx_create(int device, kifi_ctx **out);
void        kifi_ctx_destroy(kifi_ctx *ctx);
const char *kifi_last_error(const kifi_ctx *ctx);
kifi_status kifi_ctx_synchronize(kifi_ctx *ctx);
/* The HIP stream the context orders its work on (hipStream_t). */
void       *kifi_ctx_stream(kifi_ctx *ctx);

/* ---- volume: kifi::Volume (Volume.h:15-31, Volume.cpp:12-37) ---- */
kifi_status kifi_volume_create(kifi_ctx *ctx, const kifi_volume_desc *desc, kifi_volume **out);
void        kifi_volume_destroy(kifi_volume *vol);
kifi_status kifi_volume_clear(kifi_volume *vol);                  /* flat_map::clear (flat_map.h:158-162) */
kifi_status kifi_volume_size(kifi_volume *vol, size_t *n);        /* flat_map::size (flat_map.h:190-194) */
/* Volume::Data() read-out (keys_cbegin / values_cbegin, flat_map.h:86-121):
 * keys as uint32 or uint64 per key_bits, values as 8 or 4 bytes per storage.
 * Either pointer may be NULL.  Returns KIFI_ERR_CAPACITY if cap < size. */
kifi_status kifi_volume_download(kifi_volume *vol, void *keys, void *values, size_t cap, size_t *n);
/* Replaces the content with n sorted unique keys and their values. */
kifi_status kifi_volume_upload(kifi_volume *vol, const void *keys, const void *values, size_t n);
/* Checkpoint / resume: the volume's description, keys and values to a file, and
 * a new volume from one (keys checked to be strictly increasing).  No
 * counterpart in the reference; integration continues from a loaded volume as
 * from the saved one.  Synchronous. */
kifi_status kifi_volume_save(kifi_volume *vol, const char *path);
kifi_status kifi_volume_load(kifi_ctx *ctx, const char *path, kifi_volume **out);

/* ---- integration: kifi::Integrator::Integrate (Integrator.h:19-26, Integrator.cpp:26-71) ----
 * depth_mm: host pointer (kifi_integrate) or device pointer already in HBM
 * (kifi_integrate_device).  worldToEye: column storage.  Asynchronous on the
 * device: a host frame is copied into one of two device staging buffers on a
 * second stream, and the call returns once that copy has read the host buffer,
 * so the caller may refill or free it at once (the reference's Integrate takes
 * the frame by const reference and returns when done with it).  The copy of a
 * frame waits for the kernels of the frame two before it (the same staging
 * buffer), so the host runs at most two frames ahead of the device.  Device
 * frames: a call waits for the integration four before it on the same volume
 * to finish, so the host runs at most four frames ahead; the volume's
 * capacity keeps room for that many frames of growth (8 voxels per pixel
 * each), so integration does not stall the stream to check it. */
kifi_status kifi_integrate(kifi_volume *vol, const uint16_t *depth_mm, const kifi_intrinsics *cam,
                           const float worldToEye[16]);
kifi_status kifi_integrate_device(kifi_volume *vol, const uint16_t *depth_mm_dev,
                                  const kifi_intrinsics *cam, const float worldToEye[16]);
/* The same with depth in binary32 metres, used as it is: the Integrate of a
 * vector2d<float> frame (Integrator.h:19-26), e.g. DepthStream's FLOAT texels
 * (DepthStream.cpp:78-80).  A uint16 frame is the special case mm * 0.001f. */
kifi_status kifi_integrate_f32(kifi_volume *vol, const float *depth_m, const kifi_intrinsics *cam,
                               const float worldToEye[16]);
kifi_status kifi_integrate_device_f32(kifi_volume *vol, const float *depth_m_dev,
                                      const kifi_intrinsics *cam, const float worldToEye[16]);
/* Counts of the last integration: [0] K keygen keys, [1] 8-neighbourhood candidates
 * sorted (at most 8K: a half block equal to the previous pixel's is not written),
 * [2] E keys handed to the merge, [3] volume size after the merge.  Synchronous. */
kifi_status kifi_integrate_stats(kifi_volume *vol, size_t stats[4]);
/* Running totals since the volume was created: [0] integrations, [1] sum of V0,
 * [2] sum of E, [3] sum of V1, [4] sum of K, [5] sum of A (voxels whose TSDF the
 * update changed).  Synchronous. */
kifi_status kifi_integrate_totals(kifi_volume *vol, uint64_t totals[6]);

/* ---- meshing: kifi::Mesher (Mesher.h:17-18, Mesher.cpp:14-45) ----
 * kifi_mesh: Mesh(volume, verts) -- float3 vertices in generation order.
 * kifi_mesh_indexed: Mesh(volume, verts, idx) -- vertices ordered by vertex ID,
 * indices (3 per triangle) into them.  Synchronous; KIFI_ERR_CAPACITY with the
 * required counts if a buffer is too small.  *stalled (may be NULL) reports the
 * reference's join stall (an index ID without a vertex ID). */
kifi_status kifi_mesh(kifi_volume *vol, float *verts, size_t vcap, size_t *nverts);
kifi_status kifi_mesh_indexed(kifi_volume *vol, float *verts, size_t vcap, uint32_t *idx,
                              size_t icap, size_t *nverts, size_t *nidx, int *stalled);

/* kifi_mesh_async: kifi_mesh whose copy of the vertices to the host overlaps the
 * work queued after it.  Returns once the count is known (*nverts, or
 * KIFI_ERR_CAPACITY with nothing queued) and the mesh kernels and the copy are
 * queued; verts must stay valid, and is complete, after kifi_mesh_wait (or
 * kifi_ctx_synchronize, or kifi_host_free) returns.  A later mesh call waits for
 * a pending copy on the device.  The copy overlaps only into page-locked memory
 * (kifi_host_alloc).  No counterpart in the reference: Mesher::Mesh
 * (Mesher.cpp:14-26) returns a filled vector; this splits its return for a
 * caller that keeps integrating frames meanwhile. */
kifi_status kifi_mesh_async(kifi_volume *vol, float *verts, size_t vcap, size_t *nverts);
kifi_status kifi_mesh_wait(kifi_ctx *ctx);

/* ---- model prediction: kifi::Renderer::Bin (Renderer.h:22-29, Renderer.cpp:39-80) ----
 * z-buffered splat of n float3 points (host) into a width x height float3 map
 * (host, 0 = empty) with the given worldToClip. Synchronous. */
kifi_status kifi_bin(kifi_ctx *ctx, const float *points, size_t n, const float worldToClip[16],
                     int width, int height, float *out_points);

/* ---- tracking: kifi::ICP::Align (ICP.h:18-27, ICP.cpp:9-265) ----
 * The reference aligns with 4 fixed Horn point-to-point steps against a model
 * point buffer.  Here the model is a device-resident point + normal map rendered
 * from the volume (kifi_model_predict), and ICP is projective point-to-plane
 * (KIFI_ICP_POINT_TO_PLANE, 6x6 normal equations) or point-to-point with Horn's
 * closed form (KIFI_ICP_HORN) on a depth pyramid with a convergence test.
 * Sums are binary32 with compensated reduction; solves and poses are FP64. */
typedef struct kifi_model kifi_model;
typedef struct kifi_pipeline kifi_pipeline;

enum { KIFI_ICP_POINT_TO_PLANE = 0, KIFI_ICP_HORN = 1 };

typedef struct {
    int   mode;            /* KIFI_ICP_POINT_TO_PLANE or KIFI_ICP_HORN                 */
    int   levels;          /* pyramid levels, 1..3 (level 0 = full resolution)         */
    int   iters[3];        /* maximum iterations per level, index = level              */
    float dist_thresh;     /* correspondence gate at full resolution, metres           */
    float eps_rot;         /* convergence: |rotation increment| < eps_rot (rad) ...    */
    float eps_trans;       /* ... and |translation increment| < eps_trans (m)          */
    int   min_inliers;     /* fewer correspondences: stop (tracking failure)           */
    float fallback_rms;    /* > 0: re-run plain ICP from the previous pose when the    */
    int   fallback_iters;  /*       first pass failed, its rms_1m exceeds fallback_rms */
                           /*       or it used fallback_iters iterations (> 0); the    */
                           /*       fallback result replaces the first only if that    */
                           /*       failed or broke fallback_rms and the fallback's    */
                           /*       rms_1m is lower                                    */
    float level_gate_growth; /* level l gates at dist_thresh x growth^l (1: the same    */
                             /* gate on every level; 2 widens the coarse levels for    */
                             /* large initial errors)                                  */
    int   naive_sums;      /* 1: plain binary32 sums without compensation (precision   */
                           /*    study only; default 0)                                */
    float eps_level_growth; /* level l converges at eps_rot, eps_trans x growth^l;     */
                            /* default 4: a coarser level has half the resolution and */
                            /* a quarter of the samples, and stops at that precision  */
                            /* (1: the same test on every level)                      */
    int   spin_limit;      /* polls before a waiting workgroup of the one-launch      */
                           /* fallback pass gives up (0: 2^18, about 0.1 s).  A test  */
                           /* hook: a tiny limit makes the pass time out on purpose,  */
                           /* and the next state read returns KIFI_ERR_STATE          */
    int   first_iters[3];  /* with a fallback configured (fallback_rms > 0 or          */
                           /* fallback_iters > 0), the first pass starts from a        */
                           /* predicted pose and stops level l after first_iters[l]    */
                           /* iterations when that is > 0 (0: iters[l]); the fallback  */
                           /* pass from the previous pose always runs iters            */
    float motion_prior;    /* point-to-plane only; > 0: Gaussian prior on the motion   */
                           /* from the pass's start pose, weight motion_prior per      */
                           /* inlier: each solve minimises sum r^2 + motion_prior x    */
                           /* inliers x (|w|^2 + |dc|^2) over the accumulated rotation */
                           /* w (rad) and camera-centre displacement dc (m) since the  */
                           /* start.  Keeps directions the scene does not constrain    */
                           /* (a plane, a sphere on a floor) at the start instead of   */
                           /* letting them drift; 0 (default): plain Gauss-Newton      */
} kifi_icp_params;

typedef struct {
    int   iterations;      /* solves performed in the frame (both passes)              */
    int   iters[3];        /* solves per level in the frame (both passes)              */
    float rms;             /* RMS residual of the last association (m)                 */
    int   inliers;         /* correspondences of the last association                  */
    int   failed;          /* too few inliers or a singular system                     */
    int   fallback;        /* the frame re-ran plain ICP from the previous pose        */
    long long sum_iters;   /* running totals over tracked frames                       */
    long long frames;
    long long fallbacks;
    float rms_1m;          /* RMS of the point-to-plane residual scaled to 1 m depth   */
                           /* by the sensor noise model (r / d^2 x 1 m^2), metres      */
    long long sum_level_iters[3]; /* running totals of solves per pyramid level        */
    double start_pose[16]; /* eye->world start of the last frame's first pass (the     */
                           /* caller's or the learned guess, else the previous pose)   */
} kifi_icp_stats;

kifi_status kifi_icp_params_default(kifi_icp_params *p);

/* Model point/normal maps at the resolution of `cam` (device-resident). */
kifi_status kifi_model_create(kifi_ctx *ctx, const kifi_intrinsics *cam, kifi_model **out);
void        kifi_model_destroy(kifi_model *m);
/* Renders the model of `vol` seen from worldToEye (column storage). Asynchronous. */
kifi_status kifi_model_predict(kifi_model *m, kifi_volume *vol, const float worldToEye[16]);
/* Copies the maps out: W*H float4 each (xyz + validity).  Synchronous. */
kifi_status kifi_model_download(kifi_model *m, float *points, float *normals);
/* ICP::Align: align a depth frame (host, mm) to the model starting from the guess
 * eye->world pose; returns the pose and statistics.  Synchronous. */
kifi_status kifi_icp_align(kifi_model *m, const uint16_t *depth_mm, const kifi_intrinsics *cam,
                           const double guess_eyeToWorld[16], const kifi_icp_params *params,
                           double out_eyeToWorld[16], kifi_icp_stats *stats);
/* kifi_icp_align with the frame already in device memory, allocating nothing:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its work memory (depth pyramid, block partials, tracking
 * state) is a fixed 2 MiB array in the library, one per GPU, held until the
 * process ends; frames above 640 x 480 pixels return KIFI_ERR_CAPACITY.  The
 * result equals kifi_icp_align's for the same inputs.  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (from two contexts or
 * threads) race; calls in sequence are independent (the state is cleared each
 * call).  Stage timers are off during the call.  Synchronous. */
kifi_status kifi_icp_align_device_noalloc(kifi_model *m, const uint16_t *depth_mm_dev, const kifi_intrinsics *cam,
                                         const double guess_eyeToWorld[16], const kifi_icp_params *params,
                                         double out_eyeToWorld[16], kifi_icp_stats *stats);
/* ICP::Align itself (ICP.h:18-27, ICP.cpp:9-265): `iterations` window-search
 * Horn steps (the reference runs 4) of the raw frame (host, mm) against a host
 * point buffer synth_xyz (synth_width x synth_height float3 world points, all
 * zero = empty) rendered at synthEyeToWorld; poses are eye->world, column
 * storage.  The 21 x 21 window clamps to the buffer (the reference clamps to
 * 640 x 480).  With no correspondence the guess is returned and stats->failed
 * is set (the reference divides by zero).  Synchronous. */
kifi_status kifi_icp_align_points(kifi_ctx *ctx, const uint16_t *depth_mm, const kifi_intrinsics *cam,
                                  const float guess[16], const float *synth_xyz, int synth_width, int synth_height,
                                  const float synthEyeToWorld[16], int iterations, float out[16],
                                  kifi_icp_stats *stats);
/* The same with the raw frame in binary32 metres (ICP::Align's vector2d<float>). */
kifi_status kifi_icp_align_points_f32(kifi_ctx *ctx, const float *depth_m, const kifi_intrinsics *cam,
                                      const float guess[16], const float *synth_xyz, int synth_width,
                                      int synth_height, const float synthEyeToWorld[16], int iterations,
                                      float out[16], kifi_icp_stats *stats);
/* Precision study: the reduced point-to-plane normal equations of one projective
 * association of the frame (host, mm) at eyeToWorld against the model, on
 * pyramid level `level` (0..2) with gate dist_thresh (m), weights w = 1/d^4:
 * out[0..20] = sum w J_i J_j for i >= j row by row (J = [p x n, n]),
 * out[21..26] = sum w J_i r, out[27] = count, out[28] = sum r^2,
 * out[29] = sum w r^2; r = n . (q - p).  The sums are the ones the tracker
 * solves: binary32, compensated unless naive_sums, widened to FP64.  No
 * counterpart in the reference (its ICP is point-to-point, ICP.cpp:35-265).
 * Synchronous. */
kifi_status kifi_icp_normal_equations(kifi_model *m, const uint16_t *depth_mm, const kifi_intrinsics *cam,
                                      const double eyeToWorld[16], int level, float dist_thresh, int naive_sums,
                                      double out[30]);
/* Horn's closed-form rigid alignment of n correspondences (ICP.cpp:166-264):
 * compensated centroid and cross-covariance reductions, FP64 4x4 eigen solve.
 * out = the transform mapping src onto dst.  Synchronous. */
kifi_status kifi_horn_align(kifi_ctx *ctx, const float *src, const float *dst, size_t n, float out[16]);

/* ---- frame loop: kifi::Pipeline (Pipeline.h:19-41, Pipeline.cpp:26-68) ----
 * Owns a volume, a model and the tracking state.  kifi_pipeline_integrate with a
 * worldToEye integrates at that pose (the reference's #if 1 branch); with NULL
 * it tracks the frame with ICP against the model of the previous frame (the
 * #else branch), integrates at the tracked pose and renders the next model.
 * Asynchronous: no host synchronisation per frame. */
kifi_status kifi_pipeline_create(kifi_ctx *ctx, const kifi_volume_desc *vol, const kifi_intrinsics *cam,
                                 const kifi_icp_params *params, kifi_pipeline **out);
void        kifi_pipeline_destroy(kifi_pipeline *p);
kifi_volume *kifi_pipeline_volume(kifi_pipeline *p);
kifi_model  *kifi_pipeline_model(kifi_pipeline *p);
kifi_status kifi_pipeline_integrate(kifi_pipeline *p, const uint16_t *depth_mm, int depth_on_device,
                                    const float worldToEye[16]);
/* The same with depth in binary32 metres (Pipeline::Integrate of a
 * vector2d<float>, Pipeline.h:30).  A metre frame does not feed the learned
 * initial pose (it reads mm); tracking starts from the previous pose. */
kifi_status kifi_pipeline_integrate_f32(kifi_pipeline *p, const float *depth_m, int depth_on_device,
                                        const float worldToEye[16]);
/* Tracks the frame as kifi_pipeline_integrate with no pose does (the tracked
 * pose, kifi_pipeline_pose, the statistics and the network's frame sequence
 * are this frame's), then fuses it at fuseWorldToEye instead of the tracked
 * pose; the next frame is tracked from fuseWorldToEye's pose.  The first frame
 * is fused at fuseWorldToEye untracked.  For trackers evaluated against known
 * poses (volumes equal whatever the start) and callers with an external pose
 * source.  No counterpart in the reference. */
kifi_status kifi_pipeline_integrate_at(kifi_pipeline *p, const uint16_t *depth_mm, int depth_on_device,
                                       const float fuseWorldToEye[16]);
/* Initial guess (eye->world) for the next tracked frame instead of the previous
 * pose, e.g. a learned prediction; plain ICP from the previous pose remains the
 * fallback per kifi_icp_params. */
kifi_status kifi_pipeline_set_guess(kifi_pipeline *p, const double eyeToWorld[16]);
/* Back to the state after kifi_pipeline_create, keeping every allocation: the
 * volume is emptied (capacity kept), the pose is the identity, the next frame
 * is the first, the caller's guess and the network's frame history are
 * dropped.  Running totals (kifi_pipeline_stats, kifi_integrate_totals) go on.
 * Synchronous.  No counterpart in the reference (a new Pipeline there). */
kifi_status kifi_pipeline_reset(kifi_pipeline *p);
kifi_status kifi_pipeline_set_pose(kifi_pipeline *p, const double eyeToWorld[16]);
kifi_status kifi_pipeline_pose(kifi_pipeline *p, double eyeToWorld[16]);     /* EyeToWorld(), sync */
kifi_status kifi_pipeline_stats(kifi_pipeline *p, kifi_icp_stats *stats);     /* sync */

/* ---- learned initial pose (no reference counterpart: the reference starts ICP
 * from the previous pose, Pipeline.cpp:37-39) ----
 * A small CNN predicts the rigid twist between the previous and the current
 * frame; its pose (previous * exp(twist)) starts ICP, and plain ICP from the
 * previous pose is the fallback when the final rms_1m exceeds fallback_rms or the
 * iterations reach fallback_iters (kifi_icp_params); ICP judges the two starts.  Input: both frames as
 * 128 x 106 4x4-block mean depth / 4 m (bfloat16); five 3x3 stride-2
 * convolutions (16-32-64-64-128, ReLU), FC 2048-256 (ReLU)-6.  BF16 weights and
 * activations, FP32 accumulation on the matrix cores, FP32 master weights, Adam.
 * Twists are (omega [rad], v [m]) with T = [R(omega) | v], eye(t) -> eye(t-1);
 * the network's raw outputs are twist / (0.15, 0.15, 0.15, 0.06, 0.06, 0.06). */
typedef struct kifi_twistnet kifi_twistnet;
typedef struct kifi_pairset kifi_pairset;
struct kifi_scene_s;

size_t      kifi_twistnet_param_count(void);
/* seeded He-normal initialisation */
kifi_status kifi_twistnet_create(kifi_ctx *ctx, uint64_t seed, kifi_twistnet **out);
void        kifi_twistnet_destroy(kifi_twistnet *net);
/* FP32 master weights in layer order (conv1 W [16][2][3][3], conv1 b, ..., fc2 W [6][256], fc2 b).
 * set also resets the Adam state. */
kifi_status kifi_twistnet_get_params(kifi_twistnet *net, float *out);
kifi_status kifi_twistnet_set_params(kifi_twistnet *net, const float *in);
/* B <= 256 host inputs (B x 2 x 106 x 128, rounded to bfloat16) -> B x 6 raw outputs */
kifi_status kifi_twistnet_forward(kifi_twistnet *net, const float *x, int B, float *y);
/* loss 0.5/B sum (y - t)^2 against raw targets t, and its gradient (param order) */
kifi_status kifi_twistnet_loss_grad(kifi_twistnet *net, const float *x, const float *t, int B, float *loss,
                                    float *grad);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over `steps` batches drawn without
 * replacement from the valid pairs of `set` (seeded); losses[steps] may be NULL */
kifi_status kifi_twistnet_train(kifi_twistnet *net, kifi_pairset *set, int steps, int batch, float lr,
                                uint64_t seed, float *losses);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8; the same step as training) with a
 * caller's gradient in parameter order, e.g. one all-reduced over several
 * trainers.  The step count continues the network's (set_params resets it). */
kifi_status kifi_twistnet_apply_gradient(kifi_twistnet *net, const float *grad, float lr);
/* twists (physical units) predicted for pairs [first, first + n) of a set */
kifi_status kifi_twistnet_predict_set(kifi_twistnet *net, kifi_pairset *set, int first, int n, float *twists);

/* Training pairs generated on the device: per pair a random pose in the room
 * (within 1 m of its centre), a relative motion of random axis and direction,
 * angle U[0, max_rot_deg], length U[0, max_trans]; both frames rendered with
 * the sensor noise model; the target is the twist that point-to-plane ICP of
 * the second frame against the first frame's surface converges to from the
 * true pose.  A pair is valid when ICP kept >2000 inliers and stayed within
 * 2 deg / 2 cm of the truth.  scenes: kifi_scene[n_scenes] (kifi_synth.h),
 * pair i uses scenes[i % n_scenes]. */
kifi_status kifi_pairset_generate(kifi_ctx *ctx, const kifi_intrinsics *cam, const struct kifi_scene_s *scenes,
                                  int n_scenes, int n, uint64_t seed, float max_trans, float max_rot_deg,
                                  kifi_pairset **out);
/* The same with motion ranges: angle U[min_rot_deg, max_rot_deg], length
 * U[min_trans, max_trans].  walk = 0: random axis and direction as above.
 * walk = 1: the step of a walking camera (the fast synthetic trajectory's law):
 * yaw by +-U[min_rot_deg, max_rot_deg], pitch by +-U[0, max_pitch_step_deg]
 * (kept in [-20, 0] deg), the same roll, a horizontal translation of
 * U[min_trans, max_trans] in a random direction. */
typedef struct {
    float min_trans, max_trans;          /* metres  */
    float min_rot_deg, max_rot_deg;      /* degrees */
    float max_pitch_step_deg;            /* walk only */
    int   walk;
} kifi_motion_steps;
kifi_status kifi_pairset_generate_steps(kifi_ctx *ctx, const kifi_intrinsics *cam, const struct kifi_scene_s *scenes,
                                        int n_scenes, int n, uint64_t seed, const kifi_motion_steps *steps,
                                        kifi_pairset **out);
void        kifi_pairset_destroy(kifi_pairset *set);
int         kifi_pairset_size(const kifi_pairset *set);
/* inputs n x 2 x 106 x 128, raw targets n x 6, true twists n x 6, valid n; any may be NULL */
kifi_status kifi_pairset_download(kifi_pairset *set, float *inputs, float *targets, float *truth, int *valid);
/* kifi_scene_render (kifi_synth.h) on the device; out is a host or (out_on_device) device buffer */
kifi_status kifi_render(kifi_ctx *ctx, const struct kifi_scene_s *scene, const kifi_intrinsics *cam,
                        const double eyeToWorld[16], uint64_t noise_seed, double noise_k, double invalid_frac,
                        uint16_t *out, int out_on_device);
/* Attach a network to a pipeline (NULL detaches); tracked frames then start ICP
 * from its prediction, with the fallback of the pipeline's kifi_icp_params. */
kifi_status kifi_pipeline_set_twistnet(kifi_pipeline *p, kifi_twistnet *net);

/* ---- the wave64 workgroup primitives: cuda::segmented_inclusive_scan,
 * segmented_exclusive_scan (cuda/scan.h:8-21) and segmented_reduce
 * (cuda/reduce.h:8-14) of the reference's unbuilt CUDA part, with the same
 * argument meaning: `size` unsigned values, segments of `segment` values (64,
 * 256 or 1024: one workgroup each), scanned in place per segment or summed
 * into sums[ceil(size / segment)].  Here data and out are host buffers and the
 * call is synchronous (the reference takes device pointers).  They run the
 * scan and reduction every kernel of the library uses (kifi_prims.hpp). */
kifi_status kifi_segmented_inclusive_scan(kifi_ctx *ctx, const uint32_t *data, uint32_t size, uint32_t segment,
                                          uint32_t *out);
kifi_status kifi_segmented_exclusive_scan(kifi_ctx *ctx, const uint32_t *data, uint32_t size, uint32_t segment,
                                          uint32_t *out);
kifi_status kifi_segmented_reduce(kifi_ctx *ctx, const uint32_t *data, uint32_t size, uint32_t segment,
                                  uint32_t *sums);

/* ---- measurement and staging (no reference counterpart; the reference times
 * stages with a CPU stop_watch, util/stop_watch.cpp:23-58) ----
 * Stage timers bracket the named kernels with HIP events on the context's
 * stream for the stages whose bit is set in stage_mask (1 << KIFI_STAGE_*;
 * 0 disables; each event pair costs some microseconds of stream time, so time
 * only what is read); kifi_timing_read synchronises and returns the summed
 * elapsed milliseconds and the number of timed launches, then resets. */
enum {
    KIFI_STAGE_MERGE_UPDATE = 0, /* fused merge + TSDF update kernel            */
    KIFI_STAGE_SORT = 1,         /* keygen + radix sort + unique                */
    KIFI_STAGE_ICP = 2,          /* depth pyramid + all ICP iterations          */
    KIFI_STAGE_MODEL = 3,        /* model splat + resolve + normals             */
    KIFI_STAGE_MESH = 4,         /* marching cubes                              */
    KIFI_STAGE_COUNT = 5,
    KIFI_TIMING_ALL = (1 << 5) - 1
};
kifi_status kifi_timing_enable(kifi_ctx *ctx, int stage_mask);
kifi_status kifi_timing_read(kifi_ctx *ctx, int stage, double *total_ms, long long *count);
/* Device buffers for callers without their own allocator (frames resident in HBM). */
kifi_status kifi_device_alloc(kifi_ctx *ctx, size_t bytes, void **out);
kifi_status kifi_device_free(kifi_ctx *ctx, void *ptr);
/* Page-locked host buffers: frames and meshes in them cross PCIe at full rate
 * (a pageable buffer is staged through a bounce buffer by the runtime). */
kifi_status kifi_host_alloc(kifi_ctx *ctx, size_t bytes, void **out);
kifi_status kifi_host_free(kifi_ctx *ctx, void *ptr);
kifi_status kifi_memcpy_to_device(kifi_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Plain streaming copy kernel (16-byte vectors), the calibration reference for
 * HBM byte counters.  bytes must be a multiple of 16.  Asynchronous. */
kifi_status kifi_device_copy(kifi_ctx *ctx, void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
