/*
 * oracle.h -- CPU restatement of the KinectFission (kifi) fusion path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path in merge_fusion_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links it.
 *
 * Every function states the reference file:line it restates (paths relative
 * to the reference repository root, under KinectFission/kifi/).  Arithmetic is
 * plain IEEE binary32 in the reference's source order (the SSE `vector` ops of
 * util/math.h are per-lane binary32 ops); the build uses -ffp-contract=off so
 * no multiply-add is fused.
 *
 * Keys are carried as uint64_t with `kb` bits per axis: kb = 10 reproduces the
 * reference's 32-bit `z<<20 | y<<10 | x` packing (util/math.h:442-457) exactly
 * (all values stay below 2^32); kb = 21 is the 64-bit layout for resolutions
 * above 1024 (no reference counterpart).
 *
 * Pinning of each function (known answers, closed forms, brute force) lives in
 * tests/test_oracle_*.py.  Functions without such a check say "parity unpinned".
 */
#ifndef KIFI_ORACLE_H
#define KIFI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- camera / volume descriptions (plain data) ---- */
typedef struct {
    int   width, height;
    float fx, fy;      /* focal length in pixels   (DepthSensorParams.h:31) */
    float cx, cy;      /* principal point, pixels  (DepthSensorParams.h:32) */
    float znear, zfar; /* sensible range, metres   (DepthSensorParams.h:33) */
} orc_camera;

typedef struct {
    int   res;          /* voxels per axis, power of two        */
    int   kb;           /* key bits per axis: 10 or 21          */
    float side;         /* cube side length in metres           */
    float trunc;        /* truncation margin in metres          */
    /* derived exactly as Volume.cpp:12-37 */
    float vl;           /* m_tmpVoxelLenf                       */
    float vlHalfPlusMin;/* m_tmpVoxelLenOver2PlusMinf           */
    float vlInv;        /* 1 / vl                               */
    float minOverNegVl; /* vlInv * (-Minimum().x)               */
} orc_volume_desc;

/* storage layouts of a voxel */
enum { ORC_STORE_F32 = 0, ORC_STORE_F16U8 = 1, ORC_STORE_F64 = 2 /* {double sum, double weight}:
                                                                     the full-FP64 precision oracle */ };

/* ---- util/math.h restatements ---- */
void orc_mat_mul(const float m[16], const float n[16], float out[16]);      /* math.h:367-377 */
void orc_mat_vec(const float m[16], const float v[4], float out[4]);        /* math.h:362-365, 381-397 */
void orc_invert_transform(float m[16]);                                     /* math.h:615-625 */
void orc_eye_to_clip_rh(const orc_camera *cam, float out[16]);              /* DepthSensorParams.cpp:47-63 */
void orc_volume_desc_init(orc_volume_desc *d, int res, float side, float trunc, int kb); /* Volume.cpp:12-37 */
uint64_t orc_pack(uint64_t x, uint64_t y, uint64_t z, int kb);               /* math.h:442-445 */
void orc_unpack(uint64_t k, int kb, uint32_t *x, uint32_t *y, uint32_t *z);  /* math.h:452-457 */

/* ---- Integrator (Integrator.cpp) ---- */
size_t orc_keygen(const float *depth_m, const orc_camera *cam, const float eyeToWorld[16],
                  const orc_volume_desc *vol, uint64_t *out);                   /* :76-182 */
void   orc_radix_sort(uint64_t *keys, size_t n, int key_bits, uint64_t *scratch); /* algorithm.h:63-91 */
size_t orc_unique(uint64_t *keys, size_t n);                                  /* :50-55 */
size_t orc_expand(uint64_t *keys, size_t n, int kb, uint64_t *scratch, size_t cap); /* :186-250 */
size_t orc_set_union(const uint64_t *k1, size_t n1, const uint64_t *k2, size_t n2,
                     const void *v1, const void *v2_default, size_t vbytes,
                     uint64_t *kout, void *vout);                             /* algorithm.h:138-198 */
void   orc_update(const uint64_t *keys, void *values, size_t n, int storage,
                  const float *depth_m, const orc_camera *cam, const float worldToClip[16],
                  const orc_volume_desc *vol);                                /* :253-358 */
void   orc_update_f64(const uint64_t *keys, double *sum_weight, size_t n,
                      const float *depth_m, const orc_camera *cam, const double worldToClip[16],
                      const orc_volume_desc *vol);                            /* FP64 twin of :253-358 */

/* A sparse volume held by the oracle: sorted keys + values (flat_map.h:14-197). */
typedef struct {
    orc_volume_desc desc;
    int      storage;   /* ORC_STORE_* */
    size_t   size, cap;
    uint64_t *keys;
    void     *values;   /* float2 {sum, weight} or uint32 {half | u8 weight << 16} */
} orc_volume;

orc_volume *orc_volume_create(int res, float side, float trunc, int kb, int storage);
void        orc_volume_destroy(orc_volume *v);
void        orc_volume_clear(orc_volume *v);
size_t      orc_volume_size(const orc_volume *v);
/* copies out up to cap entries; returns size */
size_t      orc_volume_read(const orc_volume *v, uint64_t *keys, void *values, size_t cap);
/* replaces the content (keys must be sorted unique) */
int         orc_volume_write(orc_volume *v, const uint64_t *keys, const void *values, size_t n);

/* Integrator::Integrate (Integrator.cpp:26-71) with depth given in millimetres,
 * converted as DepthStream.cpp:89-90 (mm * 0.001f).  Returns the number of new
 * keys E handed to the merge, or (size_t)-1 on error. Optional stats[4] receive
 * K (keygen), U (unique), E (expanded), V1 (volume size after merge). */
size_t orc_integrate(orc_volume *v, const uint16_t *depth_mm, const orc_camera *cam,
                     const float worldToEye[16], size_t stats[4]);
/* the same with a frame in binary32 metres, used as it is */
size_t orc_integrate_f32(orc_volume *v, const float *depth_m, const orc_camera *cam,
                         const float worldToEye[16], size_t stats[4]);

/* ---- voxel helpers ---- */
void     orc_voxel_update_f32(float *sum_weight, float d);      /* Voxel.h:53-57 */
float    orc_voxel_distance_f32(const float *sum_weight);       /* Voxel.h:39-44 */
float    orc_f16u8_recip(uint32_t w);
void     orc_voxel_update_f16u8(uint32_t *packed, float d);     /* FP16 storage, no reference counterpart */
float    orc_voxel_distance_f16u8(uint32_t packed);
uint16_t orc_f32_to_f16(float f);                                /* IEEE binary16, round to nearest even */
float    orc_f16_to_f32(uint16_t h);

/* ---- Mesher (Mesher.cpp) ---- */
/* Generate<false>: vertices in generation order.  Returns count (<= cap*... caller sizes 3*V). */
size_t orc_mesh(const orc_volume *v, float *verts, size_t cap);                        /* :14-17, 85-229 */
/* Mesh with indices (:19-45).  verts sized >= 3*V float3, idx >= 15*V.  Returns vertex
 * count, *n_idx receives the index count; *stalled receives 1 if the index join stalled. */
size_t orc_mesh_indexed(const orc_volume *v, float *verts, size_t vcap, uint32_t *idx,
                        size_t icap, size_t *n_idx, int *stalled);
const int *orc_tri_offsets(void);   /* 257 entries, Mesher.cpp:239-261 */
const int *orc_tri_table(void);     /* 2460 entries, Mesher.cpp:264-475 */

/* ---- Renderer::Bin (Renderer.cpp:39-80) ---- */
void orc_bin(const float *points, size_t n, const float worldToClip[16], int W, int H,
             float *out_points /* W*H*3 */, float *zbuf /* W*H */);

/* ---- ICP ---- */
/* Horn closed-form step from correspondences, ICP.cpp:166-264 (Kahan centroids and
 * cross-covariance in binary32, Horn N, eigenvector of the largest eigenvalue, q->R,
 * T*R).  The eigen solve uses the oracle's own FP64 cyclic Jacobi (the reference's
 * routine is not restated). out = T*R in column storage. */
void orc_horn_step(const float *src, const float *dst, size_t n, float out[16]);
/* ICP::AlignStep in full (ICP.cpp:35-265): window-search correspondences + Horn.
 * model is mW x mH float3 (0 = invalid); scratch buffers hold H*(W/4+1) float3. */
size_t orc_align_step(const float *depth_m, const orc_camera *cam, const float guess_e2w[16],
                      const float *model, int mW, int mH, const float model_e2w[16], float out[16],
                      float *scratch_src, float *scratch_dst);
/* ICP::Align (ICP.cpp:9-31): `iters` AlignSteps (the reference runs 4). */
size_t orc_icp_align_ref(const float *depth_m, const orc_camera *cam, const float guess_e2w[16],
                         const float *model, int mW, int mH, const float model_e2w[16], int iters,
                         float out_e2w[16]);
/* Symmetric 4x4 eigen decomposition, FP64 cyclic Jacobi. vecs columns, column storage. */
void orc_jacobi4(const double a[16], double evals[4], double evecs[16]);
/* Kahan sum of binary32 values in order (numeric.h:52-74) */
float orc_kahan_sum_f32(const float *x, size_t n);
/* exclusive scan (numeric.h:78-108) */
void  orc_partial_sum_exclusive_u32(const uint32_t *in, size_t n, uint32_t *out);

/* ---- model prediction + point-to-plane ICP (oracle_icp.c; no reference counterpart) ---- */
/* binary32, bit-exact twins of the HIP model kernels */
void orc_model_splat(const orc_volume *v, const float w2e[16], const orc_camera *cam, float dmax,
                     float znear, uint64_t *zbuf);
void orc_model_resolve(const orc_volume *v, const float w2e[16], const float camPos[3],
                       const uint64_t *zbuf, int W, int H, float dmax, float znear, float *pts);
void orc_model_normals4(const float *pts, int W, int H, const float camPos[3], float disc, float *nrm);
void orc_depth_pyramid_level(const float *src, int W, int H, float *dst);
/* FP64 point-to-plane ICP on one pyramid level (the FP64 pose oracle) */
int orc_icp_level_f64(const float *depth, int Wl, int Hl, double fxl, double fyl, double cxl, double cyl,
                      const float *pts, const float *nrm, const orc_camera *cam, const double W2E[16],
                      double pose[16], int max_iters, double dist_thresh, double eps_rot,
                      double eps_trans, double stats[2]);
/* the same with the motion prior of kifi_icp_params.motion_prior: weight `prior` per
 * inlier, prior_c the pass's start camera centre, prior_acc (in/out, zero at the
 * pass's start) the accumulated motion (rotation, camera-centre displacement) */
int orc_icp_level_f64_prior(const float *depth, int Wl, int Hl, double fxl, double fyl, double cxl, double cyl,
                            const float *pts, const float *nrm, const orc_camera *cam, const double W2E[16],
                            double pose[16], int max_iters, double dist_thresh, double eps_rot,
                            double eps_trans, double prior, const double prior_c[3], double prior_acc[6],
                            double stats[2]);
/* the FP64 normal equations of one association (layout of kifi_icp_normal_equations) */
void orc_icp_sums_f64(const float *depth, int Wl, int Hl, double fxl, double fyl, double cxl, double cyl,
                      const float *pts, const float *nrm, const orc_camera *cam, const double W2E[16],
                      const double pose[16], double dist_thresh, double out[30]);
/* the same sums from the binary32 terms of the HIP association, summed in double */
void orc_icp_sums_f32terms(const float *depth, int level, const float *pts, const float *nrm,
                           const orc_camera *cam, const float w2e[16], const float pose_f[16], float gate,
                           double out[30]);

#ifdef __cplusplus
}
#endif
#endif
